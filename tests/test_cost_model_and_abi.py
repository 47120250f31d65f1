"""CPU tests: cost model (oracle restatement vs product; the oracle's own
known answers are in tests/test_oracle_cost_model.py) and the C ABI (the library loads and exports every symbol of
include/stile.h; argument validation paths that need no device)."""
import ctypes
import os
import re

import numpy as np
import pytest
import scipy.sparse

from oracle import cost_model as oc
from stile_amd import datasets
from stile_amd import gen_formats_v2 as gf
from stile_amd import my_cost_model as mc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _ell_op(A, mbs):
    op = gf.Op('spmm', [0, 1], [2], [A.shape[0], 4, A.shape[1]], None,
               [A, gf.DenseTensor(np.zeros((A.shape[1], 4), np.float32))], None)
    return gf.gen_candidate_formats(op, mbs, ([1], [1], [1]), False, gen_TC_formats=False)[0][0]


@pytest.mark.parametrize("dsize", [16, 32])
def test_product_cost_model_matches_oracle(dsize):
    mc.set_profile("a100")
    A = datasets.rmat_graph(10, seed=1)
    ell = _ell_op(A, 32)
    assert mc.lower_bound_ELL_cost(ell, 32, dsize) == pytest.approx(
        oc.lower_bound_ell([(ts[0][1], 32, ts[2][1]) for ts in gf.ell_tile_shapes(ell, 32)], dsize), rel=1e-12)
    ps = ell.position_space[0]
    for ts in gf.ell_tile_shapes(ell, 32):
        I, J, K = ts[0][1], ts[1][1], ts[2][1]
        npos = -(-ps.shape[0] // I)
        pos = np.arange(npos)
        for atomic in (False, True):
            got = mc.cost_tb_latency_given_ELL_tiles(ell, ts, pos, dsize, A.nnz, 108, atomic)
            for p in range(0, npos, max(1, npos // 7)):
                s, e = ps.indptr[I * p], ps.indptr[min(I * (p + 1), ps.shape[0])]
                cols = ps.indices[s:e]
                k_is = len(np.unique(np.concatenate([cols, [0]]))) if e - s < I * K else len(np.unique(cols))
                assert got[p] == pytest.approx(oc.ell_tile_cost(I, J, K, k_is, atomic, dsize), rel=1e-12)


def test_tc_tables_are_the_reference_values():
    p = mc.get_profile("a100")
    assert p.tc_cost[(16, 32, 16)] == 0.0002922496385872364          # my_cost_model.py:672
    assert p.tc_cost_atomic[(16, 32, 160)] == 0.0021382145583629607  # :695
    assert p.ell_cost[(8, 32, 32)] == 0.226909105 / 100              # :714
    assert p.sm_num == 108


# --- C ABI -----------------------------------------------------------------

def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "stile.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(stile_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    from stile_amd import _native
    lib = _native.load()
    syms = _header_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in include/stile.h but not exported"
        assert s in _native._SIGNATURES, f"{s} has no ctypes signature"
    assert lib.stile_abi_version() == 3


def test_abi_rejects_bad_arguments_without_a_device():
    from stile_amd import _native
    lib = _native.load()
    out = ctypes.c_void_p()
    assert lib.stile_plan_create_spmm(None, 0, ctypes.byref(out)) == 1
    assert b"pack" in lib.stile_last_error()
    pack = _native.StileSpmmPack(M=4, K=4, num_ell=0, num_tc=0)
    assert lib.stile_plan_create_spmm(ctypes.byref(pack), 7, ctypes.byref(out)) == 1
    assert b"blk_dtype" in lib.stile_last_error()
    assert lib.stile_spmm_run(None, None, 0, 4, None, 0, None) == 1
    assert lib.stile_spmm_run_noalloc(None, 0, None, 4, 4, None, 4, 1, None) == 1
    assert b"not an SpMM plan" in lib.stile_last_error()
    assert lib.stile_csr_spmm_f32(-1, 1, 1, None, None, None, None, 1, None, 1, None) == 1
    assert lib.stile_plan_destroy(None) == 0
    info = _native.StilePlanInfo()
    assert lib.stile_plan_info(None, ctypes.byref(info)) == 1


def test_abi_rejects_out_of_range_pack_before_any_device_call():
    from stile_amd import _native
    lib = _native.load()
    keep = _native._PackArrays()
    bad = dict(M=8, K=8, ell=[dict(m_blk=2, nnz_cols=2, num_tiles=1, a=np.ones(4, np.float32),
                                   indices_k=np.array([[0, 9], [1, 2]]), indices_reordered_i=np.array([0, 1]))],
               tc=[])
    st = _native._spmm_pack_struct(bad, keep)
    out = ctypes.c_void_p()
    assert lib.stile_plan_create_spmm(ctypes.byref(st), 0, ctypes.byref(out)) == 1
    assert b"column index out of range" in lib.stile_last_error()
    bad["ell"][0]["indices_k"] = np.array([[0, 1], [1, 2]])
    bad["ell"][0]["indices_reordered_i"] = np.array([0, 12])
    st = _native._spmm_pack_struct(bad, keep)
    assert lib.stile_plan_create_spmm(ctypes.byref(st), 0, ctypes.byref(out)) == 1
    assert b"row index out of range" in lib.stile_last_error()

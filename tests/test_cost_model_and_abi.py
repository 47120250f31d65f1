"""CPU tests: cost model (hand-evaluated known answers, oracle restatement vs
product) and the C ABI (the library loads and exports every symbol of
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


# --- known answers, evaluated by hand from my_cost_model.py:786-816, 526-577, 1219-1248

def test_ell_cost_known_answers():
    # non-atomic, fp32 coefficients: 8.23071690e-08 * 1000 + 3.42962663e-03
    assert oc.ell_cost_model(1000.0, None, False, 32) == pytest.approx(3.51193380e-03, rel=1e-8)
    # (8, 32, 32) tile, 20 distinct columns, fp16: memory 16 + 16 + 40 = 72 units
    # -> x = 72 * 32 = 2304 -> 8.80425616e-08 * 2304 + 4.17697078e-04
    assert oc.ell_tile_cost(8, 32, 32, 20, False, 16) == pytest.approx(6.20547140e-04, rel=1e-8)
    # atomic: x_atomic = 88 * 32 = 2816 -> * (8.80688635 * 2816 / 2304 - 7.9094305)
    assert oc.ell_tile_cost(8, 32, 32, 20, True, 16) == pytest.approx(1.77137764e-03, rel=1e-7)


def test_sddmm_1d_and_lower_bound_known_answers():
    # rows 4, cols 30, K 32, bucket 32, fp32: mem = ((4+30)*32 + 32) * 4 = 4480 bytes,
    # x = 4480 -> 5.90185816e-08 * 4480 + 1.83606878e-04
    assert oc.sddmm_1d_cost(4, 30, 32, 32, 32) == pytest.approx(4.480101236e-04, rel=1e-8)
    # mbs 32, fp16: best shape (8,32,32) / (16,32,16): memory 66 units -> x = 2112
    # -> (8.80425616e-08 * 2112 + 4.17697078e-04) / 8192
    shapes = [(256, 32, 1), (128, 32, 2), (64, 32, 4), (32, 32, 8), (16, 32, 16), (8, 32, 32)]
    assert oc.lower_bound_ell(shapes, 16) == pytest.approx(6.03642968e-04 / 8192, rel=1e-7)


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

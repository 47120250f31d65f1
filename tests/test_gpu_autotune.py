"""The autotuner's synthetic plans compute what they describe (FP64 oracle)
and its measurements return finite positive per-tile costs."""
import math

import numpy as np
import pytest
import scipy.sparse

from oracle import ops
from stile_amd import autotune
from tests.test_packing import expand_spmm_pack

pytestmark = pytest.mark.gpu


def _spmm_check(pack, blk, cuda, N=64, tol=1e-5):
    import torch
    from stile_amd import _native
    r, c, v = expand_spmm_pack(pack)
    A = scipy.sparse.csr_matrix((v, (r, c)), shape=(pack["M"], pack["K"]))
    B = np.random.default_rng(0).standard_normal((pack["K"], N), dtype=np.float32)
    plan = _native.Plan.spmm(pack, blk)
    got = plan.run_spmm(torch.from_numpy(B).to(cuda)).cpu().numpy()
    assert ops.max_rel_err(got, ops.spmm_fp64(A, B)) <= tol


def test_synthetic_plans_compute_their_matrices(cuda):
    print("autotune synthetic TC / ELL / SDDMM plans", flush=True)
    _spmm_check(autotune.tc_pack(40, 96), "float32", cuda)
    _spmm_check(autotune.tc_pack(40, 96), "bfloat16", cuda, tol=1e-2)
    pack, _ = autotune.ell_pack(512, 16, 16, 4096, 4096, np.random.default_rng(3))
    _spmm_check(pack, "float32", cuda)
    import torch
    from stile_amd import _native
    rng = np.random.default_rng(4)
    for pack in (autotune.sddmm_1d_pack(300, 4, 500, rng)[0], autotune.sddmm_tc_pack(20)):
        g1, gt = pack["g1d"], pack["tc"]
        rows = np.concatenate([g["mid_idx"] for g in g1] + [g["i_indices"][np.repeat(np.arange(g["nnzb"]), 256), g["indices_dense"] // 16] for g in gt])
        cols = np.concatenate([g["j_idx"] for g in g1] + [g["j_indices"][np.repeat(np.arange(g["nnzb"]), 256), g["indices_dense"] % 16] for g in gt])
        outs = np.concatenate([g["remap"] for g in g1] + [g["indices"] for g in gt])
        X = rng.standard_normal((pack["M"], 32), dtype=np.float32)
        Y = rng.standard_normal((pack["N"], 32), dtype=np.float32)
        plan = _native.Plan.sddmm(pack, "float32")
        got = plan.run_sddmm(torch.from_numpy(X).to(cuda), torch.from_numpy(Y).to(cuda)).cpu().numpy()
        want = np.zeros(pack["nnz"])
        want[outs] = (X[rows].astype(np.float64) * Y[cols]).sum(1)
        assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(want))


def test_measurements_are_positive(cuda):
    print("autotune small measurements", flush=True)
    rng = np.random.default_rng(5)
    vals = [autotune.measure_tc(32, "float32", W=256),
            autotune.measure_ell(32, 8, 1 << 12, rng, R=1 << 12, ncols=1 << 12)[0],
            autotune.measure_sddmm_1d(4, rng, R=1 << 12, ncols=1 << 12)[0],
            autotune.measure_sddmm_tc(32, W=256)]
    assert all(math.isfinite(v) and v > 0 for v in vals)

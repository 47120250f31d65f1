"""stile_spmm_run_noalloc (no allocation: the deterministic workspace is the
library's static device scratch) against stile_spmm_run_ex with a caller
workspace, through the C ABI.

Calls in a row on one device, each later one on a smaller input than the
first, so each finds the scratch holding the previous call's partial sums:
  1. R-MAT scale 14, FP32 blocks, N = 128 (split hub rows: pieces);
  2. a 40 x 20000 matrix of windows with more than 256 dense chunks, N = 72
     (parts, stored after the pieces' region);
  3. the first plan again at N = 40.
Bars: deterministic mode bit-identical to run_ex's deterministic mode (same
kernels, same summation order); default mode (atomics) within 1e-5
norm-wise of run_ex's default mode; a workspace need above the scratch size
is refused before any launch.
"""

import numpy as np
import pytest
import scipy.sparse

from stile_amd import _native, api, datasets

pytestmark = pytest.mark.gpu


def _check_call(plan, Bd, N, cuda, label):
    import torch
    need = plan.workspace_bytes(N, deterministic=True)
    print(f"noalloc {label} N={N} workspace={need}", flush=True)
    assert 0 < need <= _native.STILE_NOALLOC_SCRATCH_BYTES
    ws = torch.empty(need, dtype=torch.uint8, device=cuda)
    ref = plan.run_spmm(Bd, deterministic=True, workspace=ws)
    out = torch.full((plan.M, N), float("nan"), device=cuda)
    plan.run_spmm_noalloc(Bd, out, deterministic=True)
    torch.cuda.synchronize()
    assert torch.equal(out, ref), f"{label}: deterministic no-alloc run differs from run_ex"
    ref0 = plan.run_spmm(Bd)
    out.fill_(float("nan"))
    plan.run_spmm_noalloc(Bd, out)
    torch.cuda.synchronize()
    err = float((out - ref0).abs().max() / ref0.abs().max().clamp_min(1e-30))
    assert err <= 1e-5, f"{label}: default-mode no-alloc run off by {err:.3g}"


def test_noalloc_matches_run_ex_on_calls_in_a_row(cuda):
    import torch
    A1 = datasets.rmat_graph(14, seed=3)
    plan1 = api.build_spmm_plan(A1, blk_dtype="float32", N=128)
    assert plan1.info()["num_split"] > 0
    rng = np.random.default_rng(5)
    B1 = torch.from_numpy(rng.standard_normal((A1.shape[1], 128), dtype=np.float32)).to(cuda)
    _check_call(plan1, B1, 128, cuda, "rmat14 pieces")

    A2 = scipy.sparse.random(40, 20000, density=0.06, format="csr", random_state=8, dtype=np.float32)
    plan2 = api.build_spmm_plan(A2, blk_dtype="float32", opts=api.SearchOptions(only_TC=True), N=72)
    B2 = torch.from_numpy(rng.standard_normal((20000, 72), dtype=np.float32)).to(cuda)
    assert B2.numel() < B1.numel()
    _check_call(plan2, B2, 72, cuda, "huge-window parts")

    B3 = B1[:, :40].contiguous()
    _check_call(plan1, B3, 40, cuda, "rmat14 pieces, narrower")
    plan1.close()
    plan2.close()


def test_noalloc_refuses_a_workspace_above_the_scratch(cuda):
    A = datasets.rmat_graph(14, seed=3)
    plan = api.build_spmm_plan(A, blk_dtype="float32", N=128)
    per_col = plan.workspace_bytes(1, deterministic=True)
    assert per_col > 0
    N = _native.STILE_NOALLOC_SCRATCH_BYTES // per_col + 1
    assert plan.workspace_bytes(N, deterministic=True) > _native.STILE_NOALLOC_SCRATCH_BYTES
    print(f"noalloc refusal N={N}", flush=True)
    lib = _native.load()
    # refused on the size alone, before the pointers are looked at or any launch
    rc = lib.stile_spmm_run_noalloc(plan._h, _native.STILE_BLK_F32, None, N, N, None, N,
                                    _native.STILE_RUN_DETERMINISTIC, None)
    assert rc == 1
    assert b"exceeds the static scratch" in lib.stile_last_error()
    plan.close()

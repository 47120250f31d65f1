"""Parity at the benchmark workloads' full sizes.

Where the FP64 oracle is cheap it is applied to the whole output; otherwise
to a sample, plus size-independent identities checked over every row and
column in FP64 on the host:
    column sums  1^T (A B) = (1^T A) B
    row sums     (A B) 1   = A (B 1)
    attention    every output row is a convex combination of V rows, so it
                 lies between the column-wise min and max of V.
Bars: FP32 paths 1e-5 (norm-wise relative), BF16 blocks 1e-2.
"""
import numpy as np
import pytest
import scipy.sparse

from oracle import ops
from stile_amd import api, datasets

pytestmark = pytest.mark.gpu


def _t(x, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _rel(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-30))


def _sum_identities(A, B, C, tol):
    A64 = scipy.sparse.csr_matrix(A, dtype=np.float64)
    B64 = B.astype(np.float64)
    colsum_ref = np.asarray(A64.sum(axis=0)).reshape(-1) @ B64
    rowsum_ref = A64 @ B64.sum(axis=1)
    C64 = C.astype(np.float64)
    # scale: the sums of |A||B| bound what rounding can reach
    absA = abs(A64)
    col_scale = np.asarray(absA.sum(axis=0)).reshape(-1) @ np.abs(B64)
    row_scale = absA @ np.abs(B64).sum(axis=1)
    assert np.max(np.abs(C64.sum(axis=0) - colsum_ref) / np.maximum(col_scale, 1e-30)) <= tol
    assert np.max(np.abs(C64.sum(axis=1) - rowsum_ref) / np.maximum(row_scale, 1e-30)) <= tol


def test_uniform_spmm_full(cuda):
    A, B = datasets.uniform_spmm(seed=0)
    print("full size: uniform 4096^2 1% N=64 FP32", flush=True)
    plan = api.build_spmm_plan(A, blk_dtype="float32", N=64)
    C = plan.run_spmm(_t(B, cuda)).cpu().numpy()
    assert ops.max_rel_err(C, ops.spmm_fp64(A, B)) <= 1e-5
    _sum_identities(A, B, C, 1e-5)


def test_rmat_sddmm_full(cuda):
    A, X, Y = datasets.sddmm_rmat(seed=0)
    print(f"full size: R-MAT 16 SDDMM K=32 nnz={A.nnz}", flush=True)
    plan = api.build_sddmm_plan(A, Kf=32)
    got = plan.run_sddmm(_t(X, cuda), _t(Y, cuda)).cpu().numpy()
    ref = ops.sddmm_fp64(A, X, Y)
    assert _rel(got, ref) <= 1e-5
    # linearity in X: SDDMM(2X + X', Y) = 2 SDDMM(X, Y) + SDDMM(X', Y), all outputs
    X2 = np.random.default_rng(9).standard_normal(X.shape, dtype=np.float32)
    got2 = plan.run_sddmm(_t(X2, cuda), _t(Y, cuda)).cpu().numpy()
    got3 = plan.run_sddmm(_t(2 * X + X2, cuda), _t(Y, cuda)).cpu().numpy()
    assert _rel(got3, 2.0 * got + got2) <= 1e-5


def test_pruned_layer_full_n4096_bf16(cuda):
    A, B = datasets.pruned_layer(seed=0, N=4096)
    print("full size: pruned layer 3072x768 N=4096 BF16 blocks", flush=True)
    plan = api.build_spmm_plan(A, blk_dtype="bfloat16")
    C = plan.run_spmm(_t(B, cuda)).cpu().numpy()
    cols = np.random.default_rng(1).choice(4096, 256, replace=False)
    assert ops.max_rel_err(C[:, cols], ops.spmm_fp64(A, B[:, cols])) <= 1e-2
    _sum_identities(A, B, C, 1e-2)
    C2 = plan.run_spmm(_t(B, cuda)).cpu().numpy()
    assert np.array_equal(C, C2)


def test_attention_full_64_instances(cuda):
    import torch
    from stile_amd.attention import BatchedSparseAttention
    masks = datasets.attention_masks(count=64)
    print(f"full size: attention 64 masks nnz={sum(m.nnz for m in masks)}", flush=True)
    att = BatchedSparseAttention(masks, 64, blk_dtype="bfloat16", device=cuda)
    n = int(att.row_off[-1])
    g = torch.Generator(device=cuda)
    g.manual_seed(3)
    Q, K, V = (torch.randn(n, 64, device=cuda, generator=g).to(torch.bfloat16) for _ in range(3))
    out = att.forward(Q, K, V).cpu().numpy()
    Vh = V.float().cpu().numpy()
    for b, m in enumerate(masks):
        a, e = att.row_off[b], att.row_off[b + 1]
        lo, hi = Vh[a:e].min(axis=0), Vh[a:e].max(axis=0)
        slack = 1e-2 * (hi - lo)
        assert np.all(out[a:e] >= lo - slack) and np.all(out[a:e] <= hi + slack), f"instance {b}"
    for b in (0, 40):
        m = masks[b]
        a, e = att.row_off[b], att.row_off[b + 1]
        q, k, v = (x[a:e].double().cpu().numpy() for x in (Q, K, V))
        rows = np.repeat(np.arange(m.shape[0]), np.diff(m.indptr))
        s = (q[rows] * k[m.indices]).sum(1) / 8.0
        w = np.exp(s - np.maximum.reduceat(s, m.indptr[:-1])[rows])
        w /= np.add.reduceat(w, m.indptr[:-1])[rows]
        ref = np.zeros_like(q)
        np.add.at(ref, rows, w[:, None] * v[m.indices])
        assert _rel(out[a:e], ref) <= 1e-2, f"instance {b}"


def _gcn_layer_check(cuda, scale, workers=None):
    """GCN layer on R-MAT `scale`: forward rows and dvals on samples vs FP64,
    the forward and transposed products through the sum identities over
    every row and column (FP32 bar 1e-5)."""
    import torch
    from stile_amd.autograd import SparseOperand
    A, H, W = datasets.gcn_graph(scale)
    print(f"full-path: GCN scale {scale} nnz={A.nnz}", flush=True)
    op = SparseOperand(A, blk_dtype="float32", device=cuda, N=128, workers=workers)
    assert op.symmetric
    HW = (H.astype(np.float64) @ W.astype(np.float64)).astype(np.float32)
    vals = op.values.clone().requires_grad_(True)
    HWd = _t(HW, cuda).requires_grad_(True)
    Y = op.spmm(HWd, vals)
    dY = np.random.default_rng(2).standard_normal(Y.shape).astype(np.float32)
    Y.backward(_t(dY, cuda))
    Yh = Y.detach().cpu().numpy()
    rows = np.random.default_rng(0).choice(A.shape[0], 2000, replace=False)
    ref = scipy.sparse.csr_matrix(A[rows], dtype=np.float64) @ HW.astype(np.float64)
    assert _rel(Yh[rows], ref) <= 1e-5
    _sum_identities(A, HW, Yh, 1e-5)
    _sum_identities(A.T.tocsr(), dY, HWd.grad.cpu().numpy(), 1e-5)
    pos = np.random.default_rng(1).choice(A.nnz, 20000, replace=False)
    prow = np.searchsorted(A.indptr, pos, side="right") - 1
    dv_ref = (dY[prow].astype(np.float64) * HW[A.indices[pos]]).sum(1)
    assert _rel(vals.grad.cpu().numpy()[pos], dv_ref) <= 1e-5
    return A, HW, rows, ref


def test_gcn_layer_scale18(cuda):
    """GCN layer on R-MAT scale 18 (7.9 M nonzeros)."""
    _gcn_layer_check(cuda, 18, workers=8)


@pytest.mark.timeout(900)
def test_gcn_layer_scale20_full(cuda):
    """The GCN workload at its full size: R-MAT scale 20, 32.45 M nonzeros,
    F = 128, with the searches on at most 8 worker processes.  FP32 blocks
    as above, then one run with BF16 blocks (bar 1e-2: sampled rows vs FP64
    and the sum identities)."""
    A, HW, rows, ref = _gcn_layer_check(cuda, 20, workers=8)
    assert A.nnz == 32453462
    print("full-path: GCN scale 20 BF16 blocks", flush=True)
    from stile_amd import _native
    pack = api.spmm_pack_blocked(A, api.options_for("bfloat16"), 128, workers=8)
    plan = _native.Plan.spmm(pack, "bfloat16")
    C = plan.run_spmm(_t(HW, cuda)).cpu().numpy()
    assert _rel(C[rows], ref) <= 1e-2
    _sum_identities(A, HW, C, 1e-2)
    plan.close()

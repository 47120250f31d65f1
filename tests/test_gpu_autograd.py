"""Autograd pairing of hybrid SpMM and SDDMM on the GPU.

Checked against (1) torch.sparse gradients in float64 on the CPU, (2) central
finite differences of the GPU forward itself (SpMM and SDDMM are linear in
each argument, so central differences are exact up to rounding; bar 1e-3
relative), and (3) a GCN layer A.(HW) through torch autograd in float64.
FP32 plans: gradient bar 1e-5 relative (norm-wise) to float64.
"""
import numpy as np
import pytest
import scipy.sparse
import torch

from stile_amd import datasets
from stile_amd.autograd import SparseOperand

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def _graph(scale=9, seed=0):
    A = datasets.rmat_graph(scale, seed=seed)
    rng = np.random.default_rng(seed)
    A.data = rng.uniform(-1, 1, A.nnz).astype(np.float32)
    return A


def _torch_sparse(A):
    coo = A.tocoo()
    idx = torch.tensor(np.vstack([coo.row, coo.col]), dtype=torch.long)
    # A is canonical CSR: COO built from it keeps CSR order of the values
    return idx, torch.tensor(A.data, dtype=torch.float64)


@pytest.mark.parametrize("blk", ["float32", "bfloat16"])
def test_spmm_grads_match_torch_sparse(cuda, blk):
    A = _graph(9, 1)
    N = 48
    rng = np.random.default_rng(2)
    B = rng.standard_normal((A.shape[1], N)).astype(np.float32)
    G = rng.standard_normal((A.shape[0], N)).astype(np.float32)
    print(f"autograd spmm blk={blk}", flush=True)
    op = SparseOperand(A, blk_dtype=blk, device=cuda)
    vals = op.values.clone().requires_grad_(True)
    Bd = torch.from_numpy(B).to(cuda).requires_grad_(True)
    C = op.spmm(Bd, vals)
    (C * torch.from_numpy(G).to(cuda)).sum().backward()
    idx, v64 = _torch_sparse(A)
    v64.requires_grad_(True)
    B64 = torch.from_numpy(B).double().requires_grad_(True)
    S = torch.sparse_coo_tensor(idx, v64, A.shape)
    C64 = torch.sparse.mm(S, B64)
    (C64 * torch.from_numpy(G).double()).sum().backward()
    tol = 1e-5 if blk == "float32" else 1e-2
    assert _rel(C.detach().cpu(), C64.detach()) <= tol
    assert _rel(Bd.grad.cpu(), B64.grad) <= tol
    assert _rel(vals.grad.cpu(), v64.grad) <= 1e-5   # SDDMM path is FP32 in both modes


def test_sddmm_grads_match_dense_reference(cuda):
    A = _graph(9, 3)
    Kf = 32
    rng = np.random.default_rng(4)
    X = rng.standard_normal((A.shape[0], Kf)).astype(np.float32)
    Y = rng.standard_normal((A.shape[1], Kf)).astype(np.float32)
    G = rng.standard_normal(A.nnz).astype(np.float32)
    print("autograd sddmm", flush=True)
    op = SparseOperand(A, device=cuda)
    Xd = torch.from_numpy(X).to(cuda).requires_grad_(True)
    Yd = torch.from_numpy(Y).to(cuda).requires_grad_(True)
    out = op.sddmm(Xd, Yd)
    (out * torch.from_numpy(G).to(cuda)).sum().backward()
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    X64, Y64 = torch.from_numpy(X).double().requires_grad_(True), torch.from_numpy(Y).double().requires_grad_(True)
    o64 = (X64[rows] * Y64[A.indices]).sum(1)
    (o64 * torch.from_numpy(G).double()).sum().backward()
    assert _rel(out.detach().cpu(), o64.detach()) <= 1e-5
    assert _rel(Xd.grad.cpu(), X64.grad) <= 1e-5
    assert _rel(Yd.grad.cpu(), Y64.grad) <= 1e-5
    # the SpMM plans still hold A's values after the SDDMM backward used them
    B = torch.from_numpy(rng.standard_normal((A.shape[1], 16)).astype(np.float32)).to(cuda)
    C = op.spmm(B).cpu().numpy()
    assert _rel(C, A.astype(np.float64) @ B.cpu().double().numpy()) <= 1e-5


def test_finite_differences(cuda):
    A = _graph(8, 5)
    N = 16
    rng = np.random.default_rng(6)
    op = SparseOperand(A, device=cuda)
    B = torch.from_numpy(rng.standard_normal((A.shape[1], N)).astype(np.float32)).to(cuda)
    G = torch.from_numpy(rng.standard_normal((A.shape[0], N)).astype(np.float32)).to(cuda)
    vals = op.values.clone().requires_grad_(True)
    Bv = B.clone().requires_grad_(True)
    (op.spmm(Bv, vals) * G).sum().backward()
    print("autograd finite differences", flush=True)
    eps = 0.5

    def loss(v, b):
        with torch.no_grad():
            return float((op.spmm(b, v).double() * G.double()).sum())

    for _ in range(6):
        p = int(rng.integers(A.nnz))
        vp, vm = vals.detach().clone(), vals.detach().clone()
        vp[p] += eps
        vm[p] -= eps
        fd = (loss(vp, B) - loss(vm, B)) / (2 * eps)
        assert abs(fd - float(vals.grad[p])) <= 1e-3 * max(1.0, abs(fd))
        i, j = int(rng.integers(B.shape[0])), int(rng.integers(N))
        bp, bm = B.clone(), B.clone()
        bp[i, j] += eps
        bm[i, j] -= eps
        fd = (loss(vals.detach(), bp) - loss(vals.detach(), bm)) / (2 * eps)
        assert abs(fd - float(Bv.grad[i, j])) <= 1e-3 * max(1.0, abs(fd))


def test_gcn_layer_against_torch_autograd(cuda):
    A, H, W = datasets.gcn_graph(scale=11, feat=64, seed=1)
    print("autograd gcn layer", flush=True)
    op = SparseOperand(A, device=cuda)
    Hd = torch.from_numpy(H).to(cuda).requires_grad_(True)
    Wd = torch.from_numpy(W).to(cuda).requires_grad_(True)
    vals = op.values.clone().requires_grad_(True)
    out = op.spmm(Hd @ Wd, vals)
    G = torch.from_numpy(np.random.default_rng(7).standard_normal(out.shape).astype(np.float32))
    (out * G.to(cuda)).sum().backward()
    idx, v64 = _torch_sparse(A)
    v64.requires_grad_(True)
    H64 = torch.from_numpy(H).double().requires_grad_(True)
    W64 = torch.from_numpy(W).double().requires_grad_(True)
    o64 = torch.sparse.mm(torch.sparse_coo_tensor(idx, v64, A.shape), H64 @ W64)
    (o64 * G.double()).sum().backward()
    assert _rel(out.detach().cpu(), o64.detach()) <= 1e-5
    assert _rel(Hd.grad.cpu(), H64.grad) <= 1e-5
    assert _rel(Wd.grad.cpu(), W64.grad) <= 1e-5
    assert _rel(vals.grad.cpu(), v64.grad) <= 1e-5


def test_set_values_reuses_the_plan(cuda):
    A = _graph(9, 8)
    op = SparseOperand(A, device=cuda)
    rng = np.random.default_rng(9)
    B = torch.from_numpy(rng.standard_normal((A.shape[1], 64)).astype(np.float32)).to(cuda)
    print("autograd set_values", flush=True)
    for _ in range(3):
        v = rng.uniform(-2, 2, A.nnz).astype(np.float32)
        C = op.spmm(B, torch.from_numpy(v).to(cuda)).cpu().numpy()
        A2 = scipy.sparse.csr_matrix((v, A.indices, A.indptr), shape=A.shape)
        assert _rel(C, A2.astype(np.float64) @ B.cpu().double().numpy()) <= 1e-5
        Ct = op.spmm_t(B[:A.shape[0]] if A.shape[0] == A.shape[1] else B, torch.from_numpy(v).to(cuda))
        assert _rel(Ct.cpu().numpy(), A2.T.astype(np.float64) @ B.cpu().double().numpy()) <= 1e-5


def test_gcn_layer_slab_weight_gradient(cuda):
    """gcn_layer on a graph of 65,536 nodes, where H.W's weight gradient runs
    as the slab-split product; every gradient vs FP64 torch."""
    from stile_amd.autograd import gcn_layer
    A, H, W = datasets.gcn_graph(scale=16, feat=32, seed=2)
    print(f"autograd gcn layer (slab weight gradient) nnz={A.nnz}", flush=True)
    op = SparseOperand(A, device=cuda)
    Hd = torch.from_numpy(H).to(cuda).requires_grad_(True)
    Wd = torch.from_numpy(W).to(cuda).requires_grad_(True)
    vals = op.values.clone().requires_grad_(True)
    out = gcn_layer(op, Hd, Wd, vals)
    G = torch.from_numpy(np.random.default_rng(8).standard_normal(out.shape).astype(np.float32))
    (out * G.to(cuda)).sum().backward()
    idx, v64 = _torch_sparse(A)
    v64.requires_grad_(True)
    H64 = torch.from_numpy(H).double().requires_grad_(True)
    W64 = torch.from_numpy(W).double().requires_grad_(True)
    o64 = torch.sparse.mm(torch.sparse_coo_tensor(idx, v64, A.shape), H64 @ W64)
    (o64 * G.double()).sum().backward()
    assert _rel(out.detach().cpu(), o64.detach()) <= 1e-5
    assert _rel(Hd.grad.cpu(), H64.grad) <= 1e-5
    assert _rel(Wd.grad.cpu(), W64.grad) <= 1e-5
    assert _rel(vals.grad.cpu(), v64.grad) <= 1e-5

"""Differentiable hybrid SpMM / SDDMM: each is the other's adjoint.

For C = A.B with A sparse (pattern S, values a):
    dB      = A^T . dC                      (SpMM with the transposed tile set)
    da[p]   = <dC[row p], B[col p]>         (SDDMM(dC, B) on S)
For out = SDDMM(X, Y) on S (out[p] = <X[row p], Y[col p]>):
    dX      = S(dout) . Y                   (SpMM with S's tiles, values dout)
    dY      = S(dout)^T . X                 (SpMM with S^T's tiles, values dout)

A SparseOperand builds, once, the plans for S and S^T (both with value ids
mapping every packed slot to S's CSR value order) and the SDDMM plan on S;
values are pushed to the device plans with stile_plan_set_values whenever the
tensor feeding them changes, so the plans are reused across training steps.

This pairing is not part of the reference (it has no backward pass); it is
the config the GCN workload (forward A.(HW), backward A^T SpMM + SDDMM for
dA) needs.
"""
import copy
import os

import numpy as np
import scipy.sparse
import torch

from . import _native, api


def _transpose_with_ids(A):
    """A^T as CSR whose data are the ids (CSR positions) of A's nonzeros."""
    ids = scipy.sparse.csr_matrix((np.arange(A.nnz, dtype=np.float64), A.indices, A.indptr), shape=A.shape)
    T = ids.T.tocsr()
    T.sort_indices()
    return T


def _remap_ids(pack, id_map):
    """Replace the pack's value ids (positions in the packed matrix) by id_map[ids]."""
    for g in pack["ell"] + pack["tc"]:
        a = g["a_ids"]
        m = a >= 0
        a[m] = id_map[a[m]]
    return pack


class SparseOperand(object):
    """A sparse matrix prepared for differentiable SpMM and SDDMM on the GPU.

    blk_dtype: arithmetic of the MFMA blocks ("float32" keeps the whole
    computation FP32; "bfloat16" runs the dense blocks in BF16 with FP32
    accumulation).  opts: api.SearchOptions for the tile search."""

    def __init__(self, A, blk_dtype="float32", opts=None, device="cuda", N=32, workers=None,
                 nnz_per_part=1 << 21, tune=False, log=None, cache_dir=None):
        A = scipy.sparse.csr_matrix(A, dtype=np.float32)
        A.sort_indices()
        self.A = A
        self.shape = A.shape
        self.nnz = A.nnz
        self.device = torch.device(device)
        ones = scipy.sparse.csr_matrix((np.ones(A.nnz, np.float32), A.indices, A.indptr), shape=A.shape)
        T = _transpose_with_ids(A)
        id_map = T.data.astype(np.int64)
        # a symmetric pattern has A^T's CSR structure = A's: the search result
        # for A serves A^T with its value ids mapped through the transpose
        self.symmetric = (A.shape[0] == A.shape[1] and np.array_equal(T.indptr, A.indptr)
                          and np.array_equal(T.indices, A.indices))
        kw = dict(workers=workers, nnz_per_part=nnz_per_part)
        self.tuning = None
        if tune:
            # plans chosen by measured latency at width N (api.tune_plan)
            cp = (lambda name: None if cache_dir is None else os.path.join(cache_dir, name))
            self.plan, pack, rep = api.tune_plan("spmm", ones, blk_dtype, N, workers=workers, log=log,
                                                 cache=cp("spmm.npz"))
            srep = None
            if self.symmetric:
                tpack = copy.deepcopy(pack)
            else:
                tones = scipy.sparse.csr_matrix((np.ones(T.nnz, np.float32), T.indices, T.indptr), shape=T.shape)
                plan_t, tpack, _ = api.tune_plan("spmm", tones, blk_dtype, N, workers=workers, log=log,
                                                 cache=cp("spmm_t.npz"))
                plan_t.close()
            self.sddmm_plan, _, srep = api.tune_plan("sddmm", ones, "float32", N, workers=workers, log=log,
                                                     cache=cp("sddmm.npz"))
            self.tuning = {"spmm": rep, "sddmm": srep}
        else:
            opts = api.options_for(blk_dtype, opts)
            pack = api.spmm_pack_blocked(ones, opts, N, **kw)
            if self.symmetric:
                tpack = copy.deepcopy(pack)
            else:
                tones = scipy.sparse.csr_matrix((np.ones(T.nnz, np.float32), T.indices, T.indptr), shape=T.shape)
                tpack = api.spmm_pack_blocked(tones, opts, N, **kw)
            spack = api.sddmm_pack_blocked(ones, opts, 32, **kw)
            self.plan = _native.Plan.spmm(pack, blk_dtype)
            self.sddmm_plan = _native.Plan.sddmm(spack, "float32")
        self.plan_t = _native.Plan.spmm(_remap_ids(tpack, id_map), blk_dtype)
        self.values = torch.from_numpy(A.data.astype(np.float32)).to(self.device)
        # (tensor, data_ptr, version) of the values now in the SpMM plans; the
        # tensor is held so its memory cannot be reused by another tensor
        # with the same address and version
        self._loaded = None

    def _load(self, vals):
        if self._loaded is not None:
            t, ptr, ver = self._loaded
            if ptr == vals.data_ptr() and ver == vals._version and t.numel() == vals.numel():
                return
        self.plan.set_values(vals)
        self.plan_t.set_values(vals)
        self._loaded = (vals, vals.data_ptr(), vals._version)

    def spmm(self, B, values=None):
        """A.B (differentiable in B and, if given, in values)."""
        v = self.values if values is None else values
        return _SpMM.apply(v, B, self)

    def sddmm(self, X, Y):
        """out[p] = <X[row p], Y[col p]> on A's pattern (differentiable)."""
        return _SDDMM.apply(X, Y, self)

    def spmm_t(self, B, values=None):
        v = self.values if values is None else values
        self._load(v.detach())
        return self.plan_t.run_spmm(B.contiguous())


class _SpMM(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, B, op):
        B = B.contiguous()
        op._load(values.detach())
        C = op.plan.run_spmm(B)
        ctx.save_for_backward(values, B)
        ctx.op = op
        return C

    @staticmethod
    def backward(ctx, dC):
        values, B = ctx.saved_tensors
        op = ctx.op
        dC = dC.contiguous()
        dvals = dB = None
        if ctx.needs_input_grad[1]:
            op._load(values.detach())
            dB = op.plan_t.run_spmm(dC)
        if ctx.needs_input_grad[0]:
            dvals = op.sddmm_plan.run_sddmm(dC, B)
        return dvals, dB, None


class _SDDMM(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, Y, op):
        X = X.contiguous()
        Y = Y.contiguous()
        out = op.sddmm_plan.run_sddmm(X, Y)
        ctx.save_for_backward(X, Y)
        ctx.op = op
        return out

    @staticmethod
    def backward(ctx, dout):
        X, Y = ctx.saved_tensors
        op = ctx.op
        dout = dout.contiguous()
        dX = dY = None
        op._load(dout)
        if ctx.needs_input_grad[0]:
            dX = op.plan.run_spmm(Y)
        if ctx.needs_input_grad[1]:
            dY = op.plan_t.run_spmm(X)
        return dX, dY, None


def spmm(op, B, values=None):
    return op.spmm(B, values)


def sddmm(op, X, Y):
    return op.sddmm(X, Y)


def _split_k_tn(H, G, slab_rows=16384):
    """H^T G for tall H, G (M x F, M >> F).  As one GEMM the reduction over M
    leaves the library a handful of F x F output tiles for the whole GPU; cut
    into M / slab_rows slabs it is one batched product with that many times
    more tiles, then a sum over the slabs.  FP32 only (other dtypes, and M not
    a multiple of the slab, take the single GEMM)."""
    M = H.shape[0]
    if H.dtype != torch.float32 or G.dtype != torch.float32 or M % slab_rows or M < 4 * slab_rows:
        return H.t() @ G
    S = M // slab_rows
    H3 = H.contiguous().view(S, slab_rows, H.shape[1])
    G3 = G.contiguous().view(S, slab_rows, G.shape[1])
    return torch.bmm(H3.transpose(1, 2), G3).sum(0)


class _TallMatmul(torch.autograd.Function):
    """Y = H W for a tall H; the weight gradient H^T dY runs split over row
    slabs (see _split_k_tn), the input gradient dY W^T as one GEMM."""

    @staticmethod
    def forward(ctx, H, W):
        ctx.save_for_backward(H, W)
        return H @ W

    @staticmethod
    def backward(ctx, dY):
        H, W = ctx.saved_tensors
        dH = dW = None
        if ctx.needs_input_grad[0]:
            dH = dY @ W.t()
        if ctx.needs_input_grad[1]:
            dW = _split_k_tn(H, dY)
        return dH, dW


def tall_matmul(H, W):
    return _TallMatmul.apply(H, W)


def gcn_layer(op, H, W, values=None):
    """One GCN layer A.(H.W) (the GCN workload's forward); H.W is a
    tall_matmul, so the layer's weight gradient is reduced in row slabs."""
    return op.spmm(tall_matmul(H, W), values)

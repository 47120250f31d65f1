"""Batched sparse attention (SDDMM -> row softmax -> SpMM) over uneven masks,
eager and HIP-graph replayed, against a float64 CPU reference per instance.

Bars: FP32 blocks 1e-5, BF16 blocks 1e-2 (norm-wise relative, per instance).
Graph replays are checked one by one with eager work between them.
"""
import numpy as np
import pytest
import torch

from stile_amd import datasets
from stile_amd.attention import BatchedSparseAttention

pytestmark = pytest.mark.gpu


def _reference(masks, Q, K, V, offs):
    outs = []
    d = Q.shape[1]
    for b, m in enumerate(masks):
        q = Q[offs[b]:offs[b + 1]].astype(np.float64)
        k = K[offs[b]:offs[b + 1]].astype(np.float64)
        v = V[offs[b]:offs[b + 1]].astype(np.float64)
        rows = np.repeat(np.arange(m.shape[0]), np.diff(m.indptr))
        s = (q[rows] * k[m.indices]).sum(1) / np.sqrt(d)
        out = np.zeros_like(q)
        for i in range(m.shape[0]):
            a, e = m.indptr[i], m.indptr[i + 1]
            if a == e:
                continue
            w = np.exp(s[a:e] - s[a:e].max())
            w /= w.sum()
            out[i] = w @ v[m.indices[a:e]]
        outs.append(out)
    return np.vstack(outs)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _masks(count=6, seed=0):
    rng = np.random.default_rng(seed)
    lens = rng.integers(128, 1025, size=count)
    return [datasets.attention_mask(int(L), band=32, n_random=4, rng=rng) for L in lens]


@pytest.mark.parametrize("blk,tol,fused", [("float32", 1e-5, False), ("bfloat16", 1e-2, False),
                                           ("bfloat16", 1e-2, True)])
def test_batched_attention_eager_and_graph(cuda, blk, tol, fused):
    masks = _masks()
    d = 64
    att = BatchedSparseAttention(masks, d, blk_dtype=blk, device=cuda, workers=4, fused=fused)
    offs = att.row_off
    n = int(offs[-1])
    rng = np.random.default_rng(1)

    def inputs():
        return [torch.from_numpy(rng.standard_normal((n, d)).astype(np.float32)).to(torch.bfloat16) for _ in range(3)]

    Q, K, V = inputs()
    print(f"attention eager blk={blk} fused={fused} instances={len(masks)} nnz={att.nnz}", flush=True)
    ref = _reference(masks, *(t.float().numpy() for t in (Q, K, V)), offs)
    out = att.forward(Q.to(cuda), K.to(cuda), V.to(cuda)).cpu().numpy()
    assert _rel(out, ref) <= tol
    gQ, gK, gV = Q.to(cuda), K.to(cuda), V.to(cuda)
    print(f"attention graph capture blk={blk}", flush=True)
    att.capture(gQ, gK, gV)
    for rep in range(3):
        Q, K, V = inputs()
        gQ.copy_(Q.to(cuda))
        gK.copy_(K.to(cuda))
        gV.copy_(V.to(cuda))
        _ = (gQ.float() * 2).sum()          # eager work between replays
        torch.cuda.synchronize()
        got = att.replay()
        torch.cuda.synchronize()
        ref = _reference(masks, *(t.float().numpy() for t in (Q, K, V)), offs)
        assert _rel(got.cpu().numpy(), ref) <= tol, f"replay {rep}"


def test_row_softmax_kernel(cuda):
    from stile_amd import _native
    m = datasets.attention_mask(300, band=16, n_random=3, rng=np.random.default_rng(3))
    vals = np.random.default_rng(4).standard_normal(m.nnz).astype(np.float32) * 5
    print("row softmax", flush=True)
    ip = torch.from_numpy(m.indptr.astype(np.int32)).to(cuda)
    got = _native.csr_row_softmax_f32(ip, torch.from_numpy(vals).to(cuda), 0.5).cpu().numpy()
    for i in range(m.shape[0]):
        a, e = m.indptr[i], m.indptr[i + 1]
        w = np.exp(0.5 * vals[a:e].astype(np.float64) - 0.5 * vals[a:e].max())
        w /= w.sum()
        assert np.max(np.abs(got[a:e] - w)) <= 1e-6


def test_fused_attention_with_sparse_part_and_empty_rows(cuda):
    """Masks whose search leaves a sparse part (random columns) and rows
    without entries: the fused kernel's per-row path and its zero rows."""
    import scipy.sparse
    from stile_amd import api
    rng = np.random.default_rng(11)
    m1 = datasets.attention_mask(333, band=8, n_random=6, rng=rng).tolil()
    m1[5, :] = 0
    m1[200:216, :] = 0
    m1 = m1.tocsr()
    m1.eliminate_zeros()
    m2 = scipy.sparse.random(150, 150, density=0.03, format="csr", random_state=3, dtype=np.float32)
    m2.data[:] = 1.0
    masks = [m1, m2]
    att = BatchedSparseAttention(masks, 64, device=cuda, workers=1, fused=True,
                                 opts=api.SearchOptions(only_ELL=True))
    assert att.att_plan.info()["sparse_nnz"] > 0
    n = int(att.row_off[-1])
    Q, K, V = (torch.from_numpy(rng.standard_normal((n, 64)).astype(np.float32)).to(torch.bfloat16)
               for _ in range(3))
    print("fused attention sparse part + empty rows", flush=True)
    out = att.forward(Q.to(cuda), K.to(cuda), V.to(cuda)).cpu().numpy()
    ref = _reference(masks, *(t.float().numpy() for t in (Q, K, V)), att.row_off)
    assert _rel(out, ref) <= 1e-2
    assert np.all(out[5] == 0) and np.all(out[200:216] == 0)

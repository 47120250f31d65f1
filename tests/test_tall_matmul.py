"""The GCN layer's dense product H.W with its slab-split weight gradient
(stile_amd.autograd.tall_matmul), on the CPU against FP64 closed forms."""
import numpy as np
import torch

from stile_amd.autograd import _split_k_tn, tall_matmul


def test_split_k_tn_matches_the_single_product():
    rng = np.random.default_rng(0)
    for M, slab in ((8 * 64, 64), (5 * 32, 32), (100, 16)):  # split, split, not a multiple (single GEMM)
        H = rng.standard_normal((M, 12)).astype(np.float32)
        G = rng.standard_normal((M, 7)).astype(np.float32)
        got = _split_k_tn(torch.from_numpy(H), torch.from_numpy(G), slab_rows=slab).numpy()
        ref = H.astype(np.float64).T @ G.astype(np.float64)
        assert got.shape == (12, 7)
        assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(H.T.astype(np.float64)) @ np.abs(G))


def test_split_k_tn_row_slabs_are_summed_not_mixed():
    # one nonzero row per slab: a slab/feature transposition in the view
    # would move these products to the wrong output entries
    M, slab = 4 * 8, 8
    H = torch.zeros(M, 3)
    G = torch.zeros(M, 2)
    for s in range(4):
        H[s * slab + s, s % 3] = s + 1.0
        G[s * slab + s, s % 2] = 10.0
    got = _split_k_tn(H, G, slab_rows=slab)
    ref = torch.zeros(3, 2)
    for s in range(4):
        ref[s % 3, s % 2] += (s + 1.0) * 10.0
    assert torch.equal(got, ref)


def test_tall_matmul_gradients():
    rng = np.random.default_rng(1)
    M, F = 16384 * 4, 8  # long enough for the slab path at its default slab
    H = torch.from_numpy(rng.standard_normal((M, F)).astype(np.float32)).requires_grad_(True)
    W = torch.from_numpy(rng.standard_normal((F, F)).astype(np.float32)).requires_grad_(True)
    G = torch.from_numpy(rng.standard_normal((M, F)).astype(np.float32))
    (tall_matmul(H, W) * G).sum().backward()
    H64, W64, G64 = (x.detach().double() for x in (H, W, G))
    assert torch.allclose(H.grad.double(), G64 @ W64.T, rtol=0, atol=1e-4)
    dW = H64.T @ G64
    assert torch.max(torch.abs(W.grad.double() - dW)) <= 1e-5 * torch.max(torch.abs(H64).T @ torch.abs(G64))

"""Parity of the HIP kernels (through the C ABI) against the FP64 oracle and
the reference's own FP32 CPU path, on packed groups built directly from the
matrix (tests/_packs.py), so every group shape is covered independently of the
search.

Bars (written in each test):
  FP32 kernels: max |got - ref64| within the worst-case FP32 dot-product bound
  gamma_n |A||B| (bound_ratio <= 1), and norm-wise relative difference to the
  reference's scipy FP32 product <= 1e-5.
  BF16 MFMA blocks: norm-wise relative error vs FP64 <= 1e-2.
"""
import numpy as np
import pytest
import scipy.sparse

from oracle import ops
from stile_amd import api, datasets
from tests import _packs

pytestmark = pytest.mark.gpu


def _t(x, dev, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(x))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(dev)


def _check_fp32(got, A, B, ref32=None):
    ref64 = ops.spmm_fp64(A, B)
    bound = ops.spmm_abs_bound(A, B)
    terms = int(np.diff(scipy.sparse.csr_matrix(A).indptr).max(initial=1))
    ratio = ops.bound_ratio(got, ref64, bound, terms)
    assert ratio <= 1.0, f"error exceeds the FP32 summation bound ({ratio:.3g} x)"
    r32 = ops.spmm_ref_fp32(A, B) if ref32 is None else ref32
    rel = ops.max_rel_err(got, r32)
    assert rel <= 1e-5, f"relative difference to scipy FP32 {rel:.3g} > 1e-5"
    return ratio, rel


def _csr_dev(A, dev):
    import torch
    A = scipy.sparse.csr_matrix(A, dtype=np.float32)
    return (_t(A.indptr.astype(np.int32), dev), _t(A.indices.astype(np.int32), dev),
            _t(A.data.astype(np.float32), dev))


@pytest.mark.parametrize("N", [1, 3, 16, 64, 100, 128, 300, 1024])
def test_csr_spmm_uniform(cuda, N):
    from stile_amd import _native
    A, _ = datasets.uniform_spmm(seed=0)
    B = np.random.default_rng(1).standard_normal((A.shape[1], N), dtype=np.float32)
    print(f"csr_spmm uniform N={N}", flush=True)
    C = _native.csr_spmm_f32(*_csr_dev(A, cuda), _t(B, cuda)).cpu().numpy()
    _check_fp32(C, A, B)


def test_csr_spmm_edges(cuda):
    from stile_amd import _native
    import torch
    # empty rows, an empty matrix, a hub row longer than 4096 (compensated sum)
    rng = np.random.default_rng(3)
    M, K, N = 300, 9000, 64
    A = scipy.sparse.random(M, K, density=0.002, format="csr", random_state=3, dtype=np.float32)
    A = A.tolil()
    A[7, :] = rng.uniform(-1, 1, K).astype(np.float32)       # hub row, 9000 nonzeros
    A[8, :] = 0
    A = A.tocsr()
    B = rng.standard_normal((K, N), dtype=np.float32)
    print("csr_spmm edges", flush=True)
    C = _native.csr_spmm_f32(*_csr_dev(A, cuda), _t(B, cuda)).cpu().numpy()
    _check_fp32(C, A, B)
    assert np.all(C[8] == 0)
    Z = scipy.sparse.csr_matrix((5, 7), dtype=np.float32)
    C0 = _native.csr_spmm_f32(*_csr_dev(Z, cuda), _t(np.ones((7, 4), np.float32), cuda))
    assert torch.all(C0 == 0)


@pytest.mark.parametrize("blk", ["float32", "bfloat16"])
@pytest.mark.parametrize("layout", ["win1d", "rows2d", "k48"])
@pytest.mark.parametrize("N", [64, 70, 100, 256])
def test_hybrid_spmm_pruned(cuda, blk, layout, N):
    from stile_amd import _native
    A, _ = datasets.pruned_layer(seed=2, M=512, K=256, N=N)
    B = np.random.default_rng(4).standard_normal((A.shape[1], N), dtype=np.float32)
    perm = np.random.default_rng(5).permutation(A.shape[0]) if layout == "rows2d" else None
    G = 48 if layout == "k48" else 16
    pack = _packs.spmm_pack(A, min_fill=64, perm=perm, tc_group_cols=G, sentinel_rows=5)
    assert pack["tc"] and pack["ell"]
    print(f"hybrid spmm blk={blk} layout={layout} N={N}", flush=True)
    plan = _native.Plan.spmm(pack, blk)
    info = plan.info()
    assert info["block_nnz"] + info["sparse_nnz"] == A.nnz
    C = plan.run_spmm(_t(B, cuda)).cpu().numpy()
    if blk == "float32":
        _check_fp32(C, A, B)
    else:
        rel = ops.max_rel_err(C, ops.spmm_fp64(A, B))
        assert rel <= 1e-2, f"bf16 blocks: relative error {rel:.3g} > 1e-2"


def test_hybrid_spmm_full_pruned_layer(cuda):
    from stile_amd import _native
    A, B = datasets.pruned_layer(seed=0, N=128)
    pack = _packs.spmm_pack(A, min_fill=200)
    for blk, tol in (("float32", None), ("bfloat16", 1e-2)):
        print(f"hybrid spmm pruned layer 3072x768 N=128 blk={blk}", flush=True)
        plan = _native.Plan.spmm(pack, blk)
        C = plan.run_spmm(_t(B, cuda)).cpu().numpy()
        if tol is None:
            _check_fp32(C, A, B)
        else:
            assert ops.max_rel_err(C, ops.spmm_fp64(A, B)) <= tol


def test_hybrid_spmm_batched_block_diagonal(cuda):
    from stile_amd import _native
    rng = np.random.default_rng(9)
    mats = [datasets.pruned_layer(seed=s, M=16 * m, K=16 * k, N=8)[0]
            for s, (m, k) in enumerate([(5, 3), (9, 17), (2, 2), (33, 8)])]
    packs = [_packs.spmm_pack(A, min_fill=64, I=4, Kell=8) for A in mats]
    N = 72
    Bs = [rng.standard_normal((A.shape[1], N), dtype=np.float32) for A in mats]
    print("hybrid spmm batched", flush=True)
    plan = _native.Plan.spmm(packs, "float32")
    C = plan.run_spmm(_t(np.vstack(Bs), cuda)).cpu().numpy()
    bd = scipy.sparse.block_diag(mats, format="csr")
    _check_fp32(C, bd, np.vstack(Bs))


def test_hybrid_spmm_only_ell_and_hub(cuda):
    from stile_amd import _native
    rng = np.random.default_rng(11)
    M, K, N = 97, 6000, 40
    A = scipy.sparse.random(M, K, density=0.01, format="csr", random_state=11, dtype=np.float32).tolil()
    A[13, :5000] = rng.uniform(-1, 1, 5000).astype(np.float32)
    A = A.tocsr()
    B = rng.standard_normal((K, N), dtype=np.float32)
    pack = _packs.spmm_pack(A, min_fill=10 ** 9)
    assert not pack["tc"]
    print("hybrid spmm only ELL + hub row", flush=True)
    plan = _native.Plan.spmm(pack, "float32")
    assert plan.info()["hub_rows"] == 1
    C = plan.run_spmm(_t(B, cuda)).cpu().numpy()
    _check_fp32(C, A, B)


@pytest.mark.parametrize("Kf", [32, 64, 20, 7])
def test_csr_sddmm(cuda, Kf):
    from stile_amd import _native
    A = datasets.rmat_graph(scale=11, seed=1)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((A.shape[0], Kf), dtype=np.float32)
    Y = rng.standard_normal((A.shape[1], Kf), dtype=np.float32)
    print(f"csr_sddmm Kf={Kf}", flush=True)
    ip, ix, _ = _csr_dev(A, cuda)
    out = _native.csr_sddmm_f32(ip, ix, _t(X, cuda), _t(Y, cuda)).cpu().numpy()
    ref = ops.sddmm_fp64(A, X, Y)
    ratio = ops.bound_ratio(out, ref, ops.sddmm_abs_bound(A, X, Y), Kf)
    assert ratio <= 1.0
    assert ops.max_rel_err(out, ops.sddmm_ref_fp32(A, X, Y)) <= 1e-5


@pytest.mark.parametrize("Kf", [32, 64, 24])
def test_hybrid_sddmm(cuda, Kf):
    from stile_amd import _native
    A, _ = datasets.pruned_layer(seed=3, M=256, K=320, block_keep=0.2, residual=0.02)
    A = scipy.sparse.csr_matrix(A)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((A.shape[0], Kf), dtype=np.float32)
    Y = rng.standard_normal((A.shape[1], Kf), dtype=np.float32)
    pack = _packs.sddmm_pack(A, min_fill=64)
    assert pack["tc"] and pack["g1d"]
    print(f"hybrid sddmm Kf={Kf}", flush=True)
    plan = _native.Plan.sddmm(pack)
    out = plan.run_sddmm(_t(X, cuda), _t(Y, cuda)).cpu().numpy()
    ref = ops.sddmm_fp64(A, X, Y)
    assert ops.bound_ratio(out, ref, ops.sddmm_abs_bound(A, X, Y), Kf) <= 1.0
    assert ops.max_rel_err(out, ops.sddmm_ref_fp32(A, X, Y)) <= 1e-5


@pytest.mark.parametrize("parts,N", [(1, 64), (4, 100)])
def test_hyb_ell_bucket_comparator(cuda, parts, N):
    from stile_amd import hyb
    A = datasets.rmat_graph(11, seed=5)
    A.data = np.random.default_rng(1).uniform(-1, 1, A.nnz).astype(np.float32)
    B = np.random.default_rng(2).standard_normal((A.shape[1], N), dtype=np.float32)
    print(f"hyb ELL-bucket comparator parts={parts} N={N}", flush=True)
    h = hyb.HybSpMM(A, parts, device=cuda)
    C = h.run(_t(B, cuda)).cpu().numpy()
    ref64 = ops.spmm_fp64(A, B)
    assert ops.max_rel_err(C, ref64) <= 1e-5


def test_output_and_index_arguments_are_checked(cuda):
    """Caller-supplied outputs and index arrays are checked before their
    pointers reach a kernel (a wrong shape or dtype raises, nothing runs)."""
    import torch
    from stile_amd import _native
    A = datasets.uniform_spmm(seed=3, M=64, K=48, nnz=300, N=8)[0]
    print("argument checks", flush=True)
    plan = api.build_spmm_plan(A, blk_dtype="float32")
    B = torch.randn(48, 8, device=cuda)
    for bad in (torch.empty(63, 8, device=cuda), torch.empty(64, 8, device=cuda, dtype=torch.float64),
                torch.empty(8, 64, device=cuda).t(), torch.empty(64, 8)):
        with pytest.raises(_native.StileError):
            plan.run_spmm(B, out=bad)
    ip = torch.from_numpy(A.indptr.astype(np.int64)).to(cuda)
    ix = torch.from_numpy(A.indices.astype(np.int32)).to(cuda)
    vv = torch.from_numpy(A.data.astype(np.float32)).to(cuda)
    with pytest.raises(_native.StileError):
        _native.csr_spmm_f32(ip, ix, vv, B)
    ok = _native.csr_spmm_f32(ip.int(), ix, vv, B, out=torch.empty(64, 8, device=cuda))
    assert ops.max_rel_err(ok.cpu().numpy(), ops.spmm_fp64(A, B.cpu().numpy())) <= 1e-5


def test_malformed_packs_are_rejected(cuda):
    """Plan creation validates every index a kernel would follow: a column
    past K, a row past M or a non-monotone window pointer raise StileError
    (nothing is uploaded or launched)."""
    import copy
    from stile_amd import _native
    A, _ = datasets.pruned_layer(seed=4, M=128, K=64, N=16)
    pack = _packs.spmm_pack(A, min_fill=64)
    assert pack["tc"] and pack["ell"]
    print("malformed packs", flush=True)
    _native.Plan.spmm(pack, "float32").close()  # the valid pack builds

    def broken(edit):
        p = copy.deepcopy(pack)
        edit(p)
        with pytest.raises(_native.StileError):
            _native.Plan.spmm(p, "float32")

    def ell_col(p):
        g = p["ell"][0]
        i = int(np.flatnonzero(np.asarray(g["a"]).reshape(-1) != 0)[0])
        g["indices_k"].reshape(-1)[i] = A.shape[1]

    def tc_col(p):
        g = p["tc"][0]
        t = int(np.flatnonzero(np.abs(np.asarray(g["a"])).sum(axis=1).reshape(len(g["indices"]), -1).sum(1))[0])
        g["indices"][t] = A.shape[1] + 5
    def tc_indptr(p):
        g = p["tc"][0]
        g["indptr"][1], g["indptr"][0] = g["indptr"][0] + len(g["indices"]) + 1, 0
    def ell_row(p):
        g = p["ell"][0]
        g["indices_reordered_i"][0] = A.shape[0] + 1

    for edit in (ell_col, tc_col, tc_indptr, ell_row):
        broken(edit)


@pytest.mark.parametrize("N", [64, 128, 512])
def test_bf16_lds_plan_with_sparse_only_rows(cuda, N):
    """B-stationary plan whose lower half of rows has no dense block (sparse-only
    windows, 7..31 nonzeros per row): 256 rows x nC tasks ride as the LDS
    launch's tail at N = 64 and 128 (256 and 512 tasks) and get a launch of
    their own at N = 512 (2048 tasks).  vs FP64 (1e-2), run to run bit-identical."""
    from stile_amd import _native
    rng = np.random.default_rng(21)
    top, _ = datasets.pruned_layer(seed=21, M=256, K=512, N=8)
    bottom = scipy.sparse.random(256, 512, density=0.04, format="csr", random_state=22, dtype=np.float32)
    A = scipy.sparse.vstack([top, bottom], format="csr")
    pack = _packs.spmm_pack(A, min_fill=64)
    B = rng.standard_normal((512, N), dtype=np.float32)
    print(f"bf16 LDS plan with sparse-only rows N={N}", flush=True)
    plan = _native.Plan.spmm(pack, "bfloat16")
    info = plan.info()
    assert info["block_nnz"] + info["sparse_nnz"] == A.nnz
    Bd = _t(B, cuda)
    C = plan.run_spmm(Bd).cpu().numpy()
    rel = ops.max_rel_err(C, ops.spmm_fp64(A, B))
    assert rel <= 1e-2, f"relative error {rel:.3g} > 1e-2"
    # the sparse-only rows are FP32 products
    _check_fp32(C[256:], bottom, B)
    assert np.array_equal(C, plan.run_spmm(Bd).cpu().numpy())

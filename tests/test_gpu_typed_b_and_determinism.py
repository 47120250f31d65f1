"""GPU parity of the typed-B runs (B in BF16 or FP16, as the reference keeps
B in its run dtype, MY_sparse/my_fuse_formats.py:3529) and of the
deterministic mode (split rows and huge windows summed in a fixed order
through a workspace, instead of atomically), through the C ABI
(stile_spmm_run_ex).

The FP64 oracle multiplies A by B's exact values (B rounded to its type on
the host first).  Bars, written in each test: BF16 MFMA blocks 1e-2
norm-wise relative error vs FP64; FP32 sparse products within the FP32
dot-product bound gamma_n |A||B| (bound_ratio <= 1); deterministic runs
bit-identical run to run.
"""
import numpy as np
import pytest
import scipy.sparse

from oracle import ops
from stile_amd import _native, api, datasets
from tests import _packs

pytestmark = pytest.mark.gpu


def _dev(x, dev, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    return t if dtype is None else t.to(dtype)


def _exact(Bd):
    """B's exact values as float64 on the host."""
    return Bd.double().cpu().numpy()


def _fp32_bound_ok(got, A, B64):
    ref = ops.spmm_fp64(A, B64)
    terms = int(np.diff(scipy.sparse.csr_matrix(A).indptr).max(initial=1))
    r = ops.bound_ratio(got, ref, ops.spmm_abs_bound(A, B64), terms)
    assert r <= 1.0, f"error exceeds the FP32 summation bound ({r:.3g} x)"


TYPES = {"bf16": "bfloat16", "f16": "float16"}


@pytest.mark.parametrize("bt", sorted(TYPES))
@pytest.mark.parametrize("profile", ["a100", "mi355x"])
@pytest.mark.parametrize("N", [128, 1004, 4096])
def test_typed_b_pruned_layer(cuda, bt, profile, N):
    """Config-3 shape (3072 x 768, 10 % of 16x16 blocks + 1 % residual), BF16
    blocks on the B-stationary kernel with a 16-bit B: the A100-model plan
    (all blocks) and the MI355X-model plan (blocks + FP32 residual, mixed
    windows in one launch).  N = 1004 takes the unaligned staging path."""
    import torch
    A, _ = datasets.pruned_layer(seed=0, N=8)
    Bf = np.random.default_rng(N).standard_normal((A.shape[1], N), dtype=np.float32)
    Bd = _dev(Bf, cuda, getattr(torch, TYPES[bt]))
    print(f"typed B {bt} pruned layer profile={profile} N={N}", flush=True)
    plan = api.build_spmm_plan(A, blk_dtype="bfloat16", opts=api.SearchOptions(profile=profile))
    C = plan.run_spmm(Bd).cpu().numpy()
    rel = ops.max_rel_err(C, ops.spmm_fp64(A, _exact(Bd)))
    assert rel <= 1e-2, f"relative error {rel:.3g} > 1e-2"
    # the same plan on the FP32 copy of B agrees within the BF16 bar, and
    # repeated runs are bit-identical
    C32 = plan.run_spmm(Bd.float()).cpu().numpy()
    assert ops.max_rel_err(C, C32) <= 1e-2
    assert np.array_equal(C, plan.run_spmm(Bd).cpu().numpy())


@pytest.mark.parametrize("bt", sorted(TYPES))
@pytest.mark.parametrize("N", [64, 70, 256])
def test_typed_b_mixed_windows_sparse_part_is_fp32(cuda, bt, N):
    """Direct pack with dense blocks (>= 64 nonzeros) and an FP32 ELL rest on
    the same rows: the B-stationary kernel adds each dense window's sparse
    part itself.  Rows without a dense block are pure FP32 products: within
    the FP32 bound of FP64 on B's exact values."""
    import torch
    rng = np.random.default_rng(31)
    top, _ = datasets.pruned_layer(seed=31, M=256, K=512, N=8)
    bottom = scipy.sparse.random(256, 512, density=0.04, format="csr", random_state=32, dtype=np.float32)
    A = scipy.sparse.vstack([top, bottom], format="csr")
    pack = _packs.spmm_pack(A, min_fill=64)
    Bd = _dev(rng.standard_normal((512, N), dtype=np.float32), cuda, getattr(torch, TYPES[bt]))
    print(f"typed B {bt} mixed windows N={N}", flush=True)
    plan = _native.Plan.spmm(pack, "bfloat16")
    assert plan.info()["sparse_nnz"] > 0 and plan.info()["block_nnz"] > 0
    C = plan.run_spmm(Bd).cpu().numpy()
    B64 = _exact(Bd)
    assert ops.max_rel_err(C, ops.spmm_fp64(A, B64)) <= 1e-2
    _fp32_bound_ok(C[256:], bottom, B64)


@pytest.mark.parametrize("blk", ["float32", "bfloat16"])
def test_typed_b_graph_streaming_kernel(cuda, blk):
    """GCN-shaped R-MAT graph (scale 13, K > 1024: the streaming dense
    kernel) with a BF16 B.  FP32 blocks: every product FP32 on B's exact
    values (FP32 bound); BF16 blocks: 1e-2."""
    import torch
    A = datasets.rmat_graph(13, seed=5)
    Bd = _dev(np.random.default_rng(5).standard_normal((A.shape[1], 128), dtype=np.float32), cuda, torch.bfloat16)
    print(f"typed B bf16 rmat13 blk={blk}", flush=True)
    plan = api.build_spmm_plan(A, blk_dtype=blk, N=128)
    C = plan.run_spmm(Bd).cpu().numpy()
    B64 = _exact(Bd)
    if blk == "float32":
        _fp32_bound_ok(C, A, B64)
    else:
        assert ops.max_rel_err(C, ops.spmm_fp64(A, B64)) <= 1e-2


def _long_rows(seed):
    rng = np.random.default_rng(seed)
    M, K = 512, 8192
    lens = rng.choice([0, 3, 40, 300], size=M, p=[0.1, 0.3, 0.5, 0.1])
    lens[7] = 5000
    lens[300] = 2000
    rows = np.repeat(np.arange(M), lens)
    cols = np.concatenate([rng.choice(K, n, replace=False) for n in lens])
    vals = rng.uniform(-1, 1, rows.size).astype(np.float32)
    return scipy.sparse.csr_matrix((vals, (rows, cols)), shape=(M, K)), lens


@pytest.mark.parametrize("N", [64, 72, 200])
def test_deterministic_split_rows(cuda, N):
    """Rows of up to 5000 sparse nonzeros (pieces of 256): the default mode
    adds the pieces with atomics; the deterministic mode sums them through
    the workspace in piece order.  Both within the FP32 bound; the
    deterministic runs bit-identical; a too-small workspace is refused.
    N = 200: four column chunks, the last partial (the sparse and fix-up
    kernels take their waves column chunk by column chunk)."""
    import torch
    A, lens = _long_rows(41)
    B = np.random.default_rng(42).standard_normal((A.shape[1], N), dtype=np.float32)
    Bd = _dev(B, cuda)
    print(f"deterministic split rows N={N}", flush=True)
    plan = api.build_spmm_plan(A, blk_dtype="float32", opts=api.SearchOptions(only_ELL=True), N=N)
    assert plan.workspace_bytes(N) == 0
    need = plan.workspace_bytes(N, deterministic=True)
    assert need == plan.info()["num_split"] * N * 4 > 0
    Cd = plan.run_spmm(Bd).cpu().numpy()
    ws = torch.empty(need, dtype=torch.uint8, device=cuda)
    D1 = plan.run_spmm(Bd, deterministic=True, workspace=ws).cpu().numpy()
    ws.fill_(0xFF)     # the workspace holds garbage between runs: it is fully rewritten
    D2 = plan.run_spmm(Bd, deterministic=True, workspace=ws).cpu().numpy()
    assert np.array_equal(D1, D2)
    _fp32_bound_ok(D1, A, B.astype(np.float64))
    _fp32_bound_ok(Cd, A, B.astype(np.float64))
    short = lens <= 256
    assert np.array_equal(D1[short], Cd[short])       # rows without pieces: the same sums
    with pytest.raises(_native.StileError):
        plan.run_spmm(Bd, deterministic=True, workspace=ws[:need - 4])


@pytest.mark.parametrize("blk", ["float32", "bfloat16"])
def test_deterministic_huge_windows(cuda, blk):
    """Windows of more than 256 dense chunks: parts added with atomics into
    zeroed rows (default) or summed in part order through the workspace
    (deterministic: bit-identical run to run, and no zeroing pass)."""
    A = scipy.sparse.random(40, 20000, density=0.06, format="csr", random_state=8, dtype=np.float32)
    B = np.random.default_rng(8).standard_normal((20000, 72), dtype=np.float32)
    Bd = _dev(B, cuda)
    print(f"deterministic huge windows blk={blk}", flush=True)
    plan = api.build_spmm_plan(A, blk_dtype=blk, opts=api.SearchOptions(only_TC=True), N=72)
    assert plan.workspace_bytes(72, deterministic=True) > 0
    tol = 1e-5 if blk == "float32" else 1e-2
    D1 = plan.run_spmm(Bd, deterministic=True).cpu().numpy()
    D2 = plan.run_spmm(Bd, deterministic=True).cpu().numpy()
    assert np.array_equal(D1, D2)
    ref = ops.spmm_fp64(A, B)
    assert ops.max_rel_err(D1, ref) <= tol
    assert ops.max_rel_err(plan.run_spmm(Bd).cpu().numpy(), ref) <= tol


def test_deterministic_rmat_hub_rows(cuda):
    """R-MAT scale 14 (hub rows) with FP32 blocks at N = 128: the modes agree
    within the FP32 bound, and the deterministic one is bit-identical across
    runs, also inside a replayed HIP graph."""
    import torch
    A = datasets.rmat_graph(14, seed=3)
    B = np.random.default_rng(3).standard_normal((A.shape[1], 128), dtype=np.float32)
    Bd = _dev(B, cuda)
    print(f"deterministic rmat14 nnz={A.nnz}", flush=True)
    plan = api.build_spmm_plan(A, blk_dtype="float32", N=128)
    ws = torch.empty(max(plan.workspace_bytes(128, True), 16), dtype=torch.uint8, device=cuda)
    out = torch.empty((A.shape[0], 128), device=cuda)
    D1 = plan.run_spmm(Bd, deterministic=True, workspace=ws).cpu().numpy()
    _fp32_bound_ok(D1, A, B.astype(np.float64))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.run_spmm(Bd, out=out, deterministic=True, workspace=ws)
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.run_spmm(Bd, out=out, deterministic=True, workspace=ws)
    for _ in range(3):
        out.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), D1)


def test_lds_kernel_sums_long_sparse_rows_of_dense_windows_whole(cuda):
    """A BF16 B-stationary plan whose dense windows hold rows with 600 and
    3000 sparse nonzeros: the kernel sums them whole (no pieces, no
    atomics), so even the default mode is bit-identical run to run."""
    rng = np.random.default_rng(51)
    top, _ = datasets.pruned_layer(seed=51, M=256, K=1024, N=8)
    A = top.tolil()
    A[5, :600] = rng.uniform(-1, 1, 600).astype(np.float32)
    A[40, :] = 0
    A[40, rng.choice(1024, 1000, replace=False)] = rng.uniform(-1, 1, 1000).astype(np.float32)
    A = A.tocsr()
    pack = _packs.spmm_pack(A, min_fill=64)
    B = rng.standard_normal((1024, 128), dtype=np.float32)
    Bd = _dev(B, cuda)
    print("LDS kernel long sparse rows in dense windows", flush=True)
    plan = _native.Plan.spmm(pack, "bfloat16")
    assert plan.info()["num_split"] > 0
    C = plan.run_spmm(Bd).cpu().numpy()
    assert ops.max_rel_err(C, ops.spmm_fp64(A, B)) <= 1e-2
    assert np.array_equal(C, plan.run_spmm(Bd).cpu().numpy())


def test_value_refresh_reaches_the_mixed_windows_sparse_rounds(cuda):
    """stile_plan_set_values on a BF16 B-stationary plan with mixed windows:
    the dense windows' sparse nonzeros live in the kernel's round layout as
    well as in the CSR part, and both take the new values."""
    import torch
    rng = np.random.default_rng(61)
    top, _ = datasets.pruned_layer(seed=61, M=256, K=512, N=8)
    A = top.tolil()
    A[3, 500:511] = 1.0                       # sparse extras in dense windows
    A[70, 5:9] = 1.0
    A = A.tocsr()
    A.sort_indices()
    pack = _packs.add_value_ids(_packs.spmm_pack(A, min_fill=64), A)
    B = rng.standard_normal((512, 96), dtype=np.float32)
    Bd = _dev(B, cuda, torch.bfloat16)
    print("value refresh of a mixed B-stationary plan", flush=True)
    plan = _native.Plan.spmm(pack, "bfloat16")
    assert plan.info()["sparse_nnz"] > 0
    newv = rng.uniform(-1, 1, A.nnz).astype(np.float32)
    plan.set_values(_dev(newv, cuda))
    A2 = scipy.sparse.csr_matrix((newv, A.indices, A.indptr), shape=A.shape)
    C = plan.run_spmm(Bd).cpu().numpy()
    assert ops.max_rel_err(C, ops.spmm_fp64(A2, _exact(Bd))) <= 1e-2

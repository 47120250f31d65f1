"""End-to-end parity of the product path on the GPU: format generation ->
greedy search -> withdraw post-processing -> packing -> C-ABI plan -> HIP
kernels, against the FP64 oracle and the reference's FP32 CPU path.

Bars: FP32 blocks -> within the FP32 dot-product error bound of FP64 and
<= 1e-5 norm-wise relative difference to scipy FP32; BF16 blocks -> <= 1e-2
norm-wise relative error vs FP64 (the stated bound for BF16 blocks).
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


def _check_fp32(got, A, B):
    ref64 = ops.spmm_fp64(A, B)
    terms = int(np.diff(scipy.sparse.csr_matrix(A).indptr).max(initial=1))
    assert ops.bound_ratio(got, ref64, ops.spmm_abs_bound(A, B), terms) <= 1.0
    assert ops.max_rel_err(got, ops.spmm_ref_fp32(A, B)) <= 1e-5


CASES = {
    "pruned_layer": lambda: datasets.pruned_layer(seed=0, N=128),
    "uniform_1pct": lambda: datasets.uniform_spmm(seed=0, N=64),
    "logsparse": lambda: (datasets.log_sparse_mask(1024), np.random.default_rng(0).standard_normal((1024, 96), dtype=np.float32)),
    "strided": lambda: (datasets.strided_mask(1024), np.random.default_rng(1).standard_normal((1024, 32), dtype=np.float32)),
    "rmat12": lambda: (datasets.rmat_graph(12, seed=2), np.random.default_rng(2).standard_normal((4096, 128), dtype=np.float32)),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("blk", ["float32", "bfloat16"])
def test_spmm_pipeline(cuda, case, blk):
    A, B = CASES[case]()
    print(f"pipeline spmm case={case} blk={blk} nnz={A.nnz}", flush=True)
    plan = api.build_spmm_plan(A, blk_dtype=blk)
    info = plan.info()
    assert info["block_nnz"] + info["sparse_nnz"] == A.nnz
    C = plan.run_spmm(_t(B, cuda)).cpu().numpy()
    if blk == "float32":
        _check_fp32(C, A, B)
    else:
        assert ops.max_rel_err(C, ops.spmm_fp64(A, B)) <= 1e-2


@pytest.mark.parametrize("opts", ["only_ELL", "single_level", "no_withdraw"])
def test_spmm_pipeline_search_variants(cuda, opts):
    A, B = datasets.pruned_layer(seed=1, M=512, K=256, N=64)
    o = api.SearchOptions(**{opts: True} if opts != "single_level" else {"max_level_num": 1})
    print(f"pipeline spmm variant={opts}", flush=True)
    plan = api.build_spmm_plan(A, blk_dtype="float32", opts=o)
    _check_fp32(plan.run_spmm(_t(B, cuda)).cpu().numpy(), A, B)


@pytest.mark.parametrize("Kf", [32, 64])
def test_sddmm_pipeline(cuda, Kf):
    A = datasets.rmat_graph(12, seed=0)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((A.shape[0], Kf), dtype=np.float32)
    Y = rng.standard_normal((A.shape[1], Kf), dtype=np.float32)
    print(f"pipeline sddmm rmat12 Kf={Kf}", flush=True)
    plan = api.build_sddmm_plan(A, Kf=Kf)
    out = plan.run_sddmm(_t(X, cuda), _t(Y, cuda)).cpu().numpy()
    ref = ops.sddmm_fp64(A, X, Y)
    assert ops.bound_ratio(out, ref, ops.sddmm_abs_bound(A, X, Y), Kf) <= 1.0
    assert ops.max_rel_err(out, ops.sddmm_ref_fp32(A, X, Y)) <= 1e-5


def test_spmm_pipeline_empty_rows_and_tiny(cuda):
    rng = np.random.default_rng(4)
    A = scipy.sparse.random(37, 53, density=0.05, format="csr", random_state=4, dtype=np.float32).tolil()
    A[5, :] = 0
    A[36, :] = 0
    A = A.tocsr()
    B = rng.standard_normal((53, 7), dtype=np.float32)
    print("pipeline spmm tiny", flush=True)
    plan = api.build_spmm_plan(A, blk_dtype="float32")
    C = plan.run_spmm(_t(B, cuda)).cpu().numpy()
    _check_fp32(C, A, B)
    assert np.all(C[5] == 0) and np.all(C[36] == 0)


def test_spmm_pipeline_batched_uneven(cuda):
    mats = [datasets.attention_mask(L, band=32, n_random=4, rng=np.random.default_rng(L)) for L in (100, 257, 64, 512)]
    rng = np.random.default_rng(5)
    N = 64
    Bs = [rng.standard_normal((A.shape[1], N), dtype=np.float32) for A in mats]
    print("pipeline spmm batched uneven", flush=True)
    plan = api.build_spmm_plan_batched(mats, blk_dtype="float32")
    C = plan.run_spmm(_t(np.vstack(Bs), cuda)).cpu().numpy()
    _check_fp32(C, scipy.sparse.block_diag(mats, format="csr"), np.vstack(Bs))


@pytest.mark.parametrize("blk", ["float32", "bfloat16", "float16"])
@pytest.mark.parametrize("N", [64, 70, 100, 256])
def test_spmm_big_windows_split_over_a_workgroup(cuda, blk, N):
    """Windows with far more than 8 dense chunks (K = 2048, 5 % density, all
    TC tiles) run as workgroups that split the chunks over 8 waves."""
    A = scipy.sparse.random(100, 2048, density=0.05, format="csr", random_state=7, dtype=np.float32)
    B = np.random.default_rng(7).standard_normal((2048, N), dtype=np.float32)
    print(f"big windows blk={blk} N={N}", flush=True)
    plan = api.build_spmm_plan(A, blk_dtype=blk, opts=api.SearchOptions(only_TC=True), N=N)
    info = plan.info()
    assert info["num_chunks"] >= 8 * info["num_windows"]
    C = plan.run_spmm(_t(B, cuda)).cpu().numpy()
    if blk == "float32":
        _check_fp32(C, A, B)
    else:
        assert ops.max_rel_err(C, ops.spmm_fp64(A, B)) <= 1e-2
    C2 = plan.run_spmm(_t(B, cuda)).cpu().numpy()
    assert np.array_equal(C, C2), "split windows must sum in a fixed order"


def test_plan_from_saved_pack(cuda, tmp_path):
    A, B = datasets.pruned_layer(seed=3, M=512, K=256, N=64)
    pack = api.spmm_pack(A)[0]
    api.save_pack(tmp_path / "p.npz", pack)
    from stile_amd import _native
    print("plan from saved pack", flush=True)
    p0 = _native.Plan.spmm(pack, "float32")
    p1 = _native.Plan.spmm(api.load_pack(tmp_path / "p.npz"), "float32")
    Bd = _t(B, cuda)
    assert np.array_equal(p0.run_spmm(Bd).cpu().numpy(), p1.run_spmm(Bd).cpu().numpy())


def test_tuned_plan_is_correct_and_reports_every_candidate(cuda):
    A, B = datasets.pruned_layer(seed=4, M=512, K=256, N=96)
    print("tune_plan pruned 512x256", flush=True)
    plan, pack, rep = api.tune_plan("spmm", A, "float32", 96)
    assert set(rep["ms"]) == set(api.tuning_candidates("float32"))
    assert rep["choice"] == min(rep["ms"], key=rep["ms"].get)
    _check_fp32(plan.run_spmm(_t(B, cuda)).cpu().numpy(), A, B)
    X = np.random.default_rng(5).standard_normal((512, 32), dtype=np.float32)
    Y = np.random.default_rng(6).standard_normal((256, 32), dtype=np.float32)
    splan, _, srep = api.tune_plan("sddmm", A, "float32", 32)
    got = splan.run_sddmm(_t(X, cuda), _t(Y, cuda)).cpu().numpy()
    ref = ops.sddmm_fp64(A, X, Y)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))


@pytest.mark.parametrize("blk", ["float32", "bfloat16"])
def test_spmm_huge_windows_split_into_parts(cuda, blk):
    """Windows with more than 256 dense chunks (hub rows) run as several
    workgroups adding into zeroed rows with atomics."""
    A = scipy.sparse.random(40, 20000, density=0.06, format="csr", random_state=8, dtype=np.float32)
    B = np.random.default_rng(8).standard_normal((20000, 72), dtype=np.float32)
    print(f"huge windows blk={blk}", flush=True)
    plan = api.build_spmm_plan(A, blk_dtype=blk, opts=api.SearchOptions(only_TC=True), N=72)
    info = plan.info()
    assert info["num_chunks"] > 2 * 256
    C = plan.run_spmm(_t(B, cuda)).cpu().numpy()
    tol = 1e-5 if blk == "float32" else 1e-2
    assert ops.max_rel_err(C, ops.spmm_fp64(A, B)) <= tol
    # a second run: the zeroing before the atomics makes runs independent
    C2 = plan.run_spmm(_t(B, cuda)).cpu().numpy()
    assert ops.max_rel_err(C2, ops.spmm_fp64(A, B)) <= tol


@pytest.mark.parametrize("N", [64, 72, 70])
def test_sparse_rows_split_over_lane_groups(cuda, N):
    """Few sparse-only rows (the launch takes one wave per row): empty rows,
    short and medium rows, rows longer than a window keeps (pieces added by
    the fix-up kernel) and a hub row summed with compensation."""
    rng = np.random.default_rng(11)
    M, K = 512, 8192
    lens = rng.choice([0, 3, 40, 300], size=M, p=[0.1, 0.3, 0.5, 0.1])
    lens[7] = 5000
    rows = np.repeat(np.arange(M), lens)
    cols = np.concatenate([rng.choice(K, n, replace=False) for n in lens])
    vals = rng.uniform(-1, 1, rows.size).astype(np.float32)
    A = scipy.sparse.csr_matrix((vals, (rows, cols)), shape=(M, K))
    B = rng.standard_normal((K, N), dtype=np.float32)
    print(f"pipeline spmm split sparse rows N={N} nnz={A.nnz}", flush=True)
    plan = api.build_spmm_plan(A, blk_dtype="float32", opts=api.SearchOptions(only_ELL=True), N=N)
    assert plan.info()["block_nnz"] == 0
    C = plan.run_spmm(_t(B, cuda)).cpu().numpy()
    _check_fp32(C, A, B)
    # bit-identical run to run, except rows with two or more fix-up pieces
    # (more than 512 sparse nonzeros: their pieces are added atomically)
    C2 = plan.run_spmm(_t(B, cuda)).cpu().numpy()
    fixed = lens <= 512
    assert np.array_equal(C[fixed], C2[fixed])
    assert ops.max_rel_err(C2, C) <= 1e-6


@pytest.mark.parametrize("N", [128, 70, 512])
def test_batched_bf16_matrices_on_the_b_stationary_kernel(cuda, N):
    """Uneven block-pruned matrices (K <= 1024 each) in one block-diagonal
    BF16 plan: every workgroup stages the B rows of one matrix.  Each
    matrix's rows of C vs FP64 (1e-2), and run to run bit-identical."""
    import torch
    shapes = [(512, 256), (1024, 1024), (256, 768), (768, 768), (96, 1024)]
    mats = [datasets.pruned_layer(seed=20 + i, M=m, K=k, N=8)[0] for i, (m, k) in enumerate(shapes)]
    print(f"batched bf16 B-stationary N={N} mats={len(mats)}", flush=True)
    plan = api.build_spmm_plan_batched(mats, blk_dtype="bfloat16", N=N)
    Ktot = sum(k for _, k in shapes)
    B = np.random.default_rng(5).standard_normal((Ktot, N), dtype=np.float32)
    C = plan.run_spmm(_t(B, cuda)).cpu().numpy()
    r0 = c0 = 0
    for A, (m, k) in zip(mats, shapes):
        rel = ops.max_rel_err(C[r0:r0 + m], ops.spmm_fp64(A, B[c0:c0 + k]))
        assert rel <= 1e-2, f"matrix {m}x{k}: {rel:.3g}"
        r0, c0 = r0 + m, c0 + k
    assert np.array_equal(C, plan.run_spmm(_t(B, cuda)).cpu().numpy())


@pytest.mark.parametrize("count,N", [(30, 1024), (12, 4160)])
def test_batched_bf16_b_stationary_items_above_cus(cuda, count, N):
    """More (unit, column chunk) items than compute units: each persistent
    workgroup of the B-stationary kernel restages LDS for several items.  N =
    1024 uses the plan's list choice for 16 column chunks; N = 4160 (65
    chunks, the last one partial) the fallback choice.  Each matrix's rows of C
    vs FP64 (1e-2), and run to run bit-identical."""
    cycle = [(512, 256), (1024, 1024), (256, 768), (768, 512), (96, 1024), (1024, 128)]
    shapes = [cycle[i % len(cycle)] for i in range(count)]
    mats = [datasets.pruned_layer(seed=40 + i, M=m, K=k, N=8)[0] for i, (m, k) in enumerate(shapes)]
    print(f"batched bf16 B-stationary persistent N={N} mats={count}", flush=True)
    plan = api.build_spmm_plan_batched(mats, blk_dtype="bfloat16", N=N)
    Ktot = sum(k for _, k in shapes)
    B = np.random.default_rng(6).standard_normal((Ktot, N), dtype=np.float32)
    Bd = _t(B, cuda)
    C = plan.run_spmm(Bd).cpu().numpy()
    r0 = c0 = 0
    for A, (m, k) in zip(mats, shapes):
        rel = ops.max_rel_err(C[r0:r0 + m], ops.spmm_fp64(A, B[c0:c0 + k]))
        assert rel <= 1e-2, f"matrix {m}x{k}: {rel:.3g}"
        r0, c0 = r0 + m, c0 + k
    assert np.array_equal(C, plan.run_spmm(Bd).cpu().numpy())


def test_batched_bf16_xcd_grouped_launch_in_graph(cuda):
    """120 small uneven matrices at N = 256 (4 column chunks): one unit per
    matrix with dense windows already gives more (unit, chunk) items than the
    GPU has CUs, so the B-stationary kernel runs persistent with its per-XCD
    item ranges.  Eager result vs FP64 per matrix (1e-2); the same
    launch captured in a HIP graph and replayed three times, with eager work
    and a synchronisation between replays, equals the eager result bit for
    bit each time."""
    import torch
    rng = np.random.default_rng(12)
    shapes = [(int(16 * rng.integers(2, 33)), int(16 * rng.integers(4, 65))) for _ in range(120)]
    mats = [datasets.pruned_layer(seed=100 + i, M=m, K=k, N=8)[0] for i, (m, k) in enumerate(shapes)]
    N = 256
    print(f"batched bf16 XCD-grouped N={N} mats={len(mats)} + graph", flush=True)
    plan = api.build_spmm_plan_batched(mats, blk_dtype="bfloat16", N=N)
    Ktot = sum(k for _, k in shapes)
    Mtot = sum(m for m, _ in shapes)
    B = rng.standard_normal((Ktot, N), dtype=np.float32)
    Bd = _t(B, cuda)
    C_eager = plan.run_spmm(Bd).cpu().numpy()
    r0 = c0 = 0
    for A, (m, k) in zip(mats, shapes):
        rel = ops.max_rel_err(C_eager[r0:r0 + m], ops.spmm_fp64(A, B[c0:c0 + k]))
        assert rel <= 1e-2, f"matrix {m}x{k}: {rel:.3g}"
        r0, c0 = r0 + m, c0 + k
    C = torch.empty((Mtot, N), device=cuda)
    plan.run_spmm(Bd, C)  # warm up outside the capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.run_spmm(Bd, C)
    for r in range(3):
        C.fill_(float("nan"))  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(C.cpu().numpy(), C_eager), f"replay {r}"


@pytest.mark.parametrize("count,N,every", [(40, 1024, 40), (120, 256, 12)])
def test_batched_bf16_persistent_with_sparse_only_tail(cuda, count, N, every):
    """Persistent B-stationary launches (items above the CUs: round-robin at
    N = 1024, per-XCD ranges at N = 256) whose plan also has sparse-only rows:
    every `every`-th matrix gets 16 rows without a dense block (with the
    windows that hold none by chance: 3 and 14 sparse-only windows, 768 and
    896 (row, 64 columns) tasks, within the tail limit of 1,024), and the
    rows' tasks run as the launch's tail.  Each matrix's rows vs FP64
    (1e-2), the appended rows within the FP32 bound, run to run
    bit-identical."""
    rng = np.random.default_rng(30)
    mats, shapes, tails = [], [], []
    for i in range(count):
        m, k = [(256, 1024), (128, 1024), (384, 768), (64, 1024)][i % 4]
        A, _ = datasets.pruned_layer(seed=200 + i, M=m, K=k, N=8)
        if i % every == 0:
            S = scipy.sparse.random(16, k, density=0.03, format="csr", random_state=300 + i, dtype=np.float32)
            A = scipy.sparse.vstack([A, S], format="csr")
            tails.append((len(mats), S))
        mats.append(A)
        shapes.append(A.shape)
    print(f"batched bf16 persistent + sparse-only tail N={N} mats={count}", flush=True)
    # packed directly (dense 16x16 blocks as TC tiles, the rest ELL), so the
    # appended rows are sparse-only whatever a search would pick
    from stile_amd import _native
    from tests import _packs
    plan = _native.Plan.spmm([_packs.spmm_pack(A, min_fill=64) for A in mats], "bfloat16")
    Ktot = sum(k for _, k in shapes)
    B = rng.standard_normal((Ktot, N), dtype=np.float32)
    Bd = _t(B, cuda)
    C = plan.run_spmm(Bd).cpu().numpy()
    offs = np.cumsum([(0, 0)] + shapes, axis=0)
    for i, A in enumerate(mats):
        (r0, c0), (m, k) = offs[i], shapes[i]
        rel = ops.max_rel_err(C[r0:r0 + m], ops.spmm_fp64(A, B[c0:c0 + k]))
        assert rel <= 1e-2, f"matrix {i} {m}x{k}: {rel:.3g}"
    for i, S in tails:
        (r0, c0), (m, k) = offs[i], shapes[i]
        ref = ops.spmm_fp64(S, B[c0:c0 + k])
        bound = ops.spmm_abs_bound(S, B[c0:c0 + k])
        assert ops.bound_ratio(C[r0 + m - 16:r0 + m], ref, bound, int(np.diff(S.indptr).max(initial=1))) <= 1.0
    assert np.array_equal(C, plan.run_spmm(Bd).cpu().numpy())

"""GPU parity of scatter chunks (B-stationary plans whose sparse TC columns
are stored one entry per k-slot, StileSpmmOptions.scatter_max_nnz) and of
the plans' launch options (stile_plan_set_option), through the C ABI.

Bars: BF16 MFMA blocks 1e-2 norm-wise relative error vs the FP64 oracle on
B's exact values; a scatter plan against the same pack built without
scatter chunks 1e-5 (the same BF16 products, summed in another order);
launch options that only change the launch shape bit-identical (every
window is summed by one wave in chunk order whatever the shape).
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
    return Bd.double().cpu().numpy()


def _pair(pack, scatter):
    return _native.Plan.spmm(pack, "bfloat16"), _native.Plan.spmm(pack, "bfloat16", scatter_max_nnz=scatter)


@pytest.mark.parametrize("N", [128, 1004, 4096])
def test_scatter_chunks_pruned_layer(cuda, N):
    """The block-pruned layer (1% residual: TC columns of one or two
    nonzeros per window) with scatter chunks at N = 128, 1004 (a partial
    column chunk) and 4096."""
    import torch
    A, _ = datasets.pruned_layer(seed=0, N=1)
    pack, _, _ = api.spmm_pack(A, api.SearchOptions(profile="a100", dsize=16), N)
    plain, sc = _pair(pack, 2)
    assert sc.info()["num_chunks"] > plain.info()["num_chunks"]   # the residual moved to scatter chunks
    B = np.random.default_rng(N).standard_normal((A.shape[1], N), dtype=np.float32)
    Bd = _dev(B, cuda, torch.bfloat16)
    print(f"scatter chunks pruned layer N={N}", flush=True)
    C0 = plain.run_spmm(Bd).cpu().numpy()
    C1 = sc.run_spmm(Bd).cpu().numpy()
    assert ops.max_rel_err(C1, ops.spmm_fp64(A, _exact(Bd))) <= 1e-2
    assert ops.max_rel_err(C1, C0) <= 1e-5
    assert np.array_equal(C1, sc.run_spmm(Bd).cpu().numpy())
    plain.close()
    sc.close()


@pytest.mark.parametrize("scatter", [1, 2, 16])
def test_scatter_only_windows_and_row_classes(cuda, scatter):
    """A 1%-dense random matrix packed as TC blocks only: every column of a
    window holds one or two nonzeros, so whole windows become scatter chunks
    (no dense chunk), with uneven row-residue classes; FP32 B and BF16 B."""
    import torch
    A = scipy.sparse.random(512, 700, density=0.01, format="csr", random_state=5, dtype=np.float32)
    A.data = np.random.default_rng(5).uniform(-1, 1, A.nnz).astype(np.float32)
    pack = _packs.spmm_pack(A, min_fill=1)
    assert not pack["ell"]
    plain, sc = _pair(pack, scatter)
    B = np.random.default_rng(6).standard_normal((700, 200), dtype=np.float32)
    print(f"scatter-only windows scatter={scatter}", flush=True)
    for dt in (torch.float32, torch.bfloat16):
        Bd = _dev(B, cuda, dt)
        C1 = sc.run_spmm(Bd).cpu().numpy()
        assert ops.max_rel_err(C1, ops.spmm_fp64(A, _exact(Bd))) <= 1e-2
        assert ops.max_rel_err(C1, plain.run_spmm(Bd).cpu().numpy()) <= 1e-5
    plain.close()
    sc.close()


def test_scatter_chunks_with_mixed_windows_and_value_refresh(cuda):
    """Scatter chunks beside windows' FP32 sparse parts (rounds), and a value
    refresh: the scatter k-slots take the new values and keep their rows."""
    import torch
    rng = np.random.default_rng(71)
    top, _ = datasets.pruned_layer(seed=71, M=512, K=768, N=8)
    A = top.tolil()
    A[3, 700:711] = 1.0
    A[70, 5:9] = 1.0
    A = A.tocsr()
    A.sort_indices()
    pack = _packs.add_value_ids(_packs.spmm_pack(A, min_fill=1), A)
    # a few columns as ELL rows of dense windows: the windows' sparse parts
    pack2 = _packs.add_value_ids(_packs.spmm_pack(A, min_fill=3), A)
    B = rng.standard_normal((768, 160), dtype=np.float32)
    Bd = _dev(B, cuda, torch.bfloat16)
    print("scatter chunks + mixed windows + value refresh", flush=True)
    for pk in (pack, pack2):
        plan = _native.Plan.spmm(pk, "bfloat16", scatter_max_nnz=2)
        C = plan.run_spmm(Bd).cpu().numpy()
        assert ops.max_rel_err(C, ops.spmm_fp64(A, _exact(Bd))) <= 1e-2
        newv = rng.uniform(-1, 1, A.nnz).astype(np.float32)
        plan.set_values(_dev(newv, cuda))
        A2 = scipy.sparse.csr_matrix((newv, A.indices, A.indptr), shape=A.shape)
        C2 = plan.run_spmm(Bd).cpu().numpy()
        assert ops.max_rel_err(C2, ops.spmm_fp64(A2, _exact(Bd))) <= 1e-2
        plan.close()
    assert pack2["ell"]


def test_scatter_chunks_encoder_batch_and_determinism(cuda):
    """Ten matrices of the encoder batch in one plan at N = 128 with scatter
    chunks: within the bar against FP64, and the deterministic mode agrees
    bit for bit with the default one (a B-stationary plan has no atomics)."""
    import torch
    mats, shapes = datasets.pruned_encoder()
    mats = mats[:10]
    opts = api.SearchOptions(profile="a100", dsize=16)
    packs = [api.spmm_pack(A, opts, 128)[0] for A in mats]
    sc = _native.Plan.spmm(packs, "bfloat16", scatter_max_nnz=2)
    Kt = sum(A.shape[1] for A in mats)
    B = np.random.default_rng(8).standard_normal((Kt, 128), dtype=np.float32)
    Bd = _dev(B, cuda, torch.bfloat16)
    print("scatter chunks encoder batch", flush=True)
    C = sc.run_spmm(Bd).cpu().numpy()
    big = scipy.sparse.block_diag(mats, format="csr")
    assert ops.max_rel_err(C, ops.spmm_fp64(big, _exact(Bd))) <= 1e-2
    assert np.array_equal(C, sc.run_spmm(Bd, deterministic=True).cpu().numpy())
    sc.close()


def test_launch_options(cuda):
    """Unit lists and the XCD grouping change only the launch shape: results
    bit-identical to the default; the streaming kernel (another summation
    order) within 1e-5 of it; a scatter plan refuses the streaming kernel,
    and bad option values are refused."""
    import torch
    A, _ = datasets.pruned_layer(seed=3, N=1)
    pack, _, _ = api.spmm_pack(A, api.SearchOptions(profile="a100", dsize=16), 128)
    plan = _native.Plan.spmm(pack, "bfloat16")
    B = np.random.default_rng(9).standard_normal((A.shape[1], 512), dtype=np.float32)
    Bd = _dev(B, cuda, torch.bfloat16)
    print("launch options", flush=True)
    ref = plan.run_spmm(Bd).cpu().numpy()
    for N in (128, 512):
        base = plan.run_spmm(Bd[:, :N].contiguous()).cpu().numpy()
        for t in range(12):
            try:
                plan.set_option("lds_list", t)
            except _native.StileError:
                continue                      # a list without units is refused
            for x in (1, 0):
                plan.set_option("lds_xcd", x)
                assert np.array_equal(plan.run_spmm(Bd[:, :N].contiguous()).cpu().numpy(), base), (N, t, x)
        plan.set_option("lds_list", -1)
        plan.set_option("lds_xcd", 1)
    plan.set_option("kernel", 1)
    assert ops.max_rel_err(plan.run_spmm(Bd).cpu().numpy(), ref) <= 1e-5
    plan.set_option("kernel", 0)
    assert np.array_equal(plan.run_spmm(Bd).cpu().numpy(), ref)
    for name, bad in (("kernel", 2), ("lds_list", 12), ("lds_xcd", 5)):
        with pytest.raises(_native.StileError):
            plan.set_option(name, bad)
    sc = _native.Plan.spmm(pack, "bfloat16", scatter_max_nnz=2)
    with pytest.raises(_native.StileError):
        sc.set_option("kernel", 1)
    plan.close()
    sc.close()


def test_tuner_times_build_and_launch_variants(cuda, tmp_path):
    """tune_plan on a B-stationary matrix times every search setting with and
    without scatter chunks, then the work-unit lists of its choice; the plan
    rebuilt from the cache (pack + build + launch options) gives the same
    output bit for bit, and both are within the BF16 bar against FP64."""
    import torch
    A, _ = datasets.pruned_layer(seed=4, M=1024, K=512, N=1)
    cache = str(tmp_path / "tuned.npz")
    print("tuner variants", flush=True)
    plan, pack, rep = api.tune_plan("spmm", A, "bfloat16", 128, b_dtype="bfloat16", cache=cache, reps=5, rounds=3)
    assert any(k.endswith("+scatter2") for k in rep["ms"]) and rep["lds_list_ms"]
    assert "-1" in rep["lds_list_ms"]
    B = np.random.default_rng(12).standard_normal((512, 128), dtype=np.float32)
    Bd = _dev(B, cuda, torch.bfloat16)
    C1 = plan.run_spmm(Bd).cpu().numpy()
    assert ops.max_rel_err(C1, ops.spmm_fp64(A, _exact(Bd))) <= 1e-2
    plan2, _, rep2 = api.tune_plan("spmm", A, "bfloat16", 128, b_dtype="bfloat16", cache=cache)
    assert rep2["choice"] == "cached" and rep2["build"] == rep["build"]
    assert np.array_equal(plan2.run_spmm(Bd).cpu().numpy(), C1)
    plan.close()
    plan2.close()

"""Benchmark of the hybrid sparse-tile SpMM / SDDMM path on one MI355X.

Default workload (one step = one launch of the fused SpMM):
  C = A . B, A the 3072 x 768 block-pruned layer (10 % of the 16x16 blocks
  dense + 1 % unstructured residual, 257,265 nonzeros, seed 0), B 768 x 4096
  BF16 resident in HBM, C 3072 x 4096 FP32.  Tiles from the STile search
  (greedy + withdraw, cost model), the search setting picked by measured
  latency (api.tune_plan); dense blocks run BF16 MFMA with FP32
  accumulation, the sparse rest FP32.  Sub-objects of the same line: the
  layer with FP32 B (b_fp32) and at N = 128 (n128), the encoder batch below
  at N = 128 (pruned_batch_n128), and the uniform SpMM and R-MAT SDDMM below
  (uniform_spmm_n64, rmat_sddmm_k32; --no-n128, --no-batch, --no-small skip
  them).

Other workloads (--workload), each printed as its own JSON line:
  pruned_batch  120 block-pruned matrices of a 24-layer BERT-large-shaped
                encoder (Q, K, V, O 1024 x 1024, FFN-in 4096 x 1024; 21.95M
                nonzeros) x B 1024 x 128 each, in one block-diagonal launch
  uniform_spmm  A 4096^2, 167,772 uniform nonzeros U(-1,1), B 4096 x 64 FP32
  rmat_sddmm    R-MAT scale 16 (+ self loops), X, Y 65536 x 32 FP32
  gcn           GCN layer on R-MAT scale 20 (32.45M nonzeros, symmetric
                normalised), H 1M x 128, W 128 x 128 FP32: forward A.(HW),
                backward A^T.dY and SDDMM dA (one step = forward + backward);
                sub-object bf16: the forward product with H.W in BF16 and
                BF16 blocks (the configuration's BF16 variant)
  attention     64 sparse-attention instances, L ~ U[512, 8192], d = 64,
                band 128 + 16 random columns: softmax(QK^T / 8 on the mask) V
                in one fused kernel, replayed as a HIP graph (one step = one
                replay); the unfused SDDMM -> softmax -> SpMM graph beside it
  all           every workload above, one line each

Each line carries GFLOP/s (2 nnz N per sparse product), the HBM roofline of
the dominant kernel (algorithmic bytes over its HIP-event-timed launch
duration, on the stream it is launched on), the plain CSR HIP kernel and the
vendor CSR SpMM (torch.sparse, rocSPARSE) on the same operands, the error
against the FP64 oracle on a sample, and a CPU baseline timed on a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def spmm_alg_bytes(nnz, M, K, N, s_b=4):
    """nnz (value + index) + indptr + B once + C once."""
    return nnz * 8 + 4 * (M + 1) + K * N * s_b + M * N * 4


def full_blocks(A, b=16):
    """Number of completely dense b x b blocks of A (on the b-aligned grid)."""
    import numpy as np
    A = A.tocoo()
    nbc = (A.shape[1] + b - 1) // b
    key = (A.row // b).astype(np.int64) * nbc + A.col // b
    _, counts = np.unique(key, return_counts=True)
    return int(np.sum(counts == b * b))


def spmm_alg_bytes_blocks(nblocks, nnz, M, K, N, s_blk=2, s_b=4):
    """The dense-block variant: a dense 16 x 16 block costs its stored values
    (256 * s_blk) and one 4-byte index; the other nonzeros value + index; then
    indptr, B once and C once."""
    residual = nnz - 256 * nblocks
    return nblocks * (256 * s_blk + 4) + residual * 8 + 4 * (M + 1) + K * N * s_b + M * N * 4


def sddmm_alg_bytes(nnz, M, Ncols, K):
    """index + output per nonzero, indptr, X and Y once."""
    return nnz * 8 + 4 * (M + 1) + (M + Ncols) * K * 4


def measured_traffic(name):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel
    calibration factors stored with it), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return int((t["fetch_kib"] * t["fetch_factor"] + t["write_kib"] * t["write_factor"]) * 1024)


def time_launches(fn, steps, warmup):
    """ms per call of fn, HIP events on torch's current stream (the stream
    every entry point here launches on)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / steps


def roofline(alg_bytes, ms, traffic=None):
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "alg_bytes": int(alg_bytes)}


def line(metric, value, unit, ms, args, dtype, config, **extra):
    out = {"metric": metric, "value": round(value, 2), "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}
    out.update(extra)
    return out


def _repeat_for(fn, seconds, max_runs=200):
    fn()
    t0 = time.perf_counter()
    n = 0
    while True:
        fn()
        n += 1
        if time.perf_counter() - t0 >= seconds or n >= max_runs:
            break
    return (time.perf_counter() - t0) / n, n


def host_cpu():
    """The host the CPU baselines ran on: logical CPUs and model string."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def cpu_spmm_baseline(A, B, cols, seconds, label):
    """The reference's CPU path (scipy csr.astype(float32).dot(B), one core)
    on a bounded sample: the first `cols` columns of B."""
    import numpy as np
    from oracle import ops
    Bs = np.ascontiguousarray(B[:, :cols])
    dt, n = _repeat_for(lambda: ops.spmm_ref_fp32(A, Bs), seconds)
    return {"value": round(2 * A.nnz * Bs.shape[1] / dt / 1e9, 3), "unit": "GFLOP/s", "cores": 1,
            "kind": "port", "sample": f"scipy CSR({label}, nnz={A.nnz}) . B[:, :{Bs.shape[1]}] FP32, "
                                      f"{n} runs of {dt * 1e3:.1f} ms", "host": host_cpu()}


def cpu_torch_csr_baseline(A, B, cols, seconds):
    """torch's CPU CSR x dense product on every host thread, reported beside
    the one-core scipy baseline, so the host side is also measured on every core."""
    import warnings
    warnings.filterwarnings("ignore", message="Sparse CSR tensor support is in beta")
    import numpy as np
    import torch
    At = torch.sparse_csr_tensor(torch.from_numpy(A.indptr.astype(np.int64)),
                                 torch.from_numpy(A.indices.astype(np.int64)),
                                 torch.from_numpy(A.data.astype(np.float32)), size=A.shape)
    Bs = torch.from_numpy(np.ascontiguousarray(B[:, :cols]))
    dt, n = _repeat_for(lambda: torch.sparse.mm(At, Bs), seconds)
    return {"value": round(2 * A.nnz * Bs.shape[1] / dt / 1e9, 3), "unit": "GFLOP/s",
            "cores": torch.get_num_threads(), "kind": "torch CPU CSR",
            "sample": f"torch.sparse.mm(CSR A, B[:, :{Bs.shape[1]}]) FP32, {n} runs of {dt * 1e3:.1f} ms"}


def vendor_csr_spmm(A, Bd):
    """torch.sparse CSR x dense (rocSPARSE) on the same A; None if unavailable."""
    import numpy as np
    import torch
    try:
        At = torch.sparse_csr_tensor(torch.from_numpy(A.indptr.astype(np.int64)),
                                     torch.from_numpy(A.indices.astype(np.int64)),
                                     torch.from_numpy(A.data.astype(np.float32)), size=A.shape).to(Bd.device)
        return lambda: torch.sparse.mm(At, Bd)
    except Exception:                                            # noqa: BLE001
        return None


def _csr_dev(A, dev):
    import numpy as np
    import torch
    return (torch.from_numpy(A.indptr.astype(np.int32)).to(dev),
            torch.from_numpy(A.indices.astype(np.int32)).to(dev),
            torch.from_numpy(A.data.astype(np.float32)).to(dev))


def _spmm_comparators(A, Bd, steps, warmup, hyb=False):
    """ms per launch of the plain CSR HIP kernel, the vendor CSR SpMM and,
    with hyb, the SparseTIR-style ELL-bucket comparator (one column part,
    buckets 1 .. 2^ceil(log2(nnz/rows)))."""
    import torch
    from stile_amd import _native
    ip, ix, vv = _csr_dev(A, Bd.device)
    C2 = torch.empty((A.shape[0], Bd.shape[1]), device=Bd.device, dtype=torch.float32)
    ms_csr = time_launches(lambda: _native.csr_spmm_f32(ip, ix, vv, Bd, C2), steps, warmup)
    vfn = vendor_csr_spmm(A, Bd)
    ms_vendor = None
    if vfn is not None:
        try:
            ms_vendor = time_launches(vfn, steps, warmup)
        except Exception:                                        # noqa: BLE001
            ms_vendor = None
    if not hyb:
        return ms_csr, ms_vendor
    from stile_amd.hyb import HybSpMM
    h = HybSpMM(A, 1, device=Bd.device)
    ms_hyb = time_launches(lambda: h.run(Bd, C2), steps, warmup)
    return ms_csr, ms_vendor, ms_hyb


def cache_path(args, name):
    return None if args.plan_cache is None else os.path.join(args.plan_cache, name + ".npz")


def cache_dir(args, name):
    return None if args.plan_cache is None else os.path.join(args.plan_cache, name)


def progress(msg):
    """Progress of the long set-up steps (search, tuning) on stderr."""
    print(msg, file=sys.stderr, flush=True)


MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak


def mfma_use(chunks, N, ms):
    """Secondary figure for the dense blocks: MFMA work issued per launch
    (every 16x32 chunk times N columns, padding slots included) over the
    launch time, against the dense BF16 peak."""
    tf = 2.0 * chunks * 16 * 32 * N / (ms * 1e-3) / 1e12
    return {"issued_tflops": round(tf, 1), "peak_tflops": MFMA_BF16_PEAK_TFLOPS,
            "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4)}


def _gf(flops, ms):
    return None if ms is None else round(flops / (ms * 1e-3) / 1e9, 2)


# ---------------------------------------------------------------------------
# Workloads.
# ---------------------------------------------------------------------------

def run_pruned_layer(args, dev):
    """Config 3: the block-pruned layer A 3072 x 768 times B 768 x N with B in
    BF16 (the reference keeps B in its run dtype) and C FP32.  Secondary
    figures on the same plan: FP32 B, N = 128, and (unless --no-batch) the
    120-matrix encoder batch at N = 128 in one launch."""
    import numpy as np
    import torch
    from oracle import ops
    from stile_amd import api, datasets
    N = args.n
    A, B = datasets.pruned_layer(seed=0, N=N)
    M, K = A.shape
    t0 = time.perf_counter()
    plan, _, tuning = api.tune_plan("spmm", A, "bfloat16", N, log=progress, b_dtype="bfloat16",
                                    cache=cache_path(args, f"pruned_layer_n{N}"))
    t_plan = time.perf_counter() - t0
    info = plan.info()
    Bf = torch.from_numpy(B).to(dev)
    Bd = Bf.to(torch.bfloat16)
    C = torch.empty((M, N), device=dev, dtype=torch.float32)
    ms = time_launches(lambda: plan.run_spmm(Bd, C), args.steps, args.warmup)
    C_h = C.cpu().numpy()
    ms_f32 = time_launches(lambda: plan.run_spmm(Bf, C), args.steps, args.warmup)
    ms_csr, ms_vendor, ms_hyb = _spmm_comparators(A, Bf, max(args.steps // 4, 10), 5, hyb=True)
    flops = 2 * A.nnz * N
    nblocks = full_blocks(A)
    alg = spmm_alg_bytes_blocks(nblocks, A.nnz, M, K, N, s_b=2)
    n128 = None
    if not args.no_n128:
        B128f = torch.from_numpy(np.ascontiguousarray(B[:, :128])).to(dev)
        B128 = B128f.to(torch.bfloat16)
        C128 = torch.empty((M, 128), device=dev, dtype=torch.float32)
        # the plan tuned at N = 128 (a narrow B changes what pays: staging,
        # launch and the sparse rounds weigh more against the MFMA chunks)
        plan128, _, tuning128 = api.tune_plan("spmm", A, "bfloat16", 128, log=progress, b_dtype="bfloat16",
                                              cache=cache_path(args, "pruned_layer_n128"))
        ms128 = time_launches(lambda: plan128.run_spmm(B128, C128), args.steps, args.warmup)
        ms128_csr, ms128_vendor = _spmm_comparators(A, B128f, args.steps, args.warmup)
        n128 = {"hybrid_us": round(ms128 * 1e3, 2), "csr_us": round(ms128_csr * 1e3, 2),
                "vendor_us": None if ms128_vendor is None else round(ms128_vendor * 1e3, 2),
                "hybrid_over_csr": round(ms128_csr / ms128, 3), "hybrid_gflops": _gf(2 * A.nnz * 128, ms128),
                "hbm_frac": roofline(spmm_alg_bytes_blocks(nblocks, A.nnz, M, K, 128, s_b=2), ms128)["frac"],
                "b_dtype": "bf16", "tuning": tuning128}
        plan128.close()
    # back-to-back launch floor (a one-element kernel): what a launch costs
    # before any work; it bounds small-N launches
    tiny = torch.zeros(1, device=dev)
    ms_floor = time_launches(lambda: tiny.add_(1.0), 200, 20)
    # measured device-to-device copy bandwidth (the practical HBM ceiling)
    cp_src = torch.empty(64 << 20, device=dev, dtype=torch.float32)
    cp_dst = torch.empty_like(cp_src)
    ms_copy = time_launches(lambda: cp_dst.copy_(cp_src), 20, 3)
    copy_gbs = 2 * cp_src.numel() * 4 / (ms_copy * 1e-3) / 1e9
    del cp_src, cp_dst
    traffic = measured_traffic("hyb_spmm_traffic_pruned_layer.json") if N == 4096 else None
    out = line("SpMM GFLOP/s (2*nnz*N), hybrid fused vs plain CSR", flops / (ms * 1e-3) / 1e9, "GFLOP/s",
               ms, args, "bf16 B and MFMA blocks, f32 sparse products, f32 accumulate and C",
               {"workload": "hybrid SpMM, block-pruned layer A 3072x768 (10% of 16x16 blocks + 1% residual, "
                            f"nnz={A.nnz}) x B 768x{N} BF16 -> C FP32", "M": M, "K": K, "N": N,
                "nnz": int(A.nnz), "b_dtype": "bf16", "global_batch": 1, "parallelism": "single GPU"},
               roofline=roofline(alg, ms, traffic),
               plain_csr_gflops=_gf(flops, ms_csr), vendor_csr_gflops=_gf(flops, ms_vendor),
               sparsetir_hyb_gflops=_gf(flops, ms_hyb),
               hybrid_over_csr=round(ms_csr / ms, 3),
               hybrid_over_vendor=None if ms_vendor is None else round(ms_vendor / ms, 3),
               hybrid_over_sparsetir_hyb=round(ms_hyb / ms, 3),
               b_fp32={"us": round(ms_f32 * 1e3, 2), "gflops": _gf(flops, ms_f32),
                       "hbm_frac": roofline(spmm_alg_bytes_blocks(nblocks, A.nnz, M, K, N, s_b=4), ms_f32)["frac"]},
               mfma=mfma_use(info["num_chunks"], N, ms),
               n128=n128, launch_floor_us=round(ms_floor * 1e3, 2),
               hbm_copy_gbs={"value": round(copy_gbs, 1), "what": "256 MiB device copy (read + write)",
                             "achieved_over_copy": round(alg / (ms * 1e-3) / 1e9 / copy_gbs, 4)},
               plan={"windows": info["num_windows"], "mfma_chunks": info["num_chunks"],
                     "block_nnz": info["block_nnz"], "sparse_nnz": info["sparse_nnz"],
                     "search_tune_and_plan_s": round(t_plan, 3), "tuning": tuning})
    if not args.no_cpu:
        B64 = Bd.double().cpu().numpy()
        out["max_rel_err_vs_fp64"] = ops.max_rel_err(C_h[:, :256], ops.spmm_fp64(A, B64[:, :256]))
        out["cpu_baseline"] = cpu_spmm_baseline(A, B, 512, 10.0, "3072x768")
        out["cpu_torch_csr_all_threads"] = cpu_torch_csr_baseline(A, B, 512, 5.0)
    if not args.no_batch:
        out["pruned_batch_n128"] = pruned_batch_summary(args, dev)
    if not args.no_small:
        # the two small configurations, so that the default run times them too
        keep = ("metric", "ms_per_step", "value", "unit", "roofline", "hybrid_over_csr", "hybrid_over_vendor",
                "max_rel_err_vs_fp64", "config")
        for name, fn in (("uniform_spmm_n64", run_uniform_spmm), ("rmat_sddmm_k32", run_rmat_sddmm)):
            r = fn(args, dev, summary=True)
            out[name] = {k: r[k] for k in keep if k in r}
    return out


def pruned_batch_summary(args, dev):
    """The pruned encoder batch at N = 128 (BF16 B) as a sub-object of the
    default line: one launch over 120 matrices."""
    r = run_pruned_batch(args, dev, summary=True)
    keep = ("ms_per_step", "value", "roofline", "hybrid_over_csr", "plan", "max_rel_err_vs_fp64_3_matrices")
    return {k: r[k] for k in keep if k in r}


def run_pruned_batch(args, dev, summary=False):
    """Many uneven BF16-block matrices in one launch at N = 128: the
    block-pruned projections of a 24-layer BERT-large-shaped encoder (120
    matrices, K = 1024 each), one block-diagonal plan on the B-stationary
    kernel."""
    import numpy as np
    import scipy.sparse
    import torch
    from oracle import ops
    from stile_amd import api, datasets
    N = 128
    t0 = time.perf_counter()
    mats, shapes = datasets.pruned_encoder()
    progress(f"pruned_batch: {len(mats)} matrices, nnz {sum(A.nnz for A in mats)}")
    plan, _, tuning = api.tune_plan("spmm", mats, "bfloat16", N, log=progress, b_dtype="bfloat16",
                                    cache=cache_path(args, "pruned_batch_n128"))
    t_plan = time.perf_counter() - t0
    info = plan.info()
    Mt, Kt = sum(m for m, _ in shapes), sum(k for _, k in shapes)
    nnz = sum(A.nnz for A in mats)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    Bf = torch.randn(Kt, N, device=dev, generator=g)
    Bd = Bf.to(torch.bfloat16)
    C = torch.empty((Mt, N), device=dev, dtype=torch.float32)
    ms = time_launches(lambda: plan.run_spmm(Bd, C), args.steps, args.warmup)
    big = scipy.sparse.block_diag(mats, format="csr").astype(np.float32)
    ms_csr, ms_vendor = _spmm_comparators(big, Bf, max(args.steps // 4, 10), 3)
    alg = sum(spmm_alg_bytes_blocks(full_blocks(A), A.nnz, m, k, N, s_b=2) for A, (m, k) in zip(mats, shapes))
    flops = 2 * nnz * N
    out = line("SpMM GFLOP/s (2*nnz*N), many BF16-block matrices in one launch at N=128",
               flops / (ms * 1e-3) / 1e9, "GFLOP/s", ms, args, "bf16 MFMA blocks + f32 sparse, f32 accumulate",
               {"workload": f"{len(mats)} block-pruned projections of a 24-layer BERT-large-shaped encoder "
                            f"(Q, K, V, O 1024x1024 and FFN-in 4096x1024 per layer; 10% of 16x16 blocks + 1% "
                            f"residual; nnz={nnz}) x B 1024x128 BF16 each, one block-diagonal launch",
                "matrices": len(mats), "M_total": Mt, "K_total": Kt, "N": N, "nnz": int(nnz),
                "parallelism": "single GPU"},
               roofline=roofline(alg, ms, measured_traffic("hyb_spmm_traffic_pruned_batch.json")),
               plain_csr_gflops=_gf(flops, ms_csr), vendor_csr_gflops=_gf(flops, ms_vendor),
               hybrid_over_csr=round(ms_csr / ms, 3),
               hybrid_over_vendor=None if ms_vendor is None else round(ms_vendor / ms, 3),
               mfma=mfma_use(info["num_chunks"], N, ms),
               plan={"windows": info["num_windows"], "mfma_chunks": info["num_chunks"],
                     "block_nnz": info["block_nnz"], "sparse_nnz": info["sparse_nnz"],
                     "search_tune_and_plan_s": round(t_plan, 3), "tuning": tuning})
    if not args.no_cpu or summary:
        Bh = Bd.double().cpu().numpy()
        Ch = C.cpu().numpy()
        errs = []
        r0 = c0 = 0
        for i, (A, (m, k)) in enumerate(zip(mats, shapes)):
            if i in (0, 4, len(mats) - 1):
                errs.append(ops.max_rel_err(Ch[r0:r0 + m], ops.spmm_fp64(A, Bh[c0:c0 + k])))
            r0, c0 = r0 + m, c0 + k
        out["max_rel_err_vs_fp64_3_matrices"] = max(errs)
        if not summary and not args.no_cpu:
            out["cpu_baseline"] = cpu_spmm_baseline(mats[4], Bh[4 * 1024:5 * 1024].astype(np.float32), N, 10.0,
                                                    "4096x1024 FFN-in")
    return out


def run_uniform_spmm(args, dev, summary=False):
    import torch
    from oracle import ops
    from stile_amd import api, datasets
    A, B = datasets.uniform_spmm(seed=0)
    M, K = A.shape
    N = B.shape[1]
    t0 = time.perf_counter()
    plan, _, tuning = api.tune_plan("spmm", A, "float32", N, log=progress, cache=cache_path(args, "uniform_spmm"))
    t_plan = time.perf_counter() - t0
    Bd = torch.from_numpy(B).to(dev)
    C = torch.empty((M, N), device=dev, dtype=torch.float32)
    ms = time_launches(lambda: plan.run_spmm(Bd, C), args.steps, args.warmup)
    ms_csr, ms_vendor, ms_hyb = _spmm_comparators(A, Bd, args.steps, args.warmup, hyb=True)
    flops = 2 * A.nnz * N
    out = line("SpMM GFLOP/s (2*nnz*N), uniform 1% A", flops / (ms * 1e-3) / 1e9, "GFLOP/s", ms, args, "f32",
               {"workload": f"SpMM FP32, uniform A 4096x4096 nnz={A.nnz} x B 4096x{N}", "M": M, "K": K, "N": N,
                "nnz": int(A.nnz), "parallelism": "single GPU"},
               roofline=roofline(spmm_alg_bytes(A.nnz, M, K, N), ms),
               plain_csr_gflops=_gf(flops, ms_csr), vendor_csr_gflops=_gf(flops, ms_vendor),
               sparsetir_hyb_gflops=_gf(flops, ms_hyb),
               hybrid_over_csr=round(ms_csr / ms, 3),
               hybrid_over_vendor=None if ms_vendor is None else round(ms_vendor / ms, 3),
               hybrid_over_sparsetir_hyb=round(ms_hyb / ms, 3),
               plan=dict(plan.info(), search_tune_and_plan_s=round(t_plan, 3), tuning=tuning))
    if summary or not args.no_cpu:
        out["max_rel_err_vs_fp64"] = ops.max_rel_err(C.cpu().numpy(), ops.spmm_fp64(A, B))
    if not summary and not args.no_cpu:
        out["cpu_baseline"] = cpu_spmm_baseline(A, B, N, 10.0, "4096x4096")
    return out


def run_rmat_sddmm(args, dev, summary=False):
    import numpy as np
    import torch
    from oracle import ops
    from stile_amd import _native, api, datasets
    A, X, Y = datasets.sddmm_rmat(seed=0)
    M, Nc = A.shape
    Kf = X.shape[1]
    t0 = time.perf_counter()
    plan, _, tuning = api.tune_plan("sddmm", A, "float32", Kf, log=progress, cache=cache_path(args, "rmat_sddmm"))
    t_plan = time.perf_counter() - t0
    Xd, Yd = torch.from_numpy(X).to(dev), torch.from_numpy(Y).to(dev)
    out_d = torch.empty(A.nnz, device=dev, dtype=torch.float32)
    ms = time_launches(lambda: plan.run_sddmm(Xd, Yd, out_d), args.steps, args.warmup)
    ip, ix, _ = _csr_dev(A, dev)
    o2 = torch.empty_like(out_d)
    ms_csr = time_launches(lambda: _native.csr_sddmm_f32(ip, ix, Xd, Yd, o2), args.steps, args.warmup)
    flops = 2 * A.nnz * Kf
    out = line("SDDMM GFLOP/s (2*nnz*K)", flops / (ms * 1e-3) / 1e9, "GFLOP/s", ms, args, "f32",
               {"workload": f"SDDMM FP32, R-MAT scale 16 + self loops (nnz={A.nnz}), X, Y {M}x{Kf}",
                "M": M, "N": Nc, "K": Kf, "nnz": int(A.nnz), "parallelism": "single GPU"},
               roofline=roofline(sddmm_alg_bytes(A.nnz, M, Nc, Kf), ms,
                                 measured_traffic("hyb_sddmm_traffic_rmat_sddmm.json")),
               plain_csr_gflops=_gf(flops, ms_csr), hybrid_over_csr=round(ms_csr / ms, 3),
               plan=dict(plan.info(), search_tune_and_plan_s=round(t_plan, 3), tuning=tuning))
    if summary or not args.no_cpu:
        got = out_d.cpu().numpy()
        ref = ops.sddmm_fp64(A, X, Y)
        out["max_rel_err_vs_fp64"] = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    if not summary and not args.no_cpu:
        rows = np.repeat(np.arange(M), np.diff(A.indptr))
        s = min(A.nnz, 1 << 18)

        def cpu():
            return (X[rows[:s]] * Y[A.indices[:s]]).sum(1)
        dt, n = _repeat_for(cpu, 10.0)
        out["cpu_baseline"] = {"value": round(2 * s * Kf / dt / 1e9, 3), "unit": "GFLOP/s", "cores": 1,
                               "kind": "port",
                               "sample": f"numpy (X[row]*Y[col]).sum(1) over the first {s} nonzeros, "
                                         f"{n} runs of {dt * 1e3:.1f} ms", "host": host_cpu()}
        Xt, Yt = torch.from_numpy(X), torch.from_numpy(Y)
        rt = torch.from_numpy(rows.astype(np.int64))
        ct = torch.from_numpy(A.indices.astype(np.int64))
        dt2, n2 = _repeat_for(lambda: (Xt[rt] * Yt[ct]).sum(1), 5.0)
        out["cpu_torch_all_threads"] = {"value": round(2 * A.nnz * Kf / dt2 / 1e9, 3), "unit": "GFLOP/s",
                                        "cores": torch.get_num_threads(), "kind": "torch CPU",
                                        "sample": f"torch (X[row]*Y[col]).sum(1) over all {A.nnz} nonzeros, "
                                                  f"{n2} runs of {dt2 * 1e3:.1f} ms"}
    return out


def run_gcn(args, dev):
    import numpy as np
    import torch
    from stile_amd import _native, datasets
    from stile_amd.autograd import SparseOperand, gcn_layer
    A, H, W = datasets.gcn_graph(args.gcn_scale)
    M = A.shape[0]
    F = H.shape[1]
    t0 = time.perf_counter()
    progress(f"gcn: graph scale {args.gcn_scale}, nnz {A.nnz}")
    op = SparseOperand(A, blk_dtype="float32", device=dev, N=F, tune=True, log=progress,
                       cache_dir=cache_dir(args, f"gcn{args.gcn_scale}"))
    t_plan = time.perf_counter() - t0
    Hd = torch.from_numpy(H).to(dev).requires_grad_(True)
    Wd = torch.from_numpy(W).to(dev).requires_grad_(True)
    vals = op.values.clone().requires_grad_(True)
    g = torch.Generator(device=dev)
    g.manual_seed(1)
    dY = torch.randn((M, F), device=dev, generator=g)

    def step():
        Hd.grad = None
        Wd.grad = None
        vals.grad = None
        y = gcn_layer(op, Hd, Wd, vals)
        y.backward(dY)
    steps = max(args.steps // 20, 5)
    ms = time_launches(step, steps, 2)
    HW = (Hd @ Wd).detach().contiguous()
    C = torch.empty_like(HW)
    ms_fwd = time_launches(lambda: op.plan.run_spmm(HW, C), steps, 2)
    ms_t = time_launches(lambda: op.plan_t.run_spmm(dY, C), steps, 2)
    sd = torch.empty(A.nnz, device=dev)
    ms_sd = time_launches(lambda: op.sddmm_plan.run_sddmm(dY, HW, sd), steps, 2)
    ms_csr, ms_vendor, ms_hyb = _spmm_comparators(A, HW, steps, 2, hyb=True)
    ip, ix, _ = _csr_dev(A, dev)
    ms_csr_sd = time_launches(lambda: _native.csr_sddmm_f32(ip, ix, dY, HW, sd), steps, 2)
    flops = 3 * 2 * A.nnz * F
    out = line("GCN layer sparse GFLOP/s (A.(HW), A^T.dY, SDDMM dA: 3 x 2*nnz*F per fwd+bwd step)",
               flops / (ms * 1e-3) / 1e9, "GFLOP/s", ms, args, "f32",
               {"workload": f"GCN layer fwd+bwd, R-MAT scale {args.gcn_scale} symmetric normalised "
                            f"(nnz={A.nnz}), H {M}x{F}, W {F}x{F}", "M": M, "F": F, "nnz": int(A.nnz),
                "parallelism": "single GPU"},
               roofline=roofline(spmm_alg_bytes(A.nnz, M, M, F), ms_fwd),
               kernels_ms={"spmm_fwd": round(ms_fwd, 4), "spmm_transposed": round(ms_t, 4),
                           "sddmm": round(ms_sd, 4), "plain_csr_spmm": round(ms_csr, 4),
                           "vendor_csr_spmm": None if ms_vendor is None else round(ms_vendor, 4),
                           "sparsetir_hyb_spmm": round(ms_hyb, 4),
                           "plain_csr_sddmm": round(ms_csr_sd, 4)},
               plan={"spmm": op.plan.info(), "sddmm": op.sddmm_plan.info(), "symmetric_reuse": op.symmetric,
                     "search_tune_and_plan_s": round(t_plan, 3), "tuning": op.tuning})
    if not args.no_cpu:
        rng = np.random.default_rng(0)
        rows = rng.choice(M, 2000, replace=False)
        hw = HW.double().cpu().numpy()
        ref = A[rows].astype(np.float64) @ hw
        op.plan.run_spmm(HW, C)
        got = C[torch.from_numpy(rows).to(dev)].double().cpu().numpy()
        out["max_rel_err_vs_fp64_sample"] = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        HWh = np.ascontiguousarray(HW.cpu().numpy())
        out["cpu_baseline"] = cpu_spmm_baseline(A, HWh, 16, 10.0, f"{M}x{M}")
    out["bf16"] = gcn_bf16_forward(args, dev, A, HW)
    return out


def gcn_bf16_forward(args, dev, A, HW):
    """The configuration's BF16 variant (H, W in BF16): the forward product
    A.(HW) with H.W in BF16 as B and BF16 MFMA blocks (FP32 sparse products,
    accumulation and C), its own tuned plan; error against FP64 on B's BF16
    values, sampled rows."""
    import numpy as np
    import torch
    from stile_amd import api
    M, F = HW.shape
    t0 = time.perf_counter()
    plan, _, tuning = api.tune_plan("spmm", A, "bfloat16", F, log=progress, b_dtype="bfloat16", workers=8,
                                    cache=cache_path(args, f"gcn{args.gcn_scale}_bf16"))
    t_plan = time.perf_counter() - t0
    HWb = HW.to(torch.bfloat16)
    C = torch.empty_like(HW)
    steps = max(args.steps // 20, 5)
    ms = time_launches(lambda: plan.run_spmm(HWb, C), steps, 2)
    rows = np.random.default_rng(0).choice(M, 2000, replace=False)
    plan.run_spmm(HWb, C)
    got = C[torch.from_numpy(rows).to(dev)].double().cpu().numpy()
    ref = A[rows].astype(np.float64) @ HWb.double().cpu().numpy()
    info = plan.info()
    plan.close()
    return {"spmm_ms": round(ms, 4), "gflops": _gf(2 * A.nnz * F, ms),
            "roofline": roofline(spmm_alg_bytes(A.nnz, M, M, F, s_b=2), ms),
            "max_rel_err_vs_fp64_sample": float(np.max(np.abs(got - ref)) / np.max(np.abs(ref))),
            "plan": {"windows": info["num_windows"], "mfma_chunks": info["num_chunks"],
                     "block_nnz": info["block_nnz"], "sparse_nnz": info["sparse_nnz"],
                     "search_tune_and_plan_s": round(t_plan, 3),
                     "tuning": {k: tuning.get(k) for k in ("choice", "ms", "build")}}}


def run_attention(args, dev):
    import numpy as np
    import torch
    from stile_amd import datasets
    from stile_amd.attention import BatchedSparseAttention
    masks = datasets.attention_masks(count=64)
    d = 64
    t0 = time.perf_counter()
    att = BatchedSparseAttention(masks, d, blk_dtype="bfloat16", device=dev, tune=True, log=progress,
                                 cache_dir=cache_dir(args, "attention"))
    t_plan = time.perf_counter() - t0
    n = int(att.row_off[-1])
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    Q, K, V = (torch.randn(n, d, device=dev, generator=g).to(torch.bfloat16) for _ in range(3))
    eager = att.forward(Q, K, V).clone()
    att.capture(Q, K, V)
    ms = time_launches(att.replay, max(args.steps // 4, 10), 3)
    o2 = torch.empty((n, d), device=dev, dtype=torch.float32)
    ms_fused = time_launches(lambda: att.att_plan.run_attention(Q, K, V, att.scale, out=o2),
                             max(args.steps // 4, 10), 3)
    # the unfused pipeline (SDDMM -> row softmax -> value refresh -> SpMM, one
    # graph) with the default search, for comparison
    progress("attention: unfused pipeline for comparison")
    un = BatchedSparseAttention(masks, d, blk_dtype="bfloat16", device=dev, fused=False,
                                cache_dir=cache_dir(args, "attention"))
    un.capture(Q, K, V)
    ms_unfused = time_launches(un.replay, max(args.steps // 10, 5), 2)
    unfused_equal = float((un.gout - att.gout).abs().max())
    flops = 2 * 2 * att.nnz * d
    # algorithmic bytes of the fused kernel: Q, K, V (BF16) and out (FP32)
    # once, 4 bytes of column index per mask entry, window row ids
    alg = 3 * n * d * 2 + n * d * 4 + att.nnz * 4 + 4 * n
    out = line("batched sparse attention GFLOP/s (QK^T and PV on the masks, 2 x 2*nnz*d per graph replay)",
               flops / (ms * 1e-3) / 1e9, "GFLOP/s", ms, args,
               "bf16 Q/K/V and probabilities, f32 scores/softmax/accumulation",
               {"workload": f"64 attention instances, L~U[512,8192] (sum {n}), d={d}, band 128 + 16 random "
                            f"(nnz={att.nnz}), fused kernel in one HIP graph", "instances": 64, "rows": n,
                "nnz": int(att.nnz), "parallelism": "single GPU"},
               roofline=roofline(alg, ms_fused),
               kernels_ms={"fused_attention": round(ms_fused, 4), "unfused_graph": round(ms_unfused, 4)},
               unfused_max_abs_diff=unfused_equal,
               plan={"attention": att.att_plan.info(), "search_tune_and_plan_s": round(t_plan, 3),
                     "tuning": att.tuning})
    if not args.no_cpu:
        att.replay()
        torch.cuda.synchronize()
        errs = []
        for b in (0, 31, 63):
            m = masks[b]
            a, e = att.row_off[b], att.row_off[b + 1]
            qq, kk, vv = (x[a:e].double().cpu().numpy() for x in (Q, K, V))
            rows = np.repeat(np.arange(m.shape[0]), np.diff(m.indptr))
            s = (qq[rows] * kk[m.indices]).sum(1) / np.sqrt(d)
            mx = np.maximum.reduceat(s, m.indptr[:-1])
            w = np.exp(s - mx[rows])
            w /= np.add.reduceat(w, m.indptr[:-1])[rows]
            ref = np.zeros_like(qq)
            np.add.at(ref, rows, w[:, None] * vv[m.indices])
            got = att.gout[a:e].double().cpu().numpy()
            errs.append(float(np.abs(got - ref).max() / np.abs(ref).max()))
        out["max_rel_err_vs_fp64_3_instances"] = max(errs)
        out["graph_replay_equals_eager"] = bool(torch.equal(att.gout, eager))
        m = masks[0]
        qq = Q[:m.shape[0]].float().cpu().numpy()
        kk = K[:m.shape[0]].float().cpu().numpy()
        vv = V[:m.shape[0]].float().cpu().numpy()
        rows = np.repeat(np.arange(m.shape[0]), np.diff(m.indptr))

        def cpu():
            s = (qq[rows] * kk[m.indices]).sum(1) * np.float32(1 / np.sqrt(d))
            mx = np.maximum.reduceat(s, m.indptr[:-1])
            w = np.exp(s - mx[rows])
            w /= np.add.reduceat(w, m.indptr[:-1])[rows]
            P = m.copy()
            P.data = w.astype(np.float32)
            return P @ vv
        dt, nrun = _repeat_for(cpu, 10.0)
        out["cpu_baseline"] = {"value": round(2 * 2 * m.nnz * d / dt / 1e9, 3), "unit": "GFLOP/s", "cores": 1,
                               "kind": "port",
                               "sample": f"numpy/scipy SDDMM + softmax + SpMM on instance 0 (L={m.shape[0]}, "
                                         f"nnz={m.nnz}), {nrun} runs of {dt * 1e3:.1f} ms"}
    return out


WORKLOADS = {"pruned_layer": run_pruned_layer, "uniform_spmm": run_uniform_spmm,
             "rmat_sddmm": run_rmat_sddmm, "gcn": run_gcn, "attention": run_attention,
             "pruned_batch": run_pruned_batch}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096, help="dense width of the pruned-layer workload")
    ap.add_argument("--workload", default="pruned_layer", choices=sorted(WORKLOADS) + ["all"])
    ap.add_argument("--gcn-scale", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-n128", action="store_true", help="pruned layer: skip the N=128 secondary figures")
    ap.add_argument("--no-batch", action="store_true",
                    help="pruned layer: skip the 120-matrix encoder batch sub-object")
    ap.add_argument("--no-small", action="store_true",
                    help="pruned layer: skip the uniform SpMM and R-MAT SDDMM sub-objects")
    ap.add_argument("--plan-cache", default=None,
                    help="directory of tuned packs: used if present (no search), written otherwise")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the tile path would shard "
              "by output rows with B replicated; only the one-GPU share is built)", file=sys.stderr)
        sys.exit(2)
    import torch
    dev = torch.device("cuda:0")
    names = [args.workload]
    if args.workload == "all":
        names = ["pruned_layer", "pruned_batch", "uniform_spmm", "rmat_sddmm", "attention", "gcn"]
    for name in names:
        print(json.dumps(WORKLOADS[name](args, dev)), flush=True)


if __name__ == "__main__":
    main()

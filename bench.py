"""Benchmark: hybrid fused SpMM on one MI355X (pruned-layer workload).

Workload (one step): C = A . B with A the 3072 x 768 block-pruned layer
(10 % of the 16x16 blocks dense + 1 % unstructured residual, 257,265
nonzeros, seed 0), B 768 x 4096 FP32 resident in HBM, C 3072 x 4096 FP32.
The tile set comes from the STile search (greedy + withdraw, cost model);
dense blocks run BF16 MFMA with FP32 accumulation, the rest FP32.

Printed: one JSON line with GFLOP/s (2 nnz N per step), the HBM roofline of
the fused kernel (algorithmic bytes nnz*8 + 4(M+1) + K*N*4 + M*N*4 per
launch over the HIP-event-timed launch duration), the plain FP32 CSR HIP
kernel on the same A and N, and the reference's CPU path (scipy CSR x dense,
FP32, one core) timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _alg_bytes(nnz, M, K, N):
    return nnz * 8 + 4 * (M + 1) + K * N * 4 + M * N * 4


def _measured_traffic(N):
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel calibration),
    for the N it was measured at; None otherwise."""
    path = os.path.join(ROOT, "profiles", "hyb_spmm_traffic_pruned_layer.json")
    if N != 4096 or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return int((t["fetch_kib"] * t["fetch_factor"] + t["write_kib"] * t["write_factor"]) * 1024)


def _time_launches(fn, steps, warmup):
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
    return s.elapsed_time(e) / steps  # ms per launch


def cpu_baseline(A, B, seconds=10.0):
    """The reference's CPU path (scipy csr.astype(float32).dot(B), one core)
    on a bounded sample: the first 512 columns of B, repeated for ~seconds."""
    import numpy as np
    from oracle import ops
    Bs = np.ascontiguousarray(B[:, :512])
    ops.spmm_ref_fp32(A, Bs)
    t0 = time.perf_counter()
    n = 0
    while True:
        ops.spmm_ref_fp32(A, Bs)
        n += 1
        if time.perf_counter() - t0 >= seconds or n >= 200:
            break
    dt = (time.perf_counter() - t0) / n
    return {"value": round(2 * A.nnz * Bs.shape[1] / dt / 1e9, 3), "unit": "GFLOP/s", "cores": 1,
            "kind": "port",
            "sample": f"scipy CSR(3072x768, nnz={A.nnz}) . B[:, :512] FP32, {n} runs of {dt * 1e3:.1f} ms"}


def check_error(A, B, C_gpu, cols=256):
    """Max norm-wise relative error vs FP64 on the first `cols` columns."""
    import numpy as np
    from oracle import ops
    ref = ops.spmm_fp64(A, B[:, :cols])
    return ops.max_rel_err(C_gpu[:, :cols], ref)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the tile path shards by "
              "output rows, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from stile_amd import _native, api, datasets

    dev = torch.device("cuda:0")
    N = args.n
    A, B = datasets.pruned_layer(seed=0, N=N)
    M, K = A.shape
    t0 = time.perf_counter()
    plan = api.build_spmm_plan(A, blk_dtype="bfloat16")
    t_plan = time.perf_counter() - t0
    info = plan.info()
    Bd = torch.from_numpy(B).to(dev)
    C = torch.empty((M, N), device=dev, dtype=torch.float32)
    ms = _time_launches(lambda: plan.run_spmm(Bd, C), args.steps, args.warmup)
    C_h = C.cpu().numpy()

    ip = torch.from_numpy(A.indptr.astype(np.int32)).to(dev)
    ix = torch.from_numpy(A.indices.astype(np.int32)).to(dev)
    vv = torch.from_numpy(A.data.astype(np.float32)).to(dev)
    C2 = torch.empty((M, N), device=dev, dtype=torch.float32)
    ms_csr = _time_launches(lambda: _native.csr_spmm_f32(ip, ix, vv, Bd, C2), max(args.steps // 4, 10), 5)

    flops = 2 * A.nnz * N
    gflops = flops / (ms * 1e-3) / 1e9
    gflops_csr = flops / (ms_csr * 1e-3) / 1e9
    byt = _alg_bytes(A.nnz, M, K, N)
    achieved = byt / (ms * 1e-3) / 1e9

    # secondary: the north-star width N = 128
    B128 = torch.from_numpy(np.ascontiguousarray(B[:, :128])).to(dev)
    C128 = torch.empty((M, 128), device=dev, dtype=torch.float32)
    ms128 = _time_launches(lambda: plan.run_spmm(B128, C128), args.steps, args.warmup)
    ms128_csr = _time_launches(lambda: _native.csr_spmm_f32(ip, ix, vv, B128, C128), args.steps, args.warmup)

    out = {
        "metric": "SpMM GFLOP/s (2*nnz*N), hybrid fused vs plain CSR",
        "value": round(gflops, 2), "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 MFMA blocks + f32 sparse, f32 accumulate", "data": "synthetic",
        "config": {"workload": "hybrid SpMM, block-pruned layer A 3072x768 (10% of 16x16 blocks + 1% residual, "
                               f"nnz={A.nnz}) x B 768x{N} FP32", "M": M, "K": K, "N": N, "nnz": int(A.nnz),
                   "global_batch": 1, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": _measured_traffic(N),
                     "alg_bytes": int(byt)},
        "plain_csr_gflops": round(gflops_csr, 2), "hybrid_over_csr": round(gflops / gflops_csr, 3),
        "n128": {"hybrid_us": round(ms128 * 1e3, 2), "csr_us": round(ms128_csr * 1e3, 2),
                 "hybrid_gflops": round(2 * A.nnz * 128 / (ms128 * 1e-3) / 1e9, 2),
                 "hbm_frac": round(_alg_bytes(A.nnz, M, K, 128) / (ms128 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "plan": {"windows": info["num_windows"], "mfma_chunks": info["num_chunks"],
                 "block_nnz": info["block_nnz"], "sparse_nnz": info["sparse_nnz"],
                 "search_and_plan_s": round(t_plan, 3)},
    }
    if not args.no_cpu:
        out["max_rel_err_vs_fp64"] = check_error(A, B, C_h)
        out["cpu_baseline"] = cpu_baseline(A, B)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

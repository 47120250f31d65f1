"""Measure the tile search's cost-model profile on an MI355X with this
library's kernels, and write it to stile_amd/mi355x_profile.json.

The reference fits its cost model to A100 measurements of its own kernels:
  * TC tiles: a fake graph of 108*100 row windows, each holding one dense
    16 x K tile, timed at feature size 32 (MY_sparse/SPMM_get_cost_model_TC.py
    :15-25, 94-184) -> the per-tile tables of my_cost_model.py:672-761;
  * ELL tiles: selected tiles replicated to fill the GPU and timed
    (MY_sparse/SPMM_get_cost_model_ELL.py:36-100) -> a linear fit of latency
    against the roofline feature, my_cost_model.py:786-800;
  * 1D SDDMM buckets likewise (MY_sparse/SDDMM_get_cost_model_1D.py:28-85)
    -> my_cost_model.py:1238-1245.
Here every measurement is one saturating launch of the hybrid kernels on a
synthetic plan made of a single tile kind, and the cost of one tile is
    launch time (ms) x number of CUs / number of tiles
with tiles counted per 32 columns of the dense operand (the reference's
J = 32), i.e. the same unit as the A100 tables (per-SM latency of a tile).

  TC (16, 32, K), K = 16..160: W windows, window w holding one dense 16 x K
      block on its own K columns; N = 64.  Measured for FP32 blocks (dsize
      32) and BF16 blocks (dsize 16) on the streaming kernel.
  ELL (I, 32, K) for the six shapes with I*K = 256: rows of exactly K
      nonzeros on uniformly random columns of a 65536-column B (and, to
      spread the feature, columns drawn from a 512-column pool); the fit
      cost = a1 * x + b1 + c1 * I uses the cost model's own feature x plus
      a per-row term (c1 = 0 gives the reference's form).
  B-stationary kernel (BF16 blocks, B of at most LDS_MAX_K rows), in its
      own work-unit shape: W windows over one B of 1024 rows, each with a
      dense block of 32 + K random columns, and with a 32-column block plus
      r sparse entries per row (r = 1..16).  The store of a window's rows
      and the staging of B are paid whatever a window's tiles are, so the
      table tc_cost_lds holds what K columns add to a window (a line in K
      fitted to the ten added costs), and the window's sparse part is
      priced per entry plus per window (ell_lds_coeff).  B is BF16.
  1D SDDMM buckets of 32 nonzeros: rows of r nonzeros, r = 1..32, K_feat =
      32; fit of the cost model's 1D feature.
  TC SDDMM (16, J, 16): dense 16 x 16 output blocks at K_feat = J.

With these kernels an atomic tile costs what a plain one does: ELL pieces of
one row are merged into that row's sparse part by the plan builder and TC
windows never use atomics.  So the atomic tables equal the plain ones and
the atomic multiplier is (a2, b2) = (0, 1).

Run on the GPU:  python -m stile_amd.autotune [--out PATH]
"""
import argparse
import datetime
import json
import math
import os
import sys
import time

import numpy as np

CU_NUM = 256          # MI355X compute units
THREADS_PER_CU = 2048
HBM_TBS = 8.0
FP32_TFLOPS = 157.3
BF16_TFLOPS = 2500.0


# ---------------------------------------------------------------------------
# Synthetic single-kind packs.
# ---------------------------------------------------------------------------

def tc_pack(W, K):
    """W windows; window w: rows 16w..16w+15, one dense 16 x K tile on
    columns w*K .. w*K+K-1 (the reference's fake TC graph)."""
    cols = np.arange(W * K, dtype=np.int32).reshape(W, K)
    ncols = W * K
    g = dict(group_size=K, num_windows=W, nnzb=W, rows_2d=False, k_notsorted=False, real_atomic=False,
             a=np.ones((W, 16, K), np.float32), i_indices=np.arange(W, dtype=np.int32),
             indptr=np.arange(W + 1, dtype=np.int32), indices=cols, a_ids=None)
    return dict(M=16 * W, K=ncols, ell=[], tc=[g])


def ell_pack(R, I, K, ncols, pool, rng):
    """R rows of exactly K nonzeros (columns drawn without repetition from
    `pool` columns, randomly placed in [0, ncols)), grouped as ELL tiles of I
    rows."""
    T = R // I
    R = T * I
    base = rng.choice(ncols, size=pool, replace=False).astype(np.int32) if pool < ncols else None
    if K > min(pool, ncols):
        raise ValueError("rows of K distinct columns need K <= pool")

    def draw(shape):
        if base is None:
            return rng.integers(0, ncols, size=shape, dtype=np.int32)
        return base[rng.integers(0, pool, size=shape)]
    cols = np.empty((R, K), np.int32)
    for r0 in range(0, R, 4096):
        n = min(4096, R - r0)
        c = draw((n, K))
        while True:                 # redraw repeated columns until each row is distinct
            c.sort(axis=1)
            dup = np.zeros_like(c, bool)
            dup[:, 1:] = c[:, 1:] == c[:, :-1]
            if not dup.any():
                break
            c[dup] = draw(int(dup.sum()))
        cols[r0:r0 + n] = c
    g = dict(m_blk=I, nnz_cols=K, num_tiles=T, real_atomic=False, a=np.ones((T, I * K), np.float32),
             indices_k=cols, indices_reordered_i=np.arange(R, dtype=np.int32), a_ids=None)
    return dict(M=R, K=ncols, ell=[g], tc=[]), cols


def _distinct_rows(n, k, hi, rng):
    """n rows of k distinct integers in [0, hi), each row sorted."""
    c = np.sort(rng.integers(0, hi, size=(n, k)), axis=1)
    while k > 1:
        dup = np.zeros_like(c, bool)
        dup[:, 1:] = c[:, 1:] == c[:, :-1]
        if not dup.any():
            break
        c[dup] = rng.integers(0, hi, size=int(dup.sum()))
        c.sort(axis=1)
    return c


def lds_pack(W, Kd, r, rng, b_rows):
    """A B-stationary work unit shape: W windows over a B of b_rows rows;
    window w holds one dense 16 x Kd block on Kd random distinct columns and,
    with r > 0, r more entries in each of its rows on random columns outside
    the block (one ELL tile of 16 rows and r columns per window, which the
    plan turns into the window's sparse part)."""
    perm = np.argsort(rng.random((W, b_rows)), axis=1).astype(np.int32)
    cols = np.sort(perm[:, :Kd], axis=1)
    g = dict(group_size=Kd, num_windows=W, nnzb=W, rows_2d=False, k_notsorted=False, real_atomic=False,
             a=np.ones((W, 16, Kd), np.float32), i_indices=np.arange(W, dtype=np.int32),
             indptr=np.arange(W + 1, dtype=np.int32), indices=cols, a_ids=None)
    ell = []
    if r:
        idx = _distinct_rows(16 * W, r, b_rows - Kd, rng)
        rest = perm[:, Kd:]
        ecols = rest[np.repeat(np.arange(W), 16)[:, None], idx].astype(np.int32)
        ell = [dict(m_blk=16, nnz_cols=r, num_tiles=W, real_atomic=False, a=np.ones((W, 16 * r), np.float32),
                    indices_k=ecols, indices_reordered_i=np.arange(16 * W, dtype=np.int32), a_ids=None)]
    return dict(M=16 * W, K=b_rows, ell=ell, tc=[g])


def sddmm_1d_pack(R, r, ncols, rng):
    """R rows of r nonzeros each (random columns); 1D entries in CSR order."""
    cols = np.sort(rng.integers(0, ncols, size=(R, r)), axis=1).astype(np.int32)
    rows = np.repeat(np.arange(R, dtype=np.int32), r)
    nnz = R * r
    g = dict(mid_idx=rows, j_idx=cols.reshape(-1), remap=np.arange(nnz, dtype=np.int32))
    return dict(M=R, N=ncols, nnz=nnz, g1d=[g], tc=[]), rows, cols.reshape(-1)


def sddmm_tc_pack(W):
    """W dense 16 x 16 output blocks on the diagonal, every entry sampled."""
    i_ind = (np.arange(W, dtype=np.int32)[:, None] * 16 + np.arange(16, dtype=np.int32)[None, :])
    pos = np.arange(256, dtype=np.int32)
    rr, cc = pos // 16, pos % 16
    srt = np.lexsort((rr, cc))                       # CSC order inside a block
    dense = (cc[srt] + rr[srt] * 16).astype(np.int32)
    out_local = (rr[srt] * 16 + cc[srt]).astype(np.int32)
    g = dict(nnzb=W, blk_k=16, i_indices=i_ind, j_indices=i_ind.copy(),
             indptr=(np.arange(W + 1, dtype=np.int32) * 256),
             indices_dense=np.tile(dense, W),
             indices=(np.arange(W, dtype=np.int32)[:, None] * 256 + out_local[None, :]).reshape(-1))
    return dict(M=16 * W, N=16 * W, nnz=256 * W, g1d=[], tc=[g])


# ---------------------------------------------------------------------------
# Timing.
# ---------------------------------------------------------------------------

def _time_ms(fn, reps=30, warmup=5):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def measure_tc(K, blk, W=8192, N=64):
    import torch
    from . import _native
    pack = tc_pack(W, K)
    plan = _native.Plan.spmm(pack, blk)
    B = torch.randn(pack["K"], N, device="cuda")
    C = torch.empty(16 * W, N, device="cuda")
    ms = _time_ms(lambda: plan.run_spmm(B, C))
    plan.close()
    return ms * CU_NUM / (W * N / 32)


def measure_ell(I, K, pool, rng, R=1 << 18, ncols=1 << 16, N=64):
    import torch
    from . import _native
    pack, cols = ell_pack(R, I, K, ncols, pool, rng)
    plan = _native.Plan.spmm(pack, "float32")
    B = torch.randn(ncols, N, device="cuda")
    C = torch.empty(pack["M"], N, device="cuda")
    ms = _time_ms(lambda: plan.run_spmm(B, C))
    plan.close()
    T = pack["ell"][0]["num_tiles"]
    return ms * CU_NUM / (T * N / 32), cols


def measure_lds(Kd, r, rng, W=16384, N=64, b_rows=1024, reps=100, rounds=3):
    """Per-window cost (ms x CUs per window per 32 columns) of lds_pack(W,
    Kd, r) on the B-stationary kernel with BF16 B: the median of `rounds`
    timings of `reps` launches."""
    import torch
    from . import _native
    plan = _native.Plan.spmm(lds_pack(W, Kd, r, rng, b_rows), "bfloat16")
    B = torch.randn(b_rows, N, device="cuda").to(torch.bfloat16)
    C = torch.empty(16 * W, N, device="cuda")
    ms = float(np.median([_time_ms(lambda: plan.run_spmm(B, C), reps) for _ in range(rounds)]))
    plan.close()
    return ms * CU_NUM / (W * N / 32)


def measure_sddmm_1d(r, rng, R=1 << 18, ncols=1 << 16, Kf=32):
    import torch
    from . import _native
    pack, rows, cols = sddmm_1d_pack(R, r, ncols, rng)
    plan = _native.Plan.sddmm(pack, "float32")
    X = torch.randn(R, Kf, device="cuda")
    Y = torch.randn(ncols, Kf, device="cuda")
    out = torch.empty(pack["nnz"], device="cuda")
    ms = _time_ms(lambda: plan.run_sddmm(X, Y, out))
    plan.close()
    return ms * CU_NUM / (pack["nnz"] / 32), rows, cols


def measure_sddmm_tc(Kf, W=16384):
    import torch
    from . import _native
    plan = _native.Plan.sddmm(sddmm_tc_pack(W), "float32")
    X = torch.randn(16 * W, Kf, device="cuda")
    out = torch.empty(256 * W, device="cuda")
    ms = _time_ms(lambda: plan.run_sddmm(X, X, out))
    plan.close()
    return ms * CU_NUM / W


# ---------------------------------------------------------------------------
# Features (the cost model's own) and fits.
# ---------------------------------------------------------------------------

def ell_features(I, K, cols, dsize, peak):
    """x of cost_tb_latency_given_ELL_tiles for the tiles of I rows of the
    synthetic ELL matrix (every row full, so no padding column)."""
    from .my_cost_model import ell_tile_x
    T = cols.shape[0] // I
    tiles = cols[:T * I].reshape(T, I * K).astype(np.int64)
    tiles.sort(axis=1)
    uniq = 1 + (tiles[:, 1:] != tiles[:, :-1]).sum(axis=1)
    return ell_tile_x(I, 32, K, np.full(T, I * K), uniq, np.zeros(T, bool), dsize, peak)


def sddmm_1d_features(r, rows, cols, Kf, dsize, peak):
    from .my_cost_model import sddmm_1d_x
    nb = len(rows) // 32
    bid = np.repeat(np.arange(nb, dtype=np.int64), 32)
    big = int(max(rows.max(), cols.max())) + 1
    rn = np.bincount(np.unique(bid * big + rows[:nb * 32]) // big, minlength=nb)
    cn = np.bincount(np.unique(bid * big + cols[:nb * 32]) // big, minlength=nb)
    return sddmm_1d_x(rn, cn, Kf, 32, dsize, peak)


def linear_fit(x, y):
    """Least-squares y = a x + b with a, b >= 0 (a latency model must stay
    positive for the small tiles the search also prices): the free fit if
    it is non-negative, else the better of the two one-parameter fits."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    A = np.stack([x, np.ones_like(x)], axis=1)
    (a, b), *_ = np.linalg.lstsq(A, y, rcond=None)
    if a >= 0 and b >= 0:
        return float(a), float(b)
    cands = [(max(float(x @ y / (x @ x)), 0.0), 0.0), (0.0, max(float(y.mean()), 0.0))]
    err = [float(np.sum((ca * x + cb - y) ** 2)) for ca, cb in cands]
    return cands[int(np.argmin(err))]


def nonneg_fit(features, y, relative=True):
    """Least-squares y = sum_i c_i f_i with every c_i >= 0.  relative: in
    relative error (each sample weighted by 1 / y, so that small tiles count
    as much as big ones; every y must be positive); otherwise in absolute
    error (for added costs, which can be near 0 or measure slightly
    negative).  features: list of arrays (one per coefficient)."""
    from scipy.optimize import nnls
    y = np.asarray(y, np.float64)
    A = np.stack([np.asarray(f, np.float64) for f in features], axis=1)
    if relative:
        if np.any(y <= 0):
            raise ValueError("a relative fit needs positive samples")
        c, _ = nnls(A / y[:, None], np.ones_like(y))
    else:
        c, _ = nnls(A, y)
    return [float(v) for v in c]


def run(out_path=None, log=print):
    import torch
    from . import my_cost_model as mc
    rng = np.random.default_rng(0)
    dev = torch.cuda.get_device_properties(0)
    peak32 = FP32_TFLOPS / HBM_TBS
    peak16 = BF16_TFLOPS / HBM_TBS
    t0 = time.time()
    tc32, tc16 = {}, {}
    for i in range(10, 0, -1):
        K = 16 * i
        tc32[(16, 32, K)] = measure_tc(K, "float32")
        tc16[(16, 32, K)] = measure_tc(K, "bfloat16")
        log(f"TC (16,32,{K}): f32 {tc32[(16, 32, K)]:.3e}  bf16 {tc16[(16, 32, K)]:.3e} ms/tile/CU")
    ell_shapes = [(256, 1), (128, 2), (64, 4), (32, 8), (16, 16), (8, 32)]
    ell_cost, samples = {}, []
    for I, K in ell_shapes:
        for pool in (1 << 16, 512):
            c, cols = measure_ell(I, K, pool, rng)
            samples.append((I, K, cols, c))
            if pool == 1 << 16:
                ell_cost[(I, 32, K)] = c
            log(f"ELL ({I},32,{K}) pool {pool}: {c:.3e} ms/tile/CU")
    ell_coeff, ell_row, raw = {}, {}, {"ell": [], "sddmm_1d": [], "lds_tc": [], "lds_ell": []}
    for dsize, peak in ((16, peak16), (32, peak32)):
        xs, ys, Is = [], [], []
        for I, K, cols, c in samples:
            x = ell_features(I, K, cols, dsize, peak)
            xs.append(float(np.mean(x)))
            ys.append(c)
            Is.append(I)
            raw["ell"].append({"I": I, "K": K, "dsize": dsize, "x": xs[-1], "cost": c})
        # a1 x + b1 + c1 I: these kernels pay per row of a tile (its row
        # bounds, one store of its 32 columns) beyond what x counts
        a1, b1, c1 = nonneg_fit([xs, np.ones(len(xs)), Is], ys)
        ell_coeff[dsize], ell_row[dsize] = (a1, b1), c1
        log(f"ELL fit dsize {dsize}: a1={a1:.4e} b1={b1:.4e} c1={c1:.4e}")
    # B-stationary kernel (BF16 blocks, K <= LDS_MAX_K): the store of a
    # window's rows and the staging of B are paid whatever its tiles are, so
    # a TC tile is priced by what its columns add to a window that already
    # has one full chunk, and a window's sparse part by what its rounds add.
    base = measure_lds(32, 0, rng)
    ks = [16 * i for i in range(1, 11)]
    add = [measure_lds(32 + K, 0, rng) - base for K in ks]
    alpha, beta = nonneg_fit([ks, np.ones(len(ks))], add, relative=False)
    tclds = {(16, 32, K): alpha * K + beta for K in ks}
    raw["lds_tc"] = [{"K": K, "added_cost": float(a)} for K, a in zip(ks, add)]
    log(f"B-stationary TC: base window {base:.3e}, added per column {alpha:.3e} + per tile {beta:.3e}")
    rs = [1, 2, 4, 8, 16]
    sp = [measure_lds(32, r, rng) - base for r in rs]
    a_e, b_w = nonneg_fit([[16 * r for r in rs], np.ones(len(rs))], sp, relative=False)
    raw["lds_ell"] = [{"r": r, "added_cost": float(s)} for r, s in zip(rs, sp)]
    log(f"B-stationary sparse part: per entry {a_e:.3e}, per window {b_w:.3e}")
    sd_samples = []
    for r in (1, 2, 4, 8, 16, 32):
        c, rows, cols = measure_sddmm_1d(r, rng)
        sd_samples.append((r, rows, cols, c))
        log(f"1D SDDMM rows of {r}: {c:.3e} ms/bucket/CU")
    sd_coeff = {}
    for dsize, peak in ((16, peak16), (32, peak32)):
        xs = [float(np.mean(sddmm_1d_features(r, rows, cols, 32, dsize, peak))) for r, rows, cols, _ in sd_samples]
        sd_coeff[dsize] = linear_fit(np.asarray(xs), np.asarray([c for *_, c in sd_samples]))
        raw["sddmm_1d"] += [{"r": r, "dsize": dsize, "x": x, "cost": c}
                            for (r, _, _, c), x in zip(sd_samples, xs)]
        log(f"1D SDDMM fit dsize {dsize}: a={sd_coeff[dsize][0]:.4e} b={sd_coeff[dsize][1]:.4e}")
    tc_sd = {}
    for J in (16, 32, 64):
        c = measure_sddmm_tc(J)
        for i in range(1, 9):
            tc_sd[(16, J, 16 * i)] = c * i
        log(f"TC SDDMM (16,{J},16): {c:.3e} ms/block/CU")
    prof = mc.HardwareProfile(
        name="mi355x", sm_num=CU_NUM, threads_per_sm=THREADS_PER_CU,
        peak_fp32=peak32, peak_fp16=peak16, peak_tc=peak16,
        ell_coeff=ell_coeff, ell_atomic_coeff={16: (0.0, 1.0), 32: (0.0, 1.0)}, ell_row_coeff=ell_row,
        ell_lds_coeff=(a_e, b_w),
        sddmm_1d_coeff=sd_coeff, tc_cost=tc16, tc_cost_atomic=dict(tc16),
        ell_cost=ell_cost, ell_cost_atomic=dict(ell_cost), tc_sddmm_cost=tc_sd,
        tc_cost_f32=tc32, tc_cost_lds=tclds)
    meta = {"device": dev.name, "gcn_arch": getattr(dev, "gcnArchName", ""), "cus": dev.multi_processor_count,
            "date": datetime.date.today().isoformat(), "seconds": round(time.time() - t0, 1),
            "command": "python -m stile_amd.autotune", "samples": raw}
    mc.save_profile_file(prof, out_path or mc.PROFILE_PATH, meta=meta)
    return prof


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    run(args.out, log=lambda s: print(s, flush=True))


if __name__ == "__main__":
    sys.exit(main())

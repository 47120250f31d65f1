"""Restatement of the reference's analytic cost model (test infrastructure).

MY_sparse/my_cost_model.py with the A100 constants it ships.  Pinned by
tests/test_oracle_cost_model.py with values evaluated by hand.
"""
import math

import numpy as np

PEAK_FP32 = 19.49 / 1.555          # my_cost_model.py:1200-1203
PEAK_FP16 = 78 / 1.555
ELL_COEF = {16: (8.80425616e-08, 4.17697078e-04), 32: (8.23071690e-08, 3.42962663e-03)}   # :786-800
ELL_ATOMIC = {16: (8.80688635, -7.9094305), 32: (0.67944825, 0.29144282)}
SDDMM_1D = {16: (5.40647867e-08, 2.14658164e-04), 32: (5.90185816e-08, 1.83606878e-04)}  # :1238-1245


def ell_cost_model(x_nonatomic, x_atomic, is_atomic, dsize):
    """my_cost_model.py:808-816."""
    a1, b1 = ELL_COEF[dsize]
    a2, b2 = ELL_ATOMIC[dsize]
    if is_atomic:
        return (a1 * x_nonatomic + b1) * (a2 * x_atomic / x_nonatomic + b2)
    return a1 * x_nonatomic + b1


def ell_tile_cost(I, J, K, k_is, is_atomic, dsize):
    """cost_tb_latency_given_ELL_tiles for one tile (:1176-1214, memory model
    :526-577): k_is = distinct column ids of the tile (+1 when padded)."""
    cls = 32 * 8 / dsize
    flops = I * J * K * 2
    mem_a = math.ceil(I * K / cls)
    mem_b = math.ceil(J / cls) * k_is
    store = math.ceil(J * dsize / 256) * I
    mem = [mem_a + store + mem_b, mem_a + store * 2 + mem_b]
    peak = PEAK_FP32 if dsize == 32 else PEAK_FP16
    tp = [min(flops / (m * 32), peak) for m in mem]
    if is_atomic:
        return ell_cost_model(flops / tp[0], flops / tp[1], True, dsize)
    return ell_cost_model(flops / tp[0], None, False, dsize)


def lower_bound_ell(shapes, dsize):
    """lower_bound_ELL_cost_BSPMM (:820-860): min over (I, J, K) of the
    non-atomic cost of a full tile with K//2+1 distinct columns per tile."""
    best = float('inf')
    peak = PEAK_FP32 if dsize == 32 else PEAK_FP16
    for I, J, K in shapes:
        flops = I * J * K * 2
        mem = math.ceil(I * K * dsize / 256) + math.ceil(J * dsize / 256) * (K // 2 + 1) + I * math.ceil(J * dsize / 256)
        tp = min(flops / (mem * 32), peak)
        best = min(best, ell_cost_model(flops / tp, None, False, dsize) / (I * K * J))
    return best


def sddmm_1d_cost(row_nums, col_nums, Kf, bucket, dsize):
    """_cost_tb_latency_1D_tiles_given_features (:1219-1248)."""
    flops = Kf * bucket * 2
    mem = ((np.asarray(row_nums) + np.asarray(col_nums)) * Kf + bucket) * dsize / 8
    peak = PEAK_FP32 if dsize == 32 else PEAK_FP16
    tp = np.minimum(flops / mem, peak)
    a, b = SDDMM_1D[dsize]
    return a * (flops / tp) + b

"""Per-tile latency model guiding the format search (MI355X build).

Same names and formulas as the reference module MY_sparse/my_cost_model.py:
  compute_cost_tb_impl_given_tile_sizes (:81-91)
  memory_cost_tb_given_tile_sizes_and_poses (:517-582)
  cost_tb_latency_given_ELL_tiles (:1176-1214)
  ELL_non_atomic_coeff / ELL_atomic_extra_coeff / ELL_cost_model (:786-816)
  _cost_tb_latency_1D_tiles_given_features / cost_tb_latency_given_1D_tiles (:1219-1282)
  get_benchmark_cost_ / get_benchmark_cost (:639-777)
  lower_bound_ELL_cost(_BSPMM) / lower_bound_1D_cost_BSDDMM (:820-904)
  cost_tb_latency_given_tile (:1099-1147)

Model: roofline throughput R = min(FLOPs / bytes, PEAK) and latency
g = a * FLOPs / R + b, with a multiplier (a2 * x_atomic / x + b2) for atomic
ELL tiles; dense (TC) tiles use measured per-tile tables.  A measured
profile may add two terms the reference's kernels did not need:
  * a per-row term c1 * I in the ELL latency (ell_row_coeff);
  * for BF16 blocks over a B of at most LDS_MAX_K rows (the B-stationary
    kernel, where ELL entries of rows in dense windows become those windows'
    sparse parts, added in rounds of one entry per row), ELL tiles priced
    per entry as one round each plus per 16 rows (ell_lds_coeff), and TC
    tiles by what their columns add to a window (tc_cost_lds).
All constants live in a HardwareProfile:
  'a100'   the reference's constants, measured on an A100 in FP16
           (my_cost_model.py:672-761, 786-800, 1238-1245; PEAK ratios
           19.49/1.555, 78/1.555, 312/1.555; SM_num 108)
  'mi355x' constants measured with this library's kernels on an MI355X by
           stile_amd.autotune (tables written to mi355x_profile.json); the
           default once that file exists
The MLP cost model of the reference (env other_cost_model=MLP, :925-1093)
is not built.
"""
import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

from .gen_formats_v2 import get_template_str, get_tile_sizes_idx, get_tile_sizes_list


@dataclass
class HardwareProfile:
    name: str
    sm_num: int
    threads_per_sm: int
    peak_fp32: float      # flop per byte ridge of the FP32 path (PEAK in the reference)
    peak_fp16: float
    peak_tc: float
    ell_coeff: dict       # dsize -> (a1, b1)
    ell_atomic_coeff: dict  # dsize -> (a2, b2)
    sddmm_1d_coeff: dict  # dsize -> (a, b)
    tc_cost: dict         # (I, J, K) -> latency, non-atomic
    tc_cost_atomic: dict
    ell_cost: dict
    ell_cost_atomic: dict
    tc_sddmm_cost: dict
    tc_cost_f32: dict = None   # TC table for dsize 32 (FP32 blocks); tc_cost serves both if None
    tc_cost_lds: dict = None   # TC table for BF16 blocks when B has <= LDS_MAX_K rows (B-stationary kernel)
    ell_row_coeff: dict = None  # dsize -> c1, per-row term of the ELL latency (0 in the reference's form)
    ell_lds_coeff: tuple = None  # (per entry, per window) cost of dense windows' sparse parts, B-stationary kernel


# B-stationary SpMM kernel: BF16 blocks and at most this many rows of B
# (hyb_spmm_lds_bf16_kernel stages B[:, 64 columns] in LDS).
LDS_MAX_K = 1024


_A100 = HardwareProfile(
    name="a100", sm_num=108, threads_per_sm=2048,
    peak_fp32=19.49 / 1.555, peak_fp16=78 / 1.555, peak_tc=312 / 1.555,
    ell_coeff={16: (8.80425616e-08, 4.17697078e-04), 32: (8.23071690e-08, 3.42962663e-03)},
    ell_atomic_coeff={16: (8.80688635, -7.9094305), 32: (0.67944825, 0.29144282)},
    sddmm_1d_coeff={16: (5.40647867e-08, 2.14658164e-04), 32: (5.90185816e-08, 1.83606878e-04)},
    tc_cost={
        (16, 32, 16): 0.0002922496385872364, (16, 32, 32): 0.00041533440351486205,
        (16, 32, 48): 0.0005173248425126075, (16, 32, 64): 0.0006273023784160614,
        (16, 32, 80): 0.0007631871849298477, (16, 32, 96): 0.0008801279217004776,
        (16, 32, 112): 0.0009804800152778625, (16, 32, 128): 0.001063014417886734,
        (16, 32, 144): 0.0012102656066417695, (16, 32, 160): 0.001244262307882309},
    tc_cost_atomic={
        (16, 32, 16): 0.0009056255966424942, (16, 32, 32): 0.0010332158952951432,
        (16, 32, 48): 0.001133260801434517, (16, 32, 64): 0.0012712959945201873,
        (16, 32, 80): 0.0014235648512840272, (16, 32, 96): 0.0015714305639266968,
        (16, 32, 112): 0.0017159168422222137, (16, 32, 128): 0.0018328575789928437,
        (16, 32, 144): 0.0021766144037246703, (16, 32, 160): 0.0021382145583629607},
    ell_cost={(256, 32, 1): 0.342255333 / 100, (128, 32, 2): 0.27794799 / 100,
              (64, 32, 4): 0.214812706 / 100, (32, 32, 8): 0.189769996 / 100,
              (16, 32, 16): 0.170571889 / 100, (8, 32, 32): 0.226909105 / 100},
    ell_cost_atomic={(256, 32, 1): 1.324675103 / 100, (128, 32, 2): 0.866597613 / 100,
                     (64, 32, 4): 0.493691849 / 100, (32, 32, 8): 0.315331313 / 100,
                     (16, 32, 16): 0.229816495 / 100, (8, 32, 32): 0.240047723 / 100},
    tc_sddmm_cost={
        (16, 16, 16): 0.019077118486166 / 100, (16, 32, 16): 0.027914240956306458 / 100,
        (16, 64, 16): 0.04804608225822449 / 100, (16, 16, 32): 0.026531841605901718 / 100,
        (16, 32, 32): 0.0391475185751915 / 100, (16, 64, 32): 0.06515711545944214 / 100,
        (16, 16, 48): 0.03529728576540947 / 100, (16, 32, 48): 0.05030912533402443 / 100,
        (16, 64, 48): 0.07875583320856094 / 100, (16, 16, 64): 0.044257279485464096 / 100,
        (16, 32, 64): 0.06239231675863266 / 100, (16, 64, 64): 0.10551295429468155 / 100,
        (16, 16, 80): 0.05752832442522049 / 100, (16, 32, 80): 0.07265280187129974 / 100,
        (16, 64, 80): 0.118947833776474 / 100, (16, 16, 96): 0.06372351944446564 / 100,
        (16, 32, 96): 0.08369152247905731 / 100, (16, 64, 96): 0.1488281637430191 / 100,
        (16, 16, 112): 0.07803903520107269 / 100, (16, 32, 112): 0.09360384196043015 / 100,
        (16, 64, 112): 0.15593470633029938 / 100, (16, 16, 128): 0.07907328009605408 / 100,
        (16, 32, 128): 0.10522624105215073 / 100, (16, 64, 128): 0.18569216132164001 / 100},
)

_PROFILES = {"a100": _A100}
PROFILE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mi355x_profile.json")
# The measured MI355X profile is the default once it exists.
_ACTIVE = ["mi355x" if os.path.exists(PROFILE_PATH) else "a100"]


def _key_tuple(k):
    return tuple(int(x) for x in k.strip("()").split(","))


def load_profile_file(path=PROFILE_PATH):
    """Load a measured profile (written by stile_amd.autotune) and register it."""
    with open(path) as f:
        d = json.load(f)
    conv = lambda tab: {_key_tuple(k): float(v) for k, v in tab.items()}
    prof = HardwareProfile(
        tc_cost_f32=conv(d["tc_cost_f32"]) if d.get("tc_cost_f32") else None,
        tc_cost_lds=conv(d["tc_cost_lds"]) if d.get("tc_cost_lds") else None,
        ell_row_coeff={int(k): float(v) for k, v in d["ell_row_coeff"].items()} if d.get("ell_row_coeff") else None,
        ell_lds_coeff=tuple(d["ell_lds_coeff"]) if d.get("ell_lds_coeff") else None,
        name=d["name"], sm_num=int(d["sm_num"]), threads_per_sm=int(d["threads_per_sm"]),
        peak_fp32=float(d["peak_fp32"]), peak_fp16=float(d["peak_fp16"]), peak_tc=float(d["peak_tc"]),
        ell_coeff={int(k): tuple(v) for k, v in d["ell_coeff"].items()},
        ell_atomic_coeff={int(k): tuple(v) for k, v in d["ell_atomic_coeff"].items()},
        sddmm_1d_coeff={int(k): tuple(v) for k, v in d["sddmm_1d_coeff"].items()},
        tc_cost=conv(d["tc_cost"]), tc_cost_atomic=conv(d["tc_cost_atomic"]),
        ell_cost=conv(d["ell_cost"]), ell_cost_atomic=conv(d["ell_cost_atomic"]),
        tc_sddmm_cost=conv(d["tc_sddmm_cost"]))
    _PROFILES[prof.name] = prof
    return prof


def save_profile_file(prof, path=PROFILE_PATH, meta=None):
    enc = lambda tab: {str(tuple(k)): v for k, v in tab.items()}
    d = dict(meta=meta or {}, tc_cost_f32=enc(prof.tc_cost_f32) if prof.tc_cost_f32 else None,
             tc_cost_lds=enc(prof.tc_cost_lds) if prof.tc_cost_lds else None,
             ell_row_coeff={str(k): v for k, v in prof.ell_row_coeff.items()} if prof.ell_row_coeff else None,
             ell_lds_coeff=list(prof.ell_lds_coeff) if prof.ell_lds_coeff else None,
             name=prof.name, sm_num=prof.sm_num, threads_per_sm=prof.threads_per_sm,
             peak_fp32=prof.peak_fp32, peak_fp16=prof.peak_fp16, peak_tc=prof.peak_tc,
             ell_coeff={str(k): list(v) for k, v in prof.ell_coeff.items()},
             ell_atomic_coeff={str(k): list(v) for k, v in prof.ell_atomic_coeff.items()},
             sddmm_1d_coeff={str(k): list(v) for k, v in prof.sddmm_1d_coeff.items()},
             tc_cost=enc(prof.tc_cost), tc_cost_atomic=enc(prof.tc_cost_atomic),
             ell_cost=enc(prof.ell_cost), ell_cost_atomic=enc(prof.ell_cost_atomic),
             tc_sddmm_cost=enc(prof.tc_sddmm_cost))
    with open(path, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)


def get_profile(name=None):
    name = name or _ACTIVE[0]
    if name not in _PROFILES and name == "mi355x" and os.path.exists(PROFILE_PATH):
        load_profile_file(PROFILE_PATH)
    return _PROFILES[name]


def set_profile(name):
    get_profile(name)
    _ACTIVE[0] = name


def _peak(dsize, prof):
    return prof.peak_fp32 if dsize == 32 else prof.peak_fp16


# ---------------------------------------------------------------------------
# ELL tiles.
# ---------------------------------------------------------------------------

def compute_cost_tb_impl_given_tile_sizes(template_str, tile_sizes):
    if template_str == "sparse_template_ell":
        return math.prod([math.prod(i[1:]) for i in tile_sizes]) * 2
    raise AssertionError("Only support ELL tiles.")


def _tile_unique_cols(op, I, pos_is):
    """Number of distinct columns among the op's initial entries in tile rows
    [I*pos, I*(pos+1)), and whether col 0 is among them."""
    ip = op.position_space[0].indptr
    n = len(ip) - 1
    starts = ip[np.minimum(I * pos_is, n)]
    ends = ip[np.minimum(I * (pos_is + 1), n)]
    cnt = ends - starts
    cols = op.position_space[0].indices
    tid = np.repeat(np.arange(len(pos_is), dtype=np.int64), cnt)
    offs = np.arange(len(tid), dtype=np.int64) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    c = cols[np.repeat(starts.astype(np.int64), cnt) + offs].astype(np.int64)
    K = op.idx_lens[2] + 1
    key = np.unique(tid * K + c)
    uniq = np.bincount(key // K, minlength=len(pos_is))
    has0 = np.zeros(len(pos_is), bool)
    has0[(key[key % K == 0] // K)] = True
    return cnt, uniq, has0


def _ell_mem_units(I, J, K, cnt, uniq, has0, dsize):
    """[non-atomic, atomic] memory cost in 32-byte units of ELL tiles with
    cnt nonzeros on uniq distinct columns (has0: column 0 among them; a tile
    with padding reads column 0 for its padded slots)."""
    cache_line_size = 32 * 8 / dsize
    memory_cost_A = math.ceil(I * K / cache_line_size)
    pads = np.asarray(cnt) < I * K
    k_is = np.asarray(uniq) + (pads & ~np.asarray(has0, bool))
    memory_cost_B = math.ceil(J / cache_line_size) * np.asarray(k_is)
    store_cost = math.ceil(J * dsize / (32 * 8)) * I
    return [memory_cost_A + store_cost + memory_cost_B, memory_cost_A + store_cost * 2 + memory_cost_B]


def ell_tile_x(I, J, K, cnt, uniq, has0, dsize, peak):
    """The latency feature of non-atomic ELL tiles: FLOPs over the roofline
    throughput min(FLOPs / bytes, peak) (cost_tb_latency_given_ELL_tiles)."""
    compute_cost = I * J * K * 2
    mem = _ell_mem_units(I, J, K, cnt, uniq, has0, dsize)[0]
    return compute_cost / np.minimum(compute_cost / (mem * 32), peak)


def memory_cost_tb_given_tile_sizes_and_poses(tile_sizes, tile_pos_is, template_str, dsize, op):
    """Memory cost in 32-byte units of ELL tiles at positions tile_pos_is:
    [non-atomic, atomic] (reference :517-582)."""
    if template_str != "sparse_template_ell":
        raise AssertionError("We only support ELL tiles now.")
    I, J, K = tile_sizes[0][1], tile_sizes[1][1], tile_sizes[2][1]
    pos_is = np.asarray(tile_pos_is, dtype=np.int64)
    cnt, uniq, has0 = _tile_unique_cols(op, I, pos_is)
    return _ell_mem_units(I, J, K, cnt, uniq, has0, dsize)


def ELL_non_atomic_coeff(dsize, profile=None):
    p = get_profile(profile)
    return list(p.ell_coeff[16 if dsize == 16 else 32])


def ELL_atomic_extra_coeff(dsize, profile=None):
    p = get_profile(profile)
    return list(p.ell_atomic_coeff[16 if dsize == 16 else 32])


def ELL_row_coeff(dsize, profile=None):
    p = get_profile(profile)
    return p.ell_row_coeff[16 if dsize == 16 else 32] if p.ell_row_coeff else 0.0


def ELL_cost_model(x_nonatomic, x_atomic, is_atomic, dsize, profile=None, rows=0):
    """a1 x + b1 (+ c1 rows for a profile with a per-row term), times the
    atomic multiplier for atomic tiles."""
    a1, b1 = ELL_non_atomic_coeff(dsize, profile)
    a2, b2 = ELL_atomic_extra_coeff(dsize, profile)
    base = a1 * x_nonatomic + b1 + ELL_row_coeff(dsize, profile) * rows
    if is_atomic:
        return base * (a2 * x_atomic / x_nonatomic + b2)
    return base


def _b_stationary(prof, dsize, op):
    """Whether ELL entries of this op land in dense windows' sparse parts of
    the B-stationary kernel (BF16 blocks, B of at most LDS_MAX_K rows) and
    the profile prices those."""
    return (prof.ell_lds_coeff is not None and dsize == 16 and op is not None
            and getattr(op, "op_type", "spmm") == "spmm" and op.idx_lens[2] <= LDS_MAX_K)


def lds_ell_cost(cnt, I, prof):
    """Cost of ELL tiles of I rows holding cnt entries when they become
    sparse parts of B-stationary windows.  The kernel adds a window's sparse
    part in rounds, one entry of each of its 16 rows per round, at the cost
    of 16 entries whatever the round holds; the rows of an ELL tile lie in
    different windows, so each entry is priced as a round of its own (16 x
    the per-entry cost measured with full rounds), plus, per 16 rows, the
    window's sparse-part overhead."""
    a, b = prof.ell_lds_coeff
    return 16.0 * a * np.asarray(cnt, np.float64) + b * I / 16.0


def cost_tb_latency_given_ELL_tiles(op, tile_sizes, tile_pos_is, dsize, ori_tot_nnz, SM_num, is_atomic,
                                    profile=None):
    """Predicted latency of ELL tiles at positions tile_pos_is (reference :1176-1214)."""
    prof = get_profile(profile)
    I = tile_sizes[0][1]
    if _b_stationary(prof, dsize, op):
        cnt, _, _ = _tile_unique_cols(op, I, np.asarray(tile_pos_is, dtype=np.int64))
        return lds_ell_cost(np.minimum(cnt, I * tile_sizes[2][1]), I, prof)
    compute_cost = compute_cost_tb_impl_given_tile_sizes("sparse_template_ell", tile_sizes)
    mem = memory_cost_tb_given_tile_sizes_and_poses(tile_sizes, tile_pos_is, "sparse_template_ell", dsize, op)
    if not is_atomic:
        mem = (mem[0], np.array([1]))
    peak = _peak(dsize, prof)
    tp_non_atomic, tp_atomic = [np.minimum(compute_cost / (m * 32), peak) for m in mem]
    if is_atomic:
        return ELL_cost_model(compute_cost / tp_non_atomic, compute_cost / tp_atomic, True, dsize, profile, rows=I)
    return ELL_cost_model(compute_cost / tp_non_atomic, None, False, dsize, profile, rows=I)


# ---------------------------------------------------------------------------
# 1D SDDMM tiles.
# ---------------------------------------------------------------------------

def sddmm_1d_x(row_nums, col_nums, K, bucket, dsize, peak):
    """The latency feature of 1D SDDMM buckets: FLOPs over the roofline
    throughput, bytes = (distinct rows + distinct cols) * K + bucket values."""
    compute_cost = K * bucket * 2
    memory_costs = ((np.asarray(row_nums) + np.asarray(col_nums)) * K + bucket) * dsize / 8
    return compute_cost / np.minimum(compute_cost / memory_costs, peak)


def _cost_tb_latency_1D_tiles_given_features(row_nums, col_nums, K, tile_sizes, dsize, profile=None):
    prof = get_profile(profile)
    costs = sddmm_1d_x(row_nums, col_nums, K, tile_sizes[0], dsize, _peak(dsize, prof))
    a, b = prof.sddmm_1d_coeff[16 if dsize == 16 else 32]
    return a * costs + b


def cost_tb_latency_given_1D_tiles(op, tile_sizes, dsize, profile=None):
    """Latency of every 1D bucket of the op (reference :1255-1282): features are
    the numbers of distinct rows and columns in the bucket (initial space)."""
    mbs = tile_sizes[0]
    n = len(op._ent_id)
    pad_num = math.ceil(n / mbs) * mbs - n
    if len(op.row_nums_1d) == 0 or len(op.col_nums_1d) == 0:
        rows = np.asarray(op.idx_values_list[0][0])[op._ent_row]
        cols = op._ent_col
        if n:
            rows = np.concatenate([rows, np.full(pad_num, rows[-1])])
            cols = np.concatenate([cols, np.full(pad_num, cols[-1])])
        nb = len(rows) // mbs
        bid = np.repeat(np.arange(nb, dtype=np.int64), mbs)
        big = max(op.idx_lens[0], op.idx_lens[2]) + 1
        op.row_nums_1d = np.bincount(np.unique(bid * big + rows) // big, minlength=nb)
        op.col_nums_1d = np.bincount(np.unique(bid * big + cols) // big, minlength=nb)
    return _cost_tb_latency_1D_tiles_given_features(op.row_nums_1d, op.col_nums_1d, op.idx_lens[1], tile_sizes,
                                                    dsize, profile)


# ---------------------------------------------------------------------------
# Dense (TC) tiles: measured tables.
# ---------------------------------------------------------------------------

def get_benchmark_cost_(template_str, tile_sizes, dsize, is_atomic, penalty, profile=None, b_rows=None):
    """Table latency of one TC / ELL tile (reference :639-768).  The reference
    only has FP16 tables; a measured profile may carry an FP32 table (dsize
    32) and a B-stationary table (dsize 16 with b_rows, the rows of B, at
    most LDS_MAX_K)."""
    prof = get_profile(profile)
    key = (tile_sizes[0][1], tile_sizes[1][1], tile_sizes[2][1])
    if template_str == 'TensorCore_template':
        if dsize == 16 and prof.tc_cost_lds and b_rows is not None and b_rows <= LDS_MAX_K:
            return prof.tc_cost_lds[key] / penalty
        if dsize == 32 and prof.tc_cost_f32:
            return prof.tc_cost_f32[key] / penalty
        tab = prof.tc_cost_atomic if is_atomic else prof.tc_cost
        return tab[key] / penalty
    elif template_str == 'sparse_template_ell':
        tab = prof.ell_cost_atomic if is_atomic else prof.ell_cost
        return tab[key]
    elif template_str == 'TC_sddmm':
        return prof.tc_sddmm_cost[key]
    raise AssertionError("we only support TC and ELL tiles.")


def get_benchmark_cost(tile, dsize, is_atomic, penalty, profile=None):
    return get_benchmark_cost_(get_template_str(tile.op), tile.tile_sizes, dsize, is_atomic, penalty, profile,
                               b_rows=tile.op.idx_lens[2])


def cost_tb_latency_given_tile(tile, dsize, ori_tot_nnz, SM_num, is_atomic, penalty, profile=None):
    if tile.nnz == 0:
        return float('inf')
    if get_template_str(tile.op) in ['TensorCore_template', 'TC_sddmm']:
        return get_benchmark_cost(tile, dsize, is_atomic, penalty, profile)
    pos = np.asarray([tile.tile_pos[0]])
    return float(cost_tb_latency_given_ELL_tiles(tile.op, tile.tile_sizes, pos, dsize, ori_tot_nnz, SM_num,
                                                 is_atomic, profile)[0])


# ---------------------------------------------------------------------------
# Lower bounds.
# ---------------------------------------------------------------------------

def lower_bound_ELL_cost_BSPMM(sub_op, max_bucket_size, dsize, profile=None):
    """Best possible average cost of a non-atomic ELL tile (reference :820-860)."""
    prof = get_profile(profile)
    idx = get_tile_sizes_idx(sub_op, max_bucket_size, is_tb_tile=True)
    idx[1] = [(None, 32)]
    lst = get_tile_sizes_list(sub_op, idx, max_bucket_size, is_tb_tile=True)
    cacheline_bit = 32 * 8
    best = float('inf')
    for tile_sizes in lst:
        I, J, K = [i[1] for i in tile_sizes]
        compute_cost = I * J * K * 2
        memory_cost = math.ceil(I * K * dsize / cacheline_bit) + math.ceil(J * dsize / cacheline_bit) * (K // 2 + 1) \
            + I * math.ceil(J * dsize / cacheline_bit)
        if _b_stationary(prof, dsize, sub_op):
            cost = float(lds_ell_cost(I * K, I, prof))
        else:
            tp = min(compute_cost / (memory_cost * 32), _peak(dsize, prof))
            cost = ELL_cost_model(compute_cost / tp, None, False, dsize, profile, rows=I)
        best = min(best, cost / (I * K * J))
    return best


def lower_bound_1D_cost_BSDDMM(sub_op, max_bucket_size, dsize, profile=None):
    row_nums = np.concatenate([np.arange(1, nnz + 1) for nnz in range(1, max_bucket_size + 1)])
    nnzs = np.repeat(np.arange(1, max_bucket_size + 1), np.arange(1, max_bucket_size + 1))
    col_nums = nnzs / row_nums
    avg = _cost_tb_latency_1D_tiles_given_features(row_nums, col_nums, sub_op.idx_lens[1], (max_bucket_size,),
                                                   dsize, profile) / nnzs / sub_op.idx_lens[1]
    assert min(avg) > 0, "Negative lower bound."
    return float(min(avg))


def lower_bound_ELL_cost(sub_op, max_bucket_size, dsize, profile=None):
    if sub_op.op_type == 'spmm':
        return lower_bound_ELL_cost_BSPMM(sub_op, max_bucket_size, dsize, profile)
    elif sub_op.op_type == 'sddmm':
        return lower_bound_1D_cost_BSDDMM(sub_op, max_bucket_size, dsize, profile)

"""Search primitives: TC tiles by row window (with whole-row "withdraw"),
best ELL / 1D tile under the cost model, local search, next-level ops and the
withdraw post-processing (MI355X build).

Same names and decisions as MY_sparse/my_branch_and_bound.py:
  get_valid_tile_pos_list_for_ELL (:16-53), search_TC_tiles_only_by_rows
  (:238-389), search_best_TC_tile_given_csr (:403-542), search_best_TC_tile
  (:702-720) / _given_op (:558-696), search_best_tile_given_op_BSPMM
  (:728-835) / _BSDDMM (:841-878), search_best_tile (:931-977),
  gen_full_next_level_subops (:1051-1094), local_search_TC_tiles_given_csr
  (:1117-1236), local_search_TC_tiles (:1237-1405), local_search_ELL
  (:1416-1486), local_search_1D_BSDDMM (:1488-1532), local_search
  (:1537-1640), post_process_for_withdraw(_BSPMM/_BSDDMM) (:1652-1759).

The reference evaluates these on scipy 0/1 position spaces; here per-window
16-column group counts and per-piece counts are computed from the shared
coverage mask, vectorised across windows and positions.  Ties are broken as
in the reference (first index of argmax / nanargmin, strict '<' updates).
"""
import math
from fractions import Fraction

import numpy as np

from . import my_cost_model
from .gen_formats_v2 import (ComputeTile, Op, _as_csr, gen_candidate_formats, gen_position_space_for_area,
                             get_config, get_params_list, get_template_str, get_tile_sizes_idx,
                             get_tile_sizes_list, ell_tile_shapes, refresh_1D_counts)

_INF = float('inf')


# ---------------------------------------------------------------------------
# TC op helpers: per-entry window / condensed position; per-window group counts.
# ---------------------------------------------------------------------------

def _tc_prepare(op):
    if getattr(op, "_tc_ready", False):
        return
    gen_position_space_for_area(op, 0)
    W = math.ceil(len(op.idx_values_list[0][0]) / 16)
    win = op._ent_row // 16
    cpos = np.empty(len(win), np.int64)
    order = np.argsort(win, kind='stable')
    bounds = np.searchsorted(win[order], np.arange(W + 1))
    for w in range(W):
        e = order[bounds[w]:bounds[w + 1]]
        if len(e):
            cpos[e] = op.k_vals[w].pos(op._ent_col[e])
    op._ent_win = win
    op._ent_cpos = cpos
    op._W = W
    # groups of 16 condensed columns that can hold entries, per window
    gmax = np.zeros(W, np.int64)
    if len(win):
        np.maximum.at(gmax, win, cpos // 16 + 1)
    op._G = gmax
    op._gptr = np.concatenate([[0], np.cumsum(gmax)])
    op._tc_ready = True


def _group_counts(op, mask=None):
    """Flat per-window 16-column group counts of uncovered entries (or of the
    entries selected by mask)."""
    _tc_prepare(op)
    m = op.uncovered_mask() if mask is None else mask
    idx = op._gptr[op._ent_win[m]] + op._ent_cpos[m] // 16
    return np.bincount(idx, minlength=int(op._gptr[-1])).astype(np.int64)


def _block_sums(op, gc, starts_g, span):
    """For every window w and block b >= 0: sum of gc over groups
    [starts_g[w] + b*span, starts_g[w] + (b+1)*span) (clipped to the window).
    Returns (win_of_block, block_index, sums) in row-major (w, b) order."""
    G = op._G
    nblk = np.where(G > starts_g, -(-(G - starts_g) // span), 0)
    tot = int(nblk.sum())
    if tot == 0:
        z = np.zeros(0, np.int64)
        return z, z, z
    wb = np.repeat(np.arange(len(G), dtype=np.int64), nblk)
    bi = np.arange(tot, dtype=np.int64) - np.repeat(np.cumsum(nblk) - nblk, nblk)
    cs = np.concatenate([[0], np.cumsum(gc)])
    lo = op._gptr[wb] + starts_g[wb] + bi * span
    hi = np.minimum(lo + span, op._gptr[wb + 1])
    return wb, bi, cs[hi] - cs[lo]


def _first_nonempty_group(op, gc):
    """First group with uncovered entries per window (0 for empty windows)."""
    W = op._W
    res = np.zeros(W, np.int64)
    nz = np.flatnonzero(gc)
    if len(nz):
        wins = np.searchsorted(op._gptr, nz, side='right') - 1
        first = np.full(W, -1, np.int64)
        # first nonzero group per window: nz is sorted, take the first per window
        uw, idx = np.unique(wins, return_index=True)
        first[uw] = nz[idx] - op._gptr[uw]
        res = np.maximum(first, 0)
    return res


def _tc_tile_sizes(op, max_bucket_size=None):
    idx = get_tile_sizes_idx(op, None, is_tb_tile=True)
    return get_tile_sizes_list(op, idx, max_bucket_size, is_tb_tile=True)


def _tc_cost(op, tile_sizes, dsize, is_atomic, penalty):
    tpl = 'TensorCore_template' if op.op_type == 'spmm' else 'TC_sddmm'
    return my_cost_model.get_benchmark_cost_(tpl, tile_sizes, dsize, is_atomic if op.op_type == 'spmm' else False,
                                             penalty, b_rows=op.idx_lens[2])


def _make_tc_tile(op, tile_sizes, pos, is_atomic, pred_cost):
    params = {"mma_shape_str": "m16n16k16"}
    t = ComputeTile(op, tile_sizes, pos, params, not_init_nnz=True)
    t.is_atomic_tile = is_atomic
    t.pred_cost = pred_cost
    t.cost = pred_cost
    t.best_tile_sizes = t.tile_sizes
    t.best_params = t.params
    return t


def _cover(tile, entries=None, full=False):
    """Mark the tile's entries covered; record what it covered."""
    op = tile.op
    e = tile.entries() if entries is None else entries
    if not full:
        e = e[op.state.uncovered[op._ent_id[e]]]
    ids = op._ent_id[e]
    tile.covered_ids = ids
    tile.covered_rows = op._ent_row[e]
    tile.covered_cols = op._ent_col[e]
    op.state.uncovered[ids] = False
    if tile.nnz == 0:  # made without an initial count (_make_tc_tile): what it covers
        tile.nnz = len(ids)
    tile.nnz_when_selected = tile.j_num * len(ids)
    tile.nnz_uncovered = 0
    tile.set_avg_cost()
    tile.set_pred_avg_cost()
    return tile


# ---------------------------------------------------------------------------
# TC pre-pass (per row window).
# ---------------------------------------------------------------------------

def search_best_TC_tile_given_csr(sub_op, w, gc_w, dsize, tile_sizes_list, best_possible_ELL,
                                  selected_tiles_this_row, penalty):
    """Best TC choice for window w given its uncovered group counts gc_w
    (reference :403-542): one atomic tile at the densest position (16-column
    granularity) or the whole row of non-atomic tiles, whose extra cost
    is credited with the tiles already chosen in this row."""
    cfg = get_config()
    best_key, best_cost = None, _INF
    unit = {}
    tot = int(gc_w.sum())
    base_cs = np.concatenate([[0], np.cumsum(gc_w)])
    G = len(gc_w)
    J = tile_sizes_list[0][1][1]
    for tile_sizes in tile_sizes_list:
        span = tile_sizes[2][1] // 16
        if cfg.single_level:
            starts = np.arange(0, G, span)
            per = base_cs[np.minimum(starts + span, G)] - base_cs[starts]
            mp = int(np.argmax(per))
            pos = (w, 0, mp)
        else:
            hi = np.minimum(np.arange(G) + span, G)
            per = base_cs[hi] - base_cs[:G]
            mp = int(np.argmax(per))
            pos = (w, 0, mp * Fraction(16, tile_sizes[2][1]))
        nmax = per[mp]
        if sub_op.op_type == 'spmm':
            cost = _tc_cost(sub_op, tile_sizes, dsize, True, penalty)
            pred = cost / nmax / tile_sizes[1][1] if nmax > 0 else _INF
        else:
            cost = _tc_cost(sub_op, tile_sizes, dsize, False, penalty)
            pred = cost * math.ceil(sub_op.idx_lens[1] / tile_sizes[1][1]) / nmax / sub_op.idx_lens[1] \
                if nmax > 0 else _INF
        unit[tile_sizes] = [None, cost]
        if pred < best_cost and pred < best_possible_ELL:
            best_cost, best_key = pred, (sub_op.op_id, tile_sizes, [pos], True)
        if sub_op.op_type == 'sddmm':
            continue
        nb = int(sub_op.blk_nums[tile_sizes][w])
        na_cost = _tc_cost(sub_op, tile_sizes, dsize, False, penalty)
        unit[tile_sizes][0] = na_cost
        extra = na_cost * nb - (0 if cfg.no_withdraw else sum(t.pred_cost for t in selected_tiles_this_row))
        pred = extra / tot / tile_sizes[1][1] if tot > 0 else _INF
        if pred < best_cost and pred < best_possible_ELL:
            best_cost = pred
            poses = [(w, 0, i) for i in range(nb)]
            if sub_op.TC_k_notsorted:
                poses = [(w, 0, int(i)) for i in sub_op.blk_ks[tile_sizes][w]]
            best_key = (sub_op.op_id, tile_sizes, poses, False)
    return best_key, best_cost, unit


def local_search_TC_tiles_given_csr(sub_op, w, gc_w, tile_sizes, target_pred_avg_cost, dsize,
                                    no_atomic_tile_cost, atomic_tile_cost, selected_tiles_this_row):
    """Tiles of window w within target_pred_avg_cost (reference :1117-1236):
    the whole row if it qualifies, else every K-block (from the first
    nonempty 16-column group) that does."""
    cfg = get_config()
    K = tile_sizes[2][1]
    span = K // 16
    nz = np.flatnonzero(gc_w)
    neb = 0 if (cfg.single_level or len(nz) == 0) else int(nz[0])
    G = len(gc_w)
    starts = np.arange(neb, G, span)
    cs = np.concatenate([[0], np.cumsum(gc_w)])
    nnzs_update = cs[np.minimum(starts + span, G)] - cs[starts] if len(starts) else np.zeros(0, np.int64)
    if sub_op.op_type == 'spmm':
        nb = int(sub_op.blk_nums[tile_sizes][w])
        extra = no_atomic_tile_cost * nb - (0 if cfg.no_withdraw else
                                            sum(t.pred_cost for t in selected_tiles_this_row))
        tot = nnzs_update.sum()
        if tot > 0 and extra / tot / tile_sizes[1][1] <= target_pred_avg_cost:
            poses = [(w, 0, i) for i in range(nb)]
            if sub_op.TC_k_notsorted:
                poses = [(w, 0, int(i)) for i in sub_op.blk_ks[tile_sizes][w]]
            return [_make_tc_tile(sub_op, tile_sizes, p, False, no_atomic_tile_cost) for p in poses]
    kb = Fraction(16, K) * neb
    with np.errstate(divide='ignore', invalid='ignore'):
        if sub_op.op_type == 'spmm':
            ok = np.flatnonzero(atomic_tile_cost <= nnzs_update * tile_sizes[1][1] * target_pred_avg_cost)
            return [_make_tc_tile(sub_op, tile_sizes, (w, 0, int(i) + kb), True, atomic_tile_cost) for i in ok]
        f = math.ceil(sub_op.idx_lens[1] / tile_sizes[1][1])
        ok = np.flatnonzero(atomic_tile_cost * f <= nnzs_update * sub_op.idx_lens[1] * target_pred_avg_cost)
        return [_make_tc_tile(sub_op, tile_sizes, (w, 0, int(i) + kb), False, atomic_tile_cost * f) for i in ok]


def search_TC_tiles_only_by_rows(sub_op, dsize, is_tb_tile, best_possible_ELL, tot_nnz, all_TC_ops, all_sub_ops,
                                 ori_op, max_avg_cost_diff, TC_row_tile_dict, TC_row_costs, penalty,
                                 max_bucket_size, TC_is):
    """TC pre-pass, one 16-row window at a time (reference :238-389)."""
    _tc_prepare(sub_op)
    tile_sizes_list = _tc_tile_sizes(sub_op, max_bucket_size)
    selected = []
    gc = _group_counts(sub_op)
    W = sub_op._W
    row_num = len(sub_op.idx_values_list[0][0])
    # Windows whose densest possible tile cannot beat the ELL bound are skipped
    # (their loop in the reference exits on the first search).
    best_unit = min(_tc_cost(sub_op, ts, dsize, True, penalty) / (16 * ts[2][1] * ts[1][1])
                    if sub_op.op_type == 'spmm' else
                    _tc_cost(sub_op, ts, dsize, False, penalty) * math.ceil(sub_op.idx_lens[1] / ts[1][1])
                    / (16 * ts[2][1] * sub_op.idx_lens[1])
                    for ts in tile_sizes_list)
    if sub_op.op_type == 'spmm':
        best_unit = min(best_unit, min(_tc_cost(sub_op, ts, dsize, False, penalty) / (16 * ts[2][1] * ts[1][1])
                                       for ts in tile_sizes_list))
    win_tot = np.bincount(sub_op._ent_win[sub_op.uncovered_mask()], minlength=W)
    # Per-window lower bound of the first search: windows whose best possible
    # TC choice cannot beat the ELL bound end their loop at once (reference
    # :283-289), so they are skipped without building their counts.
    lb = np.full(W, _INF)
    with np.errstate(divide='ignore', invalid='ignore'):
        for ts in tile_sizes_list:
            K, J = ts[2][1], ts[1][1]
            if sub_op.op_type == 'spmm':
                ac = _tc_cost(sub_op, ts, dsize, True, penalty)
                lb = np.minimum(lb, ac / np.minimum(win_tot, 16 * K) / J)
                na = _tc_cost(sub_op, ts, dsize, False, penalty)
                lb = np.minimum(lb, na * sub_op.blk_nums[ts] / win_tot / J)
            else:
                ac = _tc_cost(sub_op, ts, dsize, False, penalty) * math.ceil(sub_op.idx_lens[1] / J)
                lb = np.minimum(lb, ac / np.minimum(win_tot, 16 * K) / sub_op.idx_lens[1])
    cand = np.flatnonzero((win_tot > 0) & (lb < best_possible_ELL))
    wbounds = np.searchsorted(sub_op._ent_win, np.arange(W + 1))
    for w in cand:
        w = int(w)
        this_row = []
        win_entries = np.arange(wbounds[w], wbounds[w + 1])
        while True:
            gc_w = _group_counts_window(sub_op, w, win_entries)
            if gc_w.sum() <= 0:
                break
            key, cost, unit = search_best_TC_tile_given_csr(sub_op, w, gc_w, dsize, tile_sizes_list,
                                                            best_possible_ELL, this_row, penalty)
            if cost == _INF:
                break
            _, tile_sizes, poses, is_atomic = key
            good = local_search_TC_tiles_given_csr(sub_op, w, gc_w, tile_sizes, cost * (1 + max_avg_cost_diff),
                                                   dsize, *(unit[tile_sizes]), this_row)
            if sub_op.op_type == 'spmm' and good and not good[0].is_atomic_tile:
                # whole row: its tiles replace the row's earlier tiles and take
                # every entry of the window (also those the earlier ones took)
                for t in this_row:
                    sub_op.state.uncovered[t.covered_ids] = True
                this_row = good
                for t in good:
                    _cover(t)
            else:
                for t in good:
                    _cover(t)
                this_row = this_row + good
            if not good:
                break
        start = len(selected)
        selected += this_row
        if sub_op.op_type == 'spmm' and this_row and not this_row[0].is_atomic_tile:
            TC_row_costs[w] = -1
        else:
            TC_row_tile_dict[w] = np.arange(start, len(selected))
            TC_row_costs[w] = sum(t.pred_cost for t in this_row)
        if this_row:
            TC_is[w * 16:min((w + 1) * 16, row_num)] = True
    return selected


def _group_counts_window(op, w, entries):
    m = op.uncovered_mask()[entries]
    g = op._ent_cpos[entries[m]] // 16
    return np.bincount(g, minlength=int(op._G[w])).astype(np.int64)


# ---------------------------------------------------------------------------
# TC search over all windows (main loop).
# ---------------------------------------------------------------------------

def search_best_TC_tile_given_op(sub_op, dsize, TC_row_tile_dict, TC_row_costs, penalty, max_bucket_size,
                                 is_tb_tile=False, lower_bound_dict=None, tile_dict=None):
    """Best single atomic TC tile over all windows (K-blocks from each window's
    first nonempty group) vs the best whole row (reference :558-696)."""
    cfg = get_config()
    _tc_prepare(sub_op)
    tile_sizes_list = _tc_tile_sizes(sub_op, max_bucket_size)
    gc = _group_counts(sub_op)
    neb = np.zeros(sub_op._W, np.int64) if cfg.single_level else _first_nonempty_group(sub_op, gc)
    win_tot = np.bincount(sub_op._ent_win[sub_op.uncovered_mask()], minlength=sub_op._W)
    best_key, best_cost = None, _INF
    for tile_sizes in tile_sizes_list:
        span = tile_sizes[2][1] // 16
        wb, bi, sums = _block_sums(sub_op, gc, neb, span)
        if len(sums) and sums.max() > 0:
            i = int(np.argmax(sums))
            nmax = sums[i]
            w, b = int(wb[i]), int(bi[i])
            if sub_op.op_type == 'spmm':
                cost = _tc_cost(sub_op, tile_sizes, dsize, True, penalty)
                pred = cost / nmax / tile_sizes[1][1]
                is_atomic = True
            else:
                cost = _tc_cost(sub_op, tile_sizes, dsize, False, penalty)
                pred = cost * math.ceil(sub_op.idx_lens[1] / tile_sizes[1][1]) / nmax / sub_op.idx_lens[1]
                is_atomic = False
            pos = (w, 0, b + int(neb[w]) * Fraction(16, tile_sizes[2][1]))
            if pred < best_cost:
                best_cost, best_key = pred, (sub_op.op_id, tile_sizes, [pos], is_atomic)
        if sub_op.op_type == 'sddmm':
            continue
        nb = sub_op.blk_nums[tile_sizes]
        na = _tc_cost(sub_op, tile_sizes, dsize, False, penalty)
        extra = na * nb - (0 if cfg.no_withdraw else TC_row_costs)
        with np.errstate(divide='ignore', invalid='ignore'):
            pred = extra / win_tot / tile_sizes[1][1]
        pred = np.where(win_tot > 0, pred, np.nan)
        if np.all(np.isnan(pred)):
            continue
        r = int(np.nanargmin(pred))
        if pred[r] < best_cost:
            best_cost = float(pred[r])
            poses = [(r, 0, i) for i in range(int(nb[r]))]
            if sub_op.TC_k_notsorted:
                poses = [(r, 0, int(i)) for i in sub_op.blk_ks[tile_sizes][r]]
            best_key = (sub_op.op_id, tile_sizes, poses, False)
    return best_key, best_cost


def search_best_TC_tile(sub_ops, dsize, TC_row_tile_dict, TC_row_costs, penalty, max_bucket_size):
    best_key, best_cost = None, _INF
    for sub_op in sub_ops:
        key, c = search_best_TC_tile_given_op(sub_op, dsize, TC_row_tile_dict, TC_row_costs, penalty,
                                              max_bucket_size)
        if c < best_cost:
            best_key, best_cost = key, c
    return best_key, best_cost


def local_search_TC_tiles(selected_tile, target_pred_avg_cost, dsize, TC_row_tile_dict, TC_row_costs,
                          cur_selected_tile_num, penalty, max_bucket_size, TC_is):
    """Whole rows within target (best tile size per row) plus atomic K-blocks
    of the selected tile's size in the other rows (reference :1237-1405)."""
    cfg = get_config()
    sub_op = selected_tile.op
    tile_sizes = selected_tile.tile_sizes
    _tc_prepare(sub_op)
    W = sub_op._W
    row_num = len(sub_op.idx_values_list[0][0])
    gc = _group_counts(sub_op)
    neb = np.zeros(W, np.int64) if cfg.single_level else _first_nonempty_group(sub_op, gc)
    span = tile_sizes[2][1] // 16
    wb, bi, sums = _block_sums(sub_op, gc, neb, span)
    win_tot = np.bincount(wb, weights=sums, minlength=W)
    good = []
    selected_rows = np.zeros(0, np.int64)
    if sub_op.op_type == 'spmm':
        lst = _tc_tile_sizes(sub_op, max_bucket_size)
        nbs = np.asarray([sub_op.blk_nums[t] for t in lst])
        na = np.asarray([_tc_cost(sub_op, t, dsize, False, penalty) for t in lst])
        extra = na.reshape(-1, 1) * nbs - (0 if cfg.no_withdraw else TC_row_costs)
        cand = np.nanargmin(extra, axis=0)
        with np.errstate(divide='ignore', invalid='ignore'):
            crit = np.nanmin(extra, axis=0) / win_tot / selected_tile.j_num
        selected_rows = np.flatnonzero((win_tot > 0) & (crit <= target_pred_avg_cost))
        TC_row_costs[selected_rows] = -1
        for r in selected_rows:
            ts = lst[cand[r]]
            poses = [(int(r), 0, i) for i in range(int(nbs[cand[r]][r]))]
            if sub_op.TC_k_notsorted:
                poses = [(int(r), 0, int(i)) for i in sub_op.blk_ks[ts][r]]
            for p in poses:
                good.append(_make_tc_tile(sub_op, ts, p, False, float(na[cand[r]])))
            TC_is[r * 16:min((r + 1) * 16, row_num)] = True
    is_atomic = sub_op.op_type == 'spmm'
    atomic_cost = _tc_cost(sub_op, tile_sizes, dsize, is_atomic, penalty)
    with np.errstate(divide='ignore', invalid='ignore'):
        if sub_op.op_type == 'spmm':
            ok = (sums > 0) & (atomic_cost / sums / selected_tile.j_num <= target_pred_avg_cost)
            pc = atomic_cost
        else:
            f = math.ceil(sub_op.idx_lens[1] / tile_sizes[1][1])
            ok = (sums > 0) & (atomic_cost * f / sums / sub_op.idx_lens[1] <= target_pred_avg_cost)
            pc = atomic_cost * f
    ok &= ~np.isin(wb, selected_rows)
    idx = np.flatnonzero(ok)
    start = cur_selected_tile_num + len(good)
    for r in np.unique(wb[idx]):
        n = int(np.count_nonzero(wb[idx] == r))
        TC_row_tile_dict[int(r)] = np.concatenate([TC_row_tile_dict[int(r)], np.arange(start, start + n)])
        TC_row_costs[r] = TC_row_costs[r] + atomic_cost * n
        start += n
        TC_is[r * 16:(r + 1) * 16] = True
    for i in idx:
        w = int(wb[i])
        pos = (w, 0, int(bi[i]) + int(neb[w]) * Fraction(16, tile_sizes[2][1]))
        good.append(_make_tc_tile(sub_op, tile_sizes, pos, is_atomic, pc))
    return good


# ---------------------------------------------------------------------------
# ELL / 1D search.
# ---------------------------------------------------------------------------

def get_valid_tile_pos_list_for_ELL(sub_op, tile_sizes, max_bucket_size):
    """Positions of (I, J, K) ELL tiles worth considering: rows with
    nnz in (K/2, K], plus the split-row region when K == max_bucket_size
    (reference :16-53)."""
    I = math.prod(tile_sizes[0][1:])
    K = math.prod(tile_sizes[2][1:])
    targets = [K, K // 2]
    t_is = [int(np.searchsorted(sub_op.hyb_getnnz_minus, -t)) + sub_op.hyb_short_begin_row for t in targets]
    list1 = list(range(math.ceil(t_is[0] / I), (t_is[1] - 1) // I + 1))
    if len(list1) > 0 and (list1[-1] + 1) * I < len(sub_op.hyb_new_rows[0]):
        list1 = list1 + [list1[-1] + 1]
    if K == max_bucket_size:
        list2 = list(range((sub_op.hyb_short_begin_row - 1) // I + 1))
        return [(i, 0, 0) for i in list2 + list1]
    return [(i, 0, 0) for i in list1]


def _ell_piece_counts(sub_op):
    return sub_op.row_uncovered_counts()


def search_best_tile_given_op_BSPMM(ori_op, sub_op, dsize, ori_tot_nnz, SM_num, max_bucket_size,
                                    best_tile_key=None, best_pred_avg_cost=_INF, second_best_pred_avg_cost=_INF,
                                    is_tb_tile=False, lower_bound_dict=None, tile_dict=None,
                                    nnzs_dict=None, pos_is_dict=None, costs_dict=None):
    """Cheapest ELL tile (average cost per uncovered point) over every valid
    position and shape of sub_op (reference :728-835)."""
    nnzs_dict = {} if nnzs_dict is None else nnzs_dict
    pos_is_dict = {} if pos_is_dict is None else pos_is_dict
    costs_dict = {} if costs_dict is None else costs_dict
    u = _ell_piece_counts(sub_op)
    npieces = len(sub_op.hyb_new_rows[0])
    sb = sub_op.hyb_short_begin_row
    ori_rownnz = np.diff(ori_op.inps[0].indptr)
    piece_orig_row = np.asarray(sub_op.idx_values_list[0][0])[sub_op.hyb_new_rows[0]]
    if sub_op.ori_row_map is not None:
        piece_orig_row = sub_op.ori_row_map[piece_orig_row]
    tail = np.arange(sb, npieces)
    atomic_rows = tail[(ori_rownnz[piece_orig_row[sb:]] > u[sb:]) & (u[sb:] > 0)]
    cs = np.concatenate([[0], np.cumsum(u)])
    for tile_sizes in ell_tile_shapes(sub_op, max_bucket_size):
        I = tile_sizes[0][1]
        pos_list = get_valid_tile_pos_list_for_ELL(sub_op, tile_sizes, max_bucket_size)
        pos_is = np.asarray([p[0] for p in pos_list], dtype=np.int64)
        atomic_cand = np.concatenate([np.arange((sb - 1) // I + 1), atomic_rows // I])
        a_pos = np.intersect1d(pos_is, atomic_cand)
        n_pos = np.setdiff1d(pos_is, a_pos, assume_unique=False)
        res = []
        for pos in (a_pos, n_pos):
            if len(pos):
                nn = cs[np.minimum(I * (pos + 1), npieces)] - cs[np.minimum(I * pos, npieces)]
                keep = nn > 0
                res.append((pos[keep], nn[keep] * tile_sizes[1][1]))
            else:
                res.append((pos, np.zeros(0, np.int64)))
        (a_pos, a_nn), (n_pos, n_nn) = res
        nnzs_dict[(sub_op.op_id, tile_sizes)] = a_nn, n_nn
        pos_is_dict[(sub_op.op_id, tile_sizes)] = a_pos, n_pos
        costs_dict[(sub_op.op_id, tile_sizes)] = []
        for pos, is_atomic, nn in ((a_pos, True, a_nn), (n_pos, False, n_nn)):
            if len(pos) == 0:
                costs_dict[(sub_op.op_id, tile_sizes)].append(np.array([]))
                continue
            costs = my_cost_model.cost_tb_latency_given_ELL_tiles(sub_op, tile_sizes, pos, dsize, ori_tot_nnz,
                                                                  SM_num, is_atomic)
            costs_dict[(sub_op.op_id, tile_sizes)].append(costs)
            avg = costs / nn
            i = int(np.nanargmin(avg))
            if avg[i] < best_pred_avg_cost:
                best_pred_avg_cost = float(avg[i])
                best_tile_key = (sub_op, tile_sizes, (int(pos[i]), 0, 0), is_atomic, int(nn[i]))
    return best_tile_key, best_pred_avg_cost, second_best_pred_avg_cost


def search_best_tile_given_op_BSDDMM(ori_op, sub_op, dsize, ori_tot_nnz, SM_num, max_bucket_size,
                                     best_tile_key=None, best_pred_avg_cost=_INF, second_best_pred_avg_cost=_INF,
                                     is_tb_tile=False, lower_bound_dict=None, tile_dict=None,
                                     nnzs_dict=None, pos_is_dict=None, costs_dict=None):
    """Cheapest 1D bucket (reference :841-878)."""
    costs_dict = {} if costs_dict is None else costs_dict
    refresh_1D_counts(sub_op)
    costs = my_cost_model.cost_tb_latency_given_1D_tiles(sub_op, (max_bucket_size,), dsize)
    nnzs = sub_op.nnz_update_1D
    with np.errstate(divide='ignore', invalid='ignore'):
        avg = costs / nnzs / sub_op.idx_lens[1]
    avg = np.where(nnzs > 0, avg, _INF)
    if len(avg):
        i = int(np.argmin(avg))
        if avg[i] < best_pred_avg_cost:
            best_pred_avg_cost = float(avg[i])
            best_tile_key = (sub_op, (max_bucket_size,), (i,), False, int(nnzs[i]) * sub_op.idx_lens[1])
    costs_dict[(sub_op.op_id, (max_bucket_size,))] = avg
    return best_tile_key, best_pred_avg_cost, second_best_pred_avg_cost


def search_best_tile_given_op(ori_op, sub_op, dsize, ori_tot_nnz, SM_num, max_bucket_size, **kw):
    if ori_op.op_type == 'spmm':
        return search_best_tile_given_op_BSPMM(ori_op, sub_op, dsize, ori_tot_nnz, SM_num, max_bucket_size, **kw)
    return search_best_tile_given_op_BSDDMM(ori_op, sub_op, dsize, ori_tot_nnz, SM_num, max_bucket_size, **kw)


def search_best_tile(ori_op, sub_ops, tile_dict, dsize, ori_tot_nnz, SM_num, max_bucket_size,
                     nnzs_dict, pos_is_dict, costs_dict, best_pred_avg_cost=_INF):
    """Best ELL/1D tile over sub_ops that beats best_pred_avg_cost (reference :931-977)."""
    best_key, second = None, _INF
    for sub_op in sub_ops:
        best_key, best_pred_avg_cost, second = search_best_tile_given_op(
            ori_op, sub_op, dsize, ori_tot_nnz, SM_num, max_bucket_size,
            best_tile_key=best_key, best_pred_avg_cost=best_pred_avg_cost, second_best_pred_avg_cost=second,
            is_tb_tile=True, tile_dict=tile_dict, nnzs_dict=nnzs_dict, pos_is_dict=pos_is_dict,
            costs_dict=costs_dict)
    if best_key is not None:
        sub_op, tile_sizes, tile_pos, is_atomic, nnz_unc = best_key
        params = get_params_list(sub_op, tile_sizes, max_bucket_size, is_tb_tile=True)[0]
        tile = ComputeTile(sub_op, tile_sizes, tile_pos, params)
        tile.is_atomic_tile = is_atomic
        tile.pred_cost = nnz_unc * best_pred_avg_cost
        tile.update_nnz_uncovered()
        tile.set_pred_avg_cost()
        key = (sub_op.op_id, tile_sizes, tile_pos, is_atomic)
        tile_dict[key] = tile
        best_key = key
    return best_key, second


def local_search_ELL(target_pred_avg_cost, ori_op, selected_tile, second_best_cost, best_pred_avg_cost,
                     ori_tot_nnz, SM_num, dsize, tile_dict, is_tb_tile, max_bucket_size, nnzs_dict, pos_is_dict,
                     costs_dict, TC_row_tile_dict, TC_row_costs, cur_selected_tile_num, reverse_idx_TC_op_row,
                     penalty, TC_is, max_avg_cost_diff=_INF):
    """Every position of the selected tile's shape within target (reference :1416-1486)."""
    sub_op = selected_tile.op
    tile_sizes = selected_tile.tile_sizes
    I = tile_sizes[0][1]
    a_nn, n_nn = nnzs_dict[(sub_op.op_id, tile_sizes)]
    a_pos, n_pos = pos_is_dict[(sub_op.op_id, tile_sizes)]
    a_c, n_c = costs_dict[(sub_op.op_id, tile_sizes)]
    good = []
    npieces = len(sub_op.hyb_new_rows[0])
    params = get_params_list(sub_op, tile_sizes, max_bucket_size, is_tb_tile=is_tb_tile)[0]
    for pos_is, is_atomic, nnzs, costs in ((a_pos, True, a_nn, a_c), (n_pos, False, n_nn, n_c)):
        if len(pos_is) == 0:
            continue
        gi = (costs / nnzs) <= target_pred_avg_cost
        for cost, p in zip(costs[gi], pos_is[gi]):
            t = ComputeTile(sub_op, tile_sizes, (int(p), 0, 0), params)
            t.is_atomic_tile = is_atomic
            t.pred_cost = float(cost)
            t.update_nnz_uncovered()
            t.set_pred_avg_cost()
            t.best_tile_sizes = selected_tile.best_tile_sizes
            t.best_params = selected_tile.best_params
            t.cost = t.pred_cost
            good.append(t)
            rows = sub_op.hyb_new_rows[0][np.arange(p * I, min((p + 1) * I, npieces))]
            rows = np.asarray(sub_op.idx_values_list[0][0])[rows]
            if sub_op.ori_row_map is not None:
                rows = sub_op.ori_row_map[rows]
            blk = np.unique(reverse_idx_TC_op_row[rows] // 16)
            if len(blk) == 1:
                b = int(blk[0])
                TC_row_tile_dict[b] = np.concatenate([TC_row_tile_dict[b], [cur_selected_tile_num + len(good) - 1]])
                TC_row_costs[b] = TC_row_costs[b] + cost
    return good


def local_search_1D_BSDDMM(target_pred_avg_cost, ori_op, selected_tile, second_best_cost, best_pred_avg_cost,
                           ori_tot_nnz, SM_num, dsize, tile_dict, is_tb_tile, max_bucket_size, nnzs_dict,
                           pos_is_dict, costs_dict, TC_row_tile_dict, TC_row_costs, cur_selected_tile_num,
                           reverse_idx_TC_op_row, penalty, TC_is, max_avg_cost_diff=_INF):
    """Every 1D bucket within target (reference :1488-1532)."""
    sub_op = selected_tile.op
    tile_sizes = selected_tile.tile_sizes
    avg = costs_dict[(sub_op.op_id, tile_sizes)]
    gi = np.flatnonzero(avg <= target_pred_avg_cost)
    params = get_params_list(sub_op, tile_sizes, max_bucket_size, is_tb_tile=is_tb_tile)[0]
    good = []
    for p in gi:
        t = ComputeTile(sub_op, tile_sizes, (int(p),), params)
        t.is_atomic_tile = False
        t.update_nnz_uncovered()
        t.pred_cost = float(avg[p]) * t.nnz_uncovered
        t.set_pred_avg_cost()
        t.best_tile_sizes = selected_tile.best_tile_sizes
        t.best_params = selected_tile.best_params
        t.cost = t.pred_cost
        good.append(t)
    return good


def local_search(ori_op, selected_tile, second_best_cost, best_pred_avg_cost, ori_tot_nnz, SM_num, dsize, tile_dict,
                 is_tb_tile, max_bucket_size, nnzs_dict, pos_is_dict, costs_dict, TC_row_tile_dict, TC_row_costs,
                 cur_selected_tile_num, reverse_idx_TC_op_row, penalty, TC_is, max_avg_cost_diff=_INF):
    """Tiles of the selected tile's op and shape whose average cost is within
    (1 + max_avg_cost_diff) of the best (reference :1537-1640)."""
    target = best_pred_avg_cost * (1 + max_avg_cost_diff) if best_pred_avg_cost >= 0 \
        else best_pred_avg_cost * (1 - max_avg_cost_diff)
    tpl = get_template_str(selected_tile.op)
    if tpl in ['TensorCore_template', 'TC_sddmm']:
        return local_search_TC_tiles(selected_tile, target, dsize, TC_row_tile_dict, TC_row_costs,
                                     cur_selected_tile_num, penalty, max_bucket_size, TC_is)
    if tpl == 'sparse_template_ell':
        return local_search_ELL(target, ori_op, selected_tile, second_best_cost, best_pred_avg_cost, ori_tot_nnz,
                                SM_num, dsize, tile_dict, is_tb_tile, max_bucket_size, nnzs_dict, pos_is_dict,
                                costs_dict, TC_row_tile_dict, TC_row_costs, cur_selected_tile_num,
                                reverse_idx_TC_op_row, penalty, TC_is, max_avg_cost_diff)
    return local_search_1D_BSDDMM(target, ori_op, selected_tile, second_best_cost, best_pred_avg_cost, ori_tot_nnz,
                                  SM_num, dsize, tile_dict, is_tb_tile, max_bucket_size, nnzs_dict, pos_is_dict,
                                  costs_dict, TC_row_tile_dict, TC_row_costs, cur_selected_tile_num,
                                  reverse_idx_TC_op_row, penalty, TC_is, max_avg_cost_diff)


# ---------------------------------------------------------------------------
# Next-level ops.
# ---------------------------------------------------------------------------

def _sub_op_from_ids(ori_op, state, ids, global_op_id, max_bucket_size):
    """A fresh row-sorted, padded ELL op over the nonzeros `ids` of the
    original matrix (the reference's Op over the uncovered position space)."""
    A = state.A
    rows = state.row_of[ids]
    cols = A.indices[ids]
    import scipy.sparse
    order = np.lexsort((cols, rows))
    ids = ids[order]
    sub = scipy.sparse.csr_matrix((np.ones(len(ids), np.float32), (rows[order], cols[order])), shape=A.shape)
    sub.sort_indices()
    new = Op(ori_op.op_type, ori_op.sidxs, ori_op.ridxs, ori_op.idx_lens, ori_op.idx_graph,
             [_as_csr(sub)] + list(ori_op.inps[1:]), None)
    new.ori_nnz_ids = ids
    new.state = state
    cands, global_op_id = gen_candidate_formats(new, max_bucket_size, ([1], [1], [1]), False,
                                                op_id_start=global_op_id, gen_TC_formats=False)
    out = []
    for c in cands:
        c.state = state
        c.ori_nnz_ids = ids
        if get_template_str(c) in ["TensorCore_template", "sparse_template", 'TC_sddmm']:
            continue
        out.append(c)
    return out, global_op_id


def gen_full_next_level_subops(TC_is, global_op_id, new_op, max_bucket_size, kernel_tile_size_options,
                               TC_k_notsorted, reorder_k_by_nnz=True, is_tb_tile=False, state=None):
    """Regenerate the ELL candidates from the uncovered nonzeros, split into
    rows touched by TC tiles and the others (reference :1051-1094)."""
    state = state if state is not None else new_op.state
    unc = np.flatnonzero(state.uncovered)
    if new_op.op_type == 'spmm':
        touched = np.asarray(TC_is)[state.row_of[unc]]
        ret = []
        for ids in (unc[touched], unc[~touched]):
            ops, global_op_id = _sub_op_from_ids(new_op, state, ids, global_op_id, max_bucket_size)
            ret += ops
        return ret, global_op_id
    ops, global_op_id = _sub_op_from_ids(new_op, state, unc, global_op_id, max_bucket_size)
    return ops, global_op_id


# ---------------------------------------------------------------------------
# Withdraw post-processing.
# ---------------------------------------------------------------------------

def post_process_for_withdraw_BSPMM(ori_op, selected_tiles, TC_row_tile_dict, TC_row_costs, reverse_idx_TC_op_row):
    """Apply whole-row TC "withdraw" (reference :1652-1702): drop the tiles of
    whole-row windows, let whole-row tiles cover every entry of their window,
    send ELL rows inside whole-row windows (and empty rows of non-atomic ELL
    tiles) to the sentinel row M, and recompute atomic flags."""
    M = ori_op.idx_lens[0]
    whole = np.flatnonzero(TC_row_costs == -1)
    to_del = np.concatenate([np.asarray(TC_row_tile_dict[b], dtype=np.int64) for b in whole] + [np.zeros(0, np.int64)])
    remain = np.setdiff1d(np.arange(len(selected_tiles)), to_del)
    remains = [selected_tiles[i] for i in remain]
    ori_rownnz = np.diff(ori_op.inps[0].indptr)
    for t in remains:
        if get_template_str(t.op) == 'TensorCore_template':
            if not t.is_atomic_tile:
                # whole-row tiles cover the window's full (initial) column range
                e = t.entries()
                op = t.op
                t.covered_ids = op._ent_id[e]
                t.covered_rows = op._ent_row[e]
                t.covered_cols = op._ent_col[e]
                t.nnz_when_selected = t.j_num * len(e)
            t.get_k_vals()
            t.tile_k_rng_sel = t.tile_k_rng
        else:
            op = t.op
            pieces = np.arange(t.tile_i_rng[0], t.tile_i_rng[1] + 1)
            out_rows = np.asarray(op.idx_values_list[0][0])[op.hyb_new_rows[0][pieces]]
            if op.ori_row_map is not None:
                out_rows = op.ori_row_map[out_rows]
            blk = reverse_idx_TC_op_row[out_rows] // 16
            inside = np.isin(blk, whole)
            t.i_vals = np.copy(out_rows)
            cov_rows = t.covered_rows - t.tile_i_rng[0] if t.covered_rows is not None else np.zeros(0, np.int64)
            cov_per_row = np.bincount(cov_rows, minlength=len(pieces))
            if inside.any():
                t.i_vals[inside] = M
                if t.is_atomic_tile:
                    keep = np.flatnonzero(~inside)
                    if ori_rownnz[out_rows[keep]].sum() == cov_per_row[keep].sum():
                        t.is_atomic_tile = False
            t.tile_i_rng_sel = t.tile_i_rng
            if not t.is_atomic_tile:
                t.i_vals[cov_per_row == 0] = M
    return remains


def post_process_for_withdraw_BSDDMM(ori_op, selected_tiles, TC_row_tile_dict, TC_row_costs, reverse_idx_TC_op_row):
    """1D tiles first, then TC tiles (reference :1709-1751); the 1D op's
    per-nonzero rows, columns and output ids are set for the packer."""
    for t in selected_tiles:
        op = t.op
        if get_template_str(op) == '1D_sddmm' and op.ori_row_ids_1d is None:
            op.ori_row_ids_1d = np.asarray(op.idx_values_list[0][0])[op._ent_row]
            op.ori_col_ids_1d = np.asarray(op.idx_values_list[0][2])[op._ent_col]
            op.nnz_id_matrix = op._ent_id
    a = [t for t in selected_tiles if get_template_str(t.op) == '1D_sddmm']
    b = [t for t in selected_tiles if get_template_str(t.op) == 'TC_sddmm']
    return a + b


def post_process_for_withdraw(ori_op, selected_tiles, TC_row_tile_dict, TC_row_costs, reverse_idx_TC_op_row):
    if ori_op.op_type == 'spmm':
        return post_process_for_withdraw_BSPMM(ori_op, selected_tiles, TC_row_tile_dict, TC_row_costs,
                                               reverse_idx_TC_op_row)
    return post_process_for_withdraw_BSDDMM(ori_op, selected_tiles, TC_row_tile_dict, TC_row_costs,
                                            reverse_idx_TC_op_row)

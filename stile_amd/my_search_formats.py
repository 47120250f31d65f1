"""Greedy tile-set search (Algorithm 2 of the STile report) for the MI355X build.

Same names, arguments and return values as MY_sparse/my_search_formats.py:
  greedy_search_use_cost_model_lower_bound (:263-833)
  update_op_after_tile_selection (:221-234)
  update_tile_after_selection (:241-256)
  update_position_space_for_ops_given_tiles (:112-211)
  estimate_max_bucket_size (:965-981)

The loop: TC pre-pass per 16-row window; then, while nonzeros are
uncovered, take the cheaper (predicted average cost per newly covered point)
of the best TC tile and the best ELL/1D tile, expand it by local search to
every tile of the same op and shape within (1 + max_avg_cost_diff) of its
cost, update coverage, and regenerate next-level ELL ops after TC picks.
TVM-only arguments (cache_set, run_params, use_faster_tuner, log_file, cuda_i,
dtype_str, zerotype) are accepted and ignored.
"""
import math

import numpy as np

from . import my_branch_and_bound, my_cost_model
from .gen_formats_v2 import (SearchState, gen_candidate_formats, gen_position_space_for_area, get_config,
                             get_template_str, set_config, SearchConfig)

global_op_id = 0


def estimate_max_bucket_size(op, sm_num=None, threads_per_sm=None):
    """Largest ELL bucket (8*2^i / 8, i = 5..2) that still gives half-occupancy
    work to every compute unit (reference :965-981); the hardware constants
    default to the active cost-model profile (A100: 108 SMs x 2048 threads;
    MI355X: 256 CUs x 2048 threads)."""
    prof = my_cost_model.get_profile()
    sm_num = prof.sm_num if sm_num is None else sm_num
    threads_per_sm = prof.threads_per_sm if threads_per_sm is None else threads_per_sm
    thread_num = 256
    blk_num_per_SM = threads_per_sm / thread_num
    thread_i = 8
    cand_workloads = [8 * (2 ** i) for i in range(2, 6)][::-1]
    cand = cand_workloads[-1]
    for cand in cand_workloads:
        if op.inps[0].nnz / (cand / 2) / sm_num >= 0.5 * blk_num_per_SM:
            return cand // thread_i
    return cand // thread_i


def update_tile_after_selection(selected_tile):
    """Record what the tile covers now and mark it covered."""
    return my_branch_and_bound._cover(selected_tile)


def update_position_space_for_ops_given_tiles(sub_ops, selected_tiles):
    """Coverage is shared (SearchState), so selecting the tiles already updated
    every op; the 1D SDDMM ops refresh their bucket counts."""
    for sub_op in sub_ops:
        if sub_op.op_type == 'sddmm' and get_template_str(sub_op) == '1D_sddmm':
            from .gen_formats_v2 import refresh_1D_counts
            refresh_1D_counts(sub_op)


def update_op_after_tile_selection(op):
    return op


def greedy_search_use_cost_model_lower_bound(op, dsize, run_params, max_bucket_size, cache_set,
                                             kernel_tile_size_options, only_TC, only_ELL, penalty, TC_k_notsorted,
                                             reorder_k_by_nnz=True, max_level_num=float('inf'),
                                             max_avg_cost_diff=float('inf'), use_faster_tuner=False,
                                             log_file=None, cuda_i=0, dtype="float16", dtype_str=None,
                                             zerotype=None, single_level=None, no_withdraw=False):
    """Select the computation tiles greedily (reference :263-833).

    Returns (selected_tiles, TC_row_tile_dict, TC_row_costs,
    reverse_idx_TC_op_row, level_num)."""
    global global_op_id
    if single_level is None:
        single_level = not (max_level_num > 1)
    cfg = get_config()
    set_config(SearchConfig(op_type=op.op_type, single_level=single_level, TC_k_notsorted=TC_k_notsorted,
                            no_withdraw=no_withdraw, TC_tile_sizes=cfg.TC_tile_sizes))
    state = SearchState(op.inps[0])
    op.state = state
    cand_ops, global_op_id = gen_candidate_formats(op, max_bucket_size, kernel_tile_size_options, TC_k_notsorted,
                                                   reorder_k_by_nnz=reorder_k_by_nnz, op_id_start=0,
                                                   gen_TC_formats=(not only_ELL))
    ori_tot_nnz = op.inps[0].getnnz()
    SM_num = my_cost_model.get_profile().sm_num
    level_num = 1
    selected_tiles = []
    tile_dict = {}
    nnzs_dict, pos_is_dict, costs_dict = {}, {}, {}
    all_subops, all_TC_ops = [], []
    for sub_op in cand_ops:
        sub_op.state = state
        gen_position_space_for_area(sub_op, 0)
        if get_template_str(sub_op) in ['TensorCore_template', 'TC_sddmm']:
            all_TC_ops.append(sub_op)
            continue
        if sub_op.op_type == 'spmm' and get_template_str(sub_op) != 'sparse_template_ell':
            continue
        all_subops.append(sub_op)
    if not only_ELL and all_TC_ops:
        reverse_idx_TC_op_row = np.argsort(all_TC_ops[0].idx_values_list[0][0])
    else:
        reverse_idx_TC_op_row = np.argsort(np.arange(op.idx_lens[0]))
    if only_TC:
        all_subops = []
    if only_ELL:
        all_TC_ops = []
    best_possible_ELL = float('inf')
    if all_subops:
        best_possible_ELL = my_cost_model.lower_bound_ELL_cost(all_subops[0], max_bucket_size, dsize)
    TC_row_num = math.ceil(op.inps[0].shape[0] / 16)
    TC_row_tile_dict = {i: np.zeros(0, np.int64) for i in range(TC_row_num)}
    TC_row_costs = np.full(TC_row_num, 0.0)
    TC_is = np.full(op.inps[0].shape[0], False)

    def regenerate():
        global global_op_id
        subs, global_op_id = my_branch_and_bound.gen_full_next_level_subops(
            TC_is, global_op_id, op, max_bucket_size, kernel_tile_size_options, TC_k_notsorted,
            reorder_k_by_nnz=reorder_k_by_nnz, is_tb_tile=True, state=state)
        return subs

    if all_TC_ops and not only_TC and not only_ELL:
        selected_tiles = my_branch_and_bound.search_TC_tiles_only_by_rows(
            all_TC_ops[0], dsize, True, best_possible_ELL, ori_tot_nnz, all_TC_ops, all_subops, op,
            max_avg_cost_diff, TC_row_tile_dict, TC_row_costs, penalty, max_bucket_size, TC_is)
        if not single_level and state.uncovered.any():
            all_subops = regenerate()
            tile_dict, nnzs_dict, pos_is_dict, costs_dict = {}, {}, {}, {}
    elif only_TC and op.op_type == 'sddmm' and all_TC_ops:
        selected_tiles = _all_nonempty_TC_tiles(all_TC_ops[0])

    no_selected_tile_round = 0
    while state.uncovered.any():
        good_TC = []
        init_avg_cost = float('inf')
        if all_TC_ops:
            key, cost_TC = my_branch_and_bound.search_best_TC_tile(all_TC_ops, dsize, TC_row_tile_dict,
                                                                   TC_row_costs, penalty, max_bucket_size)
            if key is not None:
                _, tile_sizes, poses, is_atomic = key
                tc_op = all_TC_ops[0]
                for pos in poses:
                    t = my_branch_and_bound._make_tc_tile(tc_op, tile_sizes, pos, is_atomic, None)
                    t.update_nnz_uncovered()
                    t.pred_cost = cost_TC * t.nnz_uncovered
                    t.cost = 1
                    t.set_pred_avg_cost()
                    good_TC.append(t)
                if good_TC:
                    init_avg_cost = good_TC[0].pred_avg_cost
        selected = good_TC
        second_best = float('inf')
        if not (init_avg_cost < best_possible_ELL) and all_subops:
            bkey, second_best = my_branch_and_bound.search_best_tile(op, all_subops, tile_dict, dsize, ori_tot_nnz,
                                                                     SM_num, max_bucket_size, nnzs_dict,
                                                                     pos_is_dict, costs_dict,
                                                                     best_pred_avg_cost=init_avg_cost)
            if bkey is not None:
                t = tile_dict[bkey]
                t.best_tile_sizes = t.tile_sizes
                t.best_params = t.params
                t.cost = t.pred_cost
                selected = [t]
        if not selected:
            no_selected_tile_round += 1
            if no_selected_tile_round == 2 or not all_subops:
                raise RuntimeError("greedy search stalled with uncovered nonzeros")
            all_subops = regenerate()
            tile_dict, nnzs_dict, pos_is_dict, costs_dict = {}, {}, {}, {}
            continue
        no_selected_tile_round = 0
        best_pred_avg_cost = selected[0].pred_avg_cost
        if sum(t.nnz_uncovered for t in selected) == 0:
            raise RuntimeError("greedy search selected a tile covering nothing")
        good = my_branch_and_bound.local_search(op, selected[0], second_best, best_pred_avg_cost, ori_tot_nnz,
                                                SM_num, dsize, tile_dict, True, max_bucket_size, nnzs_dict,
                                                pos_is_dict, costs_dict, TC_row_tile_dict, TC_row_costs,
                                                len(selected_tiles), reverse_idx_TC_op_row, penalty, TC_is,
                                                max_avg_cost_diff=max_avg_cost_diff)
        for t in good:
            update_tile_after_selection(t)
        selected_tiles = selected_tiles + good
        update_position_space_for_ops_given_tiles(all_subops + all_TC_ops, good)
        if not state.uncovered.any():
            break
        is_tc = get_template_str(selected[0].op) in ["TensorCore_template", "TC_sddmm"]
        if (not single_level) and level_num < max_level_num and is_tc:
            level_num += 1
            all_subops = regenerate()
            tile_dict, nnzs_dict, pos_is_dict, costs_dict = {}, {}, {}, {}
        elif is_tc or True:
            # coverage changed: cached ELL positions/costs are stale
            tile_dict, nnzs_dict, pos_is_dict, costs_dict = {}, {}, {}, {}
    return selected_tiles, TC_row_tile_dict, TC_row_costs, reverse_idx_TC_op_row, level_num


def _all_nonempty_TC_tiles(tc_op):
    """only_TC SDDMM: every nonempty TC block (reference :410-432)."""
    tile_sizes = list(tc_op.blk_nums.keys())[0]
    tiles = []
    for w in range(len(tc_op.blk_nums[tile_sizes])):
        n = int(tc_op.blk_nums[tile_sizes][w])
        poses = [(w, 0, i) for i in range(n)]
        if tc_op.TC_k_notsorted:
            poses = [(w, 0, int(i)) for i in tc_op.blk_ks[tile_sizes][w]]
        for p in poses:
            t = my_branch_and_bound._make_tc_tile(tc_op, tile_sizes, p, False, 1)
            my_branch_and_bound._cover(t)
            tiles.append(t)
    return tiles

"""Restatement of the reference's integer format construction and packers
(test infrastructure: the checker for stile_amd's vectorised versions).

Each function follows the cited reference text step by step, on scipy CSR
position spaces with 0/1 data as the reference does.  Pinned by
tests/test_oracle_formats.py with hand-computed known answers on tiny
matrices and invariants (see each docstring).
"""
import math

import numpy as np
import scipy.sparse


def _pos_space(A, idx_values):
    A01 = scipy.sparse.csr_matrix((np.ones(A.nnz), A.indices, A.indptr), shape=A.shape)
    return A01[np.asarray(idx_values[0]), :][:, np.asarray(idx_values[2])]


def reorder_idx(A, idx_values, idx_i):
    """MY_sparse/gen_formats_v2.py:820-857: argsort(-nnz, stable) of the rows
    (idx 0) or columns (idx 2) of the position space; j (idx 1) unchanged.
    Pinned: known answer test_reorder_idx_known_answer."""
    if idx_i == 1:
        return np.array(idx_values[1])
    ps = _pos_space(A, idx_values)
    nnzs = ps.getnnz(axis=1) if idx_i == 0 else ps.getnnz(axis=0)
    order = np.argsort(nnzs * (-1), kind='stable')
    return np.array(idx_values[idx_i])[order]


def pad_irregular_idx(A, idx_values, max_bucket_size):
    """gen_formats_v2.py:1065-1121 (SpMM branch), transcribed with its row loop.
    Returns hyb_new_rows, hyb_ori_row_ends, hyb_short_begin_row,
    hyb_getnnz_minus and the piece-level position space (CSR).
    Pinned: test_pad_irregular_idx_known_answer + invariants."""
    ps = _pos_space(A, idx_values)
    getnnz = ps.getnnz(axis=1)
    new_row_nums = np.maximum(np.ceil(getnnz / max_bucket_size), 1).astype('int')
    hyb_new_rows = np.repeat(np.arange(len(idx_values[0])), new_row_nums)
    hyb_ori_row_ends = np.cumsum(new_row_nums) - 1
    short_begin_row_ori = np.searchsorted(-getnnz, -max_bucket_size)
    if short_begin_row_ori == len(getnnz):
        hyb_short_begin_row = len(hyb_new_rows)
    else:
        hyb_short_begin_row = hyb_ori_row_ends[short_begin_row_ori]
    indptr = np.empty(len(hyb_new_rows) + 1)
    indptr[0] = 0
    start_idx = 0
    for i in range(len(idx_values[0])):
        indices_num = getnnz[i]
        if indices_num == 0:
            indptr[start_idx + 1] = indptr[start_idx]
        else:
            idx = slice(start_idx, hyb_ori_row_ends[i] + 1)
            start_indptr = indptr[start_idx]
            indptr[idx] = range(int(start_indptr), int(start_indptr + indices_num), max_bucket_size)
            indptr[hyb_ori_row_ends[i] + 1] = start_indptr + indices_num
        start_idx = hyb_ori_row_ends[i] + 1
    pieces = scipy.sparse.csr_matrix((ps.data, ps.indices, indptr), shape=(len(hyb_new_rows), ps.shape[1]))
    inp_getnnz = pieces.getnnz(axis=1)
    return dict(hyb_new_rows=hyb_new_rows, hyb_ori_row_ends=np.array(hyb_ori_row_ends),
                hyb_short_begin_row=int(hyb_short_begin_row),
                hyb_getnnz_minus=-inp_getnnz[hyb_short_begin_row:], position_space=pieces)


def condense_for_TC_tile(A, idx_values, blk_i_size=16, reorder_k_by_nnz=True, TC_k_notsorted=False):
    """gen_formats_v2.py:1004-1041: per window of blk_i_size rows, the column
    order argsort(-colnnz, stable) (or nonempty-first).  Returns dense arrays.
    Pinned: test_condense_known_answer."""
    row_num = len(idx_values[0])
    row_blk_num = math.ceil(row_num / blk_i_size)
    if TC_k_notsorted:
        k_vals = np.arange(len(idx_values[2]))
        return [k_vals for _ in range(row_blk_num)]
    csr = _pos_space(A, idx_values)
    if reorder_k_by_nnz:
        return [np.argsort(csr[i * blk_i_size:(i + 1) * blk_i_size, :].getnnz(axis=0) * -1, kind='stable')
                for i in range(row_blk_num)]
    return [np.argsort(csr[i * blk_i_size:(i + 1) * blk_i_size, :].getnnz(axis=0) <= 0, kind='stable')
            for i in range(row_blk_num)]


def tc_density(A, idx_values):
    """select_the_densest_tensor_for_TC_tiles, gen_formats_v2.py:1286-1309:
    nnz / (sum over windows of ceil(#nonempty cols / 16) * 256)."""
    csr = _pos_space(A, idx_values)
    row_blk_num = math.ceil(len(idx_values[0]) / 16)
    tot = sum(math.ceil(np.count_nonzero(csr[i * 16:(i + 1) * 16, :].getnnz(axis=0)) / 16)
              for i in range(row_blk_num)) * 16 * 16
    return csr.nnz / tot if tot else 0.0


def set_init_non_empty_blk_num(A, idx_values, K_list, TC_k_notsorted=False):
    """gen_formats_v2.py:1313-1335: blk_nums[K][w] = ceil(#nonempty cols / K);
    with TC_k_notsorted the number of nonempty 16-wide column blocks."""
    csr = _pos_space(A, idx_values)
    row_blk_num = math.ceil(len(idx_values[0]) / 16)
    out = {}
    for K in K_list:
        if TC_k_notsorted:
            out[K] = np.asarray([np.count_nonzero(np.add.reduceat(csr[i * 16:(i + 1) * 16, :].getnnz(axis=0),
                                                                  np.arange(0, csr.shape[1], K)))
                                 for i in range(row_blk_num)])
        else:
            out[K] = np.asarray([math.ceil(np.count_nonzero(csr[i * 16:(i + 1) * 16, :].getnnz(axis=0)) / K)
                                 for i in range(row_blk_num)])
    return out


def get_valid_tile_pos_list_for_ELL(hyb_getnnz_minus, hyb_short_begin_row, n_pieces, I, K, max_bucket_size):
    """my_branch_and_bound.py:16-53 (the 'p' protocol branch), transcribed."""
    targets = [K, K // 2]
    target_is = [np.searchsorted(hyb_getnnz_minus, -t) + hyb_short_begin_row for t in targets]
    list1 = list(range(math.ceil(target_is[0] / I), (target_is[1] - 1) // I + 1))
    if len(list1) > 0:
        if (list1[-1] + 1) * I < n_pieces:
            list1 = list1 + [list1[-1] + 1]
    if K == max_bucket_size:
        list2 = list(range((hyb_short_begin_row - 1) // I + 1))
        return [(i, 0, 0) for i in list2 + list1]
    return [(i, 0, 0) for i in list1]


def _factors(val):
    """get_factors, gen_formats_v2.py (divisors of val, ascending)."""
    return [i for i in range(1, val + 1) if val % i == 0]


def _tiles(val, split_n, base):
    """get_tiles, gen_formats_v2.py: every split of val into split_n - 1 factors
    appended to base."""
    if split_n == 2:
        return [base + (i,) for i in _factors(val)]
    out = []
    for i in _factors(val):
        out += _tiles(val // i, split_n - 1, base + (i,))
    return out


def ell_tb_tile_shapes(n_pieces, max_bucket_size):
    """Thread-block ELL tile shapes the search enumerates for an ELL op of
    n_pieces split rows, transcribed:
      get_tile_sizes_idx 'p' branch with is_tb_tile (gen_formats_v2.py:1947-1956):
        I = (None, 2^t) for t <= log2(n_pieces), K = (1, 2^t) for t <= log2(mbs);
      J fixed to (None, 32) by search_best_tile_given_op_BSPMM
        (my_branch_and_bound.py:751-758);
      itertools.product in (I, J, K) order, then tile_size_heuristics with
        is_tb_tile (gen_formats_v2.py:1188-1214): I*J a multiple of 128,
        I*J >= 256 and I*K == 8*mbs.
    Pinned: test_ell_tile_shapes_known_answer."""
    import itertools
    idx0 = [(None, 2 ** t) for t in range(int(math.log(n_pieces, 2)) + 1)]
    idx1 = [(None, 32)]
    idx2 = [(1, 2 ** t) for t in range(int(math.log(max_bucket_size, 2)) + 1)]
    out = []
    for ts in itertools.product(idx0, idx1, idx2):
        threads = ts[0][1] * ts[1][1]
        if threads % 128 != 0 or threads < 256:
            continue
        if ts[0][1] * ts[2][1] != max_bucket_size * 8:
            continue
        out.append(ts)
    return out


def tc_tile_sizes(op_type='spmm', k_notsorted=False):
    """get_TC_tile_sizes, gen_formats_v2.py:1252-1281 (no TC_tile_sizes
    override): SpMM (16, 32, 16*i) for i = 10..1 (k_notsorted: i = 1 only);
    SDDMM (16, 16*2^j, 16) for j = 0..2.  Pinned: test_tc_tile_sizes_known_answer."""
    if k_notsorted:
        return [((None, 16), (None, 32), (None, 16 * i)) for i in range(1, 2)]
    if op_type == 'sddmm':
        return [((None, 16), (None, 16 * (2 ** j)), (None, 16 * i)) for i in range(1, 2) for j in range(3)]
    return [((None, 16), (None, 32), (None, 16 * i)) for i in range(160 // 16, 0, -1)]


def sddmm_tc_tile_sizes(k_feat):
    """get_tile_sizes_idx, 'TC_sddmm' branch (gen_formats_v2.py:1966-1971):
    the SDDMM TC shapes whose j size is the largest one <= K_feat."""
    idx = tc_tile_sizes('sddmm')
    tile_j = max(j for j in set(ts[1][1] for ts in idx) if j <= k_feat)
    return [ts for ts in idx if ts[1][1] == tile_j]


def tile_params(template, tile_sizes, max_bucket_size, is_tb_tile=True):
    """get_params_list, gen_formats_v2.py:2036-2068, per template string
    ('sparse_template_ell', 'TensorCore_template', '1D_sddmm', 'TC_sddmm').
    Pinned: test_tile_params_known_answer."""
    out = []
    if template in ('sparse_template_ell', 'TensorCore_template'):
        for unroll in ([True] if is_tb_tile else [True, False]):
            params = {'use_implicit_unroll': unroll}
            if template == 'sparse_template_ell':
                params['last_bucket'] = math.prod(tile_sizes[2][1:]) == max_bucket_size
            else:
                if not unroll:
                    continue
                params['mma_shape_str'] = "m8n32k16" if tile_sizes == ((None, 8), (None, 32), (None, 16)) \
                    else "m16n16k16"
            out.append(params)
    elif template == '1D_sddmm':
        out.append({'tx': 8, 'ty': 4, 'vec_size': 4, 'group_size': 8, 'max_bucket_size': max_bucket_size})
    elif template == 'TC_sddmm':
        out.append({'mma_shape_str': "m16n16k16", 'warp_num': 1, 'vec_size': 4})
    return out


def estimate_max_bucket_size(nnz, SM_num=108, tot_thread_per_SM=2048):
    """my_search_formats.py:965-981, transcribed."""
    thread_num = 256
    blk_num_per_SM = tot_thread_per_SM / thread_num
    thread_i = 8
    cand_workloads = [8 * (2 ** i) for i in range(2, 6)][::-1]
    for cand in cand_workloads:
        if nnz / (cand / 2) / SM_num >= 0.5 * blk_num_per_SM:
            return cand // thread_i
    return cand // thread_i


def get_padded_j_TC(ori_j_len, mma_n):
    """my_fuse_formats.py:2657-2662."""
    if ori_j_len < mma_n:
        ori_j_len = mma_n
    elif ori_j_len > mma_n:
        ori_j_len = math.ceil(ori_j_len / (2 * mma_n)) * 2 * mma_n
    return ori_j_len


# ---------------------------------------------------------------------------
# Packers (my_fuse_formats.py:3321-3537 and :3544-3678), transcribed.
# They read the tiles through the reference's attribute names:
#   tile.position_space_when_selected, tile.i_vals, tile.k_vals, tile.tile_pos,
#   tile.tile_i_rng (TC), tile.op.idx_values_list, tile.op.k_vals,
#   tile.op.hyb_new_rows.
# ---------------------------------------------------------------------------

def pack_ell_group(tiles, I, K):
    """my_fuse_formats.py:3372-3419 ('sparse_template_ell'; tile_i_rng is None
    after post_process_for_withdraw, so the row ids come from tile.i_vals)."""
    a = np.zeros((I * len(tiles), K))
    indices_k = np.concatenate([np.full((I, K), tile.op.idx_values_list[0][2][0]) for tile in tiles], axis=0)
    indices_reordered_i = np.concatenate([np.full(I, tile.i_vals[-1]) for tile in tiles])
    for count, tile in enumerate(tiles):
        a_ptr = count * I
        covered_csr = tile.position_space_when_selected
        nnzs = covered_csr.getnnz(axis=1)
        idx_values = tile.op.idx_values_list[0]
        for i in range(covered_csr.shape[0]):
            idxs = slice(0, nnzs[i])
            a[a_ptr + i][idxs] = covered_csr[i].data
            indices_k[a_ptr + i][idxs] = np.asarray(idx_values[2])[covered_csr[i].indices]
        idxs = slice(a_ptr, a_ptr + covered_csr.shape[0])
        indices_reordered_i[idxs] = tile.i_vals
    a = np.reshape(a, (len(tiles), I * K))
    return dict(a=a.astype(np.float32), indices_k=indices_k.astype(np.int32),
                indices_reordered_i=indices_reordered_i.astype(np.int32))


def pack_tc_group(tiles, I, K, rows_2d, k_notsorted=False):
    """my_fuse_formats.py:3424-3504 ('TensorCore_template')."""
    tiles = sorted(tiles, key=lambda tile: tile.tile_pos)
    covered_csrs = [tile.position_space_when_selected.copy() for tile in tiles]
    for m in covered_csrs:
        m.resize((I, K))
    a = np.concatenate([m.toarray(order='C') for m in covered_csrs], axis=None)
    a = np.reshape(a, (-1, I, K))
    poi_is, counts = np.unique([tile.tile_pos[0] for tile in tiles], return_counts=True)
    indptr = np.concatenate(([0], np.cumsum(counts)))
    indices = np.concatenate([
        np.full(K, np.asarray(tile.op.idx_values_list[0][2])[tile.op.k_vals[tile.tile_pos[0]][tile.k_vals[-1]]])
        for tile in tiles])
    indices = np.reshape(indices, (-1, K))
    for count, tile in enumerate(tiles):
        indices[count][:tile.position_space_when_selected.shape[1]] = \
            np.asarray(tile.op.idx_values_list[0][2])[tile.op.k_vals[tile.tile_pos[0]][tile.k_vals]]
    if rows_2d:
        i_indices = [np.full(I, np.asarray(tiles[t].op.idx_values_list[0][0])[tiles[t].tile_i_rng[1]])
                     for t in indptr[:-1]]
        for count, t in enumerate(indptr[:-1]):
            tile = tiles[t]
            i_indices[count][:tile.position_space_when_selected.shape[0]] = \
                np.asarray(tile.op.idx_values_list[0][0])[tile.tile_i_rng[0]:tile.tile_i_rng[1] + 1]
        i_indices = np.array(i_indices)
    else:
        i_indices = poi_is
    if k_notsorted:
        indices = np.asarray([tile.tile_pos[2] for tile in tiles])
    return dict(a=a.astype(np.float32), i_indices=np.asarray(i_indices).astype(np.int32),
                indptr=indptr.astype(np.int32), indices=np.asarray(indices).astype(np.int32))


def pack_sddmm_1d_group(tiles, bucket, ty=4, group_size=8):
    """my_fuse_formats.py:3565-3598: per-nonzero original row, column and
    output id of every tile's bucket, padded to a multiple of ty*group_size."""
    mid_idx = np.concatenate([t.op.ori_row_ids_1d[t.tile_pos[0] * bucket:(t.tile_pos[0] + 1) * bucket] for t in tiles])
    j_idx = np.concatenate([t.op.ori_col_ids_1d[t.tile_pos[0] * bucket:(t.tile_pos[0] + 1) * bucket] for t in tiles])
    remap = np.concatenate([t.op.nnz_id_matrix[t.tile_pos[0] * bucket:(t.tile_pos[0] + 1) * bucket] for t in tiles])
    real = ty * group_size
    to_pad = math.ceil(len(mid_idx) / real) * real - len(mid_idx)
    mid_idx = np.concatenate([mid_idx, np.full(to_pad, mid_idx[-1])])
    j_idx = np.concatenate([j_idx, np.full(to_pad, j_idx[-1])])
    remap = np.concatenate([remap, np.full(to_pad, remap[-1])])
    return dict(mid_idx=mid_idx.astype(np.int32), j_idx=j_idx.astype(np.int32), remap=remap.astype(np.int32))


def sddmm_1d_ids(A, idx_values, sub=None):
    """post_process_for_withdraw_BSDDMM, my_branch_and_bound.py:1709-1751 (1D op):
    original row, column and 0-based output id of each nonzero in the 1D op's
    order; `sub` is the op's own pattern (op.inps[0], the uncovered part for a
    next-level op), the ids are ori_op.nnz_id_matrix.multiply(sub)."""
    A = scipy.sparse.csr_matrix(A)
    ids = scipy.sparse.csr_matrix((np.arange(1, A.nnz + 1), A.indices, A.indptr), shape=A.shape)
    if sub is not None:
        ids = ids.multiply(scipy.sparse.csr_matrix(sub) != 0).tocsr()
    m = ids[np.asarray(idx_values[0]), :][:, np.asarray(idx_values[2])]
    m.eliminate_zeros()
    m.has_sorted_indices = False
    m.sort_indices()
    nnz_ids = m.data - 1
    rows = np.repeat(np.asarray(idx_values[0]), np.diff(m.indptr))
    cols = np.asarray(idx_values[2])[m.indices]
    return rows, cols, nnz_ids


def pack_sddmm_tc_group(A, tiles, tile_sizes):
    """my_fuse_formats.py:3602-3656: block rows/cols, CSC-ordered in-block
    positions (col + row*K) and 0-based output ids of every sampled entry."""
    A = scipy.sparse.csr_matrix(A)
    sub_op = tiles[0].op
    iv = sub_op.idx_values_list[0]
    ids = scipy.sparse.csr_matrix((np.arange(1, A.nnz + 1), A.indices, A.indptr), shape=A.shape)
    nnz_id_matrix = ids[np.asarray(iv[0]), :][:, np.asarray(iv[2])]
    nnz_id_matrix.eliminate_zeros()
    K = tile_sizes[2][1]
    # The reference assumes every window and column run is full (M, N
    # multiples of 16); a run cut short by the matrix edge is padded with its
    # last index here, which no sampled entry addresses.
    def full(idx, n):
        idx = np.asarray(idx)
        return idx if len(idx) >= n else np.concatenate([idx, np.full(n - len(idx), idx[-1])])
    i_indices = np.asarray([full(np.asarray(iv[0])[t.tile_i_rng[0]:t.tile_i_rng[1] + 1], tile_sizes[0][1])
                           for t in tiles])
    max_pos_i = max(t.tile_pos[0] for t in tiles)
    j_lists = [np.asarray(iv[2])[np.asarray(sub_op.k_vals[i])] for i in range(max_pos_i + 1)]
    j_indices = np.asarray([full(j_lists[t.tile_pos[0]][t.tile_k_rng[0]:t.tile_k_rng[1] + 1], K) for t in tiles])
    per_win = [(nnz_id_matrix[i * 16:(i + 1) * 16, :][:, np.asarray(sub_op.k_vals[i])]).tocsc()
               for i in range(max_pos_i + 1)]
    cscs = [per_win[t.tile_pos[0]][:, t.tile_k_rng[0]:t.tile_k_rng[1] + 1] for t in tiles]
    for c in cscs:
        c.sort_indices()
    indicesDense = np.concatenate([np.repeat(np.arange(c.shape[1]), c.getnnz(axis=0)) + c.indices * K for c in cscs])
    indices = np.concatenate([c.data - 1 for c in cscs])
    indptr = np.cumsum([0] + [c.nnz for c in cscs])
    return dict(i_indices=i_indices.astype(np.int32), j_indices=j_indices.astype(np.int32),
                indptr=indptr.astype(np.int32), indices_dense=indicesDense.astype(np.int32),
                indices=indices.astype(np.int32))

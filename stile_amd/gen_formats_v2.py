"""Operator model, candidate sparse-tile formats and tiles (MI355X build).

Same public names and argument meaning as the reference module
MY_sparse/gen_formats_v2.py: Op (:451-525), DenseTensor, reorder_idx (:820-857),
pad_irregular_idx (:1065-1121), condense_for_TC_tile (:1004-1041),
change_to_dense (:1044-1052), gen_position_space_for_area (:900-925),
select_the_densest_tensor_for_TC_tiles (:1286-1309),
set_init_non_empty_blk_num (:1313-1335), get_1D_position_space (:1128-1147),
gen_candidate_TC_formats (:1341-1504), gen_candidate_formats(_BSPMM/_BSDDMM)
(:1510-1918), get_TC_tile_sizes (:1252-1281), tile_size_heuristics (:1188-1242),
get_tile_sizes_idx / get_tile_sizes_list / get_params_list (:1926-2068),
get_template_str (:2073-2085), ComputeTile (:2333-2485), gen_nnz_id_matrix
(:437-445).

Design differences (behaviour is the same):
  * Coverage is tracked per nonzero id (the index of a nonzero in A's CSR
    order) in one boolean array shared by every candidate op
    (SearchState.uncovered), instead of subtracting scipy 0/1 matrices from
    every op's position space.  An op's position_space_update is derived from
    it on demand.
  * A's real values travel with the ids, so the packed buffers carry A's
    values (the reference packs the 0/1 position space, i.e. all ones).
  * Global configuration is explicit (SearchConfig) instead of process-wide
    environment variables (op_type, single_level, TC_k_notsorted,
    no_withdraw, TC_tile_sizes in the reference).
"""
import copy
import itertools
import math
from dataclasses import dataclass, field
from fractions import Fraction

import numpy as np
import scipy.sparse


# ---------------------------------------------------------------------------
# Configuration (the reference's environment variables, made explicit).
# ---------------------------------------------------------------------------

@dataclass
class SearchConfig:
    op_type: str = "spmm"            # env op_type
    single_level: bool = False       # env single_level
    TC_k_notsorted: bool = False     # env TC_k_notsorted
    no_withdraw: bool = False        # env no_withdraw
    TC_tile_sizes: tuple = None      # env TC_tile_sizes (one TC shape), optional


_CONFIG = SearchConfig()


def get_config():
    return _CONFIG


def set_config(cfg):
    global _CONFIG
    _CONFIG = cfg
    return cfg


class DenseTensor(object):
    """A dense operand (reference DenseTensor): `.data` is the array."""

    def __init__(self, data):
        self.data = data


def _as_csr(A):
    A = scipy.sparse.csr_matrix(A)
    if not A.has_sorted_indices:
        A = A.copy()
        A.sort_indices()
    A.indptr = A.indptr.astype(np.int32, copy=False)
    A.indices = A.indices.astype(np.int32, copy=False)
    return A


def gen_nnz_id_matrix(op):
    """1-based id of each nonzero in A's CSR order (reference :437-445)."""
    if op.op_type == 'sddmm':
        A = op.inps[0]
        op.nnz_id_matrix = scipy.sparse.csr_matrix(
            (np.arange(1, A.nnz + 1), A.indices, A.indptr), shape=A.shape)
    else:
        raise AssertionError("Only support SDDMM now.")


class SearchState(object):
    """Coverage shared by all candidate ops of one search.

    uncovered[p] is True while nonzero p of the original A (CSR order) has not
    been covered by a selected tile."""

    def __init__(self, A):
        self.A = A
        self.nnz = A.nnz
        self.uncovered = np.ones(A.nnz, dtype=bool)
        self.row_of = np.repeat(np.arange(A.shape[0], dtype=np.int64), np.diff(A.indptr))


class Op(object):
    """Operator record (reference gen_formats_v2.py:451-525).

    inps[0] is A as CSR (op_type 'spmm' or 'sddmm'); inps[1:] are
    DenseTensor operands.  idx_lens = [M, N, K] for SpMM (i, j, k) and
    [M, K_feat, N] for SDDMM."""

    def __init__(self, op_type, sidxs, ridxs, idx_lens, idx_graph, inps, sparse_inps_transposed):
        self.op_type = op_type
        self.sidxs = sidxs
        self.ridxs = ridxs
        self.idxs = sidxs + ridxs
        self.idx_graph = idx_graph
        self.idx_lens = list(idx_lens)
        inps = list(inps)
        inps[0] = _as_csr(inps[0])
        self.inps = inps
        self.sparse_inps_transposed = sparse_inps_transposed
        self.idx_values_list = [[np.arange(max_v) for max_v in self.idx_lens]]
        self.nnz_dict = dict()
        self.loop_order = {0: self.idxs}
        if op_type in ('spmm', 'batched_spmm'):
            self.loop_protocals = {0: 'uuc'}
            self.inp_protocals = {0: ['uc', 'uu']}
        elif op_type == 'sddmm':
            self.loop_protocals = {0: 'cu'}
            self.inp_protocals = {0: ['c', 'uu', 'uu']}
        else:
            raise AssertionError(f"We do not support {op_type} currently!")
        self.position_space = dict()
        self.inp_getnnz = None
        self.pos_space_infor = dict()
        self.idx_reordered = {0: [False for _ in self.idxs]}
        self.this_area_i = 0
        self.hyb_new_rows = dict()
        self.hyb_short_begin_row = None
        self.hyb_getnnz_minus = None
        self.hyb_ori_row_ends = None
        self.ori_row_ids_1d = None
        self.ori_col_ids_1d = None
        self.row_nums_1d = list()
        self.col_nums_1d = list()
        self.nnz_id_matrix = None
        self.nnz_update_1D = None
        self.nnz_data_1D = None
        self.max_bucket_size = None
        self.k_vals = None
        self.blk_nums = dict()
        self.blk_ks = dict()
        self.TC_k_notsorted = False
        self.kernel_tile_Ns = [1 for _ in self.idxs]
        self.kernel_tile_times = 0
        self.area_pos = list()
        self.op_id = None
        # --- MI355X build: id-based position space --------------------------
        # _ent_row[e]: row (in this op's space) of entry e; _ent_id[e]: nonzero
        # id in the ORIGINAL matrix; _indptr: row pointer of the op space.
        self.state = None      # SearchState shared by the search
        self._ent_row = None
        self._ent_id = None
        self._ent_col = None   # column in this op's k space
        self._indptr = None
        self._nrows = None
        self.ori_nnz_ids = None  # ids (original matrix) this op was built from
        self.ori_row_map = None  # unused: sub-op rows are original row ids
        if self.op_type == 'sddmm':
            gen_nnz_id_matrix(self)

    # -- id-based helpers ---------------------------------------------------
    def get_key(self):
        return (self.op_id, get_template_str(self), tuple(self.idx_reordered[0]))

    @property
    def A(self):
        return self.inps[0]

    def uncovered_mask(self):
        """Boolean per entry of this op's position space: still uncovered."""
        if self.state is None:
            return np.ones(len(self._ent_id), dtype=bool)
        return self.state.uncovered[self._ent_id]

    def row_uncovered_counts(self):
        m = self.uncovered_mask()
        return np.bincount(self._ent_row[m], minlength=self._nrows).astype(np.int64)

    def _space_csr(self, mask=None, data=None):
        """scipy CSR view of (part of) the op's position space, data = 1."""
        rows = self._ent_row if mask is None else self._ent_row[mask]
        cols = self._ent_col if mask is None else self._ent_col[mask]
        vals = np.ones(len(rows)) if data is None else data
        indptr = np.zeros(self._nrows + 1, np.int64)
        np.add.at(indptr, rows + 1, 1)
        indptr = np.cumsum(indptr)
        return scipy.sparse.csr_matrix((vals, cols, indptr), shape=(self._nrows, self.idx_lens[2]))

    @property
    def position_space_update(self):
        if self._ent_id is None:
            return self.inps[0]
        return self._space_csr(self.uncovered_mask())

    @position_space_update.setter
    def position_space_update(self, value):
        # The reference assigns copies of the position space here; the id-based
        # build derives it from the shared coverage and ignores the value.
        pass


# ---------------------------------------------------------------------------
# Index reordering and position spaces.
# ---------------------------------------------------------------------------

def _reordered_view(op):
    """Rows/cols of A permuted by the op's idx_values (i and k)."""
    iv = op.idx_values_list[0]
    return op.inps[0][iv[0], :][:, iv[2]]


def reorder_idx(op, area_i, idx_i):
    """Sort index idx_i by its nonzero count, descending, stable (reference :820-857).

    j (idx 1) is never reordered."""
    if op.op_type in ['spmm', 'sddmm'] and idx_i == 1:
        op.idx_reordered[area_i][idx_i] = False
        return
    if op.op_type in ['spmm', 'sddmm']:
        idx_values = op.idx_values_list[area_i]
        A = op.inps[0]
        if idx_i == 0:
            nnzs = np.diff(A.indptr)[idx_values[0]]
            # counts restricted to the op's k set (all k here)
        else:
            nnzs = np.bincount(A.indices, minlength=A.shape[1])[idx_values[2]]
        order = np.argsort(nnzs * (-1), kind='stable')
        op.idx_values_list[area_i][idx_i] = np.array(idx_values[idx_i])[order]
        op.idx_reordered[area_i][idx_i] = True


def reorder_for_strided_pattern(op):
    """Order rows by the position of their first nonzero (reference :869-880).
    Its result is discarded by the candidate generation (reference :1541-1549)."""
    if op.op_type == 'spmm':
        A = op.inps[0]
        first = np.where(np.diff(A.indptr) > 0, A.indices[np.minimum(A.indptr[:-1], max(A.nnz - 1, 0))], np.inf)
        op.idx_values_list[0][0] = np.argsort(first, kind='stable')


def _build_entries(op, state=None):
    """Populate the op's id-based position space in its own row/col order."""
    iv = op.idx_values_list[0]
    A = op.inps[0]
    rows_perm = np.asarray(iv[0])
    inv_col = np.empty(A.shape[1], np.int64)
    inv_col[np.asarray(iv[2])] = np.arange(len(iv[2]))
    counts = np.diff(A.indptr)[rows_perm]
    starts = A.indptr[rows_perm]
    ent_row = np.repeat(np.arange(len(rows_perm), dtype=np.int64), counts)
    offs = np.arange(len(ent_row), dtype=np.int64) - np.repeat(np.cumsum(counts) - counts, counts)
    base_pos = np.repeat(starts.astype(np.int64), counts) + offs       # position in A's CSR arrays
    ent_col = inv_col[A.indices[base_pos]]
    # sort entries by (row, col) in the op space (k may be permuted)
    order = np.lexsort((ent_col, ent_row))
    ent_row, ent_col, base_pos = ent_row[order], ent_col[order], base_pos[order]
    ids = op.ori_nnz_ids[base_pos] if op.ori_nnz_ids is not None else base_pos
    op._ent_row = ent_row
    op._ent_col = ent_col
    op._ent_id = ids.astype(np.int64)
    op._nrows = len(rows_perm)
    ip = np.zeros(op._nrows + 1, np.int64)
    ip[1:] = np.cumsum(np.bincount(ent_row, minlength=op._nrows))
    op._indptr = ip
    if state is not None:
        op.state = state


def gen_position_space_for_area(op, area_i):
    """Position space of the op in its own index order (reference :900-925)."""
    if area_i in op.position_space:
        return
    if op.op_type in ['spmm', 'sddmm']:
        if op._ent_id is None:
            _build_entries(op)
        op.position_space[area_i] = op._space_csr()
        op.inp_getnnz = [np.diff(op._indptr)]


def SPMM_Gustavson_product(op, area_i):
    op.loop_order[area_i] = [0, 2, 1]
    op.loop_protocals = {area_i: 'uuc'}
    op.inp_protocals = {area_i: ['uc', 'uu']}


def SDDMM_Gustavson_product(op, area_i):
    op.loop_order[area_i] = None
    op.loop_protocals = {area_i: 'cu'}
    op.inp_protocals = {area_i: ['c', 'uu', 'uu']}


class KOrder(object):
    """Condensed column order of one 16-row window: the window's nonempty
    columns (in the order chosen by condense_for_TC_tile) followed by every
    empty column in ascending order.  Behaves like the reference's per-window
    argsort array (indexing by int, slice or int array; len = K) without
    materialising K entries per window."""

    def __init__(self, head, K):
        self.head = np.asarray(head, dtype=np.int64)
        self.K = int(K)
        self._order = np.argsort(self.head, kind='stable')
        self._sorted = self.head[self._order]
        self._gap = self._sorted - np.arange(len(self._sorted))   # empties before each head column

    def __len__(self):
        return self.K

    def _empty_at(self, j):
        j = np.asarray(j, dtype=np.int64)
        return j + np.searchsorted(self._gap, j, side='right')

    def __getitem__(self, idx):
        if isinstance(idx, slice):
            idx = np.arange(*idx.indices(self.K))
        scalar = np.ndim(idx) == 0
        idx = np.asarray(idx, dtype=np.int64)
        if np.any((idx < 0) | (idx >= self.K)):
            raise IndexError("condensed position out of range")
        n = len(self.head)
        out = np.where(idx < n, self.head[np.minimum(idx, max(n - 1, 0))] if n else 0,
                       self._empty_at(np.maximum(idx - n, 0)))
        return int(out) if scalar else out

    def __array__(self, dtype=None, copy=None):
        a = self[np.arange(self.K)]
        return a.astype(dtype) if dtype is not None else a

    def pos(self, cols):
        """Condensed position of each column in cols (inverse permutation)."""
        cols = np.asarray(cols, dtype=np.int64)
        n = len(self.head)
        if n:
            loc = np.searchsorted(self._sorted, cols)
            loc_c = np.minimum(loc, n - 1)
            hit = self._sorted[loc_c] == cols
            head_pos = self._order[loc_c]
        else:
            hit = np.zeros(len(cols), bool)
            head_pos = np.zeros(len(cols), np.int64)
        before = np.searchsorted(self._sorted, cols)          # nonempty columns < c
        return np.where(hit, head_pos, n + cols - before)


class IdentityOrder(KOrder):
    """k order of a window when columns are not condensed (TC_k_notsorted)."""

    def __init__(self, K):
        super().__init__(np.zeros(0, np.int64), K)

    def __getitem__(self, idx):
        if isinstance(idx, slice):
            return np.arange(*idx.indices(self.K))
        return idx if np.ndim(idx) == 0 else np.asarray(idx, dtype=np.int64)

    def pos(self, cols):
        return np.asarray(cols, dtype=np.int64)


def condense_for_TC_tile(op, TC_tile_sizes, TC_k_notsorted, reorder_k_by_nnz=True):
    """Per 16-row window, order the columns by their nonzero count in the
    window, descending and stable (or just nonempty-first when
    reorder_k_by_nnz is False), so nonempty columns are condensed to the front
    (reference :1004-1041).  op.k_vals[w] is a KOrder."""
    if op.op_type in ['spmm', 'sddmm']:
        row_num = len(op.idx_values_list[0][0])
        blk_i_size = TC_tile_sizes[0][1]
        row_blk_num = math.ceil(row_num / blk_i_size)
        kn = len(op.idx_values_list[0][2])
        if TC_k_notsorted:
            op.k_vals = [IdentityOrder(kn) for _ in range(row_blk_num)]
            return
        A = op.inps[0]
        iv = op.idx_values_list[0]
        inv_col = np.empty(A.shape[1], np.int64)
        inv_col[np.asarray(iv[2])] = np.arange(kn)
        rows_perm = np.asarray(iv[0])
        counts = np.diff(A.indptr)[rows_perm]
        starts = A.indptr[rows_perm].astype(np.int64)
        ent_row = np.repeat(np.arange(row_num, dtype=np.int64), counts)
        offs = np.arange(len(ent_row)) - np.repeat(np.cumsum(counts) - counts, counts)
        cols = inv_col[A.indices[np.repeat(starts, counts) + offs]]
        win = ent_row // blk_i_size
        key = win * kn + cols
        ukey, ucnt = np.unique(key, return_counts=True)
        uw, uc = ukey // kn, ukey % kn
        if reorder_k_by_nnz:
            order = np.lexsort((uc, -ucnt, uw))
        else:
            order = np.lexsort((uc, uw))
        uw, uc = uw[order], uc[order]
        bounds = np.searchsorted(uw, np.arange(row_blk_num + 1))
        op.k_vals = [KOrder(uc[bounds[w]:bounds[w + 1]], kn) for w in range(row_blk_num)]


def change_to_dense(op, area_i, TC_tile_sizes, TC_k_notsorted, reorder_k_by_nnz=True):
    op.loop_protocals[area_i] = 'u' * len(op.idxs)
    op.inp_protocals[area_i] = ['u' * len(p) for p in op.inp_protocals[area_i]]
    if op.op_type == 'sddmm':
        op.inp_protocals[area_i] = 'uuu'
    condense_for_TC_tile(op, TC_tile_sizes, TC_k_notsorted, reorder_k_by_nnz=reorder_k_by_nnz)


def pad_irregular_idx(op, area_i, idx_i, max_bucket_size):
    """Split every row of the row-sorted op into max(ceil(nnz/mbs), 1) pieces
    of at most max_bucket_size nonzeros (reference :1065-1121).

    Sets hyb_new_rows (row of each piece), hyb_ori_row_ends (last piece of
    each row), hyb_short_begin_row (first piece of the first row with
    nnz <= mbs), hyb_getnnz_minus and the piece-level position space."""
    op.loop_protocals[area_i] = op.loop_protocals[area_i][:idx_i] + 'p' + op.loop_protocals[area_i][idx_i + 1:]
    if op.op_type == 'spmm':
        gen_position_space_for_area(op, area_i)
        getnnz = np.diff(op._indptr)
        new_row_nums = np.maximum(np.ceil(getnnz / max_bucket_size), 1).astype('int')
        op.hyb_new_rows = [np.repeat(np.arange(len(op.idx_values_list[area_i][0])), new_row_nums)]
        op.hyb_ori_row_ends = np.cumsum(new_row_nums) - 1
        short_begin_row_ori = np.searchsorted(-getnnz, -max_bucket_size)
        if short_begin_row_ori == len(getnnz):
            op.hyb_short_begin_row = len(op.hyb_new_rows[0])
        else:
            op.hyb_short_begin_row = op.hyb_ori_row_ends[short_begin_row_ori]
        # piece boundaries: row r's entries start at old indptr[r] and are cut every mbs
        piece_first = np.cumsum(new_row_nums) - new_row_nums          # first piece of each row
        n_pieces = int(new_row_nums.sum())
        piece_row = op.hyb_new_rows[0]
        piece_k = np.arange(n_pieces) - piece_first[piece_row]          # piece index within its row
        start = op._indptr[piece_row] + piece_k * max_bucket_size
        end = np.minimum(op._indptr[piece_row + 1], start + max_bucket_size)
        end = np.maximum(end, start)
        indptr = np.zeros(n_pieces + 1, np.int64)
        indptr[:-1] = start
        indptr[-1] = op._indptr[-1]
        # entries keep their order; re-label their row with the piece id
        ent_piece = np.repeat(np.arange(n_pieces, dtype=np.int64), end - start)
        op._ent_row = ent_piece
        op._nrows = n_pieces
        op._indptr = indptr
        op.position_space[area_i] = op._space_csr()
        op.inp_getnnz = [np.diff(indptr)]
        op.hyb_getnnz_minus = -op.inp_getnnz[0][op.hyb_short_begin_row:]
        op.hyb_ori_row_ends = np.array(op.hyb_ori_row_ends)


def get_1D_position_space(op, max_bucket_size):
    """1D SDDMM op: nonzeros in CSR order, cut into buckets of max_bucket_size
    (reference :1128-1147); nnz_update_1D = uncovered count per bucket."""
    assert op.op_type == 'sddmm'
    gen_position_space_for_area(op, 0)
    op.nnz_data_1D = np.ones(len(op._ent_id), dtype=np.int64)
    op.nnz_update_1D = np.add.reduceat(op.nnz_data_1D, np.arange(0, len(op.nnz_data_1D), op.max_bucket_size)) \
        if len(op.nnz_data_1D) else np.zeros(0, np.int64)


def refresh_1D_counts(op):
    """Recompute nnz_data_1D / nnz_update_1D from the shared coverage."""
    op.nnz_data_1D = op.uncovered_mask().astype(np.int64)
    if len(op.nnz_data_1D):
        op.nnz_update_1D = np.add.reduceat(op.nnz_data_1D, np.arange(0, len(op.nnz_data_1D), op.max_bucket_size))
    else:
        op.nnz_update_1D = np.zeros(0, np.int64)


# ---------------------------------------------------------------------------
# Tile-size candidates.
# ---------------------------------------------------------------------------

def get_factors(val):
    return [i for i in range(1, val + 1) if val % i == 0]


def get_tiles(val, split_n, base):
    if split_n == 2:
        return [base + (i,) for i in get_factors(val)]
    ret = []
    for i in get_factors(val):
        ret += get_tiles(val // i, split_n - 1, base + (i,))
    return ret


def tile_size_heuristics(op_type, tile_sizes_list, max_bucket_size, is_tb_tile=False):
    """ELL thread-block tiles: I*32 threads a multiple of 128 and >= 256, and
    every tile the same workload I*K = 8*max_bucket_size (reference :1188-1242)."""
    if op_type != 'spmm':
        raise AssertionError(f"We do not support {op_type} now!")
    ret = []
    THREAD_num = 256
    WORKLOAD = max_bucket_size * 8
    for tile_sizes in tile_sizes_list:
        if is_tb_tile:
            thread_num = tile_sizes[0][1] * tile_sizes[1][1]
            if thread_num % 128 != 0 or thread_num < THREAD_num:
                continue
            if tile_sizes[0][1] * tile_sizes[2][1] != WORKLOAD:
                continue
            ret.append(tile_sizes)
        else:
            thread_num = tile_sizes[0][2] * tile_sizes[1][2]
            if thread_num % 32 != 0 or thread_num > 1024:
                continue
            if (tile_sizes[0][2] != 1) or (tile_sizes[1][2] != 32):
                continue
            ret.append(tile_sizes)
    return ret


def get_TC_tile_sizes():
    """TC tile shapes (16 rows, 32 columns of N, K = 16*i) for SpMM, i = 10..1;
    (16, 16*2^j, 16) for SDDMM (reference :1252-1281)."""
    cfg = get_config()
    if cfg.TC_tile_sizes is not None:
        return [tuple(tuple(j) for j in cfg.TC_tile_sizes)]
    if cfg.TC_k_notsorted:
        return [((None, 16), (None, 32), (None, 16 * i)) for i in range(1, 2)]
    ret = [((None, 16), (None, 32), (None, 16 * i)) for i in range(160 // 16, 0, -1)]
    if cfg.op_type == 'sddmm':
        ret = [((None, 16), (None, 16 * (2 ** j)), (None, 16 * i)) for i in range(1, 2) for j in range(3)]
    return ret


def get_template_str(op):
    if op.op_type in ('spmm', 'batched_spmm'):
        if op.loop_protocals[op.this_area_i][2] == 'c':
            return "sparse_template"
        elif op.loop_protocals[op.this_area_i][2] == 'p':
            return "sparse_template_ell"
        elif op.loop_protocals[op.this_area_i] == 'uuu':
            return "TensorCore_template"
    elif op.op_type == 'sddmm':
        if op.loop_protocals[0] == 'uuu':
            return "TC_sddmm"
        return "1D_sddmm"


def get_tile_sizes_idx(sub_op, max_bucket_size, is_tb_tile=False):
    """Candidate tile sizes per index (reference :1926-1977)."""
    area_i = sub_op.this_area_i
    tile_level_num = 2 if is_tb_tile else 3
    if sub_op.op_type == 'spmm':
        lp = sub_op.loop_protocals[area_i]
        if lp[2] == 'c':
            return [get_tiles(len(sub_op.idx_values_list[area_i][i]), tile_level_num, (None,)) for i in [0, 1]] \
                + [[(1, len(sub_op.idx_values_list[area_i][2]))]]
        elif lp[2] == 'p':
            tot_lens = [math.ceil(len(sub_op.hyb_new_rows[0]) / 128) * 128, len(sub_op.idx_values_list[area_i][1])]
            idx = [get_tiles(tot_lens[i], tile_level_num, (None,)) for i in [0, 1]] \
                + [[(1, 2 ** t) for t in range(int(math.log(max_bucket_size, 2)) + 1)]]
            idx[0] = [(None, 2 ** t) for t in range(int(math.log(max(len(sub_op.hyb_new_rows[0]), 1), 2)) + 1)]
            return idx
        elif lp == 'uuu':
            return get_TC_tile_sizes()
    elif sub_op.op_type == 'sddmm':
        if get_template_str(sub_op) == 'TC_sddmm':
            idx = get_TC_tile_sizes()
            tile_j = set(ts[1][1] for ts in idx)
            cands = [j for j in tile_j if j <= sub_op.idx_lens[1]]
            tile_j = max(cands) if cands else min(tile_j)
            return [ts for ts in idx if ts[1][1] == tile_j]
        return [(max_bucket_size,)]


def get_tile_sizes_list(sub_op, tile_sizes_idx, max_bucket_size, is_tb_tile=False):
    area_i = sub_op.this_area_i
    if sub_op.op_type == 'spmm':
        if sub_op.loop_protocals[area_i] != 'uuu':
            lst = list(itertools.product(*tile_sizes_idx))
            return tile_size_heuristics(sub_op.op_type, lst, max_bucket_size, is_tb_tile=is_tb_tile)
        return tile_sizes_idx
    return tile_sizes_idx


def ell_tile_shapes(sub_op, max_bucket_size):
    """The (I, 32, K) thread-block ELL shapes the search enumerates
    (get_tile_sizes_idx with j fixed to 32, then tile_size_heuristics)."""
    idx = get_tile_sizes_idx(sub_op, max_bucket_size, is_tb_tile=True)
    idx[1] = [(None, 32)]
    return get_tile_sizes_list(sub_op, idx, max_bucket_size, is_tb_tile=True)


def get_params_list(sub_op, tile_sizes, max_bucket_size, is_tb_tile=False):
    """Per-template parameters (reference :2036-2068)."""
    params_list = []
    area_i = sub_op.this_area_i
    vals = [True] if is_tb_tile else [True, False]
    if sub_op.op_type == 'spmm':
        for use_implicit_unroll in vals:
            params = {'use_implicit_unroll': use_implicit_unroll}
            lp = sub_op.loop_protocals[area_i]
            if lp[2] == 'p':
                params['last_bucket'] = math.prod(tile_sizes[2][1:]) == max_bucket_size
            elif lp == 'uuu':
                if not params['use_implicit_unroll']:
                    continue
                params['mma_shape_str'] = "m8n32k16" if tile_sizes == ((None, 8), (None, 32), (None, 16)) \
                    else "m16n16k16"
            params_list.append(params)
    elif sub_op.op_type == 'sddmm':
        if get_template_str(sub_op) == '1D_sddmm':
            params_list.append({'tx': 8, 'ty': 4, 'vec_size': 4, 'group_size': 8, 'max_bucket_size': max_bucket_size})
        else:
            params_list.append({'mma_shape_str': "m16n16k16", 'warp_num': 1, 'vec_size': 4})
    return params_list


# ---------------------------------------------------------------------------
# TC candidates.
# ---------------------------------------------------------------------------

def _tc_window_stats(op, blk_i_size=16, blk_k_size=16):
    """Per window: number of nonempty columns (initial position space)."""
    row_num = len(op.idx_values_list[0][0])
    row_blk_num = math.ceil(row_num / blk_i_size)
    win = op._ent_row // blk_i_size
    key = win * op.idx_lens[2] + op._ent_col
    u = np.unique(key)
    nonempty = np.bincount(u // op.idx_lens[2], minlength=row_blk_num)
    return nonempty, row_blk_num


def select_the_densest_tensor_for_TC_tiles(TC_ops):
    """The TC candidate with the highest density nnz / (#nonempty 16x16
    condensed blocks * 256); strict '>' keeps the first (reference :1286-1309)."""
    best_density = 0
    best_TC_op = None
    for op in TC_ops:
        gen_position_space_for_area(op, 0)
        nonempty, _ = _tc_window_stats(op)
        tot = int(np.sum(np.ceil(nonempty / 16))) * 16 * 16
        density = len(op._ent_id) / tot if tot else 0.0
        if density > best_density:
            best_TC_op = op
            best_density = density
    if best_TC_op is None:
        best_TC_op = TC_ops[0]
    return best_TC_op


def set_init_non_empty_blk_num(op):
    """blk_nums[tile_sizes][w] = ceil(#nonempty columns of window w / K)
    (reference :1313-1335); with TC_k_notsorted, the number of nonempty
    16-column blocks and their ids (blk_ks)."""
    gen_position_space_for_area(op, 0)
    for tile_sizes in get_TC_tile_sizes():
        blk_i_size = tile_sizes[0][1]
        blk_k_size = tile_sizes[2][1]
        nonempty, row_blk_num = _tc_window_stats(op, blk_i_size)
        op.blk_nums[tile_sizes] = np.ceil(nonempty / blk_k_size).astype(np.int64)
        if op.TC_k_notsorted:
            win = op._ent_row // blk_i_size
            key = np.unique(win * ((op.idx_lens[2] + blk_k_size - 1) // blk_k_size) + op._ent_col // blk_k_size)
            nb = (op.idx_lens[2] + blk_k_size - 1) // blk_k_size
            op.blk_nums[tile_sizes] = np.bincount(key // nb, minlength=row_blk_num).astype(np.int64)
            op.blk_ks[tile_sizes] = [key[(key // nb) == w] % nb for w in range(row_blk_num)]


def _copy_op(op):
    new = copy.copy(op)
    new.idx_values_list = copy.deepcopy(op.idx_values_list)
    new.idx_reordered = copy.deepcopy(op.idx_reordered)
    new.loop_protocals = copy.deepcopy(op.loop_protocals)
    new.inp_protocals = copy.deepcopy(op.inp_protocals)
    new.position_space = dict()
    new.nnz_dict = dict()
    new._ent_row = new._ent_col = new._ent_id = new._indptr = None
    new.blk_nums = dict()
    new.blk_ks = dict()
    return new


def gen_candidate_TC_formats(op, TC_k_notsorted, reorder_k_by_nnz=True):
    """TC candidates: the op in its original row order and row-sorted, each
    condensed per 16-row window (reference :1341-1504, with the kernel
    tile sizes fixed to (1, 1, 1) and the k-reordered variants dropped as
    the reference does at step 3.5)."""
    cands = [_copy_op(op)]                  # original order
    sub = _copy_op(op)
    reorder_idx(sub, 0, 0)
    cands.append(sub)                       # i reordered (stable, by nnz)
    out = []
    tc_is = []
    for sub in cands:
        if op.op_type == 'spmm':
            SPMM_Gustavson_product(sub, 0)
        else:
            SDDMM_Gustavson_product(sub, 0)
        for TC_tile_sizes in [get_TC_tile_sizes()[0]]:
            new = _copy_op(sub)
            new.k_vals = []
            change_to_dense(new, 0, TC_tile_sizes, TC_k_notsorted, reorder_k_by_nnz=reorder_k_by_nnz)
            new.TC_k_notsorted = TC_k_notsorted
            out.append(new)
    return out


def _finalize_tc_ops(TC_ops, op_id_start):
    uniq, seen = [], []
    for sub in TC_ops:
        key = list(sub.idx_values_list[0][0])
        if key in seen:
            continue
        seen.append(key)
        sub.idx_reordered = copy.deepcopy(sub.idx_reordered)
        sub.idx_reordered[0][0] = not np.array_equal(sub.idx_values_list[0][0], np.sort(sub.idx_values_list[0][0]))
        uniq.append(sub)
    TC_op = select_the_densest_tensor_for_TC_tiles(uniq)
    set_init_non_empty_blk_num(TC_op)
    TC_op.op_id = op_id_start
    return TC_op, op_id_start + 1


def gen_candidate_formats_BSPMM(op, max_bucket_size, kernel_tile_size_options, TC_k_notsorted,
                                reorder_k_by_nnz=True, op_id_start=0, gen_TC_formats=True):
    """SpMM candidates: the row-sorted op padded into max_bucket_size pieces
    (the only ELL op the search keeps, my_search_formats.py:332-333) and the
    densest TC op (reference :1510-1833)."""
    ret = []
    ell = _copy_op(op)
    reorder_idx(ell, 0, 0)
    reorder_idx(ell, 0, 1)
    SPMM_Gustavson_product(ell, 0)
    pad_irregular_idx(ell, 0, 2, max_bucket_size)
    ell.op_id = op_id_start
    op_id_start += 1
    ret.append(ell)
    if gen_TC_formats:
        tc_ops = gen_candidate_TC_formats(op, TC_k_notsorted, reorder_k_by_nnz=reorder_k_by_nnz)
        TC_op, op_id_start = _finalize_tc_ops(tc_ops, op_id_start)
        ret.append(TC_op)
    return ret, op_id_start


def gen_candidate_formats_BSDDMM(op, max_bucket_size, TC_k_notsorted,
                                 reorder_k_by_nnz=True, op_id_start=0, gen_TC_formats=True):
    """SDDMM candidates: one 1D op (buckets of max_bucket_size nonzeros in CSR
    order) and the densest TC op (reference :1837-1898)."""
    op1 = _copy_op(op)
    SDDMM_Gustavson_product(op1, 0)
    op1.max_bucket_size = max_bucket_size
    get_1D_position_space(op1, max_bucket_size)
    op1.op_id = op_id_start
    op_id_start += 1
    ret = [op1]
    if gen_TC_formats:
        tc_ops = gen_candidate_TC_formats(op, TC_k_notsorted, reorder_k_by_nnz=reorder_k_by_nnz)
        TC_op, op_id_start = _finalize_tc_ops(tc_ops, op_id_start)
        ret.append(TC_op)
    return ret, op_id_start


def gen_candidate_formats(op, max_bucket_size, kernel_tile_size_options, TC_k_notsorted,
                          reorder_k_by_nnz=True, op_id_start=0, gen_TC_formats=True):
    if op.op_type in ['spmm', 'batched_spmm']:
        return gen_candidate_formats_BSPMM(op, max_bucket_size, kernel_tile_size_options, TC_k_notsorted,
                                           reorder_k_by_nnz=reorder_k_by_nnz, op_id_start=op_id_start,
                                           gen_TC_formats=gen_TC_formats)
    elif op.op_type in ['sddmm']:
        return gen_candidate_formats_BSDDMM(op, max_bucket_size, TC_k_notsorted,
                                            reorder_k_by_nnz=reorder_k_by_nnz, op_id_start=op_id_start,
                                            gen_TC_formats=gen_TC_formats)


# ---------------------------------------------------------------------------
# Tiles.
# ---------------------------------------------------------------------------

class ComputeTile(object):
    """A selected computation tile (reference :2333-2485).

    tile_i_rng / tile_k_rng are inclusive; the k start of a TC tile may be a
    Fraction and is truncated with int() as in the reference (:2384-2387).
    covered_ids holds the original nonzero ids the tile covers once selected
    (the id-based form of position_space_when_selected)."""

    def __init__(self, op, tile_sizes, tile_pos, params, not_init_nnz=False):
        self.op = op
        self.tile_sizes = tile_sizes
        self.tile_pos = tile_pos
        self.params = params
        self.nnz = 0
        self.cost = None
        self.pred_cost = None
        self.j_num = None
        self.nnz_uncovered = 0
        self.avg_cost = None
        self.pred_avg_cost = None
        self.best_tile_sizes = None
        self.best_params = None
        self.nnz_when_selected = None
        self.tile_i_rng = None
        self.tile_k_rng = None
        self.k_vals = None
        self.i_vals = None
        self.is_atomic_tile = None
        self.remap_1d = None
        self.covered_ids = None
        self.covered_rows = None   # rows (op space) of covered_ids
        self.covered_cols = None   # cols (op k space) of covered_ids
        tpl = get_template_str(op)
        if op.op_type == 'spmm' or tpl == 'TC_sddmm':
            I = math.prod(tile_sizes[0][1:])
            K = math.prod(tile_sizes[2][1:])
            nrows = op._nrows if op._nrows is not None else op.idx_lens[0]
            self.tile_i_rng = (I * tile_pos[0], min(I * (tile_pos[0] + 1), nrows) - 1)
            kr = (K * tile_pos[2], min(K * (tile_pos[2] + 1), op.idx_lens[2]) - 1)
            self.tile_k_rng = (int(kr[0]), int(kr[1]))
            set_j_len(self)
            if not_init_nnz:
                return
            init_nnz(self)
            self.nnz_uncovered = self.nnz * self.j_num
        elif tpl == '1D_sddmm':
            set_j_len(self)
            if not_init_nnz:
                return
            init_nnz(self)
            self.nnz_uncovered = self.nnz * self.j_num

    # entries (indices into op._ent_*) of the tile in the op's position space
    def entries(self):
        op = self.op
        tpl = get_template_str(op)
        if tpl == '1D_sddmm':
            b = self.tile_sizes[0]
            s = self.tile_pos[0] * b
            return np.arange(s, min(s + b, len(op._ent_id)))
        lo, hi = op._indptr[self.tile_i_rng[0]], op._indptr[self.tile_i_rng[1] + 1]
        e = np.arange(lo, hi)
        if tpl in ('TensorCore_template', 'TC_sddmm'):
            w = self.tile_pos[0]
            cpos = getattr(op, "_ent_cpos", None)
            cp = cpos[e] if cpos is not None else _tc_condensed_pos(op, w, op._ent_col[e])
            e = e[(cp >= self.tile_k_rng[0]) & (cp <= self.tile_k_rng[1])]
        return e

    def get_k_vals(self):
        if get_template_str(self.op) == 'TensorCore_template':
            self.k_vals = self.op.idx_values_list[0][2][self.tile_k_rng[0]:self.tile_k_rng[1] + 1]

    def update_nnz_uncovered(self):
        e = self.entries()
        self.nnz_uncovered = self.j_num * int(np.count_nonzero(self.op.state.uncovered[self.op._ent_id[e]]))

    def set_avg_cost(self):
        if self.cost is None:
            return
        self.avg_cost = float('inf') if self.nnz_uncovered == 0 else self.cost / self.nnz_uncovered

    def set_pred_avg_cost(self):
        if self.pred_cost is None:
            return
        self.pred_avg_cost = self.pred_cost / np.float32(self.nnz_uncovered) \
            if self.nnz_uncovered else float('inf')

    def get_key(self, as_dict_key=False):
        if as_dict_key:
            import json
            return (self.op.op_id, tuple(self.tile_sizes), tuple(self.tile_pos), json.dumps(self.params),
                    self.is_atomic_tile)
        return (self.op.op_id, self.tile_sizes, self.tile_pos, self.params, self.is_atomic_tile)

    # -- reference-shaped views (for the packers and the oracle) -------------
    @property
    def position_space_when_selected(self):
        """scipy CSR of the covered entries over the tile's rows (op space),
        TC tiles: columns in the tile's condensed order; data = A's values."""
        return tile_covered_csr(self)


def _tc_condensed_pos(op, w, cols):
    """Condensed position, in window w's order, of each column in cols."""
    return op.k_vals[w].pos(cols)


def tile_covered_csr(tile, values=True):
    op = tile.op
    tpl = get_template_str(op)
    ids = tile.covered_ids if tile.covered_ids is not None else np.zeros(0, np.int64)
    data = op.state.A.data[ids].astype(np.float64) if values else np.ones(len(ids))
    if tpl == 'sparse_template_ell':
        nrows = tile.tile_i_rng[1] - tile.tile_i_rng[0] + 1
        r = tile.covered_rows - tile.tile_i_rng[0]
        c = tile.covered_cols
        shape = (nrows, op.idx_lens[2])
    elif tpl in ('TensorCore_template', 'TC_sddmm'):
        nrows = tile.tile_i_rng[1] - tile.tile_i_rng[0] + 1
        r = tile.covered_rows - tile.tile_i_rng[0]
        c = _tc_condensed_pos(op, tile.tile_pos[0], tile.covered_cols) - tile.tile_k_rng[0]
        shape = (nrows, tile.tile_k_rng[1] - tile.tile_k_rng[0] + 1)
    else:
        return data
    order = np.lexsort((c, r))
    return scipy.sparse.csr_matrix((data[order], (r[order], c[order])), shape=shape)


def set_j_len(tile):
    if tile.op.op_type == 'spmm':
        J = math.prod(tile.tile_sizes[1][1:])
        lo = tile.tile_pos[1] * J
        tile.j_num = min(lo + J, len(tile.op.idx_values_list[0][1])) - lo
    elif tile.op.op_type == 'sddmm':
        tile.j_num = tile.op.idx_lens[1]


def init_nnz(tile):
    """Number of nonzeros of the op's initial position space in the tile."""
    op = tile.op
    tpl = get_template_str(op)
    if tpl in ("TensorCore_template", 'TC_sddmm', 'sparse_template_ell'):
        tile.nnz = len(tile.entries())
    elif tpl == '1D_sddmm':
        b = tile.tile_sizes[0]
        tile.nnz = min((tile.tile_pos[0] + 1) * b, len(op._ent_id)) - tile.tile_pos[0] * b


# ---------------------------------------------------------------------------
# On-disk graphs.
# ---------------------------------------------------------------------------

def load_snap_graph(op_type, name="arxiv", feat_size=32, pad=True, delimiter=None, seed=0):
    """Op from a SNAP-style edge list file (gen_formats_v2.py:2629-2705).

    '.txt' files are whitespace separated ints, '.csv' comma separated, any
    other name is read as floats with `delimiter`; only the first two columns
    are used.  Duplicate edges are dropped, node ids are relabelled to
    0..n-1 in sorted order of the ids that occur, and the CSR of ones has
    shape (ceil(rows/128)*128, ceil(cols/128)*128) for spmm and
    (ceil(n/128)*128, ceil(n/160)*160) for sddmm when pad, (n, n) otherwise.
    Dense operands are U[0,1) FP32 from default_rng(seed), as in the
    reference (B: K x feat for spmm; X: M x feat and Y: K x feat for sddmm).
    """
    if ".txt" in name:
        a = np.loadtxt(name, dtype=np.int64, comments="#", ndmin=2)
    elif ".csv" in name:
        a = np.loadtxt(name, dtype=np.int64, delimiter=",", comments="#", ndmin=2)
    else:
        a = np.loadtxt(name, dtype=np.float64, delimiter=delimiter, comments="#", ndmin=2)
    a = a[:, :2].astype(np.int64)
    a = np.unique(a, axis=0).reshape(-1)
    _, ids = np.unique(a, return_inverse=True)
    starts, ends = ids.reshape(-1, 2).T
    node_num = int(max(starts.max(), ends.max())) + 1
    A = scipy.sparse.coo_matrix((np.ones(len(starts), np.float32), (starts, ends)),
                                shape=(node_num, node_num)).tocsr()
    if pad:
        rows = A.shape[0]
        cols = int(A.indices.max()) + 1 if A.nnz else 0
        if op_type == 'spmm':
            idx_lens = [math.ceil(rows / 128) * 128, feat_size, math.ceil(cols / 128) * 128]
        else:
            idx_lens = [math.ceil(node_num / 128) * 128, feat_size, math.ceil(node_num / 160) * 160]
    else:
        idx_lens = [node_num, feat_size, node_num]
    indptr = np.concatenate([A.indptr, np.full(idx_lens[0] + 1 - len(A.indptr), A.indptr[-1])])
    A = scipy.sparse.csr_matrix((np.ones(A.nnz, np.float32), A.indices, indptr), shape=(idx_lens[0], idx_lens[2]))
    A.sum_duplicates()
    A.data[:] = 1.0
    A = _as_csr(A)
    rng = np.random.default_rng(seed=seed)
    if op_type == 'spmm':
        Bs = [rng.random((idx_lens[2], feat_size)).astype(np.float32)]
    else:
        Bs = [rng.random((idx_lens[0], feat_size)).astype(np.float32),
              rng.random((idx_lens[2], feat_size)).astype(np.float32)]
    return Op(op_type, [0, 1], [2], idx_lens, None, [A] + [DenseTensor(B) for B in Bs], [A.tocsc()])

"""CPU known answers, computed by hand, for the packed buffers on the 6 x 8
matrix of tests/test_oracle_formats.py, including every padding slot:

  row0: cols 1,3    row1: (empty)    row2: cols 0..4    row3: col 7
  row4: cols 2,5,6  row5: cols 0..4  (value ids 0..15 in CSR order)

Padding rules (reference MY_sparse/my_fuse_formats.py):
  ELL indices_k padding        = the op's first column id (:3377-3381)
  ELL indices_reordered_i pad  = the tile's last row id i_vals[-1] (:3382-3390)
  TC indices padding           = the tile's last column id (:3449-3467)
  SDDMM 1D padding             = repeat of the last entry up to a multiple
                                 of ty * group_size = 32 (:3579-3588)
Both the product packers and the oracle's transcriptions are checked.  Re-
expansion skips padding slots (their value is 0), so only these answers pin
the padding values.
"""
import json
import types

import numpy as np
import scipy.sparse

from oracle import formats as of
from stile_amd import api, my_fuse_formats
from tests.test_oracle_formats import tiny


def _ell_groups(opts):
    A = tiny()
    pack, tiles, op = api.spmm_pack(A, opts, N=32)
    formats = my_fuse_formats.get_formats_from_selected_tiles(tiles, ['A'], opts.dsize)
    refs = []
    for fmt, ftiles in formats.items():
        assert fmt[0] == "sparse_template_ell"
        refs.append(of.pack_ell_group(ftiles, fmt[1][0][1], fmt[1][2][1]))
    return pack, refs


def test_ell_pack_known_answer_bucket4():
    # ELL op: rows by nnz descending (stable) 2, 5, 4, 0, 3, 1; bucket 4 splits
    # rows 2 and 5 into pieces of 4 + 1: 8 pieces, one (8, 32, 4) tile.
    pack, refs = _ell_groups(api.SearchOptions(only_ELL=True, max_bucket_size=4))
    assert len(pack["ell"]) == 1 and not pack["tc"]
    g = pack["ell"][0]
    want_k = [[0, 1, 2, 3], [4, 0, 0, 0], [0, 1, 2, 3], [4, 0, 0, 0],
              [2, 5, 6, 0], [1, 3, 0, 0], [7, 0, 0, 0], [0, 0, 0, 0]]
    want_i = [2, 2, 5, 5, 4, 0, 3, 1]
    want_a = (np.array(want_k) * 0 + np.array([[1, 1, 1, 1], [1, 0, 0, 0], [1, 1, 1, 1], [1, 0, 0, 0],
                                               [1, 1, 1, 0], [1, 1, 0, 0], [1, 0, 0, 0], [0, 0, 0, 0]]))
    want_ids = [[2, 3, 4, 5], [6, -1, -1, -1], [11, 12, 13, 14], [15, -1, -1, -1],
                [8, 9, 10, -1], [0, 1, -1, -1], [7, -1, -1, -1], [-1, -1, -1, -1]]
    assert (g["m_blk"], g["nnz_cols"], g["num_tiles"]) == (8, 4, 1)
    assert g["indices_k"].tolist() == want_k
    assert g["indices_reordered_i"].tolist() == want_i
    assert g["a"].reshape(8, 4).tolist() == want_a.tolist()
    assert g["a_ids"].reshape(8, 4).tolist() == want_ids
    r = refs[0]
    assert r["indices_k"].tolist() == want_k
    assert r["indices_reordered_i"].tolist() == want_i
    assert r["a"].reshape(8, 4).tolist() == want_a.tolist()


def test_ell_pack_known_answer_bucket2_sentinel_and_row_padding():
    # bucket 2: pieces 2,2,2 | 5,5,5 | 4,4 | 0 | 3 | 1 = 11 pieces, two (8, 32, 2)
    # tiles.  The empty row 1 is sent to the sentinel row M = 6
    # (my_branch_and_bound.py:1682-1700); the second tile holds 3 rows and
    # its 5 padding rows repeat its last row id (6).
    pack, refs = _ell_groups(api.SearchOptions(only_ELL=True, max_bucket_size=2))
    assert [g["num_tiles"] for g in pack["ell"]] == [1, 1]
    g0, g1 = pack["ell"]
    assert g0["indices_k"].tolist() == [[0, 1], [2, 3], [4, 0], [0, 1], [2, 3], [4, 0], [2, 5], [6, 0]]
    assert g0["indices_reordered_i"].tolist() == [2, 2, 2, 5, 5, 5, 4, 4]
    assert g1["indices_k"].tolist() == [[1, 3], [7, 0]] + [[0, 0]] * 6
    assert g1["indices_reordered_i"].tolist() == [0, 3, 6, 6, 6, 6, 6, 6]
    assert g1["a"].reshape(8, 2).tolist() == [[1, 1], [1, 0]] + [[0, 0]] * 6
    for mine, ref in zip(pack["ell"], refs):
        for k in ("a", "indices_k", "indices_reordered_i"):
            assert np.array_equal(mine[k], ref[k]), k


def test_tc_pack_known_answer():
    # one (16, 32, 16) tile on window 0; columns condensed by column nnz
    # descending (stable): 1, 2, 3 (3 each), 0, 4 (2), 5, 6, 7 (1); the 8
    # padding columns repeat the last column id, 7.
    A = tiny()
    opts = api.SearchOptions(only_TC=True)
    pack, tiles, op = api.spmm_pack(A, opts, N=32)
    assert len(pack["tc"]) == 1 and not pack["ell"]
    g = pack["tc"][0]
    assert g["indices"].tolist() == [[1, 2, 3, 0, 4, 5, 6, 7] + [7] * 8]
    assert (g["group_size"], g["num_windows"], g["nnzb"], g["rows_2d"]) == (16, 1, 1, False)
    assert g["i_indices"].tolist() == [0] and g["indptr"].tolist() == [0, 1]
    a = np.zeros((16, 16))
    for r, cs in {0: [1, 3], 2: [0, 1, 2, 3, 4], 3: [7], 4: [2, 5, 6], 5: [0, 1, 2, 3, 4]}.items():
        for c in cs:
            a[r, [1, 2, 3, 0, 4, 5, 6, 7].index(c)] = 1.0
    assert np.array_equal(g["a"][0], a)
    formats = my_fuse_formats.get_formats_from_selected_tiles(tiles, ['A'], opts.dsize)
    (fmt, ftiles), = formats.items()
    ref = of.pack_tc_group(ftiles, 16, 16, False)
    assert ref["indices"].tolist() == g["indices"].tolist()
    assert np.array_equal(ref["a"][0], a)


def test_sddmm_pack_known_answer():
    A = tiny()
    pack, _, _ = api.sddmm_pack(A, api.SearchOptions(only_ELL=True), Kf=16)
    (g,) = pack["g1d"]
    # 16 entries in CSR order, padded to 32 = ty * group_size by repeating the
    # last entry (row 5, column 4, output id 15)
    assert g["mid_idx"].tolist() == [0, 0, 2, 2, 2, 2, 2, 3, 4, 4, 4, 5, 5, 5, 5, 5] + [5] * 16
    assert g["j_idx"].tolist() == [1, 3, 0, 1, 2, 3, 4, 7, 2, 5, 6, 0, 1, 2, 3, 4] + [4] * 16
    assert g["remap"].tolist() == list(range(16)) + [15] * 16
    # one 16x16 TC block: rows 0..5 padded with 5, condensed columns padded
    # with 7; sampled entries in CSC order at position col + 16 * row
    pack, _, _ = api.sddmm_pack(A, api.SearchOptions(only_TC=True), Kf=16)
    (t,) = pack["tc"]
    assert t["i_indices"].tolist() == [[0, 1, 2, 3, 4, 5] + [5] * 10]
    assert t["j_indices"].tolist() == [[1, 2, 3, 0, 4, 5, 6, 7] + [7] * 8]
    assert t["indptr"].tolist() == [0, 16]
    assert t["indices_dense"].tolist() == [0, 32, 80, 33, 65, 81, 2, 34, 82, 35, 83, 36, 84, 69, 70, 55]
    assert t["indices"].tolist() == [0, 3, 12, 4, 8, 13, 1, 5, 14, 2, 11, 6, 15, 9, 10, 7]


def _fake_ell_tile(kvals, rows, covered):
    """A stand-in tile for the packers: an ELL op whose column order is
    kvals (not the identity) and a covered position space over it."""
    A = scipy.sparse.csr_matrix(covered, dtype=np.float32)
    op = types.SimpleNamespace(idx_values_list=[[np.arange(len(rows)), np.arange(1), np.asarray(kvals)]])
    return types.SimpleNamespace(op=op, i_vals=np.asarray(rows), position_space_when_selected=A)


def test_ell_padding_fill_is_the_ops_first_column():
    # the op's column order starts with column 5: every padding slot of
    # indices_k holds 5 (not 0), and the 2 padding rows of the 4-row tile
    # repeat the tile's last row id 9
    t = _fake_ell_tile([5, 3, 1], [7, 9], [[1, 0, 1], [0, 1, 0]])
    r = of.pack_ell_group([t], 4, 3)
    assert r["indices_k"].tolist() == [[5, 1, 5], [3, 5, 5], [5, 5, 5], [5, 5, 5]]
    assert r["indices_reordered_i"].tolist() == [7, 9, 9, 9]
    assert r["a"].reshape(4, 3).tolist() == [[1, 1, 0], [1, 0, 0], [0, 0, 0], [0, 0, 0]]


def test_group_keys_carry_the_reference_params():
    # the format keys (template, tile sizes, json params) of the bucket-4 plan
    A = tiny()
    opts = api.SearchOptions(only_ELL=True, max_bucket_size=4)
    _, tiles, _ = api.spmm_pack(A, opts, N=32)
    formats = my_fuse_formats.get_formats_from_selected_tiles(tiles, ['A'], opts.dsize)
    (fmt,) = formats.keys()
    params = json.loads(fmt[2])
    assert fmt[0] == "sparse_template_ell" and fmt[1] == ((None, 8), (None, 32), (1, 4))
    assert params["last_bucket"] is True and params["use_implicit_unroll"] is True
    assert params["real_atomic"] is True       # rows 2 and 5 are split over two tile rows

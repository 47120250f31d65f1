"""CPU tests of the oracle's format restatement (hand-computed known answers)
and of the product's vectorised format construction against it (bit-exact).

Tiny matrix used for the known answers (6 x 8):
  row0: cols 1,3        row1: (empty)        row2: cols 0,1,2,3,4
  row3: col 7           row4: cols 2,5,6     row5: cols 0,1,2,3,4
row nnz = [2,0,5,1,3,5]; column nnz = [2,3,3,3,2,1,1,1].
"""
import numpy as np
import pytest
import scipy.sparse

from oracle import formats as of
from stile_amd import datasets
from stile_amd import gen_formats_v2 as gf
from stile_amd import my_branch_and_bound as bb
from stile_amd import my_search_formats as sf


def tiny():
    rows = [0, 0, 2, 2, 2, 2, 2, 3, 4, 4, 4, 5, 5, 5, 5, 5]
    cols = [1, 3, 0, 1, 2, 3, 4, 7, 2, 5, 6, 0, 1, 2, 3, 4]
    return scipy.sparse.csr_matrix((np.ones(16, np.float32), (rows, cols)), shape=(6, 8))


def ident(A):
    return [np.arange(A.shape[0]), np.arange(1), np.arange(A.shape[1])]


# --- known answers (computed by hand from the reference's definitions) -------

def test_reorder_idx_known_answer():
    A = tiny()
    assert of.reorder_idx(A, ident(A), 0).tolist() == [2, 5, 4, 0, 3, 1]
    assert of.reorder_idx(A, ident(A), 2).tolist() == [1, 2, 3, 0, 4, 5, 6, 7]
    assert of.reorder_idx(A, ident(A), 1).tolist() == [0]


def test_pad_irregular_idx_known_answer():
    A = tiny()
    iv = [np.array([2, 5, 4, 0, 3, 1]), np.arange(1), np.arange(8)]
    r = of.pad_irregular_idx(A, iv, 2)
    assert r["hyb_new_rows"].tolist() == [0, 0, 0, 1, 1, 1, 2, 2, 3, 4, 5]
    assert r["hyb_ori_row_ends"].tolist() == [2, 5, 7, 8, 9, 10]
    assert r["hyb_short_begin_row"] == 8
    assert r["position_space"].indptr.tolist() == [0, 2, 4, 5, 7, 9, 10, 12, 13, 15, 16, 16]
    assert r["hyb_getnnz_minus"].tolist() == [-2, -1, 0]


def test_condense_and_blocks_known_answer():
    A = tiny()
    kv = of.condense_for_TC_tile(A, ident(A))
    assert len(kv) == 1 and kv[0].tolist() == [1, 2, 3, 0, 4, 5, 6, 7]
    kv2 = of.condense_for_TC_tile(A, ident(A), reorder_k_by_nnz=False)
    assert kv2[0].tolist() == [0, 1, 2, 3, 4, 5, 6, 7]
    assert of.set_init_non_empty_blk_num(A, ident(A), [16, 32])[16].tolist() == [1]
    assert of.tc_density(A, ident(A)) == pytest.approx(16 / 256)


def test_valid_ell_positions_known_answer():
    # pieces of the tiny matrix with mbs = 2: shape (8, 32, 2) only
    assert of.get_valid_tile_pos_list_for_ELL(np.array([-2, -1, 0]), 8, 11, 8, 2, 2) == [(0, 0, 0), (1, 0, 0)]
    # K = 1 (not the bucket size): rows with nnz in (0, 1] start at piece 9;
    # the first 8-piece tile starting at or after it is ceil(9/8) = 2, beyond
    # the last such row (piece 9 -> (10-1)//8 + 1 = 2): no position
    assert of.get_valid_tile_pos_list_for_ELL(np.array([-2, -1, 0]), 8, 11, 8, 1, 2) == []


def test_bucket_and_padding_known_answer():
    assert of.estimate_max_bucket_size(16) == 4
    assert of.estimate_max_bucket_size(167772) == 32          # uniform 4096^2 1 % workload
    assert of.get_padded_j_TC(100, 16) == 128
    assert of.get_padded_j_TC(8, 16) == 16
    assert of.get_padded_j_TC(16, 16) == 16


# --- product vs oracle (bit-exact) -------------------------------------------

def _mats():
    rng = np.random.default_rng(7)
    out = [("tiny", tiny())]
    A = scipy.sparse.random(70, 90, density=0.08, format="csr", random_state=1, dtype=np.float32).tolil()
    A[3, :] = 0
    A[10, :60] = 1.0                                   # a long row (split into pieces)
    out.append(("random_split", A.tocsr()))
    out.append(("pruned", datasets.pruned_layer(seed=4, M=128, K=96, N=8)[0]))
    out.append(("logsparse", datasets.log_sparse_mask(100)))
    out.append(("rmat8", datasets.rmat_graph(8, seed=3)))
    return out


@pytest.mark.parametrize("name,A", _mats())
@pytest.mark.parametrize("mbs", [2, 8, 32])
def test_ell_op_matches_oracle(name, A, mbs):
    op = gf.Op('spmm', [0, 1], [2], [A.shape[0], 4, A.shape[1]], None,
               [A, gf.DenseTensor(np.zeros((A.shape[1], 4), np.float32))], None)
    ops, _ = gf.gen_candidate_formats(op, mbs, ([1], [1], [1]), False, gen_TC_formats=False)
    ell = ops[0]
    iv0 = of.reorder_idx(A, ident(A), 0)
    assert np.array_equal(np.asarray(ell.idx_values_list[0][0]), iv0)
    r = of.pad_irregular_idx(A, [iv0, np.arange(1), np.arange(A.shape[1])], mbs)
    assert np.array_equal(ell.hyb_new_rows[0], r["hyb_new_rows"])
    assert np.array_equal(ell.hyb_ori_row_ends, r["hyb_ori_row_ends"])
    assert ell.hyb_short_begin_row == r["hyb_short_begin_row"]
    assert np.array_equal(ell.hyb_getnnz_minus, r["hyb_getnnz_minus"])
    assert np.array_equal(ell.position_space[0].indptr, r["position_space"].indptr.astype(np.int64))
    assert np.array_equal(ell.position_space[0].indices, r["position_space"].indices)
    for ts in gf.ell_tile_shapes(ell, mbs):
        I, K = ts[0][1], ts[2][1]
        assert bb.get_valid_tile_pos_list_for_ELL(ell, ts, mbs) == of.get_valid_tile_pos_list_for_ELL(
            r["hyb_getnnz_minus"], r["hyb_short_begin_row"], len(r["hyb_new_rows"]), I, K, mbs)


@pytest.mark.parametrize("name,A", _mats())
@pytest.mark.parametrize("by_nnz", [True, False])
def test_tc_op_matches_oracle(name, A, by_nnz):
    op = gf.Op('spmm', [0, 1], [2], [A.shape[0], 4, A.shape[1]], None,
               [A, gf.DenseTensor(np.zeros((A.shape[1], 4), np.float32))], None)
    gf.set_config(gf.SearchConfig())
    ops, _ = gf.gen_candidate_formats(op, 8, ([1], [1], [1]), False, reorder_k_by_nnz=by_nnz)
    tc = ops[-1]
    ivs = [ident(A)[0], of.reorder_idx(A, ident(A), 0)]
    dens = [of.tc_density(A, [iv, np.arange(1), np.arange(A.shape[1])]) for iv in ivs]
    pick = 1 if dens[1] > dens[0] and not np.array_equal(ivs[0], ivs[1]) else 0
    assert np.array_equal(np.asarray(tc.idx_values_list[0][0]), ivs[pick])
    iv = [ivs[pick], np.arange(1), np.arange(A.shape[1])]
    kv = of.condense_for_TC_tile(A, iv, reorder_k_by_nnz=by_nnz)
    assert len(kv) == len(tc.k_vals)
    for w in range(len(kv)):
        assert np.array_equal(np.asarray(tc.k_vals[w]), kv[w])
    K_list = [ts[2][1] for ts in gf.get_TC_tile_sizes()]
    bn = of.set_init_non_empty_blk_num(A, iv, K_list)
    for ts in gf.get_TC_tile_sizes():
        assert np.array_equal(tc.blk_nums[ts], bn[ts[2][1]])


@pytest.mark.parametrize("nnz", [16, 1000, 60000, 167772, 2_000_000])
def test_estimate_max_bucket_size_matches_oracle(nnz):
    A = scipy.sparse.csr_matrix((np.ones(1), ([0], [0])), shape=(1, 1))
    op = gf.Op('spmm', [0, 1], [2], [1, 1, 1], None, [A, gf.DenseTensor(np.zeros((1, 1)))], None)
    op.inps[0] = scipy.sparse.random(1, max(nnz, 1), density=1.0, format="csr") if nnz <= 60000 else op.inps[0]
    if nnz > 60000:
        class _F:  # only .nnz is read
            pass
        f = _F()
        f.nnz = nnz
        op.inps[0] = f
    assert sf.estimate_max_bucket_size(op, sm_num=108, threads_per_sm=2048) == of.estimate_max_bucket_size(nnz)


def test_kvals_lazy_order_is_a_permutation():
    ko = gf.KOrder([5, 2, 9], 12)
    arr = np.asarray(ko)
    assert sorted(arr.tolist()) == list(range(12))
    assert arr[:3].tolist() == [5, 2, 9]
    assert np.array_equal(ko.pos(arr), np.arange(12))


def test_compute_tile_known_answers():
    """ComputeTile ranges and counts on the 6 x 8 matrix (N = 40): a TC tile
    (16, 32, 16) at j-block 0 spans rows 0..5, columns 0..7, all 16 nonzeros,
    32 output columns (uncovered work 16 * 32); at j-block 1 only 40 - 32 = 8
    columns remain.  The ELL op (bucket 4) splits the two 5-nonzero rows, so it
    has 8 rows; an (8, 32, 4) tile spans them all and holds every nonzero."""
    A = tiny()
    op = gf.Op('spmm', [0, 1], [2], [6, 40, 8], None, [A, gf.DenseTensor(np.zeros((8, 40), np.float32))], None)
    gf.set_config(gf.SearchConfig())
    ops, _ = gf.gen_candidate_formats(op, 4, ([1], [1], [1]), False, reorder_k_by_nnz=True)
    ell, tc = ops[0], ops[-1]
    ts = ((None, 16), (None, 32), (None, 16))
    t = gf.ComputeTile(tc, ts, (0, 0, 0), {"mma_shape_str": "m16n16k16"})
    assert (t.tile_i_rng, t.tile_k_rng, t.nnz, t.j_num, t.nnz_uncovered) == ((0, 5), (0, 7), 16, 32, 512)
    t = gf.ComputeTile(tc, ts, (0, 1, 0), {"mma_shape_str": "m16n16k16"})
    assert (t.nnz, t.j_num, t.nnz_uncovered) == (16, 8, 128)
    assert len(ell.idx_values_list[0][0]) == 6 and ell.position_space[0].shape[0] == 8
    te = gf.ComputeTile(ell, ((None, 8), (None, 32), (1, 4)), (0, 0, 0), {})
    assert (te.tile_i_rng, te.tile_k_rng, te.nnz, te.j_num, te.nnz_uncovered) == ((0, 7), (0, 3), 16, 32, 512)


# --- tile-shape candidates (gen_formats_v2.py:1188-1281, 1926-2068) ----------

def _shape(I, K):
    return ((None, I), (None, 32), (1, K))


def test_ell_tile_shapes_known_answer():
    # mbs = 32: I*K = 256 with K = 1..32, I*32 a multiple of 128 and >= 256;
    # product order is I ascending (itertools.product over I, then K)
    assert of.ell_tb_tile_shapes(1000, 32) == [_shape(8, 32), _shape(16, 16), _shape(32, 8), _shape(64, 4),
                                               _shape(128, 2), _shape(256, 1)]
    # fewer than 256 split rows: I <= 2^floor(log2(100)) = 64
    assert of.ell_tb_tile_shapes(100, 32) == [_shape(8, 32), _shape(16, 16), _shape(32, 8), _shape(64, 4)]
    # the tiny matrix's ELL ops: 8 pieces (mbs 4) and 11 pieces (mbs 2)
    assert of.ell_tb_tile_shapes(8, 4) == [_shape(8, 4)]
    assert of.ell_tb_tile_shapes(11, 2) == [_shape(8, 2)]
    # mbs 8: I*K = 64 with K = 1, 2, 4, 8; I = 64 / K <= 64 pieces
    assert of.ell_tb_tile_shapes(64, 8) == [_shape(8, 8), _shape(16, 4), _shape(32, 2), _shape(64, 1)]
    # 63 pieces: I <= 32, so K = 1 drops out
    assert of.ell_tb_tile_shapes(63, 8) == [_shape(8, 8), _shape(16, 4), _shape(32, 2)]


def test_tc_tile_sizes_known_answer():
    spmm = of.tc_tile_sizes('spmm')
    assert [ts[2][1] for ts in spmm] == [160, 144, 128, 112, 96, 80, 64, 48, 32, 16]
    assert all(ts[0] == (None, 16) and ts[1] == (None, 32) for ts in spmm)
    assert of.tc_tile_sizes('spmm', k_notsorted=True) == [((None, 16), (None, 32), (None, 16))]
    assert of.tc_tile_sizes('sddmm') == [((None, 16), (None, 16), (None, 16)), ((None, 16), (None, 32), (None, 16)),
                                         ((None, 16), (None, 64), (None, 16))]
    assert of.sddmm_tc_tile_sizes(32) == [((None, 16), (None, 32), (None, 16))]
    assert of.sddmm_tc_tile_sizes(512) == [((None, 16), (None, 64), (None, 16))]
    assert of.sddmm_tc_tile_sizes(16) == [((None, 16), (None, 16), (None, 16))]


def test_tile_params_known_answer():
    assert of.tile_params('sparse_template_ell', _shape(8, 32), 32) == [
        {'use_implicit_unroll': True, 'last_bucket': True}]
    assert of.tile_params('sparse_template_ell', _shape(16, 16), 32) == [
        {'use_implicit_unroll': True, 'last_bucket': False}]
    assert of.tile_params('TensorCore_template', ((None, 16), (None, 32), (None, 64)), 32) == [
        {'use_implicit_unroll': True, 'mma_shape_str': 'm16n16k16'}]
    assert of.tile_params('1D_sddmm', (32,), 32) == [
        {'tx': 8, 'ty': 4, 'vec_size': 4, 'group_size': 8, 'max_bucket_size': 32}]
    assert of.tile_params('TC_sddmm', ((None, 16), (None, 32), (None, 16)), 32) == [
        {'mma_shape_str': 'm16n16k16', 'warp_num': 1, 'vec_size': 4}]


@pytest.mark.parametrize("name,A", _mats())
@pytest.mark.parametrize("mbs", [2, 4, 8, 32])
def test_product_tile_shapes_match_oracle(name, A, mbs):
    op = gf.Op('spmm', [0, 1], [2], [A.shape[0], 4, A.shape[1]], None,
               [A, gf.DenseTensor(np.zeros((A.shape[1], 4), np.float32))], None)
    gf.set_config(gf.SearchConfig())
    ops, _ = gf.gen_candidate_formats(op, mbs, ([1], [1], [1]), False)
    ell, tc = ops[0], ops[-1]
    shapes = gf.ell_tile_shapes(ell, mbs)
    assert shapes == of.ell_tb_tile_shapes(len(ell.hyb_new_rows[0]), mbs)
    for ts in shapes:
        assert gf.get_params_list(ell, ts, mbs, is_tb_tile=True) == of.tile_params('sparse_template_ell', ts, mbs)
    assert gf.get_tile_sizes_idx(tc, mbs, is_tb_tile=True) == of.tc_tile_sizes('spmm')
    for ts in of.tc_tile_sizes('spmm'):
        assert gf.get_params_list(tc, ts, mbs, is_tb_tile=True) == of.tile_params('TensorCore_template', ts, mbs)


@pytest.mark.parametrize("kf", [16, 32, 64, 128])
def test_product_sddmm_tile_shapes_match_oracle(kf):
    A = tiny()
    op = gf.Op('sddmm', [0, 1], [2], [A.shape[0], kf, A.shape[1]], None,
               [A, gf.DenseTensor(np.zeros((A.shape[0], kf), np.float32)),
                gf.DenseTensor(np.zeros((A.shape[1], kf), np.float32))], None)
    cfg = gf.SearchConfig()
    cfg.op_type = 'sddmm'
    gf.set_config(cfg)
    try:
        ops, _ = gf.gen_candidate_formats(op, 32, ([1], [1], [1]), False)
        tc = [o for o in ops if gf.get_template_str(o) == 'TC_sddmm'][0]
        one_d = [o for o in ops if gf.get_template_str(o) == '1D_sddmm'][0]
        assert gf.get_tile_sizes_idx(tc, 32) == of.sddmm_tc_tile_sizes(kf)
        assert gf.get_params_list(one_d, (32,), 32) == of.tile_params('1D_sddmm', (32,), 32)
        assert gf.get_params_list(tc, of.sddmm_tc_tile_sizes(kf)[0], 32) == of.tile_params(
            'TC_sddmm', of.sddmm_tc_tile_sizes(kf)[0], 32)
    finally:
        gf.set_config(gf.SearchConfig())

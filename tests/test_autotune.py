"""The synthetic single-kind packs the cost-model autotuner times: each must
describe exactly the matrix it is meant to (checked by re-expansion), and
the refactored cost-model features must equal the search's own cost path."""
import numpy as np
import pytest
import scipy.sparse

from stile_amd import autotune, datasets, my_cost_model as mc
from tests.test_packing import expand_spmm_pack


def test_tc_pack_is_block_diagonal_dense():
    pack = autotune.tc_pack(5, 48)
    r, c, v = expand_spmm_pack(pack)
    got = scipy.sparse.coo_matrix((v, (r, c)), shape=(pack["M"], pack["K"])).toarray()
    want = np.kron(np.eye(5, dtype=np.float32), np.ones((16, 48), np.float32))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("I,K,pool", [(128, 2, 1 << 12), (8, 32, 64), (32, 8, 1 << 12)])
def test_ell_pack_rows_of_exactly_K(I, K, pool):
    rng = np.random.default_rng(1)
    pack, cols = autotune.ell_pack(1000, I, K, 1 << 12, pool, rng)
    r, c, v = expand_spmm_pack(pack)
    A = scipy.sparse.coo_matrix((v, (r, c)), shape=(pack["M"], pack["K"])).tocsr()
    assert A.nnz == pack["M"] * K                  # distinct columns in every row
    assert np.all(np.diff(A.indptr) == K)
    assert np.array_equal(np.sort(cols, axis=1), A.indices.reshape(-1, K))


def test_sddmm_packs_cover_each_output_once():
    rng = np.random.default_rng(2)
    pack, rows, cols = autotune.sddmm_1d_pack(100, 4, 777, rng)
    g = pack["g1d"][0]
    assert np.array_equal(np.sort(g["remap"]), np.arange(pack["nnz"]))
    pack = autotune.sddmm_tc_pack(3)
    g = pack["tc"][0]
    assert np.array_equal(np.sort(g["indices"]), np.arange(pack["nnz"]))
    # entry s of block b is out = 256 b + 16 row + col, at in-block position col + 16 row
    b = np.repeat(np.arange(3), 256)
    pos = g["indices_dense"]
    assert np.array_equal(g["indices"], 256 * b + 16 * (pos // 16) + pos % 16)


def test_ell_feature_matches_search_cost_path():
    """ell_tile_x on a real ELL op's tiles equals what
    cost_tb_latency_given_ELL_tiles feeds its linear model."""
    from stile_amd import api
    from stile_amd.gen_formats_v2 import gen_candidate_formats
    A = datasets.rmat_graph(9, seed=1)
    op = api.spmm_op(A, 32)
    ops, _ = gen_candidate_formats(op, 32, ([1], [1], [1]), False, True, gen_TC_formats=False)
    ell = ops[0]
    p = mc.get_profile("a100")
    a1, b1 = p.ell_coeff[16]
    I, K = 32, 8
    ts = ((None, I), (None, 32), (None, K))
    pos = np.arange(4)
    got = mc.cost_tb_latency_given_ELL_tiles(ell, ts, pos, 16, A.nnz, 108, False, profile="a100")
    cnt, uniq, has0 = mc._tile_unique_cols(ell, I, pos)
    x = mc.ell_tile_x(I, 32, K, cnt, uniq, has0, 16, mc._peak(16, p))
    assert np.allclose(got, a1 * x + b1)

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


@pytest.mark.parametrize("Kd,r", [(32, 0), (48, 3), (160, 16)])
def test_lds_pack_windows(Kd, r):
    """Window w: rows 16w..16w+15, each with the same Kd block columns plus r
    distinct columns of its own outside the block, all below b_rows."""
    rng = np.random.default_rng(3)
    W, b_rows = 6, 256
    pack = autotune.lds_pack(W, Kd, r, rng, b_rows)
    rr, c, v = expand_spmm_pack(pack)
    A = scipy.sparse.coo_matrix((v, (rr, c)), shape=(pack["M"], pack["K"])).tocsr()
    assert A.nnz == len(v)                          # no entry given twice
    assert np.all(np.diff(A.indptr) == Kd + r)
    blk = pack["tc"][0]["indices"]
    for w in range(W):
        block = set(blk[w].tolist())
        assert len(block) == Kd
        for i in range(16):
            row = set(A.indices[A.indptr[16 * w + i]:A.indptr[16 * w + i + 1]].tolist())
            assert block <= row and len(row - block) == r


def test_nonneg_fit_recovers_coefficients():
    x = np.array([1e4, 2e4, 3e4, 1.5e4, 2.5e4, 5e3])
    rows = np.array([256, 128, 64, 32, 16, 8])
    y = 2e-7 * x + 3e-4 + 5e-6 * rows
    got = autotune.nonneg_fit([x, np.ones(len(x)), rows], y)
    assert np.allclose(got, [2e-7, 3e-4, 5e-6], rtol=1e-6)
    # a negative best coefficient is clamped to 0 and the rest refitted
    got = autotune.nonneg_fit([x, np.ones(len(x))], 1e-7 * x - 1e-4 + 1e-6)
    assert got[1] == 0.0 and got[0] > 0
    # added costs: absolute error, samples near 0 or slightly negative allowed,
    # and they do not pull the slope to 0 (as 1/y weights would)
    e = np.array([16.0, 32, 64, 128, 256])
    y = 1.1e-6 * e - 1e-5 + np.array([-1e-6, 2e-6, 0, -3e-6, 1e-6])
    a, b = autotune.nonneg_fit([e, np.ones(len(e))], y, relative=False)
    assert b == 0.0 and abs(a - 1.1e-6) < 0.1e-6
    with pytest.raises(ValueError):
        autotune.nonneg_fit([e], np.array([-1e-6, 1e-6, 2e-6, 3e-6, 4e-6]), relative=True)


def _profile_with(name, **kw):
    import dataclasses
    p = dataclasses.replace(mc.get_profile("a100"), name=name, **kw)
    mc._PROFILES[name] = p
    return p


def test_ell_pricing_terms():
    """The per-row term adds c1 * I; on a BF16 op whose B has at most
    LDS_MAX_K rows the B-stationary pricing replaces the roofline form by
    16 a * entries + b * I / 16 (each entry a round of its own), and the
    lower bound follows it."""
    from stile_amd import api
    from stile_amd.gen_formats_v2 import gen_candidate_formats
    A = datasets.rmat_graph(9, seed=1)
    op = api.spmm_op(A, 32)
    ops, _ = gen_candidate_formats(op, 32, ([1], [1], [1]), False, True, gen_TC_formats=False)
    ell = ops[0]
    assert ell.idx_lens[2] <= mc.LDS_MAX_K
    I, K = 32, 8
    ts = ((None, I), (None, 32), (None, K))
    pos = np.arange(4)
    p = mc.get_profile("a100")
    base = mc.cost_tb_latency_given_ELL_tiles(ell, ts, pos, 16, A.nnz, 108, False, profile="a100")
    _profile_with("t_rows", ell_row_coeff={16: 1e-5, 32: 2e-5})
    got = mc.cost_tb_latency_given_ELL_tiles(ell, ts, pos, 16, A.nnz, 108, False, profile="t_rows")
    assert np.allclose(got, base + 1e-5 * I)
    _profile_with("t_lds", ell_lds_coeff=(1e-6, 2e-5))
    got = mc.cost_tb_latency_given_ELL_tiles(ell, ts, pos, 16, A.nnz, 108, False, profile="t_lds")
    cnt, _, _ = mc._tile_unique_cols(ell, I, pos)
    assert np.allclose(got, 16 * 1e-6 * np.minimum(cnt, I * K) + 2e-5 * I / 16)
    # FP32 blocks (dsize 32) never run on the B-stationary kernel
    got32 = mc.cost_tb_latency_given_ELL_tiles(ell, ts, pos, 32, A.nnz, 108, False, profile="t_lds")
    assert np.allclose(got32, mc.cost_tb_latency_given_ELL_tiles(ell, ts, pos, 32, A.nnz, 108, False,
                                                                 profile="a100"))
    # lower bound: min over the ELL shapes of (16 a I K + b I / 16) / (I K 32): the widest shape
    lb = mc.lower_bound_ELL_cost_BSPMM(ell, 32, 16, profile="t_lds")
    assert np.isclose(lb, (16e-6 + 2e-5 / (16 * 32)) / 32)
    assert p.ell_row_coeff is None and p.ell_lds_coeff is None


def test_profile_file_round_trip(tmp_path):
    p = _profile_with("t_file", ell_row_coeff={16: 1.5e-6, 32: 2.5e-6}, ell_lds_coeff=(3e-7, 4e-5),
                      tc_cost_lds={(16, 32, 16): 1e-5}, tc_cost_f32={(16, 32, 16): 2e-4})
    path = tmp_path / "prof.json"
    mc.save_profile_file(p, str(path))
    q = mc.load_profile_file(str(path))
    assert q.ell_row_coeff == p.ell_row_coeff and tuple(q.ell_lds_coeff) == p.ell_lds_coeff
    assert q.tc_cost_lds == p.tc_cost_lds and q.tc_cost == p.tc_cost and q.ell_coeff == {
        k: tuple(v) for k, v in p.ell_coeff.items()}

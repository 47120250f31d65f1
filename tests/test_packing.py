"""CPU tests: the product's packed buffers (prepare_inputs_BSPMM/BSDDMM
layouts) against the oracle's transcription of the reference packers,
bit-exact, on tile sets from the product search; and the packs' invariants
(re-expanding them reproduces A exactly, each nonzero once)."""
import json

import numpy as np
import pytest
import scipy.sparse

from oracle import formats as of
from stile_amd import api, datasets, my_fuse_formats
from stile_amd.gen_formats_v2 import get_template_str


def _cases():
    out = []
    A = datasets.pruned_layer(seed=5, M=256, K=128, N=8)[0]
    out.append(("pruned_values", A, api.SearchOptions()))
    out.append(("pruned_ones", scipy.sparse.csr_matrix((np.ones(A.nnz, np.float32), A.indices, A.indptr),
                                                      shape=A.shape), api.SearchOptions()))
    out.append(("pruned_only_ell", A, api.SearchOptions(only_ELL=True)))
    out.append(("pruned_single_level", A, api.SearchOptions(max_level_num=1)))
    out.append(("rmat9", datasets.rmat_graph(9, seed=2), api.SearchOptions()))
    out.append(("logsparse", datasets.log_sparse_mask(300), api.SearchOptions()))
    B = scipy.sparse.random(200, 150, density=0.05, format="csr", random_state=3, dtype=np.float32).tolil()
    B[7, :] = 1.0
    B[9, :] = 0.0
    out.append(("random_hub", B.tocsr(), api.SearchOptions(max_bucket_size=8)))
    out.append(("random_nowithdraw", B.tocsr(), api.SearchOptions(no_withdraw=True, max_bucket_size=8)))
    return out


def expand_spmm_pack(pack):
    """COO of every nonzero value in the packed groups (ELL rows at the
    sentinel M excluded)."""
    M, K = pack["M"], pack["K"]
    rows, cols, vals = [], [], []
    for g in pack["ell"]:
        I, Kc, T = g["m_blk"], g["nnz_cols"], g["num_tiles"]
        a = g["a"].reshape(T * I, Kc)
        r = np.repeat(g["indices_reordered_i"], Kc).reshape(T * I, Kc)
        m = (a != 0) & (r != M)
        rows.append(r[m]); cols.append(g["indices_k"][m]); vals.append(a[m])
    for g in pack["tc"]:
        T, Kg = g["nnzb"], g["group_size"]
        win_of_tile = np.repeat(np.arange(g["num_windows"]), np.diff(g["indptr"]))
        for t in range(T):
            s = win_of_tile[t]
            rr = g["i_indices"][s] if g["rows_2d"] else g["i_indices"][s] * 16 + np.arange(16)
            cc = g["indices"][t] * 16 + np.arange(16) if g["k_notsorted"] else g["indices"][t]
            m = g["a"][t] != 0
            ri, ci = np.nonzero(m)
            rows.append(np.asarray(rr)[ri]); cols.append(np.asarray(cc)[ci]); vals.append(g["a"][t][m])
    rows, cols, vals = (np.concatenate(x) if x else np.zeros(0) for x in (rows, cols, vals))
    return rows, cols, vals


@pytest.mark.parametrize("name,A,opts", _cases())
def test_spmm_pack_matches_oracle_and_reexpands(name, A, opts):
    pack, tiles, op = api.spmm_pack(A, opts)
    formats = my_fuse_formats.get_formats_from_selected_tiles(tiles, ['A'], opts.dsize)
    gi = {"ell": 0, "tc": 0}
    for fmt, ftiles in formats.items():
        params = json.loads(fmt[2])
        I, K = fmt[1][0][1], fmt[1][2][1]
        if fmt[0] == "sparse_template_ell":
            mine = pack["ell"][gi["ell"]]
            gi["ell"] += 1
            ref = of.pack_ell_group(ftiles, I, K)
            assert np.array_equal(mine["a"], ref["a"])
            assert np.array_equal(mine["indices_k"], ref["indices_k"])
            assert np.array_equal(mine["indices_reordered_i"], ref["indices_reordered_i"])
        else:
            mine = pack["tc"][gi["tc"]]
            gi["tc"] += 1
            rows_2d = bool(params["idx_reordered"][0] or params["real_atomic"])
            ref = of.pack_tc_group(ftiles, I, K, rows_2d)
            assert mine["rows_2d"] == rows_2d
            for k in ("a", "i_indices", "indptr", "indices"):
                assert np.array_equal(mine[k], ref[k]), k
    # every nonzero exactly once, with its value
    r, c, v = expand_spmm_pack(pack)
    got = scipy.sparse.coo_matrix((v, (r, c)), shape=A.shape)
    assert len(r) == A.nnz
    key = np.asarray(r, np.int64) * A.shape[1] + np.asarray(c, np.int64)
    assert len(np.unique(key)) == A.nnz
    diff = (got.tocsr() - scipy.sparse.csr_matrix(A, dtype=np.float32))
    assert abs(diff).max() == 0 if diff.nnz else True


@pytest.mark.parametrize("scale,Kf", [(9, 32), (10, 64), ("mask891", 64), ("gcn11_dsize32", 32)])
def test_sddmm_pack_matches_oracle(scale, Kf):
    opts = None
    if scale == "mask891":
        # 891 rows: the last 16-row window is cut short by the matrix edge
        A = datasets.attention_mask(891, band=32, n_random=4, rng=np.random.default_rng(0))
    elif scale == "gcn11_dsize32":
        # one 1D group holds tiles of two 1D ops (the first level and a next level)
        A = datasets.gcn_graph(scale=11, feat=8, seed=1)[0]
        A.data[:] = 1.0
        opts = api.SearchOptions(dsize=32)
    else:
        A = datasets.rmat_graph(scale, seed=scale)
    pack, tiles, op = api.sddmm_pack(A, opts, Kf=Kf)
    formats = my_fuse_formats.get_formats_from_selected_tiles(tiles, ['A'], 16)
    seen = np.zeros(A.nnz, bool)
    i1 = it = 0
    for fmt, ftiles in formats.items():
        if fmt[0] == "1D_sddmm":
            op1 = ftiles[0].op
            rows, cols, ids = of.sddmm_1d_ids(A, op1.idx_values_list[0], op1.inps[0])
            assert np.array_equal(op1.ori_row_ids_1d, rows)
            assert np.array_equal(op1.ori_col_ids_1d, cols)
            assert np.array_equal(op1.nnz_id_matrix, ids)
            ref = of.pack_sddmm_1d_group(ftiles, fmt[1][0])
            mine = pack["g1d"][i1]
            i1 += 1
            for k in ("mid_idx", "j_idx", "remap"):
                assert np.array_equal(mine[k], ref[k]), k
            seen[mine["remap"]] = True
            assert np.array_equal(A.indices[mine["remap"]], mine["j_idx"])
        else:
            ref = of.pack_sddmm_tc_group(A, ftiles, fmt[1])
            mine = pack["tc"][it]
            it += 1
            for k in ("i_indices", "j_indices", "indptr", "indices_dense", "indices"):
                assert np.array_equal(mine[k], ref[k]), k
            seen[mine["indices"]] = True
    assert seen.all(), "some output entries are computed by no tile"


def test_set_params_scalars():
    A = datasets.pruned_layer(seed=5, M=256, K=128, N=8)[0]
    pack, tiles, op = api.spmm_pack(A, N=100)
    formats = my_fuse_formats.get_formats_from_selected_tiles(tiles, ['A'], 16)
    p = my_fuse_formats.set_params_BSPMM(formats, 16, op)
    # m_tot = max(ceil(M/16)*16, M+1) and n_tot = padded N (reference my_fuse_formats.py:2797-2822)
    assert p["m_tot"] == 257
    assert p["n_tot"] == of.get_padded_j_TC(100, 16) == 128
    assert p["k_tot"] == 128
    for g in p["groups"]:
        if g["template"] == "TensorCore_template":
            assert g["tile_size"] == 16 and g["funit"] == 32 and g["nb"] == -(-128 // g["group_size"])


@pytest.mark.parametrize("kind", ["spmm", "sddmm"])
def test_row_blocked_packs_stack_to_A(kind):
    """Row-blocked search: the stacked pack covers every nonzero of A exactly
    once, at its own (row, col), with value ids = A's CSR positions."""
    A = datasets.rmat_graph(11, seed=4)
    A.data = np.random.default_rng(0).standard_normal(A.nnz).astype(np.float32)
    row_of = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    cuts = api.row_blocks(A, 3)
    assert len(cuts) == 4 and all(c % 16 == 0 for c in cuts[1:-1])
    if kind == "spmm":
        pack = api.spmm_pack_blocked(A, parts=3, workers=2)
        r, c, v = expand_spmm_pack(pack)
        got = scipy.sparse.coo_matrix((v, (r, c)), shape=A.shape).tocsr()
        assert len(r) == A.nnz
        assert abs(got - A).max() == 0
        for g in pack["ell"]:
            ids = g["a_ids"].reshape(-1, g["nnz_cols"])
            rr = np.repeat(g["indices_reordered_i"], g["nnz_cols"]).reshape(ids.shape)
            m = ids >= 0
            assert np.array_equal(row_of[ids[m]], rr[m])
            assert np.array_equal(A.indices[ids[m]], g["indices_k"][m])
    else:
        pack = api.sddmm_pack_blocked(A, parts=3, workers=2)
        seen = np.zeros(A.nnz, np.int64)
        for g in pack["g1d"]:
            assert np.array_equal(row_of[g["remap"]], g["mid_idx"])
            assert np.array_equal(A.indices[g["remap"]], g["j_idx"])
            seen[np.unique(g["remap"])] += 1
        for g in pack["tc"]:
            for b in range(g["nnzb"]):
                s0, s1 = g["indptr"][b], g["indptr"][b + 1]
                pos, out = g["indices_dense"][s0:s1], g["indices"][s0:s1]
                assert np.array_equal(row_of[out], g["i_indices"][b][pos // 16])
                assert np.array_equal(A.indices[out], g["j_indices"][b][pos % 16])
                np.add.at(seen, out, 1)
        assert seen.min() >= 1


@pytest.mark.parametrize("kind", ["spmm", "sddmm"])
def test_pack_save_load_roundtrip(tmp_path, kind):
    A = datasets.pruned_layer(seed=2, M=128, K=96, N=8)[0]
    pack = api.spmm_pack(A)[0] if kind == "spmm" else api.sddmm_pack(A)[0]
    path = tmp_path / "pack.npz"
    api.save_pack(path, pack)
    back = api.load_pack(path)
    assert set(back) == set(pack)
    for k, v in pack.items():
        if isinstance(v, list) and k != "order":
            assert len(back[k]) == len(v)
            for g0, g1 in zip(v, back[k]):
                assert set(g0) == set(g1)
                for f in g0:
                    if isinstance(g0[f], np.ndarray):
                        assert g0[f].dtype == g1[f].dtype and np.array_equal(g0[f], g1[f]), f
                    else:
                        assert g0[f] == g1[f], f
        else:
            assert back[k] == v


def test_pack_list_save_load_roundtrip(tmp_path):
    mats = [datasets.attention_mask(L, band=16, n_random=2, rng=np.random.default_rng(L)) for L in (40, 70)]
    packs = [api.spmm_pack(m)[0] for m in mats]
    api.save_pack(tmp_path / "l.npz", packs)
    back = api.load_pack(tmp_path / "l.npz")
    assert len(back) == 2
    for p0, p1 in zip(packs, back):
        for k in ("ell", "tc"):
            for g0, g1 in zip(p0[k], p1[k]):
                for f in g0:
                    if isinstance(g0[f], np.ndarray):
                        assert np.array_equal(g0[f], g1[f])


def test_load_snap_graph_relabels_dedupes_and_pads(tmp_path):
    from stile_amd.gen_formats_v2 import load_snap_graph
    p = tmp_path / "g.txt"
    p.write_text("# comment\n10 20\n20 30\n10 20\n30 10\n")
    op = load_snap_graph('spmm', str(p), feat_size=8)
    A = op.inps[0]
    # ids 10, 20, 30 -> 0, 1, 2; the duplicate edge dropped; 128-padded
    assert op.idx_lens == [128, 8, 128]
    assert A.shape == (128, 128) and A.nnz == 3
    assert sorted(zip(*A.nonzero())) == [(0, 1), (1, 2), (2, 0)]
    assert op.inps[1].data.shape == (128, 8)
    op2 = load_snap_graph('sddmm', str(p), feat_size=8)
    assert op2.idx_lens == [128, 8, 160]
    op3 = load_snap_graph('spmm', str(p), feat_size=8, pad=False)
    assert op3.idx_lens == [3, 8, 3]


def test_pack_cache_refuses_another_pattern(tmp_path):
    A = datasets.pruned_layer(seed=1, M=64, K=48, N=8)[0]
    B = A.copy().tolil()
    B[0, 0] = 0.0 if A[0, 0] != 0 else 1.0           # one nonzero moved in or out
    B = scipy.sparse.csr_matrix(B)
    B.eliminate_zeros()
    pack = api.spmm_pack(A, api.SearchOptions())[0]
    fa = api.matrix_fingerprint("spmm", A)
    api.save_pack(tmp_path / "p.npz", pack, fingerprint=fa)
    assert api.load_pack(tmp_path / "p.npz", expect=fa)["M"] == 64
    # same shape, a different pattern: refused
    with pytest.raises(api.StalePackError):
        api.load_pack(tmp_path / "p.npz", expect=api.matrix_fingerprint("spmm", B))
    # same pattern, other operation: refused
    with pytest.raises(api.StalePackError):
        api.load_pack(tmp_path / "p.npz", expect=api.matrix_fingerprint("sddmm", A))
    # a file saved without a fingerprint is refused when one is expected
    api.save_pack(tmp_path / "q.npz", pack)
    with pytest.raises(api.StalePackError):
        api.load_pack(tmp_path / "q.npz", expect=fa)
    # values do not enter the fingerprint (plans refresh values by id)
    A2 = A.copy()
    A2.data = A2.data * 3
    assert api.matrix_fingerprint("spmm", A2) == fa
    # build options travel with the pack; files without them load as {}
    api.save_pack(tmp_path / "r.npz", pack, fingerprint=fa, build={"scatter_max_nnz": 2})
    pk, build = api.load_pack(tmp_path / "r.npz", expect=fa, with_build=True)
    assert build == {"scatter_max_nnz": 2} and pk["M"] == 64
    assert api.load_pack(tmp_path / "p.npz", expect=fa, with_build=True)[1] == {}


def test_plan_variants():
    """The tuner times scatter chunks off and on only for plans the
    B-stationary kernel runs (BF16 blocks, every K <= 1024)."""
    small = datasets.pruned_layer(seed=1, M=64, K=48, N=8)[0]
    wide = scipy.sparse.random(32, 2000, density=0.01, format="csr", random_state=0)
    assert set(api.plan_variants("spmm", "bfloat16", small)) == {"", "+scatter2"}
    assert set(api.plan_variants("spmm", "bfloat16", [small, small])) == {"", "+scatter2"}
    assert set(api.plan_variants("spmm", "bfloat16", [small, wide])) == {""}
    assert set(api.plan_variants("spmm", "float32", small)) == {""}
    assert set(api.plan_variants("sddmm", "bfloat16", small)) == {""}


def test_sampled_error_checks_detect_a_wrong_row():
    import torch
    mats = [datasets.pruned_layer(seed=s, M=48, K=32, N=8)[0] for s in (2, 3)]
    big = scipy.sparse.block_diag(mats, format="csr")
    rng = np.random.default_rng(0)
    B = torch.from_numpy(rng.standard_normal((64, 5)).astype(np.float32))
    C = torch.from_numpy(np.asarray(big @ B.numpy().astype(np.float64)).astype(np.float32))
    assert api.spmm_sample_error(mats, B, C, samples=96) < 1e-6
    C[70, 3] += 1.0
    assert api.spmm_sample_error(mats, B, C, samples=96) > 1e-2
    X = torch.from_numpy(rng.standard_normal((96, 4)).astype(np.float32))
    Y = torch.from_numpy(rng.standard_normal((64, 4)).astype(np.float32))
    rows = np.repeat(np.arange(96), np.diff(big.indptr))
    out = torch.from_numpy((X.numpy()[rows] * Y.numpy()[big.indices]).sum(1))
    assert api.sddmm_sample_error(mats, X, Y, out, samples=big.nnz) < 1e-6
    out[big.nnz - 1] += 1.0
    assert api.sddmm_sample_error(mats, X, Y, out, samples=big.nnz) > 1e-2


def test_spmm_plan_signature():
    """Packs that group the same TC nonzeros into tiles differently give one
    plan, hence one signature; moving nonzeros between TC and ELL tiles, or
    another matrix, changes it."""
    from tests import _packs
    A = datasets.pruned_layer(seed=2, M=128, K=96, N=8)[0]
    p16 = _packs.spmm_pack(A, min_fill=64, tc_group_cols=16)
    p32 = _packs.spmm_pack(A, min_fill=64, tc_group_cols=32)
    assert [g["group_size"] for g in p16["tc"]] != [g["group_size"] for g in p32["tc"]]
    assert api.spmm_plan_signature(p16) == api.spmm_plan_signature(p32)
    p_more_ell = _packs.spmm_pack(A, min_fill=257)       # every nonzero in ELL tiles
    assert api.spmm_plan_signature(p_more_ell) != api.spmm_plan_signature(p16)
    B = datasets.pruned_layer(seed=3, M=128, K=96, N=8)[0]
    assert api.spmm_plan_signature(_packs.spmm_pack(B, min_fill=64)) != api.spmm_plan_signature(p16)
    assert api.spmm_plan_signature([p16, p16]) != api.spmm_plan_signature([p16])
    assert api.spmm_plan_signature([p16]) == api.spmm_plan_signature(p16)

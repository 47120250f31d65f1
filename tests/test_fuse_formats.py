"""Known answers for the grouping helpers of the packing step
(stile_amd/my_fuse_formats.py), computed by hand from the reference's
formulas (MY_sparse/my_fuse_formats.py:57-111, 2856-2968)."""
import json

import numpy as np

from stile_amd import api, datasets, my_fuse_formats as mff


def test_smem_padding_pattern_known_answers():
    f = mff.comp_SMEM_padding_pattern
    # 32 threads along j: one row per warp, no bank conflict
    assert f('spmm', ((None, 8), (None, 1, 32), (None, 4)), 32) == (1, 0, 4)
    # 4 rows per warp, one 16-bit element per row: rows share a bank, pad each row
    # by (32 - 16) // 16 = 1 unit
    assert f('spmm', ((None, 8), (None, 1, 8), (None, 1)), 16) == (1, 1, 1)
    # 4 rows per warp, a row spans 2 banks: 16 rows before a conflict >= 4, none
    assert f('spmm', ((None, 8), (None, 1, 8), (None, 2)), 32) == (1, 0, 2)
    # 32 rows per warp, a row spans 4 banks: every 8 rows collide, pad 1 bank width
    assert f('spmm', ((None, 8), (None, 1, 1), (None, 4)), 32) == (8, 1, 4)
    # the row length is the product of the k tile sizes; the bank step uses the first
    assert f('spmm', ((None, 8), (None, 1, 1), (None, 4, 2)), 32) == (8, 1, 8)


def test_c_store_atomics_known_answer():
    ts = ((None, 8), (None, 2, 16), (None, 4))
    formats = {
        ("TensorCore_template", ts, json.dumps({"real_atomic": False})): [],
        ("sparse_template_ell", ts, json.dumps({"real_atomic": False, "use_implicit_unroll": True})): [],
        ("sparse_template_ell", ts, json.dumps({"real_atomic": True, "use_implicit_unroll": False})): [],
        ("sparse_template_ell", ts, json.dumps({"real_atomic": True, "use_implicit_unroll": True})): [],
    }
    # TC: its flag; non-atomic ELL: nothing; explicit unroll: one per j1 step (2)
    assert mff.get_C_store_atomics(formats) == [False, True, True, True]


def test_sddmm_group_scalars():
    A = datasets.rmat_graph(9, seed=1)
    pack, tiles, op = api.sddmm_pack(A, None, 32)
    formats = mff.get_formats_from_selected_tiles(tiles, ['A'], 32)
    groups = mff.set_params_BSDDMM(formats, 32, op)["groups"]
    one_d = [g for g in groups if g["template"] == "1D_sddmm"]
    tc = [g for g in groups if g["template"] == "TC_sddmm"]
    assert one_d and tc
    # every output entry is counted once across the groups
    assert sum(g["nnz_real"] for g in one_d) + sum(g["nnz"] for g in tc) == A.nnz
    for g in one_d:  # NNZ padded to whole buckets of ty * group_size = 4 * 8
        assert g["bucket"] == 32 and g["nnz"] % 32 == 0 and 0 <= g["nnz"] - g["nnz_real"] < 32
    for g in tc:  # 16 x 16 output blocks; nonzeros = the pack's sampled entries
        assert (g["mb"], g["nb"]) == (16, 16)
        assert g["nnz"] == sum(int(x["indptr"][-1]) for x in pack["tc"])
        assert g["nnzb"] == sum(int(x["nnzb"]) for x in pack["tc"])


def test_order_output_is_identity_for_scattered_kernels():
    out = np.arange(7, dtype=np.float32)
    assert mff.order_output_BSDDMM(None, out, 7) is out

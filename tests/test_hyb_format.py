"""The SparseTIR-style hyb comparator format: column parts x power-of-two
ELL buckets, long rows cut into pieces of the largest bucket.  Re-expanding
the buckets must give A exactly; every ELL row must fit its bucket."""
import numpy as np
import pytest
import scipy.sparse

from stile_amd import datasets, hyb


def _expand(buckets, shape):
    r, c, v = [], [], []
    for b in buckets:
        rr = np.repeat(b["row"], b["width"])
        r.append(rr); c.append(b["cols"].reshape(-1)); v.append(b["vals"].reshape(-1))
    return scipy.sparse.coo_matrix((np.concatenate(v), (np.concatenate(r), np.concatenate(c))), shape=shape).tocsr()


@pytest.mark.parametrize("parts,buckets", [(1, None), (3, [1, 2, 4, 8]), (2, [1, 4, 16, 64])])
def test_hyb_buckets_reproduce_A(parts, buckets):
    A = datasets.rmat_graph(10, seed=3)
    A.data = np.random.default_rng(0).standard_normal(A.nnz).astype(np.float32)
    bk = hyb.column_part_hyb(A, parts, buckets)
    sizes = buckets or hyb.default_bucket_sizes(A)
    for b in bk:
        assert b["width"] in sizes
        assert b["cols"].shape == b["vals"].shape == (len(b["row"]), b["width"])
    got = _expand(bk, A.shape)
    assert abs(got - A).max() == 0
    # a row's pieces in one part: all but the last are full largest buckets
    assert sum(int((b["vals"] != 0).sum()) for b in bk) == A.nnz


def test_hyb_known_answer():
    # row 0: 5 entries in part 0 -> pieces 4 + 1 (buckets [1, 2, 4]); row 1: 2 -> bucket 2
    A = scipy.sparse.csr_matrix(np.array([[1, 2, 3, 4, 5, 0], [0, 6, 0, 7, 0, 0]], np.float32))
    bk = {b["width"]: b for b in hyb.column_part_hyb(A, 1, [1, 2, 4])}
    assert sorted(bk) == [1, 2, 4]
    assert bk[4]["row"].tolist() == [0] and bk[4]["cols"].tolist() == [[0, 1, 2, 3]]
    assert bk[1]["row"].tolist() == [0] and bk[1]["cols"].tolist() == [[4]]
    assert bk[2]["row"].tolist() == [1] and bk[2]["cols"].tolist() == [[1, 3]]
    assert bk[2]["vals"].tolist() == [[6, 7]]

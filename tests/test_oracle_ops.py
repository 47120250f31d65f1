"""CPU pins of the oracle's closed forms (oracle/ops.py) against brute-force
Python loops on inputs of at most 16 x 16, plus special cases fixed by the
mathematics (identity, a single nonzero, empty rows, sign flips)."""
import numpy as np
import pytest
import scipy.sparse

from oracle import ops


def _rand_sparse(m, n, density, seed):
    rng = np.random.default_rng(seed)
    mask = rng.random((m, n)) < density
    vals = rng.standard_normal((m, n)) * mask
    return scipy.sparse.csr_matrix(vals), vals


def _spmm_loops(D, B):
    m, k = len(D), len(D[0])
    n = len(B[0])
    C = [[0.0] * n for _ in range(m)]
    for i in range(m):
        for kk in range(k):
            if D[i][kk] != 0.0:
                for j in range(n):
                    C[i][j] += float(D[i][kk]) * float(B[kk][j])
    return np.array(C)


@pytest.mark.parametrize("m,k,n,density,seed", [(16, 16, 5, 0.2, 0), (7, 13, 3, 0.5, 1), (1, 16, 16, 1.0, 2),
                                                (16, 1, 4, 0.7, 3), (9, 9, 1, 0.0, 4)])
def test_spmm_fp64_matches_loops(m, k, n, density, seed):
    A, D = _rand_sparse(m, k, density, seed)
    B = np.random.default_rng(seed + 10).standard_normal((k, n))
    got = ops.spmm_fp64(A, B)
    ref = _spmm_loops(D.tolist(), B.tolist())
    assert got.dtype == np.float64
    assert np.allclose(got, ref, rtol=1e-14, atol=1e-14)
    # the FP32 reference path: same closed form within FP32 rounding of
    # k-term sums (gamma_k |A||B|)
    f32 = ops.spmm_ref_fp32(A, B)
    assert f32.dtype == np.float32
    absb = _spmm_loops(np.abs(D).tolist(), np.abs(B).tolist())
    gam = k * 2.0 ** -24 / (1 - k * 2.0 ** -24)
    assert np.all(np.abs(f32 - ref) <= 2 * gam * absb + 2.0 ** -24 * np.abs(ref) + 1e-30)


def test_spmm_special_cases():
    I = scipy.sparse.identity(6, format="csr")
    B = np.arange(18, dtype=np.float64).reshape(6, 3)
    assert np.array_equal(ops.spmm_fp64(I, B), B)
    one = scipy.sparse.csr_matrix(([2.5], ([3], [1])), shape=(5, 4))
    Bq = np.arange(8, dtype=np.float64).reshape(4, 2)
    want = np.zeros((5, 2))
    want[3] = 2.5 * Bq[1]
    assert np.array_equal(ops.spmm_fp64(one, Bq), want)
    # -A gives -C exactly (a dropped sign or transposed operand fails here)
    A, _ = _rand_sparse(8, 5, 0.5, 7)
    Bm = np.random.default_rng(7).standard_normal((5, 3))
    assert np.array_equal(ops.spmm_fp64(-A, Bm), -ops.spmm_fp64(A, Bm))
    assert ops.spmm_fp64(A, Bm).shape == (8, 3)


def _sddmm_loops(D, X, Y):
    out = []
    for i in range(len(D)):
        for j in range(len(D[0])):
            if D[i][j] != 0.0:
                out.append(sum(float(X[i][f]) * float(Y[j][f]) for f in range(len(X[0]))))
    return np.array(out)


@pytest.mark.parametrize("m,n,kf,density,seed", [(16, 16, 8, 0.25, 0), (5, 11, 3, 0.6, 1), (16, 2, 16, 1.0, 2),
                                                 (3, 3, 1, 0.9, 3)])
def test_sddmm_fp64_matches_loops(m, n, kf, density, seed):
    A, D = _rand_sparse(m, n, density, seed)
    rng = np.random.default_rng(seed + 20)
    X = rng.standard_normal((m, kf))
    Y = rng.standard_normal((n, kf))
    got = ops.sddmm_fp64(A, X, Y)
    ref = _sddmm_loops(D.tolist(), X.tolist(), Y.tolist())
    assert got.shape == (A.nnz,)
    assert np.allclose(got, ref, rtol=1e-14, atol=1e-14)
    f32 = ops.sddmm_ref_fp32(A, X, Y)
    absb = _sddmm_loops(D.tolist(), np.abs(X).tolist(), np.abs(Y).tolist())
    gam = kf * 2.0 ** -24 / (1 - kf * 2.0 ** -24)
    assert np.all(np.abs(f32 - ref) <= 2 * gam * absb + 2.0 ** -24 * np.abs(ref) + 1e-30)
    assert np.allclose(ops.sddmm_abs_bound(A, X, Y), absb, rtol=1e-14, atol=1e-14)


def test_sddmm_order_is_csr_and_not_transposed():
    # out follows A's CSR order; X pairs with rows and Y with columns
    A = scipy.sparse.csr_matrix(([1.0, 1.0, 1.0], ([0, 0, 2], [2, 0, 1])), shape=(3, 3))
    X = np.array([[1.0, 0.0], [0.0, 0.0], [0.0, 2.0]])
    Y = np.array([[3.0, 0.0], [0.0, 5.0], [7.0, 0.0]])
    assert ops.sddmm_fp64(A, X, Y).tolist() == [3.0, 7.0, 10.0]      # (0,0), (0,2), (2,1)


def test_error_metrics_known_answers():
    ref = np.array([[1.0, -4.0], [2.0, 0.5]])
    got = ref + np.array([[0.0, 0.02], [-0.01, 0.0]])
    assert ops.max_rel_err(got, ref) == pytest.approx(0.02 / 4.0, rel=1e-15)
    assert ops.max_rel_err(np.zeros(0), np.zeros(0)) == 0.0
    assert ops.max_rel_err(np.array([1e-3]), np.array([0.0])) == pytest.approx(1e-3)
    # bound_ratio: error / (gamma_n |A||B|) with gamma_n = n u / (1 - n u)
    n, u = 8, 2.0 ** -24
    gam = n * u / (1 - n * u)
    absb = np.array([2.0, 3.0])
    err = np.array([0.5, 1.0]) * gam * absb
    r = ops.bound_ratio(np.array([1.0, 1.0]) + err, np.array([1.0, 1.0]), absb, n)
    assert r == pytest.approx(1.0, rel=1e-9)
    assert ops.bound_ratio(np.zeros(0), np.zeros(0), np.zeros(0), 4) == 0.0


def test_abs_bound_and_rows():
    A, D = _rand_sparse(6, 7, 0.5, 11)
    B = np.random.default_rng(11).standard_normal((7, 2))
    assert np.allclose(ops.spmm_abs_bound(A, B), _spmm_loops(np.abs(D).tolist(), np.abs(B).tolist()), rtol=1e-14)
    rows = ops.csr_rows(A)
    assert rows.tolist() == [i for i in range(6) for _ in range(int((D[i] != 0).sum()))]

"""FP64 products and the reference's own FP32 CPU path (test infrastructure).

spmm_fp64 / sddmm_fp64 are the closed forms C = A.B and out_p = <X_i, Y_j>
evaluated in float64 with scipy/numpy (textbook routines; pinned by
tests/test_oracle_ops.py against brute-force loops on tiny inputs).

spmm_ref_fp32 is the reference's golden path, scipy
`ori_op.inps[0].astype(dtype).dot(B)` (MY_sparse/my_fuse_formats.py:4194-4197)
with dtype float32; it is also the CPU baseline timed by bench.py.
sddmm_ref_fp32 replaces the reference's GPU golden dgl.ops.u_dot_v
(my_fuse_formats.py:4331-4341), absent here, with the same per-nonzero dot in
float32 numpy.
"""
import numpy as np
import scipy.sparse


def spmm_fp64(A, B):
    A64 = scipy.sparse.csr_matrix(A, dtype=np.float64)
    return np.asarray(A64 @ np.asarray(B, dtype=np.float64))


def spmm_ref_fp32(A, B):
    """scipy CSR x dense in float32, single-threaded (the reference's CPU path)."""
    return scipy.sparse.csr_matrix(A).astype(np.float32).dot(np.asarray(B, dtype=np.float32))


def spmm_abs_bound(A, B):
    """|A|.|B| in float64: the scale of the FP32 summation error bound."""
    Aa = abs(scipy.sparse.csr_matrix(A, dtype=np.float64))
    return np.asarray(Aa @ np.abs(np.asarray(B, dtype=np.float64)))


def csr_rows(A):
    A = scipy.sparse.csr_matrix(A)
    return np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))


def sddmm_fp64(A, X, Y):
    """out[p] = <X[row(p)], Y[col(p)]> for the nonzeros p of A in CSR order."""
    A = scipy.sparse.csr_matrix(A)
    rows = csr_rows(A)
    X64 = np.asarray(X, dtype=np.float64)
    Y64 = np.asarray(Y, dtype=np.float64)
    out = np.empty(A.nnz, dtype=np.float64)
    step = 1 << 20
    for s in range(0, A.nnz, step):
        r = rows[s:s + step]
        c = A.indices[s:s + step]
        out[s:s + step] = np.einsum("ij,ij->i", X64[r], Y64[c])
    return out


def sddmm_abs_bound(A, X, Y):
    return sddmm_fp64(A, np.abs(np.asarray(X, dtype=np.float64)), np.abs(np.asarray(Y, dtype=np.float64)))


def sddmm_ref_fp32(A, X, Y):
    A = scipy.sparse.csr_matrix(A)
    rows = csr_rows(A)
    X32 = np.asarray(X, dtype=np.float32)
    Y32 = np.asarray(Y, dtype=np.float32)
    return (X32[rows] * Y32[A.indices]).sum(axis=1, dtype=np.float32)


def max_rel_err(got, ref64):
    """Norm-wise relative error max|got-ref| / max|ref| (0 when ref is 0)."""
    got = np.asarray(got, dtype=np.float64)
    ref64 = np.asarray(ref64, dtype=np.float64)
    if ref64.size == 0:
        return 0.0
    scale = np.max(np.abs(ref64))
    diff = np.max(np.abs(got - ref64))
    return float(diff / scale) if scale > 0 else float(diff)


def bound_ratio(got, ref64, absbound, terms, unit=2.0 ** -24):
    """max |got-ref| / (gamma_n * |A||B|) with gamma_n = n*u/(1-n*u).

    A value <= 1 means every element is within the worst-case FP32
    summation error bound for n-term dot products (n = terms).
    """
    n = float(max(terms, 1))
    gamma = n * unit / (1.0 - n * unit)
    got = np.asarray(got, dtype=np.float64)
    err = np.abs(got - np.asarray(ref64, dtype=np.float64))
    tol = gamma * np.asarray(absbound, dtype=np.float64) + 1e-30
    return float(np.max(err / tol)) if err.size else 0.0

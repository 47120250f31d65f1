"""Seeded synthetic inputs for the five benchmark workloads, plus the
reference's deterministic attention-mask patterns and an edge-list loader.

All generators use numpy.random.default_rng(seed) and return scipy CSR
matrices with int32 indices, sorted column indices and no duplicates.

Workloads (sizes as used by bench.py and the tests):
  uniform_spmm      4096 x 4096, 167,772 uniformly drawn positions,
                    values U(-1,1); B 4096 x 64 N(0,1)
  rmat_graph        R-MAT (a,b,c,d) = (0.57, 0.19, 0.19, 0.05), scale s,
                    edge factor 16, deduplicated, self-loops added
  pruned_layer      3072 x 768: 10 % of the 16x16 blocks dense N(0, 0.02)
                    plus a 1 % unstructured residual N(0, 0.02)
  gcn_graph         R-MAT scale 20, symmetrised + self-loops,
                    values D^-1/2 (A+I) D^-1/2
  attention_masks   64 masks, L ~ U[512, 8192], band 128 + 16 random columns

Reference counterparts: the loaders of MY_sparse/gen_formats_v2.py
(load_snap_graph :2629-2705, test_LogSparse :2926-2963, test_Strided
:2970-3003), which build the same CSR-with-int32 structure (and set every
value to 1, :2692-2696; `ones=True` reproduces that).
"""
import numpy as np
import scipy.sparse


def _csr(rows, cols, vals, shape):
    A = scipy.sparse.csr_matrix((vals, (rows, cols)), shape=shape)
    A.sum_duplicates()
    A.sort_indices()
    A.indptr = A.indptr.astype(np.int32)
    A.indices = A.indices.astype(np.int32)
    return A


def uniform_spmm(seed=0, M=4096, K=4096, nnz=167772, N=64, ones=False):
    """Sparse A (M x K, nnz positions drawn without replacement, U(-1,1)) and B (K x N, N(0,1))."""
    rng = np.random.default_rng(seed)
    pos = rng.choice(M * K, size=nnz, replace=False)
    rows, cols = pos // K, pos % K
    vals = np.ones(nnz, np.float32) if ones else rng.uniform(-1.0, 1.0, nnz).astype(np.float32)
    A = _csr(rows, cols, vals, (M, K))
    B = rng.standard_normal((K, N), dtype=np.float32)
    return A, B


def rmat_edges(scale, edge_factor=16, abcd=(0.57, 0.19, 0.19, 0.05), seed=0):
    """R-MAT edge list (rows, cols) of 2^scale nodes and edge_factor * 2^scale edges."""
    rng = np.random.default_rng(seed)
    n_edges = edge_factor << scale
    rows = np.zeros(n_edges, np.int64)
    cols = np.zeros(n_edges, np.int64)
    p = np.asarray(abcd, np.float64)
    cum = np.cumsum(p / p.sum())
    chunk = 1 << 22
    for s in range(0, n_edges, chunk):
        e = min(n_edges, s + chunk)
        r = np.zeros(e - s, np.int64)
        c = np.zeros(e - s, np.int64)
        for _ in range(scale):
            u = rng.random(e - s)
            q = np.searchsorted(cum, u, side="right")
            q = np.minimum(q, 3)
            r = (r << 1) | (q >> 1)
            c = (c << 1) | (q & 1)
        rows[s:e] = r
        cols[s:e] = c
    return rows, cols


def rmat_graph(scale=16, edge_factor=16, seed=0, self_loops=True):
    """R-MAT pattern (deduplicated, + self loops), values 1."""
    n = 1 << scale
    r, c = rmat_edges(scale, edge_factor, seed=seed)
    if self_loops:
        r = np.concatenate([r, np.arange(n)])
        c = np.concatenate([c, np.arange(n)])
    A = _csr(r, c, np.ones(len(r), np.float32), (n, n))
    A.data[:] = 1.0
    return A


def sddmm_rmat(seed=0, scale=16, K=32):
    """SDDMM workload: R-MAT pattern and X, Y (2^scale x K, N(0,1))."""
    A = rmat_graph(scale, 16, seed)
    rng = np.random.default_rng(seed + 1)
    n = A.shape[0]
    X = rng.standard_normal((n, K), dtype=np.float32)
    Y = rng.standard_normal((n, K), dtype=np.float32)
    return A, X, Y


def pruned_layer(seed=0, M=3072, K=768, block=16, block_keep=0.10, residual=0.01, std=0.02,
                 N=4096, ones=False):
    """Block-pruned layer: a fraction of the 16x16 blocks dense plus a sparse residual."""
    rng = np.random.default_rng(seed)
    mb, kb = M // block, K // block
    nblk = mb * kb
    keep = rng.choice(nblk, size=int(round(block_keep * nblk)), replace=False)
    br, bc = keep // kb, keep % kb
    ii, jj = np.meshgrid(np.arange(block), np.arange(block), indexing="ij")
    rows = (br[:, None] * block + ii.reshape(1, -1)).reshape(-1)
    cols = (bc[:, None] * block + jj.reshape(1, -1)).reshape(-1)
    dense_mask = np.zeros((M, K), bool)
    dense_mask[rows, cols] = True
    free = np.flatnonzero(~dense_mask.reshape(-1))
    nres = int(round(residual * free.size))
    res = rng.choice(free, size=nres, replace=False)
    rr, rc = res // K, res % K
    all_r = np.concatenate([rows, rr])
    all_c = np.concatenate([cols, rc])
    if ones:
        vals = np.ones(len(all_r), np.float32)
    else:
        vals = rng.normal(0.0, std, len(all_r)).astype(np.float32)
        vals[vals == 0] = std  # keep every position a stored nonzero
    A = _csr(all_r, all_c, vals, (M, K))
    B = rng.standard_normal((K, N), dtype=np.float32)
    return A, B


def pruned_encoder(layers=24, d=1024, ffn=4096, seed=0):
    """Block-pruned projections of a BERT-large-shaped encoder, as a batch of
    uneven matrices: per layer Q, K, V, O (d x d) and the FFN input projection
    (ffn x d), each like pruned_layer (10 % dense 16x16 blocks + 1 %
    residual).  The FFN output projection (d x ffn) is left out: its K = ffn
    is above the 1024 B rows the B-stationary kernel stages.  Returns the
    matrices (CSR) and their (M, K) shapes."""
    mats, shapes = [], []
    for layer in range(layers):
        for j, (m, k) in enumerate(((d, d), (d, d), (d, d), (d, d), (ffn, d))):
            A, _ = pruned_layer(seed=seed + 1000 * layer + j, M=m, K=k, N=1)
            mats.append(A)
            shapes.append((m, k))
    return mats, shapes


def gcn_graph(scale=20, edge_factor=16, seed=0, feat=128):
    """Symmetric-normalised R-MAT adjacency D^-1/2 (A+I) D^-1/2, H and W."""
    n = 1 << scale
    r, c = rmat_edges(scale, edge_factor, seed=seed)
    rr = np.concatenate([r, c, np.arange(n)])
    cc = np.concatenate([c, r, np.arange(n)])
    A = _csr(rr, cc, np.ones(len(rr), np.float32), (n, n))
    A.data[:] = 1.0
    deg = np.asarray(A.sum(axis=1)).reshape(-1)
    dinv = 1.0 / np.sqrt(deg)
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    A.data = (dinv[rows] * dinv[A.indices]).astype(np.float32)
    rng = np.random.default_rng(seed + 1)
    H = rng.standard_normal((n, feat), dtype=np.float32)
    W = (rng.standard_normal((feat, feat), dtype=np.float32) / np.sqrt(feat)).astype(np.float32)
    return A, H, W


def attention_mask(L, band=128, n_random=16, rng=None):
    """Band (columns i-band/2 .. i+band/2-1) plus n_random uniform columns per row."""
    rng = rng if rng is not None else np.random.default_rng(0)
    i = np.arange(L)
    off = np.arange(-band // 2, band // 2)
    br = np.repeat(i, band)
    bc = (i[:, None] + off[None, :]).reshape(-1)
    ok = (bc >= 0) & (bc < L)
    rr = np.repeat(i, n_random)
    rc = rng.integers(0, L, size=L * n_random)
    rows = np.concatenate([br[ok], rr])
    cols = np.concatenate([bc[ok], rc])
    A = _csr(rows, cols, np.ones(len(rows), np.float32), (L, L))
    A.data[:] = 1.0
    return A


def attention_masks(count=64, lmin=512, lmax=8192, band=128, n_random=16, seed=0):
    rng = np.random.default_rng(seed)
    lens = rng.integers(lmin, lmax + 1, size=count)
    return [attention_mask(int(L), band, n_random, rng) for L in lens]


def log_sparse_mask(m, n=None):
    """LogSparse attention mask of test_LogSparse (gen_formats_v2.py:2926-2963):
    row i holds i - 2^j (j = floor(log2 i) .. 0), i, and i + 2^j
    (j < floor(log2(n-i-1))); values 1."""
    import math
    n = m if n is None else n
    indices = []
    for i in range(m):
        if i == 0:
            tmp = [i] + [i + 2 ** j for j in range(int(math.log(n - i - 1, 2)))]
        elif i == n - 1:
            tmp = [i - 2 ** j for j in range(int(math.log(i, 2)), -1, -1)] + [i]
        else:
            tmp = ([i - 2 ** j for j in range(int(math.log(i, 2)), -1, -1)] + [i]
                   + [i + 2 ** j for j in range(int(math.log(n - i - 1, 2)))])
        indices.append(tmp)
    indptr = np.cumsum([0] + [len(t) for t in indices]).astype(np.int32)
    idx = np.concatenate(indices).astype(np.int32)
    A = scipy.sparse.csr_matrix((np.ones(len(idx), np.float32), idx, indptr), shape=(m, n))
    return A


def strided_mask(m, n=None):
    """Strided attention mask of test_Strided (gen_formats_v2.py:2970-3003):
    s = floor(sqrt(n)); row i holds i-s+1 .. i+s-1 and i +- j*s for
    1 <= j < floor((n-1-i)/s) (resp. floor(i/s)); values 1."""
    n = m if n is None else n
    s = int(n ** 0.5)
    indices = []
    for i in range(m):
        tmp = (list(range(max(i - s + 1, 0), min(i + s, n)))
               + [i + j * s for j in range(1, int((n - 1 - i) / s))]
               + [i - j * s for j in range(1, int(i / s))])
        indices.append(np.unique(tmp))
    indptr = np.cumsum([0] + [len(t) for t in indices]).astype(np.int32)
    idx = np.concatenate(indices).astype(np.int32)
    A = scipy.sparse.csr_matrix((np.ones(len(idx), np.float32), idx, indptr), shape=(m, n))
    return A


def load_edge_list(path, one_based=False, symmetric=False):
    """SNAP-style edge list ('#' comments, two integer columns) -> CSR of ones
    (the on-disk format read by load_snap_graph, gen_formats_v2.py:2629-2705)."""
    data = np.loadtxt(path, comments="#", dtype=np.int64, ndmin=2)
    r, c = data[:, 0], data[:, 1]
    if one_based:
        r, c = r - 1, c - 1
    if symmetric:
        r, c = np.concatenate([r, c]), np.concatenate([c, r])
    n = int(max(r.max(), c.max())) + 1 if len(r) else 0
    A = _csr(r, c, np.ones(len(r), np.float32), (n, n))
    A.data[:] = 1.0
    return A

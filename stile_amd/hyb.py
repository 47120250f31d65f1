"""The reference's SparseTIR "hyb" comparator (SPMM/SPMMbench_spmm_hyb.py:156-316).

SparseTIR (uwsampl/sparsetir-artifact; no version pinned in the reference,
not vendored) converts A with `column_part_hyb(m, n, indptr, indices,
num_col_parts, bucket_sizes)`: the columns are cut into num_col_parts equal
ranges; inside each range every row goes to the smallest bucket whose size
holds its nonzeros there, and a row longer than the largest bucket is cut
into pieces of that size (the last piece into the bucket that fits it).
Each bucket is an ELL matrix (row ids, column ids, values; padding slots
carry value 0) and is multiplied by its own kernel that adds into C.  The
driver's default buckets are 1, 2, 4, ... up to 2^ceil(log2(nnz / rows))
with one column part (:325-365).

Restated here from that published description; the kernels are this
library's (stile_ell_buckets_spmm_f32).
"""
import math

import numpy as np
import scipy.sparse

from . import _native


def default_bucket_sizes(A):
    A = scipy.sparse.csr_matrix(A)
    avg = max(A.nnz / max(A.shape[0], 1), 1.0)
    return [2 ** i for i in range(int(math.ceil(math.log2(avg))) + 1)]


def column_part_hyb(A, num_col_parts=1, bucket_sizes=None):
    """ELL buckets of A: a list over (column part, bucket) of dicts
    {width, row[rows], cols[rows, width], vals[rows, width]}."""
    A = scipy.sparse.csr_matrix(A, dtype=np.float32)
    A.sort_indices()
    M, K = A.shape
    bucket_sizes = sorted(bucket_sizes or default_bucket_sizes(A))
    bmax = bucket_sizes[-1]
    width = int(math.ceil(K / num_col_parts)) if K else 0
    rows_all = np.repeat(np.arange(M), np.diff(A.indptr))
    out = []
    for part in range(num_col_parts):
        sel = (A.indices >= part * width) & (A.indices < (part + 1) * width)
        r, c, v = rows_all[sel], A.indices[sel], A.data[sel]
        if len(r) == 0:
            continue
        # position of every entry inside its row (within this part)
        starts = np.searchsorted(r, np.arange(M))
        pos = np.arange(len(r)) - starts[r]
        # pieces of bmax entries; each piece is one ELL row
        piece = pos // bmax
        key = r.astype(np.int64) * (np.max(piece) + 1) + piece
        uniq, inv, plen = np.unique(key, return_inverse=True, return_counts=True)
        prow = (uniq // (np.max(piece) + 1)).astype(np.int32)
        slot = pos - piece * bmax
        bidx = np.searchsorted(np.asarray(bucket_sizes), plen)   # smallest bucket >= length
        for b, bs in enumerate(bucket_sizes):
            ids = np.nonzero(bidx == b)[0]
            if len(ids) == 0:
                continue
            local = np.full(len(uniq), -1, np.int64)
            local[ids] = np.arange(len(ids))
            take = local[inv] >= 0
            cols = np.zeros((len(ids), bs), np.int32)
            vals = np.zeros((len(ids), bs), np.float32)
            # padding slots repeat the piece's first column with value 0
            first = np.zeros(len(ids), np.int32)
            li = local[inv[take]]
            cols[li, slot[take]] = c[take]
            vals[li, slot[take]] = v[take]
            first[li[slot[take] == 0]] = c[take][slot[take] == 0]
            pad = vals == 0
            pad &= np.arange(bs)[None, :] >= plen[ids][:, None]
            cols[pad] = np.repeat(first[:, None], bs, axis=1)[pad]
            out.append(dict(width=bs, row=prow[ids], cols=cols, vals=vals))
    return out


class HybSpMM(object):
    """Device copy of A's ELL buckets; run(B) = A . B (FP32)."""

    def __init__(self, A, num_col_parts=1, bucket_sizes=None, device="cuda"):
        import torch
        A = scipy.sparse.csr_matrix(A, dtype=np.float32)
        self.shape = A.shape
        self.buckets = column_part_hyb(A, num_col_parts, bucket_sizes)
        for b in self.buckets:
            if b["cols"].size and (b["cols"].min() < 0 or b["cols"].max() >= A.shape[1]):
                raise ValueError("bucket column out of range")
        dev = torch.device(device)
        self._dev = [(torch.from_numpy(b["row"]).to(dev), torch.from_numpy(np.ascontiguousarray(b["cols"])).to(dev),
                      torch.from_numpy(np.ascontiguousarray(b["vals"])).to(dev), b["width"]) for b in self.buckets]

    def run(self, B, out=None):
        import torch
        M, K = self.shape
        if B.shape[0] != K or B.dtype != torch.float32:
            raise _native.StileError("B must be float32 with A's column count of rows")
        B = B.contiguous()
        C = out if out is not None else torch.empty((M, B.shape[1]), device=B.device, dtype=torch.float32)
        return _native.ell_buckets_spmm_f32(self._dev, M, B, C)

"""Direct builders of reference-layout packed groups (the prepare_inputs_BSPMM /
prepare_inputs_BSDDMM layouts) from a CSR matrix and a chosen set of dense
16x16 blocks.  Test helpers: they let the kernel tests exercise every group
shape (window-id / 2-D row windows, split ELL rows, sentinel rows, padding)
without going through the search."""
import numpy as np
import scipy.sparse


def dense_blocks(A, min_fill):
    """(window, block column) pairs of 16x16 blocks holding >= min_fill nonzeros."""
    A = scipy.sparse.csr_matrix(A)
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    key = (rows // 16).astype(np.int64) * ((A.shape[1] + 15) // 16) + A.indices // 16
    u, c = np.unique(key, return_counts=True)
    sel = u[c >= min_fill]
    nbc = (A.shape[1] + 15) // 16
    return sel // nbc, sel % nbc


def spmm_pack(A, min_fill=64, I=8, Kell=32, perm=None, tc_group_cols=16, sentinel_rows=0, seed=0):
    """Split A into TC tiles (dense blocks) and ELL tiles (the rest).

    perm: optional row order for the TC windows (window s holds rows
    perm[16s:16s+16], i_indices 2-D); otherwise windows are row ids (1-D).
    tc_group_cols: K of the TC group (16 or a multiple): consecutive selected
    block columns of a window are packed into one tile.
    sentinel_rows: number of extra ELL tile rows pointing at the sentinel row
    M with nonzero garbage values (must be discarded)."""
    A = scipy.sparse.csr_matrix(A, dtype=np.float32)
    M, K = A.shape
    A.sort_indices()
    rows = np.repeat(np.arange(M), np.diff(A.indptr))
    cols = A.indices.astype(np.int64)
    vals = A.data.astype(np.float32)
    if perm is None:
        prow = rows
        inv = None
    else:
        perm = np.asarray(perm)
        inv = np.argsort(perm)
        prow = inv[rows]                      # position of the row in TC order
    wins = prow // 16
    nbc = (K + 15) // 16
    Ap = scipy.sparse.csr_matrix((np.ones(len(rows)), (prow, cols)), shape=(M, K))
    bw, bc = dense_blocks(Ap, min_fill)
    blk_key = set((int(w) * nbc + int(c)) for w, c in zip(bw, bc))
    key = wins * nbc + cols // 16
    in_tc = np.array([k in blk_key for k in key.tolist()], dtype=bool) if len(key) else np.zeros(0, bool)

    # TC group
    tc = []
    if len(bw):
        G = tc_group_cols
        per_win = {}
        for w, c in zip(bw.tolist(), bc.tolist()):
            per_win.setdefault(w, []).append(c)
        win_ids = sorted(per_win)
        tiles_a, tiles_idx, indptr = [], [], [0]
        for w in win_ids:
            cl = sorted(per_win[w])
            colsets = [np.arange(c * 16, min(c * 16 + 16, K)) for c in cl]
            allc = np.concatenate(colsets)
            for t0 in range(0, len(allc), G):
                tc_cols = allc[t0:t0 + G]
                a = np.zeros((16, G), np.float32)
                sel = in_tc & (wins == w) & np.isin(cols, tc_cols)
                r_in = prow[sel] - 16 * w
                c_in = np.searchsorted(tc_cols, cols[sel])
                a[r_in, c_in] = vals[sel]
                idx = np.full(G, tc_cols[-1], np.int32)
                idx[:len(tc_cols)] = tc_cols
                tiles_a.append(a)
                tiles_idx.append(idx)
            indptr.append(len(tiles_a))
        if perm is None:
            i_ind = np.asarray(win_ids, np.int32)
        else:
            i_ind = np.zeros((len(win_ids), 16), np.int32)
            for s, w in enumerate(win_ids):
                rr = perm[16 * w:16 * w + 16]
                i_ind[s, :len(rr)] = rr
                i_ind[s, len(rr):] = rr[-1]
        tc.append(dict(group_size=G, num_windows=len(win_ids), nnzb=len(tiles_a),
                       rows_2d=perm is not None, k_notsorted=False, real_atomic=True,
                       a=np.stack(tiles_a), i_indices=i_ind,
                       indptr=np.asarray(indptr, np.int32), indices=np.stack(tiles_idx)))

    # ELL group: remaining nonzeros, rows split into pieces of Kell.
    er, ec, ev = rows[~in_tc], cols[~in_tc], vals[~in_tc]
    pieces_r, pieces_c, pieces_v = [], [], []
    order = np.argsort(er, kind="stable")
    er, ec, ev = er[order], ec[order], ev[order]
    starts = np.searchsorted(er, np.arange(M))
    ends = np.searchsorted(er, np.arange(M), side="right")
    for r in range(M):
        for s in range(starts[r], ends[r], Kell):
            e = min(ends[r], s + Kell)
            pieces_r.append(r)
            pieces_c.append(ec[s:e])
            pieces_v.append(ev[s:e])
    rng = np.random.default_rng(seed)
    for _ in range(sentinel_rows):
        pieces_r.append(M)
        pieces_c.append(rng.integers(0, K, size=3))
        pieces_v.append(np.full(3, 7.0, np.float32))
    ell = []
    if pieces_r:
        T = (len(pieces_r) + I - 1) // I
        a = np.zeros((T * I, Kell), np.float32)
        ik = np.zeros((T * I, Kell), np.int32)
        ir = np.full(T * I, pieces_r[-1], np.int32)
        for q, (r, c, v) in enumerate(zip(pieces_r, pieces_c, pieces_v)):
            a[q, :len(c)] = v
            ik[q, :len(c)] = c
            ir[q] = r
        ell.append(dict(m_blk=I, nnz_cols=Kell, num_tiles=T, real_atomic=True,
                        a=a.reshape(T, I * Kell), indices_k=ik, indices_reordered_i=ir))
    return dict(M=M, K=K, ell=ell, tc=tc)


def sddmm_pack(A, min_fill=64, bucket=32):
    """1D buckets for most nonzeros, 16x16 TC blocks for dense blocks."""
    A = scipy.sparse.csr_matrix(A)
    A.sort_indices()
    M, N = A.shape
    rows = np.repeat(np.arange(M), np.diff(A.indptr))
    cols = A.indices.astype(np.int64)
    ids = np.arange(A.nnz)
    nbc = (N + 15) // 16
    bw, bc = dense_blocks(A, min_fill)
    blk = dict(((int(w), int(c)), q) for q, (w, c) in enumerate(zip(bw, bc)))
    key_w, key_c = rows // 16, cols // 16
    in_tc = np.array([(int(w), int(c)) in blk for w, c in zip(key_w, key_c)], bool) if A.nnz else np.zeros(0, bool)
    g1 = []
    sel = ~in_tc
    if sel.any():
        mid, jj, rm = rows[sel], cols[sel], ids[sel]
        pad = (-len(mid)) % bucket
        mid = np.concatenate([mid, np.full(pad, mid[-1])])
        jj = np.concatenate([jj, np.full(pad, jj[-1])])
        rm = np.concatenate([rm, np.full(pad, rm[-1])])
        g1.append(dict(mid_idx=mid, j_idx=jj, remap=rm))
    tc = []
    if len(bw):
        i_ind, j_ind, indptr, dense, outi = [], [], [0], [], []
        for (w, c) in sorted(blk):
            rr = np.minimum(np.arange(16 * w, 16 * w + 16), M - 1)
            cc = np.minimum(np.arange(16 * c, 16 * c + 16), N - 1)
            i_ind.append(rr)
            j_ind.append(cc)
            s = in_tc & (key_w == w) & (key_c == c)
            # CSC order within the block (reference: nnz_id_matrix[...].tocsc())
            r_in, c_in, o = rows[s] - 16 * w, cols[s] - 16 * c, ids[s]
            order = np.lexsort((r_in, c_in))
            dense.append(c_in[order] + 16 * r_in[order])
            outi.append(o[order])
            indptr.append(indptr[-1] + int(s.sum()))
        tc.append(dict(nnzb=len(i_ind), blk_k=16, i_indices=np.stack(i_ind), j_indices=np.stack(j_ind),
                       indptr=np.asarray(indptr), indices_dense=np.concatenate(dense),
                       indices=np.concatenate(outi)))
    return dict(M=M, N=N, nnz=A.nnz, g1d=g1, tc=tc)


def add_value_ids(pack, A):
    """Set a_ids of every group of an SpMM pack built from A: the position of
    each packed value in A's CSR value order (-1 for padding)."""
    A = scipy.sparse.csr_matrix(A)
    A.sort_indices()
    M, K = A.shape
    keys = np.repeat(np.arange(M, dtype=np.int64), np.diff(A.indptr)) * K + A.indices

    def ids_of(rows, cols, present):
        q = np.asarray(rows, np.int64) * K + np.asarray(cols, np.int64)
        pos = np.searchsorted(keys, np.where(present, q, 0))
        return np.where(present, pos, -1).astype(np.int32)

    for g in pack["ell"]:
        Kc = g["nnz_cols"]
        r = np.repeat(g["indices_reordered_i"], Kc).reshape(-1, Kc)
        a = g["a"].reshape(-1, Kc)
        g["a_ids"] = ids_of(r, g["indices_k"], (a != 0) & (r < M)).reshape(g["a"].shape)
    for g in pack["tc"]:
        ii = g["i_indices"] if g["rows_2d"] else g["i_indices"][:, None] * 16 + np.arange(16)[None, :]
        win_of = np.repeat(np.arange(g["num_windows"]), np.diff(g["indptr"]))
        rows = np.broadcast_to(np.asarray(ii)[win_of][:, :, None], g["a"].shape)
        cols = np.broadcast_to(g["indices"][:, None, :], g["a"].shape)
        g["a_ids"] = ids_of(rows, cols, g["a"] != 0)
    return pack

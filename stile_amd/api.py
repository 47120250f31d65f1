"""High-level entry points: the reference drivers' flow as functions.

The reference's do_profile (SPMM/SPMMbench_our_method.py:161-342,
SDDMM/SDDMMbench_our_method.py:158-359) builds an Op, estimates the ELL
bucket size, runs greedy_search_use_cost_model_lower_bound, applies
post_process_for_withdraw and measures the fused kernel.  These helpers run
the same sequence and return a device plan (stile_amd._native.Plan) that can
be reused for any dense width N.
"""
import os
from dataclasses import dataclass

import numpy as np
import scipy.sparse

from . import my_branch_and_bound, my_cost_model, my_fuse_formats, my_search_formats
from .gen_formats_v2 import DenseTensor, Op, SearchConfig, _as_csr, set_config


@dataclass
class SearchOptions:
    """Driver constants of SPMM/SPMMbench_our_method.py:198-205 (and the SDDMM
    driver), made explicit."""
    dsize: int = 16
    max_avg_cost_diff: float = 0.2
    max_level_num: float = float('inf')
    only_TC: bool = False
    only_ELL: bool = False
    TC_k_notsorted: bool = False
    reorder_k_by_nnz: bool = True
    no_withdraw: bool = False
    penalty: float = 1.0
    max_bucket_size: int = None       # SpMM: estimate_max_bucket_size; SDDMM: 32
    profile: str = None               # cost-model profile ('a100' or 'mi355x')


class SparseMatrix(object):
    """A CSR matrix (int32 indices, sorted) with its values."""

    def __init__(self, A):
        self.A = _as_csr(scipy.sparse.csr_matrix(A, dtype=np.float32))

    @property
    def shape(self):
        return self.A.shape


def _csr_of(m):
    return m.A if isinstance(m, SparseMatrix) else _as_csr(scipy.sparse.csr_matrix(m, dtype=np.float32))


def spmm_op(A, N=32, B=None):
    """Op('spmm', idx_lens = [M, N, K]) as built by the reference loaders."""
    A = _csr_of(A)
    M, K = A.shape
    Bd = B if B is not None else np.zeros((K, N), np.float32)
    return Op('spmm', [0, 1], [2], [M, Bd.shape[1], K], None, [A, DenseTensor(Bd)], [A.tocsc()])


def sddmm_op(A, Kf=32, X=None, Y=None):
    """Op('sddmm', idx_lens = [M, K_feat, N])."""
    A = _csr_of(A)
    M, N = A.shape
    Xd = X if X is not None else np.zeros((M, Kf), np.float32)
    Yd = Y if Y is not None else np.zeros((N, Kf), np.float32)
    return Op('sddmm', [0, 1], [2], [M, Xd.shape[1], N], None, [A, DenseTensor(Xd), DenseTensor(Yd)], [A.tocsc()])


def search_tiles(op, opts=None):
    """Greedy search + withdraw post-processing; returns the final tile list."""
    opts = opts or SearchOptions()
    prev = my_cost_model.get_profile().name
    if opts.profile:
        my_cost_model.set_profile(opts.profile)
    try:
        if op.op_type == 'spmm':
            mbs = opts.max_bucket_size or my_search_formats.estimate_max_bucket_size(op)
        else:
            mbs = opts.max_bucket_size or 32
        max_avg = float('inf') if opts.only_ELL else opts.max_avg_cost_diff
        res = my_search_formats.greedy_search_use_cost_model_lower_bound(
            op, opts.dsize, {}, mbs, ['A'], ([1], [1], [1]), opts.only_TC, opts.only_ELL, opts.penalty,
            opts.TC_k_notsorted, reorder_k_by_nnz=opts.reorder_k_by_nnz, max_level_num=opts.max_level_num,
            max_avg_cost_diff=max_avg, no_withdraw=opts.no_withdraw)
    finally:
        my_cost_model.set_profile(prev)
    tiles, row_dict, row_costs, rev, _levels = res
    tiles = my_branch_and_bound.post_process_for_withdraw(op, tiles, row_dict, row_costs, rev)
    return tiles


def spmm_pack(A, opts=None, N=32):
    """Run the search on A and return the packed groups (reference layouts)."""
    op = spmm_op(A, N)
    tiles = search_tiles(op, opts)
    formats = my_fuse_formats.get_formats_from_selected_tiles(tiles, ['A'], (opts or SearchOptions()).dsize)
    return my_fuse_formats.prepare_inputs_BSPMM(op, 'spmm', 1, formats), tiles, op


def sddmm_pack(A, opts=None, Kf=32):
    op = sddmm_op(A, Kf)
    tiles = search_tiles(op, opts)
    formats = my_fuse_formats.get_formats_from_selected_tiles(tiles, ['A'], (opts or SearchOptions()).dsize)
    return my_fuse_formats.prepare_inputs_BSDDMM(op, 'sddmm', 1, formats), tiles, op


# ---------------------------------------------------------------------------
# Row-partitioned search for large matrices.  The search is a greedy pass over
# 16-row windows and row buckets; cutting A into row blocks of whole windows,
# searching each block in its own process and stacking the packs gives one
# plan for A.  ELL tiles then group rows of similar length within a block
# rather than across the whole matrix (blocks hold ~10^5 rows).
# ---------------------------------------------------------------------------

def parallel_map(fn, jobs, workers=None, log=None):
    """fn over jobs in worker processes: forked while this process has not
    started the HIP runtime, spawned otherwise (a forked copy of an
    initialised HIP runtime is not usable).  The runtime counts as started once
    torch has initialised CUDA or this package has created a device plan.
    Spawned workers re-import the main module, so a script calling this after
    GPU work needs the `if __name__ == "__main__"` guard.  The pool is closed
    and joined (not terminated) when every job is done."""
    import multiprocessing as mp
    import os
    workers = min(len(jobs), workers or min(os.cpu_count() or 1, 16))
    if workers <= 1:
        return [fn(j) for j in jobs]
    from . import _native
    spawn = _native.hip_started()
    try:
        import torch
        spawn = spawn or torch.cuda.is_initialized()
    except ImportError:
        pass
    ctx = mp.get_context("spawn" if spawn else "fork")
    pool = ctx.Pool(workers)
    try:
        if log is None:
            out = pool.map(fn, jobs, chunksize=1)
        else:
            import time as _time
            out, t_last = [None] * len(jobs), _time.time()
            for done, (i, r) in enumerate(pool.imap_unordered(_indexed, [(fn, i, j) for i, j in enumerate(jobs)]),
                                          1):
                out[i] = r
                if _time.time() - t_last > 20 or done == len(jobs):
                    log(f"  {done}/{len(jobs)} jobs done")
                    t_last = _time.time()
        pool.close()
        pool.join()
        return out
    except BaseException:
        pool.terminate()
        pool.join()
        raise


def _indexed(args):
    fn, i, job = args
    return i, fn(job)


def row_blocks(A, parts):
    """Row cut points (multiples of 16) giving `parts` blocks of about equal nnz."""
    A = _csr_of(A)
    M = A.shape[0]
    if parts <= 1 or M <= 16:
        return np.array([0, M])
    targets = np.linspace(0, A.nnz, parts + 1)[1:-1]
    cuts = np.searchsorted(A.indptr, targets)
    cuts = np.unique(np.clip((cuts // 16) * 16, 16, (M - 1) // 16 * 16))
    return np.concatenate([[0], cuts[cuts > 0], [M]]).astype(np.int64)


def _block_job(args):
    kind, A, opts, width = args
    if kind == "spmm":
        return spmm_pack(A, opts, width)[0]
    return sddmm_pack(A, opts, width)[0]


def _shift_ids(a, off):
    a = np.array(a, np.int32, copy=True)
    a[a >= 0] += off
    return a


def stack_spmm_packs(packs, cuts, nnz_offs, M, K):
    """One SpMM pack for A from the packs of its row blocks A[cuts[b]:cuts[b+1]]
    (rows shifted, the block sentinel row -> M, value ids shifted by the
    block's first CSR position)."""
    ell, tc = [], []
    for pk, r0, r1, z0 in zip(packs, cuts[:-1], cuts[1:], nnz_offs):
        Mb = int(r1 - r0)
        for g in pk["ell"]:
            g = dict(g)
            rows = np.asarray(g["indices_reordered_i"], np.int64)
            g["indices_reordered_i"] = np.where(rows == Mb, M, rows + r0).astype(np.int32)
            if g.get("a_ids") is not None:
                g["a_ids"] = _shift_ids(g["a_ids"], z0)
            ell.append(g)
        for g in pk["tc"]:
            g = dict(g)
            ii = np.asarray(g["i_indices"], np.int64)
            if g["rows_2d"]:
                g["i_indices"] = np.where(ii == Mb, M, ii + r0).astype(np.int32)
            else:
                g["i_indices"] = (ii + r0 // 16).astype(np.int32)
            if g.get("a_ids") is not None:
                g["a_ids"] = _shift_ids(g["a_ids"], z0)
            tc.append(g)
    return dict(M=int(M), K=int(K), ell=ell, tc=tc)


def stack_sddmm_packs(packs, cuts, nnz_offs, M, N, nnz):
    """One SDDMM pack for A from the packs of its row blocks."""
    g1d, tc = [], []
    for pk, r0, z0 in zip(packs, cuts[:-1], nnz_offs):
        for g in pk["g1d"]:
            g1d.append(dict(mid_idx=(np.asarray(g["mid_idx"]) + r0).astype(np.int32), j_idx=g["j_idx"],
                            remap=(np.asarray(g["remap"]) + z0).astype(np.int32)))
        for g in pk["tc"]:
            g = dict(g)
            g["i_indices"] = (np.asarray(g["i_indices"]) + r0).astype(np.int32)
            g["indices"] = (np.asarray(g["indices"]) + z0).astype(np.int32)
            tc.append(g)
    return dict(M=int(M), N=int(N), nnz=int(nnz), g1d=g1d, tc=tc)


def _blocked_pack(kind, A, opts, width, parts, workers, nnz_per_part, log=None):
    A = _csr_of(A)
    if parts is None:
        parts = max(1, int(round(A.nnz / nnz_per_part)))
    cuts = row_blocks(A, parts)
    if len(cuts) == 2:
        return (spmm_pack if kind == "spmm" else sddmm_pack)(A, opts, width)[0]
    blocks = [A[int(a):int(b)] for a, b in zip(cuts[:-1], cuts[1:])]
    packs = parallel_map(_block_job, [(kind, B, opts, width) for B in blocks], workers, log)
    offs = A.indptr[cuts[:-1]].astype(np.int64)
    if kind == "spmm":
        return stack_spmm_packs(packs, cuts, offs, A.shape[0], A.shape[1])
    return stack_sddmm_packs(packs, cuts, offs, A.shape[0], A.shape[1], A.nnz)


def spmm_pack_blocked(A, opts=None, N=32, parts=None, workers=None, nnz_per_part=1 << 21, log=None):
    """spmm_pack for large A: row blocks of ~nnz_per_part nonzeros searched in
    parallel processes and stacked into one pack (value ids included)."""
    return _blocked_pack("spmm", A, opts, N, parts, workers, nnz_per_part, log)


def sddmm_pack_blocked(A, opts=None, Kf=32, parts=None, workers=None, nnz_per_part=1 << 21, log=None):
    return _blocked_pack("sddmm", A, opts, Kf, parts, workers, nnz_per_part, log)


def options_for(blk_dtype, opts=None):
    """Search options for plans of the given block type: the cost model's
    element size (dsize) follows the blocks (32 for FP32, 16 for BF16/FP16)
    unless opts sets it."""
    if opts is not None:
        return opts
    return SearchOptions(dsize=32 if blk_dtype in ("float32", "f32", 0) else 16)


def build_spmm_plan(A, blk_dtype="bfloat16", opts=None, N=32):
    """Search + pack + device plan for C = A.B (any N at run time)."""
    from . import _native
    pack, _, _ = spmm_pack(A, options_for(blk_dtype, opts), N)
    return _native.Plan.spmm(pack, blk_dtype)


def build_sddmm_plan(A, opts=None, Kf=32):
    from . import _native
    pack, _, _ = sddmm_pack(A, opts, Kf)
    return _native.Plan.sddmm(pack, "float32")


def build_spmm_plan_batched(mats, blk_dtype="bfloat16", opts=None, N=32):
    """One plan over many (uneven) matrices: their block-diagonal composition."""
    from . import _native
    packs = [spmm_pack(A, options_for(blk_dtype, opts), N)[0] for A in mats]
    return _native.Plan.spmm(packs, blk_dtype)


def build_sddmm_plan_batched(mats, opts=None, Kf=32):
    from . import _native
    packs = [sddmm_pack(A, opts, Kf)[0] for A in mats]
    return _native.Plan.sddmm(packs, "float32")


# ---------------------------------------------------------------------------
# Pack serialisation (.npz).  A pack is what the search produces; saving it
# skips the search on the next run (a plan is rebuilt from it in
# milliseconds).
# ---------------------------------------------------------------------------

_GROUP_LISTS = ("ell", "tc", "g1d")


def _pack_to_arrays(pack, prefix, arrays):
    meta = {"scalars": {}, "groups": {}}
    for k, v in pack.items():
        if k in _GROUP_LISTS:
            meta["groups"][k] = []
            for i, g in enumerate(v):
                gm = {}
                for f, x in g.items():
                    if x is None:
                        gm[f] = None
                    elif isinstance(x, np.ndarray):
                        arrays[f"{prefix}{k}{i}.{f}"] = x
                        gm[f] = "array"
                    else:
                        gm[f] = x.item() if isinstance(x, np.generic) else x
                meta["groups"][k].append(gm)
        elif k == "order":
            meta["order"] = [list(o) for o in v]
        else:
            meta["scalars"][k] = int(v)
    return meta


def _pack_from_arrays(meta, prefix, z):
    pack = dict(meta["scalars"])
    for k, groups in meta["groups"].items():
        pack[k] = []
        for i, gm in enumerate(groups):
            pack[k].append({f: (np.array(z[f"{prefix}{k}{i}.{f}"]) if x == "array" else x) for f, x in gm.items()})
    if "order" in meta:
        pack["order"] = [tuple(o) for o in meta["order"]]
    return pack


def matrix_fingerprint(kind, mats):
    """Identity of the sparse operand(s) a pack was built for: the operation,
    each matrix's shape and nnz, and a SHA-1 over each matrix's CSR indptr
    and indices (int64).  Values are not hashed: a pack carries value ids, and
    plans refresh values through them."""
    import hashlib
    lst = mats if isinstance(mats, (list, tuple)) else [mats]
    h = hashlib.sha1()
    shapes, nnzs = [], []
    for A in lst:
        A = _csr_of(A)
        shapes.append([int(A.shape[0]), int(A.shape[1])])
        nnzs.append(int(A.nnz))
        h.update(np.ascontiguousarray(A.indptr, dtype=np.int64).tobytes())
        h.update(np.ascontiguousarray(A.indices, dtype=np.int64).tobytes())
    return {"kind": kind, "count": len(lst), "shapes": shapes, "nnz": nnzs, "sha1": h.hexdigest()}


class StalePackError(ValueError):
    """A saved pack was built for a different matrix than the one given."""


def save_pack(path, pack, fingerprint=None, build=None):
    """Write an SpMM or SDDMM pack (the dicts of spmm_pack / sddmm_pack /
    the blocked variants), or a list of them (a batched plan's packs), to an
    .npz file: array fields as arrays, scalar fields and group counts as a
    JSON header, with the matrix fingerprint (matrix_fingerprint) and the
    plan build options (e.g. {"scatter_max_nnz": 2}) if given."""
    import json
    arrays = {}
    if isinstance(pack, (list, tuple)):
        meta = {"list": [_pack_to_arrays(p, f"p{i}/", arrays) for i, p in enumerate(pack)]}
    else:
        meta = _pack_to_arrays(pack, "", arrays)
    if fingerprint is not None:
        meta["fingerprint"] = fingerprint
    if build:
        meta["build"] = dict(build)
    arrays["__meta__"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez(path, **arrays)


def load_pack(path, expect=None, with_build=False):
    """Read a pack (or list of packs) written by save_pack.  expect: the
    fingerprint of the matrix the caller will run the pack on; a file saved
    with another fingerprint, or with none, raises StalePackError.
    with_build: return (pack, build options dict) instead."""
    import json
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(bytes(z["__meta__"]).decode())
        if expect is not None and meta.get("fingerprint") != expect:
            got = meta.get("fingerprint")
            raise StalePackError(f"{path}: pack built for {_fp_str(got)}, not for {_fp_str(expect)}")
        if "list" in meta:
            pack = [_pack_from_arrays(m, f"p{i}/", z) for i, m in enumerate(meta["list"])]
        else:
            pack = _pack_from_arrays(meta, "", z)
    return (pack, dict(meta.get("build") or {})) if with_build else pack


def _fp_str(fp):
    if not fp:
        return "an unrecorded matrix"
    return f"{fp['kind']} x{fp['count']} shapes {fp['shapes'][:2]} nnz {fp['nnz'][:2]} sha1 {fp['sha1'][:12]}"


# ---------------------------------------------------------------------------
# Measured-latency tuning.  The cost model ranks tiles in isolation; the
# fused kernels' real cost also depends on B-row reuse in the caches and on
# how dense and sparse parts share windows.  The tuner therefore searches the
# matrix under a few settings of the search, builds each plan and keeps the
# one that runs fastest at the given width (the reference likewise times its
# candidate formats, my_fuse_formats.measure_latency).
# ---------------------------------------------------------------------------

def tuning_candidates(blk_dtype, kind="spmm"):
    """Search settings the tuner tries: the measured MI355X cost model (if
    present), the reference's A100 model, and the two single-kind extremes."""
    dsize = 32 if blk_dtype in ("float32", "f32", 0) else 16
    cands = {}
    if my_cost_model.get_profile().name == "mi355x" or os.path.exists(my_cost_model.PROFILE_PATH):
        cands["mi355x"] = SearchOptions(dsize=dsize, profile="mi355x")
    cands["a100"] = SearchOptions(dsize=dsize, profile="a100")
    if kind == "spmm":
        # (an all-TC SDDMM covers every output with 16 x 16 blocks: a slow
        # search that has not beaten the others on any workload measured)
        cands["only_tc"] = SearchOptions(dsize=dsize, only_TC=True)
    cands["only_sparse"] = SearchOptions(dsize=dsize, only_ELL=True)
    return cands


def plan_variants(kind, blk_dtype, mats):
    """Build options the tuner times for every searched pack: scatter chunks
    off and on (at most 2 nonzeros per column) for SpMM plans that run on the
    B-stationary kernel (BF16 blocks, every matrix K <= LDS_MAX_K), else only
    the default build."""
    if kind == "spmm" and blk_dtype in ("bfloat16", "bf16") and \
            max(_csr_of(A).shape[1] for A in _as_list(mats)) <= my_cost_model.LDS_MAX_K:
        return {"": {}, "+scatter2": {"scatter_max_nnz": 2}}
    return {"": {}}


def spmm_plan_signature(packs):
    """SHA-1 of what a device plan built from SpMM packs (one or a list)
    depends on: per matrix the (row, column) of every nonzero held by TC
    tiles, the row sets of the TC windows, and the (row, column) of every
    nonzero held by ELL tiles.  Packs that differ only in how the TC
    nonzeros are grouped into tiles have the same signature and give the same
    plan: the plan builder merges a window's tiles into one column list."""
    import hashlib
    h = hashlib.sha1()
    for pk in _as_list(packs):
        M, K = int(pk["M"]), int(pk["K"])
        tc, wins, ell = [], [], []
        for g in pk["ell"]:
            I, Kc, T = int(g["m_blk"]), int(g["nnz_cols"]), int(g["num_tiles"])
            a = np.asarray(g["a"]).reshape(T * I, Kc)
            r = np.repeat(np.asarray(g["indices_reordered_i"], np.int64), Kc).reshape(T * I, Kc)
            m = (a != 0) & (r < M)
            if g.get("a_ids") is not None:
                m |= (np.asarray(g["a_ids"]).reshape(T * I, Kc) >= 0) & (r < M)
            ell.append(r[m] * K + np.asarray(g["indices_k"], np.int64).reshape(T * I, Kc)[m])
        for g in pk["tc"]:
            T, Kg, W = int(g["nnzb"]), int(g["group_size"]), int(g["num_windows"])
            ii = np.asarray(g["i_indices"], np.int64)
            rr = ii.reshape(W, 16) if g["rows_2d"] else ii[:, None] * 16 + np.arange(16)[None, :]
            rr = np.where(rr < M, rr, -1)
            wins.append(rr)
            win_of = np.repeat(np.arange(W), np.diff(np.asarray(g["indptr"], np.int64)))
            a = np.asarray(g["a"]).reshape(T, 16, Kg)
            present = a != 0
            if g.get("a_ids") is not None:
                present |= np.asarray(g["a_ids"]).reshape(T, 16, Kg) >= 0
            if g["k_notsorted"]:
                cols = np.asarray(g["indices"], np.int64)[:, None] * 16 + np.arange(16)[None, :]
            else:
                cols = np.asarray(g["indices"], np.int64).reshape(T, Kg)
            t, r, k = np.nonzero(present)
            tc.append(rr[win_of[t], r] * K + cols[t, k])
        for arr in (np.unique(np.concatenate(tc)) if tc else np.zeros(0, np.int64),
                    np.unique(np.concatenate(wins), axis=0) if wins else np.zeros((0, 16), np.int64),
                    np.unique(np.concatenate(ell)) if ell else np.zeros(0, np.int64)):
            h.update(np.ascontiguousarray(arr, np.int64).tobytes())
            h.update(b"|")
    return h.hexdigest()


def _pack_any(kind, mats, opts, width, workers, log=None):
    if isinstance(mats, (list, tuple)):
        job = [(kind, A, opts, width) for A in mats]
        return parallel_map(_block_job, job, workers, log)
    if kind == "spmm":
        return spmm_pack_blocked(mats, opts, width, workers=workers, log=log)
    return sddmm_pack_blocked(mats, opts, width, workers=workers, log=log)


def _as_list(mats):
    return list(mats) if isinstance(mats, (list, tuple)) else [mats]


def _offsets(mats, axis):
    return np.concatenate([[0], np.cumsum([_csr_of(A).shape[axis] for A in _as_list(mats)])]).astype(np.int64)


def sample_rows(mats, rows):
    """Rows `rows` (sorted global row ids) of the block-diagonal composition of
    mats (one matrix or a list), as a float64 CSR over the global columns."""
    lst = [_csr_of(A) for A in _as_list(mats)]
    roff, coff = _offsets(lst, 0), _offsets(lst, 1)
    parts = []
    mi = np.searchsorted(roff, rows, side="right") - 1
    for i in np.unique(mi):
        local = np.asarray(rows)[mi == i] - roff[i]
        sub = lst[i][local].tocsr()
        parts.append(scipy.sparse.csr_matrix((sub.data.astype(np.float64), sub.indices.astype(np.int64) + coff[i],
                                              sub.indptr), shape=(len(local), int(coff[-1]))))
    return scipy.sparse.vstack(parts, format="csr")


def spmm_sample_error(mats, B, C, samples=256, seed=0):
    """Norm-wise relative error max|C - A.B| / max|A.B| over `samples` random
    rows of C = A.B (A: mats, B and C device tensors), the products in
    float64 on the host (the reference checks its kernel the same way before
    timing it, my_fuse_formats.py:4194-4217)."""
    import torch
    M = int(C.shape[0])
    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(M, min(M, samples), replace=False))
    S = sample_rows(mats, rows)
    cols = np.unique(S.indices)
    if len(cols) == 0:
        got = C[torch.from_numpy(rows).to(C.device)].double().cpu().numpy()
        return float(np.max(np.abs(got))) if got.size else 0.0
    Bs = B[torch.from_numpy(cols).to(B.device)].double().cpu().numpy()
    Sc = scipy.sparse.csr_matrix((S.data, np.searchsorted(cols, S.indices), S.indptr), shape=(len(rows), len(cols)))
    ref = np.asarray(Sc @ Bs)
    got = C[torch.from_numpy(rows).to(C.device)].double().cpu().numpy()
    scale = np.max(np.abs(ref))
    diff = float(np.max(np.abs(got - ref)))
    return diff / scale if scale > 0 else diff


def sddmm_sample_error(mats, X, Y, out, samples=4096, seed=0):
    """Norm-wise relative error of the SDDMM output on `samples` random
    nonzeros of the block-diagonal pattern, dot products in float64 on the
    host (the reference's check against dgl, my_fuse_formats.py:4331-4354)."""
    import torch
    lst = [_csr_of(A) for A in _as_list(mats)]
    roff, coff = _offsets(lst, 0), _offsets(lst, 1)
    zoff = np.concatenate([[0], np.cumsum([A.nnz for A in lst])]).astype(np.int64)
    nnz = int(zoff[-1])
    if nnz == 0:
        return 0.0
    rng = np.random.default_rng(seed)
    p = np.sort(rng.choice(nnz, min(nnz, samples), replace=False))
    mi = np.searchsorted(zoff, p, side="right") - 1
    grow, gcol = np.empty(len(p), np.int64), np.empty(len(p), np.int64)
    for i in np.unique(mi):
        sel = mi == i
        q = p[sel] - zoff[i]
        grow[sel] = np.searchsorted(lst[i].indptr, q, side="right") - 1 + roff[i]
        gcol[sel] = lst[i].indices[q] + coff[i]
    dev = X.device
    xr = X[torch.from_numpy(grow).to(dev)].double().cpu().numpy()
    yc = Y[torch.from_numpy(gcol).to(dev)].double().cpu().numpy()
    ref = np.einsum("ij,ij->i", xr, yc)
    got = out[torch.from_numpy(p).to(dev)].double().cpu().numpy()
    scale = np.max(np.abs(ref))
    diff = float(np.max(np.abs(got - ref)))
    return diff / scale if scale > 0 else diff


# the reference's accuracy bar for a candidate (assert_allclose rtol = atol =
# 1e-2, my_fuse_formats.py:4204-4208, 4350-4354)
TUNE_TOL = 1e-2


def tune_plan(kind, mats, blk_dtype="bfloat16", width=128, candidates=None, reps=20, workers=None, log=None,
              cache=None, b_dtype="float32", rounds=5):
    """Build the plan of `mats` (one matrix, or a list for a batched plan)
    under each candidate search setting, check each plan's output against a
    float64 product on a sample (relative error <= TUNE_TOL, else the
    candidate is dropped), time the rest and return (fastest plan, its pack,
    report).  Timing: `rounds` turns over the candidates of `reps` launches
    each, compared by median.  kind: 'spmm' (width = N, B of type b_dtype: float32, bfloat16
    or float16, the type the plan will be run with) or 'sddmm' (width =
    K_feat, FP32 blocks).  Every searched pack is timed under each build
    variant (plan_variants: scatter chunks off / on), named search + variant;
    for a B-stationary choice every work-unit list is then timed too and the
    fastest kept as a launch option (build["launch"]).
    report: {"choice": name, "build": build options of the choice, "ms":
    {name: ms per run}, "search_s": {search: seconds}, "err": {name: sampled
    relative error}}.  cache: an .npz path; if it exists and was saved for
    this very pattern (matrix_fingerprint) its pack and build options are
    used without search or timing (report {"choice": "cached"}); a cache
    saved for another pattern raises StalePackError.  Otherwise the chosen
    pack is saved there with its build options."""
    import time as _time
    import torch
    from . import _native
    if kind == "sddmm":
        blk_dtype = "float32"
    fp = matrix_fingerprint(kind, mats)
    def make(pk, bo):
        if kind != "spmm":
            return _native.Plan.sddmm(pk, "float32")
        plan = _native.Plan.spmm(pk, blk_dtype, **{k: v for k, v in bo.items() if k != "launch"})
        for name, value in (bo.get("launch") or {}).items():
            plan.set_option(name, value)
        return plan
    M, K = int(_offsets(mats, 0)[-1]), int(_offsets(mats, 1)[-1])
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    if kind == "spmm":
        Bm = torch.randn(K, width, device=dev, generator=g).to(getattr(torch, str(b_dtype).replace("torch.", "")))
        out = torch.empty(M, width, device=dev)
    else:
        X = torch.randn(M, width, device=dev, generator=g)
        Y = torch.randn(K, width, device=dev, generator=g)

    def check(plan, pack):
        if kind == "spmm":
            plan.run_spmm(Bm, out)
            return spmm_sample_error(mats, Bm, out)
        o = torch.empty(pack_nnz(pack), device=dev)
        plan.run_sddmm(X, Y, o)
        return sddmm_sample_error(mats, X, Y, o)

    if cache is not None and os.path.exists(cache):
        pack, build = load_pack(cache, expect=fp, with_build=True)
        plan = make(pack, build)
        err = check(plan, pack)
        if not err <= TUNE_TOL:
            plan.close()
            raise StalePackError(f"{cache}: plan error {err:.3g} against float64 exceeds {TUNE_TOL}")
        return plan, pack, {"choice": "cached", "cache": str(cache), "build": build, "err": err}
    cands = candidates or tuning_candidates(blk_dtype, kind)
    packs, search_s = {}, {}
    for name, opts in cands.items():          # every search before the GPU is used here
        t0 = _time.perf_counter()
        packs[name] = _pack_any(kind, mats, opts, width, workers, log)
        search_s[name] = round(_time.perf_counter() - t0, 3)
        if log:
            log(f"tune {kind}: searched {name} in {search_s[name]} s")
    variants = plan_variants(kind, blk_dtype, mats)
    # searches that produced the same plan (the same nonzeros in TC tiles,
    # windows and ELL tiles) are built and timed once, under the first name
    same_as = {}
    if kind == "spmm":
        first = {}
        for sname, pack in packs.items():
            same_as[sname] = first.setdefault(spmm_plan_signature(pack), sname)
    times, errs, plans, runs, origin = {}, {}, {}, {}, {}
    for (sname, pack), (vname, build) in ((sp, v) for sp in packs.items() for v in variants.items()):
        name = sname + vname
        origin[name] = (sname, build)
        if same_as.get(sname, sname) != sname:
            continue
        plan = make(pack, build)
        err = check(plan, pack)
        errs[name] = float(f"{err:.4g}")
        if not err <= TUNE_TOL:
            if log:
                log(f"tune {kind} {name}: dropped, error {err:.3g} against float64")
            times[name] = None
            plan.close()
            continue
        plans[name] = plan
        if kind == "spmm":
            runs[name] = (lambda pl=plan: pl.run_spmm(Bm, out))
        else:
            o = torch.empty(pack_nnz(pack), device=dev)
            runs[name] = (lambda pl=plan, o=o: pl.run_sddmm(X, Y, o))
    # the candidates are timed in turn, `rounds` times over, and compared by
    # their median: a drift of the clocks during tuning then shifts every
    # candidate alike instead of favouring the one timed first
    samples = {name: [] for name in plans}
    names = list(runs)
    for rd in range(rounds):
        # each round starts at another candidate, so none always follows the
        # same (possibly very different) plan
        for name in names[rd % len(names):] + names[:rd % len(names)]:
            run = runs[name]
            for _ in range(10):
                run()
            torch.cuda.synchronize()
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            for _ in range(reps):
                run()
            e.record()
            torch.cuda.synchronize()
            samples[name].append(s.elapsed_time(e) / reps)
    best, best_ms, choice = None, float("inf"), None
    for name in plans:
        ms = float(np.median(samples[name]))
        times[name] = round(ms, 5)
        if log:
            log(f"tune {kind} {name}: {ms * 1e3:.2f} us (search {search_s[origin[name][0]]} s, "
                f"error {errs[name]:.2e})")
        if ms < best_ms:
            best, best_ms, choice = plans[name], ms, name
    for name, plan in plans.items():
        if name != choice:
            plan.close()
    for name, (sname, build) in origin.items():       # duplicates report their twin's figures
        twin = same_as.get(sname, sname) + name[len(sname):]
        if twin != name and twin in times:
            times[name] = times[twin]
            errs[name] = errs.get(twin)
    if best is None:
        raise RuntimeError(f"tune {kind}: no candidate plan passed the float64 check ({errs})")
    sname, build = origin[choice]
    build = dict(build)
    lists = _time_unit_lists(best, runs[choice], reps, rounds) if len(variants) > 1 else {}
    if lists:
        t_best = min(lists, key=lists.get)
        best.set_option("lds_list", t_best)
        if t_best >= 0:
            build["launch"] = {"lds_list": t_best}
        if log:
            log(f"tune {kind} {choice}: work-unit list {t_best} ({lists[t_best] * 1e3:.2f} us; "
                f"the plan's pick {lists.get(-1, float('nan')) * 1e3:.2f} us)")
    if cache is not None:
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        save_pack(cache, packs[sname], fingerprint=fp, build=build)
    return best, packs[sname], {"choice": choice, "build": build, "ms": times, "search_s": search_s, "err": errs,
                                "same_plan_as": {k: v for k, v in same_as.items() if k != v},
                                "lds_list_ms": {str(k): round(v, 5) for k, v in lists.items()}}


def _time_unit_lists(plan, run, reps, rounds):
    """Median ms per run of a B-stationary plan under each work-unit list it
    accepts (-1: the plan's own pick), timed in turn over `rounds` rounds.
    The launch option is left at -1."""
    import torch
    from ._native import StileError
    lists = []
    for t in [-1] + list(range(12)):
        try:
            plan.set_option("lds_list", t)
        except StileError:
            continue
        lists.append(t)
    samples = {t: [] for t in lists}
    for rd in range(rounds):
        for t in lists[rd % len(lists):] + lists[:rd % len(lists)]:
            plan.set_option("lds_list", t)
            for _ in range(10):
                run()
            torch.cuda.synchronize()
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            for _ in range(reps):
                run()
            e.record()
            torch.cuda.synchronize()
            samples[t].append(s.elapsed_time(e) / reps)
    plan.set_option("lds_list", -1)
    return {t: float(np.median(v)) for t, v in samples.items()}


def pack_nnz(pack):
    """Number of outputs of an SDDMM pack (or list of packs)."""
    if isinstance(pack, (list, tuple)):
        return sum(int(p["nnz"]) for p in pack)
    return int(pack["nnz"])

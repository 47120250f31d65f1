"""From a selected tile set to device buffers, the fused HIP kernel and its
timing (MI355X build).

Same names as MY_sparse/my_fuse_formats.py:
  get_C_store_atomics (:57-78), comp_SMEM_padding_pattern (:82-111),
  get_formats_from_selected_tiles (:121-192), get_padded_j_TC (:2657-2662),
  set_params_BSPMM (:2670-2849) / set_params_BSDDMM (:2856-2968),
  prepare_inputs_BSPMM (:3321-3537) / prepare_inputs_BSDDMM (:3544-3678),
  measure_seleted_formats (:4755-4847), measure_latency (:4385-4389),
  order_output_BSDDMM (:4247-4288).

The reference turns each tile group into TVMScript, compiles one CUDA kernel
per (matrix, tile set, N) and text-patches its stores into atomics.  Here the
groups are packed into the same integer/value buffers (bit-exact layouts of
prepare_inputs_*), handed to libstile_hip.so through its C ABI, which builds a
window-major plan once, and one fixed set of gfx950 kernels runs every tile
set and every N.
"""
import json
import math
import time

import numpy as np

from .gen_formats_v2 import get_template_str


# ---------------------------------------------------------------------------
# Grouping.
# ---------------------------------------------------------------------------

def get_C_store_atomics(formats):
    out = []
    for fmt in formats:
        template_str, tile_sizes, params = fmt
        params = json.loads(params)
        if template_str == "TensorCore_template":
            out.append(params['real_atomic'])
            continue
        if params['real_atomic']:
            if params.get('use_implicit_unroll', True):
                out.append(params['real_atomic'])
            else:
                out += [params['real_atomic']] * tile_sizes[1][1]
    return out


def comp_SMEM_padding_pattern(op_type, tile_sizes, dsize):
    """(rows per pad, pad units, row length) of the ELL A tile in shared memory
    (reference :82-111).  The HIP kernels do not stage ELL tiles in LDS; the
    pattern is kept for the set_params scalars (max_shared)."""
    warp_size, bank_width, bank_num = 32, 32, 32
    if op_type == 'spmm':
        thread_j = tile_sizes[1][2] if len(tile_sizes[1]) > 2 else 32
        i_num_per_warp = math.ceil(warp_size / thread_j)
        K = math.prod(tile_sizes[2][1:])
        if i_num_per_warp == 1:
            return (1, 0, K)
        second_bank_i = dsize * tile_sizes[2][1] // bank_width
        if second_bank_i == 0:
            return (1, (bank_width - dsize * tile_sizes[2][1]) // dsize, K)
        i_conflict_num = int(math.ceil(bank_num / second_bank_i))
        if i_conflict_num >= i_num_per_warp:
            return (1, 0, K)
        return (i_conflict_num, bank_width // dsize, K)


def _best_tile_sizes(tile):
    """The 32-thread schedule measure_tiles assigns to SpMM/SDDMM tiles
    (reference my_fuse_formats.py:5094-5102)."""
    if tile.best_tile_sizes is not None:
        return tile.best_tile_sizes
    return tile.tile_sizes


def get_formats_from_selected_tiles(selected_tiles, cache_set, dsize, set_atomic=None):
    """Group tiles by (template, tile sizes, params + atomic/reorder flags),
    keeping first-appearance order (reference :121-192)."""
    formats = {}
    if not selected_tiles:
        return formats
    for tile in selected_tiles:
        template_str = get_template_str(tile.op)
        params = dict(tile.best_params if tile.best_params is not None else tile.params)
        params['real_atomic'] = bool(tile.is_atomic_tile) if set_atomic is None else set_atomic
        params['atomic'] = False
        params["idx_reordered"] = [bool(x) for x in tile.op.idx_reordered[0]]
        params['SMEM_in1'] = 'A' in (cache_set or [])
        bts = _best_tile_sizes(tile)
        if template_str == '1D_sddmm':
            key = (template_str, tuple(bts), json.dumps(params))
        else:
            key = (template_str, tuple(tuple(i) for i in bts), json.dumps(params))
        formats.setdefault(key, []).append(tile)
    return formats


def get_padded_j_TC(ori_j_len, mma_n):
    if ori_j_len < mma_n:
        return mma_n
    if ori_j_len > mma_n:
        return math.ceil(ori_j_len / (2 * mma_n)) * 2 * mma_n
    return ori_j_len


def set_params_BSPMM(formats, dsize, op=None):
    """Shape scalars of every group and the globals m_tot, n_tot, k_tot,
    max_shared (reference :2670-2849).  Returned as a dict; the HIP kernels
    take N at run time, so nothing is specialised."""
    groups = []
    max_i, max_j, a_ell_offset = -1, -1, 1
    tiles = None
    for fmt, tiles in formats.items():
        if fmt[0] == "sparse_template_ell":
            m_blk = math.prod(fmt[1][0][1:])
            nnz_cols = math.prod(fmt[1][2][1:])
            n_blk = math.prod(fmt[1][1][1:])
            n = tiles[0].op.idx_lens[1]
            pad = comp_SMEM_padding_pattern('spmm', fmt[1], dsize)
            a_ell_offset = max(a_ell_offset, m_blk * nnz_cols + m_blk * pad[1])
            groups.append(dict(template=fmt[0], m_num=len(tiles), n_num=math.ceil(n / n_blk), nnz_cols=nnz_cols,
                               m_blk=m_blk, n_blk=n_blk))
        elif fmt[0] == "TensorCore_template":
            tile_size, group_size = fmt[1][0][1], fmt[1][2][1]
            feat = get_padded_j_TC(tiles[0].op.idx_lens[1], 16)
            funit = fmt[1][1][1]
            groups.append(dict(template=fmt[0], mb=len(set(t.tile_pos[0] for t in tiles)),
                               nb=math.ceil(tiles[0].op.idx_lens[2] / group_size), nnzb=len(tiles),
                               fb=feat // funit, funit=funit, tile_size=tile_size, group_size=group_size))
            max_i = max(max_i, math.ceil(tiles[0].op.idx_lens[0] / tile_size) * tile_size)
            max_j = max(max_j, feat)
            a_ell_offset = max(a_ell_offset, group_size * 16)
    ref = op if op is not None else tiles[0].op
    max_i = max(max_i, ref.idx_lens[0] + 1)
    max_j = max(max_j, ref.idx_lens[1])
    return dict(groups=groups, m_tot=int(max_i), n_tot=int(max_j), k_tot=int(ref.idx_lens[2]),
                max_shared=int(a_ell_offset))


def set_params_BSDDMM(formats, dsize, op=None):
    """Shape scalars per group (reference :2856-2968): a 1D group's NNZ is its
    nonzero count padded to whole buckets of ty * group_size; a TC group has
    MB = block rows, NB = block columns, NNZB = blocks, NNZ = nonzeros."""
    groups = []
    for fmt, tiles in formats.items():
        if fmt[0] == '1D_sddmm':
            prm = json.loads(fmt[2])
            bucket = prm['ty'] * prm['group_size']
            nnz = int(sum(t.nnz for t in tiles))
            groups.append(dict(template=fmt[0], nnz=-(-nnz // bucket) * bucket, nnz_real=nnz, bucket=bucket))
        else:
            groups.append(dict(template=fmt[0], mb=int(fmt[1][0][1]), nb=int(fmt[1][2][1]), nnzb=len(tiles),
                               nnz=int(sum(t.nnz for t in tiles)), tile_sizes=fmt[1]))
    return dict(groups=groups)


# ---------------------------------------------------------------------------
# Packing (prepare_inputs_*): the bit-exact integer/value buffers.
# ---------------------------------------------------------------------------

def _rank_in_group(keys_sorted):
    """0-based rank of each element within runs of equal keys (sorted)."""
    n = len(keys_sorted)
    if n == 0:
        return np.zeros(0, np.int64)
    start = np.ones(n, bool)
    start[1:] = keys_sorted[1:] != keys_sorted[:-1]
    first = np.maximum.accumulate(np.where(start, np.arange(n), 0))
    return np.arange(n) - first


def _values(tile, ids):
    return tile.op.state.A.data[ids].astype(np.float32)


def pack_ell_group(tiles, tile_sizes):
    """One 'sparse_template_ell' group (reference :3372-3419)."""
    I = math.prod(tile_sizes[0][1:])
    K = math.prod(tile_sizes[2][1:])
    T = len(tiles)
    a = np.zeros((T * I, K), np.float32)
    a_ids = np.full((T * I, K), -1, np.int32)
    indices_k = np.empty((T * I, K), np.int32)
    ir = np.empty(T * I, np.int32)
    t_of, r_of, c_of, v_of, i_of = [], [], [], [], []
    for c, t in enumerate(tiles):
        iv = t.op.idx_values_list[0]
        indices_k[c * I:(c + 1) * I] = iv[2][0]
        i_vals = np.asarray(t.i_vals)
        ir[c * I:(c + 1) * I] = i_vals[-1]
        ir[c * I:c * I + len(i_vals)] = i_vals
        ids = t.covered_ids
        t_of.append(np.full(len(ids), c, np.int64))
        r_of.append(t.covered_rows - t.tile_i_rng_sel[0])
        c_of.append(np.asarray(iv[2])[t.covered_cols])
        v_of.append(_values(t, ids))
        i_of.append(np.asarray(ids, np.int64))
    if T:
        tt, rr, cc, vv, ii = (np.concatenate(x) for x in (t_of, r_of, c_of, v_of, i_of))
        row = tt * I + rr
        order = np.lexsort((cc, row))
        row, cc, vv, ii = row[order], cc[order], vv[order], ii[order]
        slot = _rank_in_group(row)
        if len(slot) and slot.max() >= K:
            raise ValueError("an ELL tile row holds more nonzeros than its K")
        a[row, slot] = vv
        a_ids[row, slot] = ii
        indices_k[row, slot] = cc
    return dict(m_blk=I, nnz_cols=K, num_tiles=T, a=a.reshape(T, I * K), indices_k=indices_k,
                indices_reordered_i=ir, a_ids=a_ids.reshape(T, I * K))


def pack_tc_group(tiles, tile_sizes, params):
    """One 'TensorCore_template' group (reference :3424-3504)."""
    tiles = sorted(tiles, key=lambda t: t.tile_pos)
    I = tile_sizes[0][1]
    K = tile_sizes[2][1]
    T = len(tiles)
    a = np.zeros((T, I, K), np.float32)
    a_ids = np.full((T, I, K), -1, np.int32)
    indices = np.empty((T, K), np.int32)
    for c, t in enumerate(tiles):
        op = t.op
        w = t.tile_pos[0]
        iv = op.idx_values_list[0]
        kv = np.asarray(t.k_vals)
        cols = np.asarray(iv[2])[op.k_vals[w][kv]]
        indices[c] = cols[-1]
        indices[c, :len(cols)] = cols
        if t.covered_ids is not None and len(t.covered_ids):
            r = t.covered_rows - t.tile_i_rng[0]
            cp = op.k_vals[w].pos(t.covered_cols) - t.tile_k_rng_sel[0]
            a[c, r, cp] = _values(t, t.covered_ids)
            a_ids[c, r, cp] = t.covered_ids
    poi_is, counts = np.unique([t.tile_pos[0] for t in tiles], return_counts=True)
    indptr = np.concatenate(([0], np.cumsum(counts))).astype(np.int32)
    rows_2d = bool(params['idx_reordered'][0] or params['real_atomic'])
    if rows_2d:
        i_ind = np.empty((len(poi_is), I), np.int32)
        for s, first in enumerate(indptr[:-1]):
            t = tiles[first]
            iv0 = np.asarray(t.op.idx_values_list[0][0])
            i_ind[s] = iv0[t.tile_i_rng[1]]
            rr = iv0[t.tile_i_rng[0]:t.tile_i_rng[1] + 1]
            i_ind[s, :len(rr)] = rr
    else:
        i_ind = poi_is.astype(np.int32)
    k_notsorted = bool(tiles[0].op.TC_k_notsorted) if T else False
    if k_notsorted:
        indices = np.asarray([int(t.tile_pos[2]) for t in tiles], np.int32)
    return dict(group_size=K, num_windows=len(poi_is), nnzb=T, rows_2d=rows_2d, k_notsorted=k_notsorted,
                real_atomic=bool(params['real_atomic']), a=a, i_indices=i_ind, indptr=indptr, indices=indices,
                a_ids=a_ids)


def prepare_inputs_BSPMM(ori_op, op_type, batch_num, formats, out_shape=None, dtype="float32", dev=None):
    """Pack every group into the reference's argument layout; returns the pack
    (dict with M, K, ell, tc) consumed by stile_plan_create_spmm."""
    ell, tc = [], []
    order = []
    for fmt, tiles in formats.items():
        params = json.loads(fmt[2])
        if fmt[0] == "sparse_template_ell":
            g = pack_ell_group(tiles, fmt[1])
            g["real_atomic"] = params['real_atomic']
            ell.append(g)
            order.append(("ell", len(ell) - 1))
        elif fmt[0] == "TensorCore_template":
            tc.append(pack_tc_group(tiles, fmt[1], params))
            order.append(("tc", len(tc) - 1))
        else:
            raise ValueError(f"unsupported template {fmt[0]}")
    return dict(M=int(ori_op.idx_lens[0]), K=int(ori_op.idx_lens[2]), ell=ell, tc=tc, order=order)


def _pad_run(idx, n):
    idx = np.asarray(idx)
    if len(idx) >= n:
        return idx[:n]
    return np.concatenate([idx, np.full(n - len(idx), idx[-1], idx.dtype)])


def prepare_inputs_BSDDMM(ori_op, op_type, batch_num, formats, out_shape=None, dtype="float32", dev=None):
    """1D groups: per-nonzero (mid_idx, j_idx, remap) padded to a multiple of
    ty*group_size by repeating the last entry; TC groups: block rows/cols,
    indptr, in-block positions (CSC order) and output ids (reference :3544-3678)."""
    g1d, tc = [], []
    for fmt, tiles in formats.items():
        if fmt[0] == "1D_sddmm":
            # tiles of one group may come from different 1D ops (next levels)
            b = fmt[1][0]
            mid = np.concatenate([t.op.ori_row_ids_1d[t.tile_pos[0] * b:(t.tile_pos[0] + 1) * b] for t in tiles])
            jj = np.concatenate([t.op.ori_col_ids_1d[t.tile_pos[0] * b:(t.tile_pos[0] + 1) * b] for t in tiles])
            rm = np.concatenate([t.op.nnz_id_matrix[t.tile_pos[0] * b:(t.tile_pos[0] + 1) * b] for t in tiles])
            params = json.loads(fmt[2])
            real = params['ty'] * params['group_size']
            pad = math.ceil(len(mid) / real) * real - len(mid)
            mid = np.concatenate([mid, np.full(pad, mid[-1])]).astype(np.int32)
            jj = np.concatenate([jj, np.full(pad, jj[-1])]).astype(np.int32)
            rm = np.concatenate([rm, np.full(pad, rm[-1])]).astype(np.int32)
            g1d.append(dict(mid_idx=mid, j_idx=jj, remap=rm))
        elif fmt[0] == "TC_sddmm":
            ts = fmt[1]
            op = tiles[0].op
            if any(t.op is not op for t in tiles):
                raise AssertionError("TC_sddmm tiles of one group must share their op")
            iv = op.idx_values_list[0]
            # a window or column run cut short by the matrix edge is padded
            # with its last index: no sampled entry points into the padding
            i_ind = np.asarray([_pad_run(np.asarray(iv[0])[t.tile_i_rng[0]:t.tile_i_rng[1] + 1], ts[0][1])
                                for t in tiles], np.int32)
            j_ind = np.asarray([_pad_run(np.asarray(iv[2])[op.k_vals[t.tile_pos[0]][t.tile_k_rng[0]:t.tile_k_rng[1] + 1]],
                                         ts[2][1]) for t in tiles], np.int32)
            dense, outs, cnt = [], [], [0]
            for t in tiles:
                e = t.entries()
                r = op._ent_row[e] - t.tile_i_rng[0]
                c = op.k_vals[t.tile_pos[0]].pos(op._ent_col[e]) - t.tile_k_rng[0]
                o = op._ent_id[e]
                srt = np.lexsort((r, c))           # CSC order
                dense.append(c[srt] + r[srt] * ts[2][1])
                outs.append(o[srt])
                cnt.append(len(e))
            tc.append(dict(nnzb=len(tiles), blk_k=ts[2][1], i_indices=i_ind, j_indices=j_ind,
                           indptr=np.cumsum(cnt).astype(np.int32),
                           indices_dense=np.concatenate(dense).astype(np.int32),
                           indices=np.concatenate(outs).astype(np.int32)))
    A = ori_op.inps[0]
    return dict(M=int(A.shape[0]), N=int(A.shape[1]), nnz=int(A.nnz), g1d=g1d, tc=tc)


def prepare_inputs(ori_op, op_type, batch_num, formats, out_shape=None, dtype="float32", dev=None):
    if op_type in ('spmm', 'batched_spmm'):
        return prepare_inputs_BSPMM(ori_op, op_type, batch_num, formats, out_shape, dtype, dev)
    return prepare_inputs_BSDDMM(ori_op, op_type, batch_num, formats, out_shape, dtype, dev)


def order_output_BSDDMM(c_indices, out, nnz):
    """The HIP kernel scatters through the remap ids itself; the output is
    already in CSR order (the reference scatters private buffers on the host)."""
    return out


# ---------------------------------------------------------------------------
# Build + measure.
# ---------------------------------------------------------------------------

class FusedKernel(object):
    """What measure_seleted_formats returns in place of the TVM PackedFunc:
    the device plan plus its pack and group scalars."""

    def __init__(self, plan, pack, params, op_type):
        self.plan = plan
        self.pack = pack
        self.params = params
        self.op_type = op_type

    def __call__(self, *args):
        if self.op_type in ('spmm', 'batched_spmm'):
            return self.plan.run_spmm(*args)
        return self.plan.run_sddmm(*args)


def measure_seleted_formats(ori_op, op_type, batch_num, selected_tiles, cuda_i, cache_set, dsize,
                            gened_inputs=None, dtype="float32", dtype_str=None, zerotype=None,
                            blk_dtype=None):
    """Group, pack and build the device plan (reference :4755-4847).  Returns
    (ori_op, op_type, batch_num, f, args, out_shape, dtype, dev, dev_th,
    cuda_i, c_indices) like the reference; f is a FusedKernel."""
    import torch
    from . import _native
    formats = get_formats_from_selected_tiles(selected_tiles, cache_set, dsize)
    if blk_dtype is None:
        blk_dtype = "bfloat16" if dtype in ("float16", "bfloat16") else "float32"
    if op_type in ('spmm', 'batched_spmm'):
        pack = prepare_inputs_BSPMM(ori_op, op_type, batch_num, formats)
        params = set_params_BSPMM(formats, dsize, ori_op)
        plan = _native.Plan.spmm(pack, blk_dtype)
        out_shape = (ori_op.idx_lens[0], ori_op.idx_lens[1])
    else:
        pack = prepare_inputs_BSDDMM(ori_op, op_type, batch_num, formats)
        params = set_params_BSDDMM(formats, dsize, ori_op)
        plan = _native.Plan.sddmm(pack, "float32")
        out_shape = (ori_op.inps[0].nnz,)
    dev_th = torch.device(f"cuda:{cuda_i}")
    f = FusedKernel(plan, pack, params, op_type)
    if gened_inputs is not None:
        gened_inputs.append(pack)
    return ori_op, op_type, batch_num, f, None, out_shape, dtype, dev_th, dev_th, cuda_i, None


def measure_latency(ori_op, op_type, batch_num, f, args_nd, out_shape, dtype, dev, dev_th, cuda_i, c_indices,
                    warmup=10, iters=100, check=True):
    """Run the fused kernel, check it against the reference's CPU golden
    (scipy A.dot(B) for SpMM, per-nonzero dot for SDDMM; the reference's
    rtol = atol = 1e-2, my_fuse_formats.py:4204-4208, 4350-4354) and return
    the mean latency in ms from HIP events (reference :4152-4380)."""
    import torch
    torch.cuda.set_device(dev_th)
    if op_type in ('spmm', 'batched_spmm'):
        Bh = np.asarray(ori_op.inps[1].data, dtype=np.float32)
        B = torch.from_numpy(Bh).to(dev_th)
        C = torch.empty(out_shape, device=dev_th, dtype=torch.float32)
        run = lambda: f(B, C)
        if check:
            run()
            torch.cuda.synchronize()
            ref = ori_op.inps[0].astype(np.float32).dot(Bh)
            if not np.allclose(C.cpu().numpy(), ref, rtol=1e-2, atol=1e-2):
                raise AssertionError("fused SpMM result differs from scipy beyond rtol=atol=1e-2")
    else:
        X = torch.from_numpy(np.asarray(ori_op.inps[1].data, np.float32)).to(dev_th)
        Y = torch.from_numpy(np.asarray(ori_op.inps[2].data, np.float32)).to(dev_th)
        out = torch.empty(out_shape, device=dev_th, dtype=torch.float32)
        run = lambda: f(X, Y, out)
        if check:
            run()
            torch.cuda.synchronize()
            A = ori_op.inps[0]
            rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
            Xh, Yh = ori_op.inps[1].data, ori_op.inps[2].data
            ref = np.einsum("ij,ij->i", Xh[rows], Yh[A.indices])
            if not np.allclose(out.cpu().numpy(), ref, rtol=1e-2, atol=1e-2):
                raise AssertionError("fused SDDMM result differs from the per-nonzero dot beyond 1e-2")
    for _ in range(warmup):
        run()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for _ in range(iters):
        run()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / iters

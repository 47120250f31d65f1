"""High-level entry points: the reference drivers' flow as functions.

The reference's do_profile (SPMM/SPMMbench_our_method.py:161-342,
SDDMM/SDDMMbench_our_method.py:158-359) builds an Op, estimates the ELL
bucket size, runs greedy_search_use_cost_model_lower_bound, applies
post_process_for_withdraw and measures the fused kernel.  These helpers run
the same sequence and return a device plan (stile_amd._native.Plan) that can
be reused for any dense width N.
"""
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
    if opts.profile:
        my_cost_model.set_profile(opts.profile)
    if op.op_type == 'spmm':
        mbs = opts.max_bucket_size or my_search_formats.estimate_max_bucket_size(op)
    else:
        mbs = opts.max_bucket_size or 32
    max_avg = float('inf') if opts.only_ELL else opts.max_avg_cost_diff
    res = my_search_formats.greedy_search_use_cost_model_lower_bound(
        op, opts.dsize, {}, mbs, ['A'], ([1], [1], [1]), opts.only_TC, opts.only_ELL, opts.penalty,
        opts.TC_k_notsorted, reorder_k_by_nnz=opts.reorder_k_by_nnz, max_level_num=opts.max_level_num,
        max_avg_cost_diff=max_avg, no_withdraw=opts.no_withdraw)
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


def build_spmm_plan(A, blk_dtype="bfloat16", opts=None, N=32):
    """Search + pack + device plan for C = A.B (any N at run time)."""
    from . import _native
    pack, _, _ = spmm_pack(A, opts, N)
    return _native.Plan.spmm(pack, blk_dtype)


def build_sddmm_plan(A, opts=None, Kf=32):
    from . import _native
    pack, _, _ = sddmm_pack(A, opts, Kf)
    return _native.Plan.sddmm(pack, "float32")


def build_spmm_plan_batched(mats, blk_dtype="bfloat16", opts=None, N=32):
    """One plan over many (uneven) matrices: their block-diagonal composition."""
    from . import _native
    packs = [spmm_pack(A, opts, N)[0] for A in mats]
    return _native.Plan.spmm(packs, blk_dtype)


def build_sddmm_plan_batched(mats, opts=None, Kf=32):
    from . import _native
    packs = [sddmm_pack(A, opts, Kf)[0] for A in mats]
    return _native.Plan.sddmm(packs, "float32")

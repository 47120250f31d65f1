"""The reference drivers' call sequence, unchanged, on the MI355X path
(SPMM/SPMMbench_our_method.py:225-299, SDDMM/SDDMMbench_our_method.py:233-335):
Op -> estimate_max_bucket_size -> greedy_search_use_cost_model_lower_bound ->
post_process_for_withdraw -> measure_seleted_formats -> measure_latency.
measure_latency checks the result against the reference's own golden (scipy
A.dot(B) / per-nonzero dot, rtol = atol = 1e-2) before timing; the FP32 result
is additionally compared with the FP64 oracle at 1e-5."""
import numpy as np
import pytest

from oracle import ops
from stile_amd import datasets
from stile_amd.gen_formats_v2 import DenseTensor, Op
from stile_amd import my_branch_and_bound, my_fuse_formats, my_search_formats

pytestmark = pytest.mark.gpu


def _search(op):
    mbs = my_search_formats.estimate_max_bucket_size(op)
    tiles, row_dict, row_costs, rev, _ = my_search_formats.greedy_search_use_cost_model_lower_bound(
        op, 32, {}, mbs, ['A'], ([1], [1], [1]), False, False, 1, False,
        reorder_k_by_nnz=True, max_level_num=float('inf'), max_avg_cost_diff=0.2)
    return my_branch_and_bound.post_process_for_withdraw(op, tiles, row_dict, row_costs, rev)


def test_reference_flow_spmm(cuda):
    import torch
    A, B = datasets.pruned_layer(seed=3, M=512, K=256, N=64)
    M, K = A.shape
    op = Op('spmm', [0, 1], [2], [M, B.shape[1], K], None, [A, DenseTensor(B)], [A.tocsc()])
    tiles = _search(op)
    print(f"reference flow spmm tiles={len(tiles)}", flush=True)
    res = my_fuse_formats.measure_seleted_formats(op, 'spmm', 1, tiles, 0, ['A'], 32)
    ms = my_fuse_formats.measure_latency(*res, warmup=2, iters=5)
    assert ms > 0
    f = res[3]
    C = f(torch.from_numpy(B).to(cuda)).cpu().numpy()
    assert ops.max_rel_err(C, ops.spmm_fp64(A, B)) <= 1e-5


def test_reference_flow_sddmm(cuda):
    import torch
    A, X, Y = datasets.sddmm_rmat(seed=2, scale=11)
    op = Op('sddmm', [0, 1], [2], [A.shape[0], X.shape[1], A.shape[1]], None,
            [A, DenseTensor(X), DenseTensor(Y)], [A.tocsc()])
    tiles = _search(op)
    print(f"reference flow sddmm tiles={len(tiles)}", flush=True)
    res = my_fuse_formats.measure_seleted_formats(op, 'sddmm', 1, tiles, 0, ['A'], 32)
    ms = my_fuse_formats.measure_latency(*res, warmup=2, iters=5)
    assert ms > 0
    out = res[3](torch.from_numpy(X).to(cuda), torch.from_numpy(Y).to(cuda)).cpu().numpy()
    ref = ops.sddmm_fp64(A, X, Y)
    assert float(np.max(np.abs(out - ref)) / np.max(np.abs(ref))) <= 1e-5

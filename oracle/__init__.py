"""CPU oracle for STile's hybrid sparse-tile SpMM / SDDMM.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this package, and only as the checker
or the timed CPU baseline; the product (stile_amd/) never imports it.

Contents
  ops.py         FP64 SpMM/SDDMM, the reference's own FP32 CPU path
                 (scipy csr.astype(float32).dot(B), MY_sparse/my_fuse_formats.py:4194-4197),
                 error metrics.
  formats.py     restatement of the reference's integer format construction
                 (MY_sparse/gen_formats_v2.py, my_branch_and_bound.py,
                 my_fuse_formats.py packers), following the reference text.
  cost_model.py  restatement of the analytic cost model (MY_sparse/my_cost_model.py).

Each function's docstring cites the reference file:line it follows and names
the check that pins it (tests/test_oracle_*.py).

Parity unpinned: the exact tile set the greedy search selects
(my_search_formats.py:263-833 and the search functions of
my_branch_and_bound.py) has no restatement here.  It depends on float
tie-breaks and scipy internals, and the reference cannot be imported to
produce golden tile lists.  Product packs are checked through invariants
instead: every nonzero is covered once, re-expanding a pack reproduces A,
and every SDDMM output id is produced.
"""

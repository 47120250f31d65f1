"""MI355X-native hybrid sparse-tile SpMM / SDDMM (STile's method on gfx950).

Modules mirror the reference's: gen_formats_v2 (operator + candidate formats),
my_cost_model, my_branch_and_bound, my_search_formats (greedy search),
my_fuse_formats (packing + fused kernel); api has the driver flow, _native the
C-ABI binding of libstile_hip.so, autograd the paired SpMM/SDDMM adjoints.
"""
__version__ = "0.1.0"

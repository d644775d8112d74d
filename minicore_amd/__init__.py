"""minicore_amd: the minicore clustering hot path on MI355X (gfx950).

Kernels and drivers live in libmcg.so (C ABI: include/mcg.h); `_lib` is its
ctypes binding.  The pyminicore-compatible entry points (hcluster, scluster,
cmp, kmeanspp, CoresetSampler) are in the compiled extension `_pyminicore`.
"""
__all__ = ["_lib"]

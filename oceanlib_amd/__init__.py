"""MI355X implementation of oceanlib's per-frame wave synthesis.

The product is the C-ABI shared library liboceanlib_amd.so (HIP kernels plus
C++20 host code, built in-tree by oceanlib_amd/build.py); this package only
locates and binds it (capi.py) and wraps it for Python callers (ocean.py).
"""

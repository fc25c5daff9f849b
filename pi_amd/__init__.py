"""gnt: Goldilocks-NTT big-integer multiply for AMD Instinct MI355X (gfx950).

The product is libgnt.so (HIP kernels + C ABI, include/gnt.h); ``pi_amd.gnt``
is its Python binding.
"""
from .gnt import Context, DeviceBuf, GntError, load, mul_mag, mul_mag_ntt, sqr  # noqa: F401

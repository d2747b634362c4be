"""gpumlib_amd — MI355X (gfx950) NMF multiplicative-update library.

The product is libgpumlib_nmf.so (C-ABI: include/gpumlib_nmf.h) with the C++
drop-in headers under gpumlib_amd/src/ (GPUMLib's HostMatrix/DeviceMatrix
containers and the NMF_Multiplicative* classes).  This Python package is a
harness over that C-ABI for tests and benchmarks.
"""
from . import capi  # noqa: F401

__version__ = "0.1.0"

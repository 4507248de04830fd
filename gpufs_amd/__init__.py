"""MI355X-native GPUfs file I/O path.

The product is the native library libgpufs_amd.so (HIP device API + host
runtime behind the C ABI in include/gpufs_amd.h).  This package only binds it
(ctypes) and generates synthetic workloads; it holds no compute path.
"""
from . import _lib as native  # noqa: F401
from ._lib import Context, GpufsError, O_GRDONLY, O_GWRONLY, O_GRDWR, O_GCREAT, O_GWRONCE  # noqa: F401

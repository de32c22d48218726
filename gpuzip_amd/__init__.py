"""gpuzip_amd: one raw Deflate stream decoded in parallel on an AMD Instinct MI355X (gfx950).

The product is the native library lib/libgpuinflate.so behind the reference's C interface
(include/gpuinflate.h); `gpuinflate.Gpu` binds it from Python.
"""
from .gpuinflate import (ABI_VERSION, AbortedError, ArgError, DataError, Error, GpuError, Gpu,  # noqa: F401
                         TruncatedError, load_library)

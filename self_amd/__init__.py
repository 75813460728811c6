"""self_amd: MI355X-native spectral-element hot path of SELF.

The compute path is libself_amd.so (hand-written HIP for gfx950 behind the C
ABI in include/self_amd.h).  The Python modules are host-side mirrors of the
reference's Fortran types for tests, benchmarks and scripting; they call the
library through ctypes and fail loudly if it is not built.
"""
from . import _lib
from ._lib import SelfError, lib

__all__ = ["SelfError", "lib", "_lib"]

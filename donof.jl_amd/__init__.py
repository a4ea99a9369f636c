"""MI355X-native orbital-optimization hot path of DoNOF.jl.

The compute lives in libdonofhip.so (HIP kernels for gfx950 behind a C ABI, declared in
include/donof_hip.h).  This package is the host-side mirror of the reference's Julia
interface for that path: computeJKH_MO, compute_Lagrange2, rotate_orbital, the ADAM
family of orbital optimisers, the occupation functions and the energy driver, with the
reference's names and argument meaning.  There is no CPU fallback: every entry point
raises if the HIP library is missing or no GPU is visible.

The directory name contains a dot, so it is loaded by path (see
__graft_entry__.load_package); the module name is `donof_amd`.
"""
from .param import Param, set_ncwo  # noqa: F401
from . import synthetic  # noqa: F401
from ._lib import LIB_PATH, load_library, library_symbols, DnError  # noqa: F401

__all__ = ["Param", "set_ncwo", "synthetic", "LIB_PATH", "load_library",
           "library_symbols", "DnError"]
# The device-facing API lives in donof_amd.api (imported explicitly; it loads the HIP
# library and opens a device context on first use).

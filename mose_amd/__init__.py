"""MI355X-native explicit residual-and-update path of the MOSE compressible-flow solver.

The compute path is libmose_hip.so (mose_amd/csrc, hand-written HIP for gfx950) behind the
C ABI in include/mose_hip.h; `mose_amd.lib` binds it with ctypes and `mose_amd.cases`
builds the workloads.  The Fortran host (mose_amd/fortran) calls the same ABI through
ISO_C_BINDING.
"""
from . import abi, cases  # noqa: F401

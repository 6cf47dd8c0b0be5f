"""ctypes mirrors of the C-ABI types declared in include/mose_hip.h.

The struct layouts are the drop-in boundary: the Fortran host (mose_amd/fortran), the
Python harness and the tests all build these same records.  Field meanings follow the
reference's configuration keys (docs/user/registry.md: [MOSE-Numerics]) and the bc.txt
record (src/lib/io/IO_BC.f90:136-414).
"""
import ctypes as C

import numpy as np

LIM_MINMOD, LIM_VANLEER, LIM_VANALBADA, LIM_FIRST_ORDER, LIM_MC, LIM_SUPERBEE = 0, 1, 2, 3, 4, 5
RS_HLLC, RS_AUSMP, RS_HLLE, RS_SLAU, RS_SLAU2, RS_HLLEPP, RS_HLLCP = 0, 1, 2, 3, 4, 5, 6
PREC_FP64, PREC_FP32_FLUX = 0, 1

E_OK, E_ARG, E_HIP, E_STATE, E_INTERFACE, E_METRIC, E_BC, E_NOGPU, E_SIZE = range(9)

LIMITERS = {"minmod": LIM_MINMOD, "vanleer": LIM_VANLEER, "vanalbada": LIM_VANALBADA,
            "first-order": LIM_FIRST_ORDER, "mc": LIM_MC, "superbee": LIM_SUPERBEE}
RIEMANN = {"HLLC": RS_HLLC, "AUSM+": RS_AUSMP, "HLLE": RS_HLLE, "SLAU": RS_SLAU, "SLAU2": RS_SLAU2,
           "HLLE++": RS_HLLEPP, "HLLC+": RS_HLLCP}

NPRIM = 5
GC = 2


class MoseCfg(C.Structure):
    _fields_ = [("gamma", C.c_double), ("Rgas", C.c_double), ("cfl", C.c_double),
                ("limiter", C.c_int32), ("riemann", C.c_int32), ("n_rk", C.c_int32),
                ("time_accurate", C.c_int32), ("rampa_cfl_iter", C.c_int32),
                ("precision", C.c_int32), ("muscl_sd", C.c_int32), ("prim_vars", C.c_int32),
                ("reserved", C.c_int32 * 2),
                ("irs_beta", C.c_double)]


class MoseBcEntry(C.Structure):
    _fields_ = [("b", C.c_int32), ("i", C.c_int32), ("j", C.c_int32), ("k", C.c_int32),
                ("f", C.c_int32), ("type", C.c_int32),
                ("bs", C.c_int32), ("is_", C.c_int32), ("js", C.c_int32), ("ks", C.c_int32),
                ("fs", C.c_int32),
                ("d11", C.c_int32), ("d12", C.c_int32), ("d21", C.c_int32), ("d22", C.c_int32),
                ("pad_", C.c_int32),
                ("mach", C.c_double), ("T", C.c_double), ("p", C.c_double),
                ("alpha", C.c_double), ("beta", C.c_double), ("rel_fac", C.c_double)]


class MoseErr(C.Structure):
    _fields_ = [("code", C.c_int32), ("block", C.c_int32), ("i", C.c_int32),
                ("j", C.c_int32), ("k", C.c_int32)]


BC_DTYPE = np.dtype([("b", "<i4"), ("i", "<i4"), ("j", "<i4"), ("k", "<i4"), ("f", "<i4"),
                     ("type", "<i4"), ("bs", "<i4"), ("is", "<i4"), ("js", "<i4"), ("ks", "<i4"),
                     ("fs", "<i4"), ("d11", "<i4"), ("d12", "<i4"), ("d21", "<i4"), ("d22", "<i4"),
                     ("pad_", "<i4"), ("mach", "<f8"), ("T", "<f8"), ("p", "<f8"),
                     ("alpha", "<f8"), ("beta", "<f8"), ("rel_fac", "<f8")])
assert BC_DTYPE.itemsize == C.sizeof(MoseBcEntry) == 112


def make_cfg(gamma=1.4, Rgas=1.0 / 1.4, cfl=0.5, limiter="minmod", riemann="HLLC", n_rk=3,
             time_accurate=True, rampa_cfl_iter=0, precision=PREC_FP64, muscl_sd=False, irs_beta=0.0,
             prim_vars=False):
    """Build a MoseCfg from the reference's key values (flux-limiter, riemann-solver, ...);
    muscl_sd: space-reconstruction = MUSCL-SD; irs_beta: implicit residual smoothing;
    prim_vars: integration-variables = prim."""
    lim = LIMITERS[limiter] if isinstance(limiter, str) else int(limiter)
    rs = RIEMANN[riemann] if isinstance(riemann, str) else int(riemann)
    cfg = MoseCfg(gamma, Rgas, cfl, lim, rs, n_rk, int(bool(time_accurate)), rampa_cfl_iter,
                  precision, int(bool(muscl_sd)))
    cfg.irs_beta = float(irs_beta)
    cfg.prim_vars = int(bool(prim_vars))
    return cfg


def padded_shape(dims):
    """numpy shape (k, j, i, v) of blk%P(nprim, 1-gc:ni+gc, 1-gc:nj+gc, 1-gc:nk+gc)."""
    ni, nj, nk = dims
    return (nk + 2 * GC, nj + 2 * GC, ni + 2 * GC, NPRIM)


def interior(a):
    """Interior view (k, j, i, v) of a padded state array."""
    return a[GC:-GC, GC:-GC, GC:-GC, :]

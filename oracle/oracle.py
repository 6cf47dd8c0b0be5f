"""ctypes driver of oracle/libmose_oracle.so (test infrastructure only).

Mirrors the method names of mose_amd.lib.GpuSolver so parity tests read the same for both
sides.  Build the library with `make -C oracle` (done by __graft_entry__.build()).
"""
import ctypes as C
import os

import numpy as np

from mose_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}
# "check": the parity checker (no fast-math); "release": the same source with the reference's
# Release-style flags, used only to time the CPU baseline
_FILES = {"check": "libmose_oracle.so", "release": "libmose_oracle_release.so"}


def lib(variant="check"):
    if variant not in _LIBS:
        path = os.path.join(_HERE, _FILES[variant])
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        P = C.POINTER
        d = C.c_void_p
        L.ob_create.restype = d
        L.ob_create.argtypes = [P(abi.MoseCfg), C.c_int32]
        L.ob_add_block.argtypes = [d, C.c_int32, P(C.c_int32), P(C.c_double)]
        L.ob_set_bc.argtypes = [d, C.c_void_p, C.c_int64]
        L.ob_setup.argtypes = [d]
        L.ob_set_viscous.argtypes = [d, C.c_int32, P(C.c_double)]
        L.ob_put_state.argtypes = [d, C.c_int32, P(C.c_double)]
        L.ob_get_state.argtypes = [d, C.c_int32, P(C.c_double)]
        L.ob_fill_ghost.argtypes = [d]
        L.ob_residual.argtypes = [d, C.c_int32, P(C.c_double)]
        L.ob_compute_dt.argtypes = [d, C.c_int32, P(C.c_double)]
        L.ob_step.argtypes = [d, P(C.c_double), P(C.c_double), P(C.c_double)]
        L.ob_get_time.argtypes = [d, P(C.c_double), P(C.c_int32)]
        L.ob_set_time.argtypes = [d, C.c_double, C.c_int32]
        L.ob_get_metrics.argtypes = [d, C.c_int32] + [P(C.c_double)] * 5
        L.ob_get_dtlocal.argtypes = [d, C.c_int32, P(C.c_double)]
        L.ob_last_error.argtypes = [d, P(abi.MoseErr)]
        L.ob_ndir.argtypes = [d]
        L.ob_num_threads.argtypes = []
        L.ob_destroy.argtypes = [d]
        L.ob_limiter.restype = C.c_double
        L.ob_limiter.argtypes = [C.c_int32, C.c_double, C.c_double]
        L.ob_reconstruct.argtypes = [C.c_int32, P(C.c_double)] + [C.c_double] * 5 + [P(C.c_double)] * 2
        L.ob_riemann.argtypes = [P(abi.MoseCfg), P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double)]
        L.ob_riemann_beta.argtypes = [P(abi.MoseCfg), P(C.c_double), P(C.c_double), P(C.c_double), C.c_double,
                                      P(C.c_double)]
        L.ob_irs_smooth.argtypes = [C.c_void_p, C.c_int32, C.c_double, P(C.c_double)]
        L.ob_sd_tramel.restype = C.c_double
        L.ob_sd_tramel.argtypes = [C.c_double] * 7
        L.ob_prim2cons.argtypes = [P(abi.MoseCfg), P(C.c_double), P(C.c_double)]
        L.ob_cons2prim.argtypes = [P(abi.MoseCfg), P(C.c_double), P(C.c_double)]
        _LIBS[variant] = L
    return _LIBS[variant]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleError(RuntimeError):
    def __init__(self, err):
        self.err = err
        super().__init__("oracle error %d at block %d cell (%d,%d,%d)" % (err.code, err.block, err.i, err.j, err.k))


class OracleSolver:
    """The CPU oracle domain for one Case."""

    def __init__(self, case, cfg=None, variant="check"):
        self.L = L = lib(variant)
        self.case = case
        self.cfg = cfg if cfg is not None else case.cfg
        self.h = L.ob_create(C.byref(self.cfg), len(case.blocks))
        self.dims = [tuple(b.dims) for b in case.blocks]
        for q, blk in enumerate(case.blocks, start=1):
            dims = (C.c_int32 * 3)(*blk.dims)
            nodes = np.ascontiguousarray(blk.nodes, dtype=np.float64)
            self._check(L.ob_add_block(self.h, q, dims, _dp(nodes)))
        bc = np.ascontiguousarray(case.bc)
        self._check(L.ob_set_bc(self.h, bc.ctypes.data, len(bc)))
        for b, prm in case.meta.get("viscous", []):
            prm = np.ascontiguousarray(prm, dtype=np.float64)
            self._check(L.ob_set_viscous(self.h, b, _dp(prm)))
        self._check(L.ob_setup(self.h))
        for q, blk in enumerate(case.blocks, start=1):
            self.put_state(q, blk.P)

    def _check(self, rc):
        if rc:
            e = abi.MoseErr()
            self.L.ob_last_error(self.h, C.byref(e))
            raise OracleError(e)

    def close(self):
        if self.h:
            self.L.ob_destroy(self.h)
            self.h = None

    __del__ = close

    def put_state(self, b, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        assert P.shape == abi.padded_shape(self.dims[b - 1])
        self._check(self.L.ob_put_state(self.h, b, _dp(P)))

    def get_state(self, b):
        P = np.zeros(abi.padded_shape(self.dims[b - 1]))
        self._check(self.L.ob_get_state(self.h, b, _dp(P)))
        return P

    def fill_ghost(self):
        self._check(self.L.ob_fill_ghost(self.h))

    def residual(self, b):
        R = np.zeros(abi.padded_shape(self.dims[b - 1]))
        self._check(self.L.ob_residual(self.h, b, _dp(R)))
        return R

    def irs_smooth(self, b, R, beta):
        """Residual_Smoothing of the field R (padded blk%R layout) of block b, the other
        blocks' residuals zero; returns the smoothed field."""
        R = np.ascontiguousarray(R, dtype=np.float64).copy()
        self._check(self.L.ob_irs_smooth(self.h, b, float(beta), _dp(R)))
        return R

    def residuals(self):
        return [self.residual(b) for b in range(1, len(self.dims) + 1)]

    def compute_dt(self, it=1):
        dt = C.c_double()
        self._check(self.L.ob_compute_dt(self.h, it, C.byref(dt)))
        return dt.value

    def step(self):
        dt, t = C.c_double(), C.c_double()
        norms = np.zeros(5)
        self._check(self.L.ob_step(self.h, C.byref(dt), C.byref(t), _dp(norms)))
        return dt.value, t.value, norms

    def time(self):
        t, it = C.c_double(), C.c_int32()
        self.L.ob_get_time(self.h, C.byref(t), C.byref(it))
        return t.value, it.value

    def metrics(self, b):
        npad = int(np.prod(abi.padded_shape(self.dims[b - 1])[:3]))
        fn = np.zeros((3, npad, 3)); fa = np.zeros((3, npad)); dl = np.zeros((npad, 3))
        M = np.zeros((npad, 9)); vol = np.zeros(npad)
        self.L.ob_get_metrics(self.h, b, _dp(fn), _dp(fa), _dp(dl), _dp(M), _dp(vol))
        sh = abi.padded_shape(self.dims[b - 1])[:3]
        return dict(fn=fn.reshape((3,) + sh + (3,)), fa=fa.reshape((3,) + sh), dl=dl.reshape(sh + (3,)),
                    M=M.reshape(sh + (3, 3)), vol=vol.reshape(sh))

    def dtlocal(self, b):
        sh = abi.padded_shape(self.dims[b - 1])[:3]
        dt = np.zeros(int(np.prod(sh)))
        self.L.ob_get_dtlocal(self.h, b, _dp(dt))
        return dt.reshape(sh)

    def ndir(self):
        return self.L.ob_ndir(self.h)


def limiter(kind, x, y):
    return lib().ob_limiter(kind, x, y)


def riemann(cfg, PL, PR, n, beta=1.0):
    """Face flux per unit area; beta is the shock-detector weight the HLLE++ / HLLC+ solvers take."""
    PL = np.ascontiguousarray(PL, dtype=np.float64); PR = np.ascontiguousarray(PR, dtype=np.float64)
    n = np.ascontiguousarray(n, dtype=np.float64)
    F = np.zeros(5)
    lib().ob_riemann_beta(C.byref(cfg), _dp(PL), _dp(PR), _dp(n), float(beta), _dp(F))
    return F


def sd_tramel(p, p1, p2, p3, p4, p5, p6):
    return lib().ob_sd_tramel(p, p1, p2, p3, p4, p5, p6)


def reconstruct(kind, q, l0, l1, l2, lm, lp):
    q = np.ascontiguousarray(q, dtype=np.float64)
    ql, qr = C.c_double(), C.c_double()
    lib().ob_reconstruct(kind, _dp(q), l0, l1, l2, lm, lp, C.byref(ql), C.byref(qr))
    return ql.value, qr.value


def prim2cons(cfg, P):
    P = np.ascontiguousarray(P, dtype=np.float64); U = np.zeros(5)
    lib().ob_prim2cons(C.byref(cfg), _dp(P), _dp(U))
    return U


def cons2prim(cfg, U):
    U = np.ascontiguousarray(U, dtype=np.float64); P = np.zeros(5)
    lib().ob_cons2prim(C.byref(cfg), _dp(U), _dp(P))
    return P


def num_threads(variant="check"):
    """OpenMP threads the oracle runs its sweeps on."""
    return int(lib(variant).ob_num_threads())

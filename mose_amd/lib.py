"""ctypes binding of libmose_hip.so (the C ABI of include/mose_hip.h).

`GpuSolver(case)` drives the HIP path for one Case with the same method names as the CPU
oracle's driver, so parity tests read the same on both sides.  The library must be built
in-tree (python -m mose_amd.build); there is no CPU fallback: every method calls the HIP
library and raises when it is missing or returns an error.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
# MOSE_HIP_LIB may name another build of the same library (profiling of kernel variants)
LIB_PATH = os.environ.get("MOSE_HIP_LIB") or os.path.join(_HERE, "libmose_hip.so")
_LIB = None

EXPORTS = ["mose_create", "mose_add_block", "mose_set_viscous", "mose_set_bc", "mose_finalize", "mose_put_state",
           "mose_get_state", "mose_get_state_noalloc", "mose_compute_dt", "mose_step", "mose_advance", "mose_residual",
           "mose_get_time", "mose_set_time", "mose_time_kernels", "mose_autotune", "mose_get_tiling", "mose_put_residual_bar",
           "mose_residual_vjp", "mose_get_state_bar", "mose_residual_vjp_geometry", "mose_get_viscous_bar", "mose_get_node_bar",
           "mose_get_inflow_bar", "mose_put_objective_bar", "mose_adjoint_solve", "mose_get_residual_bar",
           "mose_unsteady_adjoint_ke", "mose_time_residual_vjp", "mose_num_cells", "mose_last_error", "mose_strerror",
           "mose_destroy"]


def load():
    """Load libmose_hip.so and declare the ABI; raises if the library was not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libmose_hip.so is not built: run `python -m mose_amd.build`")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    v = C.c_void_p
    L.mose_create.argtypes = [P(abi.MoseCfg), P(v)]
    L.mose_add_block.argtypes = [v, C.c_int32, P(C.c_int32), P(C.c_double)]
    L.mose_set_bc.argtypes = [v, C.c_void_p, C.c_int64]
    L.mose_set_viscous.argtypes = [v, C.c_int32, P(C.c_double)]
    L.mose_finalize.argtypes = [v]
    L.mose_put_state.argtypes = [v, C.c_int32, P(C.c_double)]
    L.mose_get_state.argtypes = [v, C.c_int32, P(C.c_double)]
    L.mose_get_state_noalloc.argtypes = [v, C.c_int32, P(C.c_double)]
    L.mose_compute_dt.argtypes = [v, C.c_int32, P(C.c_double)]
    L.mose_step.argtypes = [v, P(C.c_double), P(C.c_double), P(C.c_double)]
    L.mose_advance.argtypes = [v, C.c_int32, P(C.c_double), P(C.c_double)]
    L.mose_residual.argtypes = [v, C.c_int32, P(C.c_double)]
    L.mose_get_time.argtypes = [v, P(C.c_double), P(C.c_int32)]
    L.mose_set_time.argtypes = [v, C.c_double, C.c_int32]
    L.mose_time_kernels.argtypes = [v, C.c_int32, P(C.c_double)]
    L.mose_autotune.argtypes = [v, C.c_int32, P(C.c_int32), P(C.c_double)]
    L.mose_get_tiling.argtypes = [v, P(C.c_int32)]
    L.mose_put_residual_bar.argtypes = [v, C.c_int32, P(C.c_double)]
    L.mose_residual_vjp.argtypes = [v]
    L.mose_get_state_bar.argtypes = [v, C.c_int32, P(C.c_double)]
    L.mose_time_residual_vjp.argtypes = [v, C.c_int32, P(C.c_double)]
    L.mose_residual_vjp_geometry.argtypes = [v]
    L.mose_get_viscous_bar.argtypes = [v, C.c_int32, P(C.c_double)]
    L.mose_get_node_bar.argtypes = [v, C.c_int32, P(C.c_double)]
    L.mose_get_inflow_bar.argtypes = [v, C.c_int32, C.c_int32, P(C.c_double)]
    L.mose_put_objective_bar.argtypes = [v, C.c_int32, P(C.c_double)]
    L.mose_adjoint_solve.argtypes = [v, C.c_int32, C.c_double, C.c_int32, P(C.c_int32), P(C.c_double)]
    L.mose_get_residual_bar.argtypes = [v, C.c_int32, P(C.c_double)]
    L.mose_unsteady_adjoint_ke.argtypes = [v, C.c_int32, C.c_double, C.c_int32, P(C.c_double), P(C.c_double)]
    L.mose_num_cells.argtypes = [v]
    L.mose_num_cells.restype = C.c_int64
    L.mose_last_error.argtypes = [v, P(abi.MoseErr)]
    L.mose_strerror.argtypes = [C.c_int32]
    L.mose_strerror.restype = C.c_char_p
    L.mose_destroy.argtypes = [v]
    L.mose_destroy.restype = None
    _LIB = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class MoseError(RuntimeError):
    def __init__(self, code, err=None):
        self.code = code
        self.err = err
        msg = load().mose_strerror(code).decode()
        if err is not None and err.block:
            msg += " at b, i, j, k: %d %d %d %d" % (err.block, err.i, err.j, err.k)
        super().__init__("mose error %d: %s" % (code, msg))


class GpuSolver:
    """One HIP context (device buffers, graph) for a Case."""

    def __init__(self, case, cfg=None):
        L = load()
        self.case = case
        self.cfg = cfg if cfg is not None else case.cfg
        self.h = C.c_void_p()
        rc = L.mose_create(C.byref(self.cfg), C.byref(self.h))
        if rc:
            self.h = None
            raise MoseError(rc)
        self.dims = [tuple(int(x) for x in b.dims) for b in case.blocks]
        for q, blk in enumerate(case.blocks, start=1):
            dims = (C.c_int32 * 3)(*blk.dims)
            nodes = np.ascontiguousarray(blk.nodes, dtype=np.float64)
            self._check(L.mose_add_block(self.h, q, dims, _dp(nodes)))
        bc = np.ascontiguousarray(case.bc)
        self._check(L.mose_set_bc(self.h, bc.ctypes.data, len(bc)))
        for b, prm in case.meta.get("viscous", []):
            prm = np.ascontiguousarray(prm, dtype=np.float64)
            self._check(L.mose_set_viscous(self.h, b, _dp(prm)))
        self._check(L.mose_finalize(self.h))
        for q, blk in enumerate(case.blocks, start=1):
            self.put_state(q, blk.P)

    def _check(self, rc):
        if rc:
            e = abi.MoseErr()
            load().mose_last_error(self.h, C.byref(e))
            raise MoseError(rc, e)

    def close(self):
        if getattr(self, "h", None):
            load().mose_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def put_state(self, b, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        assert P.shape == abi.padded_shape(self.dims[b - 1])
        self._check(load().mose_put_state(self.h, b, _dp(P)))

    def get_state(self, b):
        P = np.zeros(abi.padded_shape(self.dims[b - 1]))
        self._check(load().mose_get_state(self.h, b, _dp(P)))
        return P

    def get_state_noalloc(self, b):
        """mose_get_state through the fixed device staging array (no device allocation)."""
        P = np.zeros(abi.padded_shape(self.dims[b - 1]))
        self._check(load().mose_get_state_noalloc(self.h, b, _dp(P)))
        return P

    def residual(self, b):
        R = np.zeros(abi.padded_shape(self.dims[b - 1]))
        self._check(load().mose_residual(self.h, b, _dp(R)))
        return R

    def residual_vjp(self, Rbars):
        """Discrete adjoint of the residual: Rbars, one blk%R-layout array per block ->
        Pbar = (dR/dP)^T Rbar, one blk%P-layout array per block (ghost entries 0)."""
        L = load()
        for b, Rb in enumerate(Rbars, start=1):
            Rb = np.ascontiguousarray(Rb, dtype=np.float64)
            assert Rb.shape == abi.padded_shape(self.dims[b - 1])
            self._check(L.mose_put_residual_bar(self.h, b, _dp(Rb)))
        self._check(L.mose_residual_vjp(self.h))
        out = []
        for b in range(1, len(self.dims) + 1):
            Pb = np.zeros(abi.padded_shape(self.dims[b - 1]))
            self._check(L.mose_get_state_bar(self.h, b, _dp(Pb)))
            out.append(Pb)
        return out

    def viscous_bar(self, b):
        """After residual_vjp: d(Rbar . R)/d mu_ref of block b (0 for an inviscid block)."""
        x = C.c_double()
        self._check(load().mose_get_viscous_bar(self.h, b, C.byref(x)))
        return x.value

    def residual_vjp_geometry(self, Rbars):
        """Adjoint of the residual with respect to the state, the block nodes and the inflow
        values: returns (Pbar per block, Xbar per block in the node layout (nk+1, nj+1, ni+1, 3)
        of Case blocks, {(b, f): d/d(mach, T, p, alpha, beta)} for every 405 face)."""
        L = load()
        for b, Rb in enumerate(Rbars, start=1):
            Rb = np.ascontiguousarray(Rb, dtype=np.float64)
            assert Rb.shape == abi.padded_shape(self.dims[b - 1])
            self._check(L.mose_put_residual_bar(self.h, b, _dp(Rb)))
        self._check(L.mose_residual_vjp_geometry(self.h))
        return self._geometry_bars()

    def _geometry_bars(self):
        L = load()
        Pbars, Xbars = [], []
        for b in range(1, len(self.dims) + 1):
            Pb = np.zeros(abi.padded_shape(self.dims[b - 1]))
            self._check(L.mose_get_state_bar(self.h, b, _dp(Pb)))
            Pbars.append(Pb)
            ni, nj, nk = self.dims[b - 1]
            Xb = np.zeros((nk + 1, nj + 1, ni + 1, 3))
            self._check(L.mose_get_node_bar(self.h, b, _dp(Xb)))
            Xbars.append(Xb)
        inflow = {}
        for row in np.asarray(self.case.bc):
            if int(row["type"]) == 405:
                key = (int(row["b"]), int(row["f"]))
                if key not in inflow:
                    out = np.zeros(5)
                    self._check(L.mose_get_inflow_bar(self.h, key[0], key[1], _dp(out)))
                    inflow[key] = out
        return Pbars, Xbars, inflow

    def adjoint_solve(self, gPs, max_iter=20000, rtol=1e-10, check_every=100):
        """Steady adjoint: gPs (dJ/dP per block, blk%P layout) -> psi per block solving
        (dR/dP)^T psi = dJ/dP at the current (converged) state; returns (psi list, iterations,
        |Q_0|, |Q|).  psi stays in place as the residual adjoint for residual_vjp_geometry_inplace."""
        L = load()
        for b, gP in enumerate(gPs, start=1):
            gP = np.ascontiguousarray(gP, dtype=np.float64)
            assert gP.shape == abi.padded_shape(self.dims[b - 1])
            self._check(L.mose_put_objective_bar(self.h, b, _dp(gP)))
        it = C.c_int32()
        res = np.zeros(2)
        self._check(L.mose_adjoint_solve(self.h, max_iter, rtol, check_every, C.byref(it), _dp(res)))
        psi = []
        for b in range(1, len(self.dims) + 1):
            x = np.zeros(abi.padded_shape(self.dims[b - 1]))
            self._check(L.mose_get_residual_bar(self.h, b, _dp(x)))
            psi.append(x)
        return psi, it.value, float(res[0]), float(res[1])

    def march_kinetic_energy(self, nsteps, dt, stride=0, adjoint=True):
        """nsteps fixed-dt steps on the device from the current state (left unchanged): the
        final kinetic energy per block and, with adjoint=True, (dJ/dmu_ref per block, dJ/dP0
        per block) for J = the total kinetic energy."""
        nb = len(self.dims)
        ke = np.zeros(nb)
        mub = np.zeros(nb) if adjoint else None
        self._check(load().mose_unsteady_adjoint_ke(self.h, nsteps, dt, stride, _dp(ke),
                                                    _dp(mub) if adjoint else None))
        if not adjoint:
            return ke
        P0b = []
        for b in range(1, nb + 1):
            x = np.zeros(abi.padded_shape(self.dims[b - 1]))
            self._check(load().mose_get_state_bar(self.h, b, _dp(x)))
            P0b.append(x)
        return ke, mub, P0b

    def residual_vjp_geometry_inplace(self):
        """residual_vjp_geometry with the residual adjoint already on the device (after
        adjoint_solve): (Pbar, Xbar, inflow) as residual_vjp_geometry returns them."""
        L = load()
        self._check(L.mose_residual_vjp_geometry(self.h))
        return self._geometry_bars()

    def time_residual_vjp(self, nrep=10, geometry=False):
        """Average device ms of one residual and one adjoint evaluation (and, with
        geometry=True, of one geometry adjoint evaluation: 0 on rectilinear layouts)."""
        ms = np.zeros(3)
        self._check(load().mose_time_residual_vjp(self.h, nrep, _dp(ms)))
        if geometry:
            return float(ms[0]), float(ms[1]), float(ms[2])
        return float(ms[0]), float(ms[1])

    def residuals(self):
        return [self.residual(b) for b in range(1, len(self.dims) + 1)]

    def compute_dt(self, it=1):
        dt = C.c_double()
        self._check(load().mose_compute_dt(self.h, it, C.byref(dt)))
        return dt.value

    def step(self):
        dt, t = C.c_double(), C.c_double()
        norms = np.zeros(5)
        self._check(load().mose_step(self.h, C.byref(dt), C.byref(t), _dp(norms)))
        return dt.value, t.value, norms

    def advance(self, nsteps):
        t = C.c_double()
        norms = np.zeros(5)
        self._check(load().mose_advance(self.h, nsteps, C.byref(t), _dp(norms)))
        return t.value, norms

    def time(self):
        t, it = C.c_double(), C.c_int32()
        self._check(load().mose_get_time(self.h, C.byref(t), C.byref(it)))
        return t.value, it.value

    def time_kernels(self, nrep=20):
        """Average ms per launch: [CFL dt, ghost fill, RK stage, last RK stage]."""
        ms = np.zeros(4)
        self._check(load().mose_time_kernels(self.h, nrep, _dp(ms)))
        return ms

    def tiling(self):
        """The tiling in use: {"tile": (TX, TY), "kchunk": kc, "xcd_order": bool}."""
        ch = np.zeros(3, dtype=np.int32)
        self._check(load().mose_get_tiling(self.h, ch.ctypes.data_as(C.POINTER(C.c_int32))))
        return {"tile": (int(ch[0]) // 1000, int(ch[0]) % 1000), "kchunk": int(ch[1]), "xcd_order": bool(ch[2] & 1),
                "persistent": not bool(ch[2] & 2)}

    def autotune(self, nrep=3):
        """Pick the fastest tile shape / k chunk / tile order by timing; returns
        ({"tile": (TX, TY), "kchunk": kc, "xcd_order": bool}, ms of the three stage launches)."""
        ch = np.zeros(3, dtype=np.int32)
        ms = np.zeros(1)
        self._check(load().mose_autotune(self.h, nrep, ch.ctypes.data_as(C.POINTER(C.c_int32)), _dp(ms)))
        return {"tile": (int(ch[0]) // 1000, int(ch[0]) % 1000), "kchunk": int(ch[1]),
                "xcd_order": bool(ch[2] & 1), "persistent": not bool(ch[2] & 2)}, float(ms[0])

    def num_cells(self):
        return load().mose_num_cells(self.h)

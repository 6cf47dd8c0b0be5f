"""Adjoint of the explicit time integration with the time step held fixed: sensitivities of a
functional of the state after K SSP-RK steps (e.g. the kinetic energy of the Taylor-Green
boxes at t = 1) with respect to the initial state and to each block's Sutherland reference
viscosity mu_ref (mu_ref = 1/Re, so d/dRe = -mu_ref^2 d/dmu_ref).

The forward march stores every stage state and repeats the reference's update
(Lib_Newstate.f90:134-160 with RKcoeff, Lib_RK.f90:8-27):
    U_s = U_n + c_s (U_{s-1} - U_n - dt/vol R(P(U_{s-1}))),   U_0 = U_n,
c = (1), (1, 1/2), (1, 1/4, 2/3), where R is the device residual (mose_residual, ghost fill
included).  The reverse sweep walks the stages backwards with the device adjoint
(mose_residual_vjp, viscous terms included, and mose_get_viscous_bar):
    U_n_bar  += (1 - c_s) U_s_bar,      U_{s-1}_bar += c_s U_s_bar,
    R_bar     = -c_s dt/vol U_s_bar  ->  P_bar = (dR/dP)^T R_bar  ->  U_{s-1}_bar += (dP/dU)^T P_bar.
The per-stage state storage is host memory (5 doubles per cell per stage per step), which is
meant for sensitivity studies on small grids; the residual and its adjoint run on the GPU.
Rectilinear blocks (cell volume = product of the node spacings), global time step."""
import numpy as np

from . import abi

RK = {1: (1.0,), 2: (1.0, 0.5), 3: (1.0, 0.25, 2.0 / 3.0)}


def _volumes(case):
    out = []
    for blk in case.blocks:
        x = blk.nodes
        dx = np.diff(x[0, 0, :, 0])
        dy = np.diff(x[0, :, 0, 1])
        dz = np.diff(x[:, 0, 0, 2])
        if not (np.allclose(x[..., 0], x[0:1, 0:1, :, 0]) and np.allclose(x[..., 1], x[0:1, :, 0:1, 1])
                and np.allclose(x[..., 2], x[:, 0:1, 0:1, 2])):
            raise ValueError("the time-integration adjoint handles rectilinear blocks only")
        out.append(dz[:, None, None] * dy[None, :, None] * dx[None, None, :])
    return out


def prim2cons(P, gamma):
    rho = P[..., 0]
    U = np.empty_like(P)
    U[..., 0] = rho
    U[..., 1:4] = rho[..., None] * P[..., 1:4]
    U[..., 4] = P[..., 4] / (gamma - 1.0) + 0.5 * rho * np.sum(P[..., 1:4] ** 2, axis=-1)
    return U


def cons2prim(U, gamma):
    rho = U[..., 0]
    P = np.empty_like(U)
    P[..., 0] = rho
    P[..., 1:4] = U[..., 1:4] / rho[..., None]
    P[..., 4] = (gamma - 1.0) * (U[..., 4] - 0.5 * np.sum(U[..., 1:4] * P[..., 1:4], axis=-1))
    return P


def cons_bar(P, Pbar, gamma):
    """(dP/dU)^T Pbar for P = (rho, u, v, w, p), U = (rho, rho u, rho v, rho w, E)."""
    rho = P[..., 0]
    u = P[..., 1:4]
    g1 = gamma - 1.0
    Ub = np.empty_like(Pbar)
    Ub[..., 0] = Pbar[..., 0] - np.sum(u * Pbar[..., 1:4], axis=-1) / rho + g1 * 0.5 * np.sum(u * u, axis=-1) * Pbar[..., 4]
    Ub[..., 1:4] = Pbar[..., 1:4] / rho[..., None] - g1 * u * Pbar[..., 4:5]
    Ub[..., 4] = g1 * Pbar[..., 4]
    return Ub


def _padded(interior_arr, dims):
    out = np.zeros(abi.padded_shape(dims))
    abi.interior(out)[...] = interior_arr
    return out


class FixedStepMarch:
    """K explicit steps of a fixed dt on a GpuSolver, with the stage states kept for the
    reverse sweep."""

    def __init__(self, solver, case, dt):
        self.g, self.case, self.dt = solver, case, float(dt)
        self.gamma = case.cfg.gamma
        self.nrk = case.cfg.n_rk
        self.vol = _volumes(case)
        self.dims = [tuple(int(x) for x in b.dims) for b in case.blocks]
        self.tape = []            # per step: list of stage input conserved states (per block)

    def _residual(self, U):
        for b, (Ub, d) in enumerate(zip(U, self.dims), start=1):
            self.g.put_state(b, _padded(cons2prim(Ub, self.gamma), d))
        return [abi.interior(self.g.residual(b)) for b in range(1, len(U) + 1)]

    def run(self, P0, nsteps):
        """March from the interior primitive states P0 (one (nk, nj, ni, 5) array per block);
        returns the interior primitive states after nsteps."""
        Un = [prim2cons(P, self.gamma) for P in P0]
        self.tape = []
        for _ in range(nsteps):
            stages = []
            Us = Un
            for c in RK[self.nrk]:
                stages.append(Us)
                R = self._residual(Us)
                Us = [un + c * (us - un - self.dt / v[..., None] * r) for un, us, r, v in zip(Un, Us, R, self.vol)]
            self.tape.append(stages)
            Un = Us
        return [cons2prim(U, self.gamma) for U in Un]

    def reverse(self, Ubar_final):
        """Adjoint sweep from dJ/dU of the final state: returns (dJ/dP0 per block,
        dJ/dmu_ref per block)."""
        nb = len(self.dims)
        Ub = [x.copy() for x in Ubar_final]
        mub = np.zeros(nb)
        coef = RK[self.nrk]
        for stages in reversed(self.tape):
            Un_b = [np.zeros_like(x) for x in Ub]
            Us_b = Ub
            for s in range(len(coef) - 1, -1, -1):
                c = coef[s]
                Uin = stages[s]
                prev_b = [c * x for x in Us_b]
                for q in range(nb):
                    Un_b[q] += (1.0 - c) * Us_b[q]
                Rbar = [_padded(-c * self.dt / v[..., None] * x, d) for x, v, d in zip(Us_b, self.vol, self.dims)]
                for b, (U, d) in enumerate(zip(Uin, self.dims), start=1):
                    self.g.put_state(b, _padded(cons2prim(U, self.gamma), d))
                Pbar = self.g.residual_vjp(Rbar)
                for q in range(nb):
                    mub[q] += self.g.viscous_bar(q + 1)
                    prev_b[q] += cons_bar(cons2prim(Uin[q], self.gamma), abi.interior(Pbar[q]), self.gamma)
                if s == 0:      # U_0 is U_n
                    for q in range(nb):
                        Un_b[q] += prev_b[q]
                else:
                    Us_b = prev_b
            Ub = Un_b
        if not self.tape:
            raise ValueError("run() first")
        P0 = [cons2prim(U, self.gamma) for U in self.tape[0][0]]
        return [cons_bar_inverse_t(P, x, self.gamma) for P, x in zip(P0, Ub)], mub


def cons_bar_inverse_t(P, Ubar, gamma):
    """(dU/dP)^T Ubar: the adjoint with respect to the initial primitive state."""
    rho = P[..., 0]
    u = P[..., 1:4]
    Pb = np.empty_like(Ubar)
    Pb[..., 0] = Ubar[..., 0] + np.sum(u * Ubar[..., 1:4], axis=-1) + 0.5 * np.sum(u * u, axis=-1) * Ubar[..., 4]
    Pb[..., 1:4] = rho[..., None] * Ubar[..., 1:4] + rho[..., None] * u * Ubar[..., 4:5]
    Pb[..., 4] = Ubar[..., 4] / (gamma - 1.0)
    return Pb


def kinetic_energy(case, P):
    """Sum over the cells of 1/2 rho |u|^2 vol (one value for all blocks)."""
    return float(sum(np.sum(0.5 * p[..., 0] * np.sum(p[..., 1:4] ** 2, axis=-1) * v)
                     for p, v in zip(P, _volumes(case))))


def kinetic_energy_ubar(case, P, gamma):
    """dKE/dU per block for the state P."""
    out = []
    for p, v in zip(P, _volumes(case)):
        u = p[..., 1:4]
        Ub = np.zeros_like(p)
        Ub[..., 0] = -0.5 * np.sum(u * u, axis=-1) * v
        Ub[..., 1:4] = u * v[..., None]
        out.append(Ub)
    return out

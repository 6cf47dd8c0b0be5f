"""Steady adjoint gradient of the bump channel's wall-pressure integral.

Objective (the supersonic bump channel of cases.bump_channel, 2-D, nk = 1):
    J = sum over the bump's wall faces (j-min faces whose centre lies in 0 <= x <= 1) of
        p(cell next to the wall) * A(face),
with A the face area: the segment length in the x-y plane (the mean of the two node planes)
times the extrusion depth at the face's high node column.  Its gradient with respect to the
control-point heights h and the
405 inflow record (mach, T, p, alpha, beta) at a converged state U* of the local-time-step
iteration (Mod_Explicit.f90:45-85) is
    dJ/dh = dJ/dX|_P . dX/dh - psi^T dR/dX . dX/dh,   dJ/d(inflow) = -psi^T dR/d(inflow),
with psi from the device adjoint solve (dR/dP)^T psi = dJ/dP (mose_adjoint_solve) and the
residual's geometry adjoint (mose_residual_vjp_geometry).  The wall faces are the symmetry
faces of BC_Symmetry_Eul (Lib_BC_Fluxes_Symmetry.f90:14-125).
"""
import numpy as np

from .cases import bump_control_shape


def converge(solver, max_steps=200000, tol=1e-13, chunk=500, stall=8):
    """Step until the largest step norm (L2 of the primitive change over one step) drops below
    tol, or stops falling for `stall` chunks.  Returns (steps, last norms)."""
    steps, best, flat = 0, np.inf, 0
    norms = np.full(5, np.inf)
    while steps < max_steps:
        n = min(chunk, max_steps - steps)
        if hasattr(solver, "advance"):
            _, norms = solver.advance(n)
        else:                                       # the oracle steps one at a time
            for _ in range(n):
                _, _, norms = solver.step()
        steps += n
        m = float(np.max(norms))
        if m < tol:
            break
        if m < 0.5 * best:
            best, flat = m, 0
        else:
            flat += 1
            if flat >= stall:
                break
    return steps, norms


def _wall_faces(case):
    xs = case.meta["xs"]
    xc = 0.5 * (xs[:-1] + xs[1:])
    return np.nonzero((xc >= 0.0) & (xc <= 1.0))[0] + 1      # 1-based i of the bump's wall cells


def wall_objective(case, P):
    """J, dJ/dP (blk%P layout of block 1) and dJ/dX at fixed P (node layout (2, nj+1, ni+1, 3))."""
    X = case.blocks[0].nodes
    ii = _wall_faces(case)
    dJdP = np.zeros_like(P)
    dJdX = np.zeros_like(X)
    J = 0.0
    for i in ii:
        p = P[2, 2, i + 1, 4]                       # cell (i, 1, 1): padded index + 1
        area, dA = 0.0, []
        for k in (0, 1):
            a, b = X[k, 0, i - 1, :2], X[k, 0, i, :2]
            seg = b - a
            L = float(np.hypot(seg[0], seg[1]))
            area += 0.5 * L
            dA.append((k, seg / L))
        dz = X[1, 0, i, 2] - X[0, 0, i, 2]           # extrusion depth at node column i
        J += p * area * dz
        dJdP[2, 2, i + 1, 4] = area * dz
        for k, u in dA:                              # d(0.5 L_k dz)/d(node) along the segment
            dJdX[k, 0, i, :2] += p * 0.5 * dz * u
            dJdX[k, 0, i - 1, :2] -= p * 0.5 * dz * u
        dJdX[1, 0, i, 2] += p * area
        dJdX[0, 0, i, 2] -= p * area
    return J, dJdP, dJdX


def height_gradient_of_nodes(case, Xbar):
    """sum_nodes Xbar . dX/dh_m for every control point m: the nodes are linear in the
    heights, Y(k, j, i) = yb(x_i) (1 - eta_j) + eta_j H with yb linear in h."""
    eta = case.meta["eta"]
    W = np.einsum("kji,j->i", Xbar[..., 1], 1.0 - eta)      # per node column i
    G = np.stack([bump_control_shape(case.meta["xs"], cx, case.meta["width"]) for cx in case.meta["ctrl_x"]])
    return G @ W


def gradient(case, solver, max_iter=200000, rtol=1e-11, check_every=200):
    """At the solver's current (converged) state: J, dJ/dh, dJ/d(inflow record) and the
    adjoint solve's (iterations, |Q_0|, |Q|)."""
    P = solver.get_state(1)
    J, dJdP, dJdX = wall_objective(case, P)
    _, it, r0, r = solver.adjoint_solve([dJdP], max_iter=max_iter, rtol=rtol, check_every=check_every)
    _, Xbar, inflow = solver.residual_vjp_geometry_inplace()
    dJdh = height_gradient_of_nodes(case, dJdX) - height_gradient_of_nodes(case, Xbar[0])
    dJdin = -inflow[(1, 1)]
    return {"J": J, "dJdh": dJdh, "dJdinflow": dJdin, "adjoint_iterations": it, "adjoint_res": (r0, r)}


def objective_at(case, solver_cls, warm=None, **conv):
    """Converge a fresh solver on `case` (from `warm`, a blk%P state, when given) and return
    (J, steps, norms)."""
    if warm is not None:
        case.blocks[0].P = warm.copy()
    s = solver_cls(case)
    try:
        steps, norms = converge(s, **conv)
        J, _, _ = wall_objective(case, s.get_state(1))
    finally:
        if hasattr(s, "close"):
            s.close()
    return J, steps, norms

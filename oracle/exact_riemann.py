"""Exact solution of the 1-D Euler Riemann problem (ideal gas) — test infrastructure.

The reference's verify scripts take this from the absent ExactPack submodule
(test/1D/Sod79/verify.py:7,45-64; ExactPack IGEOS_Solver, version unpinned because the
submodule is empty).  This is the published algorithm of Toro, "Riemann Solvers and
Numerical Methods for Fluid Dynamics", ch. 4: Newton iteration on the star pressure with
the two-rarefaction initial guess, then self-similar sampling at x/t.  It is pinned by the
reference's committed golden file test/1D/Sod79/reference/exact.txt (tests/golden/).
"""
import numpy as np


def _fk(p, rho, pk, ak, gamma):
    if p > pk:                       # shock
        A = 2.0 / ((gamma + 1.0) * rho)
        B = (gamma - 1.0) / (gamma + 1.0) * pk
        f = (p - pk) * np.sqrt(A / (p + B))
        df = np.sqrt(A / (B + p)) * (1.0 - 0.5 * (p - pk) / (B + p))
    else:                            # rarefaction
        f = 2.0 * ak / (gamma - 1.0) * ((p / pk) ** ((gamma - 1.0) / (2.0 * gamma)) - 1.0)
        df = 1.0 / (rho * ak) * (p / pk) ** (-(gamma + 1.0) / (2.0 * gamma))
    return f, df


def star_state(left, right, gamma=1.4, tol=1e-14, maxit=100):
    """(p*, u*) of the Riemann problem with primitive states (rho, u, p)."""
    rl, ul, pl = left
    rr, ur, pr = right
    al, ar = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)
    du = ur - ul
    z = (gamma - 1.0) / (2.0 * gamma)
    p = ((al + ar - 0.5 * (gamma - 1.0) * du) / (al / pl ** z + ar / pr ** z)) ** (1.0 / z)
    for _ in range(maxit):
        fl, dfl = _fk(p, rl, pl, al, gamma)
        fr, dfr = _fk(p, rr, pr, ar, gamma)
        pn = max(p - (fl + fr + du) / (dfl + dfr), 1e-14)
        if abs(pn - p) / (0.5 * (pn + p)) < tol:
            p = pn
            break
        p = pn
    fl, _ = _fk(p, rl, pl, al, gamma)
    fr, _ = _fk(p, rr, pr, ar, gamma)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, left, right, x0=0.5, gamma=1.4):
    """Primitive (rho, u, p) at positions x (array) and time t > 0."""
    rl, ul, pl = left
    rr, ur, pr = right
    al, ar = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)
    ps, us = star_state(left, right, gamma)
    g1 = (gamma - 1.0) / (gamma + 1.0)
    x = np.asarray(x, dtype=float)
    s = (x - x0) / t
    rho = np.empty_like(s); u = np.empty_like(s); p = np.empty_like(s)
    for n, si in enumerate(s):
        if si <= us:                                   # left of the contact
            if ps > pl:                                # left shock
                sl = ul - al * np.sqrt((gamma + 1.0) / (2 * gamma) * ps / pl + (gamma - 1.0) / (2 * gamma))
                if si <= sl:
                    rho[n], u[n], p[n] = rl, ul, pl
                else:
                    rho[n] = rl * (ps / pl + g1) / (g1 * ps / pl + 1.0)
                    u[n], p[n] = us, ps
            else:                                      # left rarefaction
                shl = ul - al
                astar = al * (ps / pl) ** ((gamma - 1.0) / (2 * gamma))
                stl = us - astar
                if si <= shl:
                    rho[n], u[n], p[n] = rl, ul, pl
                elif si >= stl:
                    rho[n] = rl * (ps / pl) ** (1.0 / gamma)
                    u[n], p[n] = us, ps
                else:
                    c = 2.0 / (gamma + 1.0) + g1 / al * (ul - si)
                    rho[n] = rl * c ** (2.0 / (gamma - 1.0))
                    u[n] = 2.0 / (gamma + 1.0) * (al + 0.5 * (gamma - 1.0) * ul + si)
                    p[n] = pl * c ** (2.0 * gamma / (gamma - 1.0))
        else:                                          # right of the contact
            if ps > pr:                                # right shock
                sr = ur + ar * np.sqrt((gamma + 1.0) / (2 * gamma) * ps / pr + (gamma - 1.0) / (2 * gamma))
                if si >= sr:
                    rho[n], u[n], p[n] = rr, ur, pr
                else:
                    rho[n] = rr * (ps / pr + g1) / (g1 * ps / pr + 1.0)
                    u[n], p[n] = us, ps
            else:                                      # right rarefaction
                shr = ur + ar
                astar = ar * (ps / pr) ** ((gamma - 1.0) / (2 * gamma))
                str_ = us + astar
                if si >= shr:
                    rho[n], u[n], p[n] = rr, ur, pr
                elif si <= str_:
                    rho[n] = rr * (ps / pr) ** (1.0 / gamma)
                    u[n], p[n] = us, ps
                else:
                    c = 2.0 / (gamma + 1.0) - g1 / ar * (ur - si)
                    rho[n] = rr * c ** (2.0 / (gamma - 1.0))
                    u[n] = 2.0 / (gamma + 1.0) * (-ar + 0.5 * (gamma - 1.0) * ur + si)
                    p[n] = pr * c ** (2.0 * gamma / (gamma - 1.0))
    return rho, u, p


SOD_LEFT = (1.0, 0.0, 1.0)
SOD_RIGHT = (0.125, 0.0, 0.1)


def sod(x, t=0.2, gamma=1.4):
    """Sod problem of test/1D/Sod79: returns (rho, u, p, sie)."""
    rho, u, p = sample(x, t, SOD_LEFT, SOD_RIGHT, 0.5, gamma)
    return rho, u, p, p / ((gamma - 1.0) * rho)


def probe_check(xc, rho, u, p, sie, exact_table, threshold=1.0):
    """The reference's acceptance rule (test/1D/Sod79/check.py:46-82): interpolate the exact
    table (x, rho, u, p, sie) to the cell centres and require <= threshold percent error in
    rho, u, p, sie at the cells int(n*f), f in (0.2, 0.5, 0.6, 0.8, 0.9).  Returns the list
    of (quantity, cell, percent error) that fail."""
    xe = exact_table[:, 0]
    ref = {name: np.interp(xc, xe, exact_table[:, col]) for name, col in
           (("rho", 1), ("u", 2), ("p", 3), ("sie", 4))}
    cfd = {"rho": rho, "u": u, "p": p, "sie": sie}
    n = len(rho)
    fails = []
    for name in ("rho", "u", "p", "sie"):
        for f in (0.20, 0.50, 0.60, 0.80, 0.90):
            i = int(n * f)
            denom = ref[name][i] + (1e-10 if name == "u" else 0.0)
            err = abs((cfd[name][i] - ref[name][i]) / denom) * 100.0
            if err > threshold:
                fails.append((name, i, err))
    return fails


def probe_fails(cells, cfd, exact, thresholds=None):
    """The reference's per-probe acceptance rule of its 1-D verify scripts
    (test/1D/Ein91/verify.py:141-215, test/1D/Noh87/verify.py:141-221): at each probe cell,
    the percent error |cfd - exact| / exact of rho, u, p and sie must not exceed 1, 1, 1.5
    and 1 percent; the velocity's denominator is exact + 1e-10.  cfd and exact map those
    names to arrays over the same cells.  Returns the failing (quantity, cell, percent)."""
    th = {"rho": 1.0, "u": 1.0, "p": 1.5, "sie": 1.0}
    th.update(thresholds or {})
    fails = []
    for name in ("rho", "u", "p", "sie"):
        for i in cells:
            denom = exact[name][i] + (1e-10 if name == "u" else 0.0)
            err = abs((cfd[name][i] - exact[name][i]) / denom) * 100.0
            if not err <= th[name]:
                fails.append((name, i, err))
    return fails


EIN91_LEFT = (1.0, -2.0, 0.4)
EIN91_RIGHT = (1.0, 2.0, 0.4)


def ein91(x, t=0.15, gamma=1.4):
    """Einfeldt's 1-2-3 problem of test/1D/Ein91 (two rarefactions, verify.py:46-49):
    returns (rho, u, p, sie)."""
    rho, u, p = sample(x, t, EIN91_LEFT, EIN91_RIGHT, 0.5, gamma)
    return rho, u, p, p / ((gamma - 1.0) * rho)


def prandtl_meyer_nu(M, gamma=1.4):
    """Prandtl-Meyer function nu(M) in radians (textbook closed form)."""
    g = (gamma + 1.0) / (gamma - 1.0)
    return np.sqrt(g) * np.arctan(np.sqrt((M * M - 1.0) / g)) - np.arctan(np.sqrt(M * M - 1.0))


def prandtl_meyer_mach(M1, theta_deg, gamma=1.4):
    """Mach number after an isentropic expansion of a Mach M1 flow by theta degrees: the root of
    nu(M2) = nu(M1) + theta by bisection on [1, 30] (the reference's check uses the same
    relation, test/2D/euler/prandtl-meyer/check.py)."""
    target = prandtl_meyer_nu(M1, gamma) + theta_deg * np.pi / 180.0
    lo, hi = 1.0, 30.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if prandtl_meyer_nu(mid, gamma) < target:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


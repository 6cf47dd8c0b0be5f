"""Shared helpers of the parity tests."""
import os

import numpy as np

from mose_amd import abi


def interior(a):
    return abi.interior(a)


def rel_l2(a, b):
    """||a - b||_2 / ||b||_2 over every interior cell and variable."""
    a = interior(a).ravel()
    b = interior(b).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def max_rel(a, b, scale):
    """max |a - b| / scale over interior cells and variables."""
    return float(np.max(np.abs(interior(a) - interior(b))) / scale)


def flux_scale(P):
    """A face-flux magnitude of the state: max over cells of rho |u|^2 + p, rho a."""
    Pi = interior(P)
    rho, p = Pi[..., 0], Pi[..., 4]
    q2 = Pi[..., 1] ** 2 + Pi[..., 2] ** 2 + Pi[..., 3] ** 2
    return float(np.max(rho * q2 + p))


def perturb(case, amp=0.02, seed=0):
    """Add a seeded smooth-random perturbation to every block's interior primitive state so
    that the residual is non-trivial in every direction and variable."""
    rng = np.random.default_rng(seed)
    for blk in case.blocks:
        Pi = blk.P[2:-2, 2:-2, 2:-2, :]
        noise = rng.standard_normal(Pi.shape)
        Pi[..., 0] *= 1.0 + amp * noise[..., 0]
        Pi[..., 1:4] += amp * noise[..., 1:4]
        Pi[..., 4] *= 1.0 + amp * noise[..., 4]
    return case


def skewed_nodes(dims, amp=0.08, hi=(1.0, 0.8, 0.6)):
    """Smoothly distorted (curvilinear) grid of a box: node (3, 0:ni, 0:nj, 0:nk) array in
    numpy (k, j, i, 3) order."""
    from mose_amd import cases
    nodes = cases.cartesian_nodes(dims, (0, 0, 0), hi)
    X, Y, Z = nodes[..., 0].copy(), nodes[..., 1].copy(), nodes[..., 2].copy()
    nodes[..., 0] = X + amp * np.sin(2 * np.pi * Y) * np.sin(np.pi * Z)
    nodes[..., 1] = Y + amp * np.sin(2 * np.pi * X) * np.cos(np.pi * Z)
    nodes[..., 2] = Z + 0.5 * amp * np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)
    return np.ascontiguousarray(nodes)


GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def oblique_shock_relations(M1, delta, gam):
    """Weak oblique shock of a wedge delta at Mach M1: the shock angle from the
    theta-beta-M relation tan(delta) = 2 cot(b) (M1^2 sin^2 b - 1) / (M1^2 (gam + cos 2b) + 2)
    (the first root above the Mach angle, by bisection), then the normal-shock jumps of
    M1 sin(b): returns (M2, p2/p1, T2/T1, rho2/rho1)."""
    f = lambda b: 2.0 / np.tan(b) * (M1 * M1 * np.sin(b) ** 2 - 1.0) / (M1 * M1 * (gam + np.cos(2 * b)) + 2.0) \
        - np.tan(delta)
    bs = np.linspace(np.arcsin(1.0 / M1) + 1e-9, 0.5 * np.pi - 1e-9, 20001)
    k = int(np.argmax(f(bs) > 0.0))
    lo, hi = bs[k - 1], bs[k]
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if f(mid) < 0.0 else (lo, mid)
    b = 0.5 * (lo + hi)
    mn1 = M1 * np.sin(b)
    r21 = (gam + 1) * mn1 ** 2 / ((gam - 1) * mn1 ** 2 + 2)
    p21 = 1 + 2 * gam / (gam + 1) * (mn1 ** 2 - 1)
    mn2 = np.sqrt((mn1 ** 2 + 2 / (gam - 1)) / (2 * gam / (gam - 1) * mn1 ** 2 - 1))
    return mn2 / np.sin(b - delta), p21, p21 / r21, r21


def oblique_shock_verify(Solver, rs):
    """test/2D/euler/oblique-shock with its input.ini (M = 4 inflow, two blocks, MUSCL-SD + van
    Albada, RK3, local dt, CFL 0.9 ramped over 10 iterations, irs-beta 0.5, 3000 iterations),
    for the four Riemann solvers its MOSE.sh runs: the pre-shock (block 1, cells 5..14 x 5..14)
    and post-shock (block 2, cells 14..34 x 5..49) means satisfy its verify.py rule, p2/p1,
    T2/T1, rho2/rho1 and M2 within 1.5 % of the oblique-shock relations for the wedge angle of
    block 2's wall nodes."""
    from mose_amd import formats
    ic = os.path.join(GOLD, "oblique_shock_ic.tec.gz")
    blocks = formats.read_tecplot(ic)
    R = float(blocks[0]["prim"][0, 0, 0, 4] / blocks[0]["prim"][0, 0, 0, 0])
    cfg = abi.make_cfg(1.4, R, 0.9, "vanalbada", rs, 3, False, rampa_cfl_iter=10, muscl_sd=True, irs_beta=0.5)
    case = formats.load_case(ic, os.path.join(GOLD, "oblique_shock_bc.txt.gz"), cfg)
    g = Solver(case)
    for _ in range(3000):
        g.step()
    P1, P2 = interior(g.get_state(1)), interior(g.get_state(2))
    g.close()
    pre = P1[0, 5:15, 5:15]
    post = P2[0, 5:50, 14:35]
    p1, r1 = pre[..., 4].mean(), pre[..., 0].mean()
    T1 = (pre[..., 4] / (pre[..., 0] * R)).mean()
    p2, r2 = post[..., 4].mean(), post[..., 0].mean()
    T2 = (post[..., 4] / (post[..., 0] * R)).mean()
    u2, v2 = post[..., 1].mean(), post[..., 2].mean()
    m2 = np.hypot(u2, v2) / np.sqrt(1.4 * R * T2)
    wall = case.blocks[1].nodes[0, 0]
    delta = np.arctan2(wall[-1, 1] - wall[0, 1], wall[-1, 0] - wall[0, 0])
    M2, p21, t21, r21 = oblique_shock_relations(4.0, delta, 1.4)
    for got, ref in ((p2 / p1, p21), (T2 / T1, t21), (r2 / r1, r21), (m2, M2)):
        assert abs(got - ref) / ref * 100 <= 1.5, (got, ref)

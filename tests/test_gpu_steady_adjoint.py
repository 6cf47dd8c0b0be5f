"""Steady adjoint gradient of the bump channel's wall-pressure integral (minmod + HLLC, local
time steps, Mach 2 inflow 405, symmetry walls 300, extrapolation outlet 400): the device
adjoint solve (mose_adjoint_solve) against its own linear system, and dJ/d(control-point
heights) and dJ/d(inflow record) against central differences (step 1e-6) of J at re-converged
states, on the GPU and on the oracle (the CPU restatement of the reference).  The loop whose
transpose the solve takes is the local-dt SSP-RK iteration (Mod_Explicit.f90:45-85); the wall
faces are BC_Symmetry_Eul's (Lib_BC_Fluxes_Symmetry.f90:14-125).

The inflow angle is 0.03 rad.  At exactly 0 the flow upstream of the bump is uniform, every
slope there is exactly zero, and a change of the angle creates new waves in those cells, where
minmod has no derivative (its response is minmod(dx, dy), not linear in the perturbation): the
adjoint, which takes the zero subgradient there, then differs from the central difference of
the angle by 4 %, while the heights, Mach number and pressure, which leave the upstream slopes
zero, still agree to 1e-7."""
import copy

import numpy as np
import pytest

from mose_amd import abi, cases, steady_adjoint as sa
from mose_amd.lib import GpuSolver
from oracle.oracle import OracleSolver

pytestmark = pytest.mark.gpu

NC = 8
H0 = np.full(NC, 0.002)
ALPHA0 = 0.03
EPS = 1e-6
TOL = 1e-6


def _case(ni, nj, heights=H0, limiter="minmod"):
    return cases.bump_channel(ni=ni, nj=nj, ncontrol=NC, heights=heights, limiter=limiter, alpha=ALPHA0)


def _converged(case):
    g = GpuSolver(case)
    steps, norms = sa.converge(g, max_steps=60000, tol=5e-15, chunk=1000)
    print("primal: %d steps, step norms %s" % (steps, np.array2string(norms, precision=2)), flush=True)
    assert np.max(norms) < 1e-13
    return g


@pytest.fixture(scope="module")
def base():
    case = _case(48, 24)
    g = _converged(case)
    res = sa.gradient(case, g, rtol=1e-12, check_every=200)
    P = g.get_state(1)
    print("J %.12e, adjoint %d iterations, |Q| %.2e -> %.2e" % ((res["J"], res["adjoint_iterations"])
                                                               + res["adjoint_res"]), flush=True)
    yield case, g, res, P
    g.close()


def test_adjoint_solution_satisfies_its_system(base):
    case, g, res, P = base
    assert res["adjoint_res"][1] <= 1e-12 * res["adjoint_res"][0]
    _, dJdP, _ = sa.wall_objective(case, P)
    psi = g.adjoint_solve([dJdP], max_iter=200000, rtol=1e-12, check_every=200)[0]
    Pbar = g.residual_vjp(psi)[0]
    inner = (slice(2, -2), slice(2, -2), slice(2, -2))
    err = np.linalg.norm(Pbar[inner] - dJdP[inner]) / np.linalg.norm(dJdP[inner])
    print("|(dR/dP)^T psi - dJ/dP| / |dJ/dP| = %.2e" % err)
    assert err < 1e-9


def _fd_heights(base_case, P, v, solver_cls):
    """Central difference of J along the height direction v, re-converged from P."""
    Js = []
    for s in (1.0, -1.0):
        c = _case(*base_case.blocks[0].dims[:2], heights=H0 + s * EPS * v)
        J, steps, norms = sa.objective_at(c, solver_cls, warm=P, max_steps=60000, tol=5e-15, chunk=1000)
        assert np.max(norms) < 1e-13, (steps, norms)
        Js.append(J)
    return (Js[0] - Js[1]) / (2 * EPS)


@pytest.mark.parametrize("m", [0, 3, 6])
def test_height_gradient_against_central_differences(base, m):
    case, g, res, P = base
    v = np.zeros(NC)
    v[m] = 1.0
    fd = _fd_heights(case, P, v, GpuSolver)
    adj = float(res["dJdh"][m])
    print("dJ/dh[%d]: adjoint %.10e  fd %.10e" % (m, adj, fd))
    assert abs(adj - fd) <= TOL * max(abs(fd), 1e-2 * np.max(np.abs(res["dJdh"])))


def test_height_gradient_random_direction(base):
    case, g, res, P = base
    v = np.random.default_rng(5).standard_normal(NC)
    fd = _fd_heights(case, P, v, GpuSolver)
    adj = float(res["dJdh"] @ v)
    print("dJ/dh . v: adjoint %.10e  fd %.10e" % (adj, fd))
    assert abs(adj - fd) <= TOL * abs(fd)


@pytest.mark.parametrize("q,key", [(0, "mach"), (1, "T"), (2, "p"), (3, "alpha")])
def test_inflow_gradient_against_central_differences(base, q, key):
    case, g, res, P = base
    Js = []
    for s in (1.0, -1.0):
        c = copy.deepcopy(case)
        sel = (c.bc["b"] == 1) & (c.bc["f"] == 1)
        c.bc[key][sel] += s * EPS
        J, steps, norms = sa.objective_at(c, GpuSolver, warm=P, max_steps=60000, tol=5e-15, chunk=1000)
        assert np.max(norms) < 1e-13
        Js.append(J)
    fd = (Js[0] - Js[1]) / (2 * EPS)
    adj = float(res["dJdinflow"][q])
    print("dJ/d%s: adjoint %.10e  fd %.10e" % (key, adj, fd))
    assert abs(adj - fd) <= TOL * max(abs(fd), 1e-2 * np.max(np.abs(res["dJdinflow"])))


def test_height_gradient_against_oracle_differences():
    """The GPU adjoint gradient on a 32 x 16 bump against central differences of J from the
    oracle's converged states (the CPU restatement of the reference's loop)."""
    case = _case(32, 16)
    g = _converged(case)
    res = sa.gradient(case, g, rtol=1e-12, check_every=200)
    P = g.get_state(1)
    g.close()
    v = np.random.default_rng(9).standard_normal(NC)
    fd = _fd_heights(case, P, v, OracleSolver)
    adj = float(res["dJdh"] @ v)
    print("32x16, dJ/dh . v: GPU adjoint %.10e  oracle fd %.10e" % (adj, fd))
    assert abs(adj - fd) <= TOL * abs(fd)


def test_height_gradient_128x64_16_controls():
    """A larger grid (128 x 64, 16 control points): the adjoint gradient along a random height
    direction against central differences (step 1e-7) of re-converged GPU states."""
    nc, eps = 16, 1e-7
    h0 = np.full(nc, 0.002)
    mk = lambda h: cases.bump_channel(ni=128, nj=64, ncontrol=nc, heights=h, limiter="minmod", alpha=ALPHA0)
    case = mk(h0)
    g = GpuSolver(case)
    steps, norms = sa.converge(g, max_steps=150000, tol=5e-15, chunk=1000)
    assert np.max(norms) < 1e-13, (steps, norms)
    res = sa.gradient(case, g, rtol=1e-12, check_every=500, max_iter=400000)
    P = g.get_state(1)
    g.close()
    v = np.random.default_rng(13).standard_normal(nc)
    Js = []
    for sgn in (1.0, -1.0):
        J, steps, norms = sa.objective_at(mk(h0 + sgn * eps * v), GpuSolver, warm=P, max_steps=150000, tol=5e-15,
                                          chunk=1000)
        assert np.max(norms) < 1e-13, (steps, norms)
        Js.append(J)
    fd = (Js[0] - Js[1]) / (2 * eps)
    adj = float(res["dJdh"] @ v)
    print("128x64, dJ/dh . v: adjoint %.10e  fd %.10e" % (adj, fd))
    assert abs(adj - fd) <= TOL * abs(fd)


def test_adjoint_solve_refuses_bad_arguments():
    case = _case(16, 8)
    g = GpuSolver(case)
    with pytest.raises(Exception):
        g.adjoint_solve([np.zeros(abi.padded_shape(case.blocks[0].dims))], max_iter=0)
    g.close()

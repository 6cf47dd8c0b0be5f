"""GPU parity of the subsonic inlet (401: stagnation T0, p0, flow direction, p0 relaxation) and
the outlet (406: back pressure) against the oracle: residuals and steps on rectilinear and
curvilinear boxes, with 'normal' and explicit angles, faces whose cells mix in- and outflow
(symmetry fallback), subsonic and supersonic exits, together with the shock detector and
residual smoothing; the steady duct flow; the refusals."""
import copy

import numpy as np
import pytest

from mose_amd import abi, cases, formats
from mose_amd.lib import GpuSolver, MoseError
from oracle.oracle import OracleSolver

from helpers import flux_scale, interior, perturb, rel_l2, skewed_nodes

pytestmark = pytest.mark.gpu

LIMS = ["minmod", "vanleer", "vanalbada", "mc", "superbee", "first-order"]
GAM, RG, DELTA, EXP = 1.4, 1.0 / 1.4, 0.2, 3.5


def _with(case, **kw):
    c = copy.deepcopy(case)
    f = {n: getattr(case.cfg, n) for n, _ in abi.MoseCfg._fields_ if n != "reserved"}
    f.update(kw)
    c.cfg = abi.MoseCfg(**f)
    return c


def _box(dims, faces, u=0.45, v=0.0, w=0.0, nodes=None, amp=0.03, seed=3, **cfgkw):
    nodes = cases.cartesian_nodes(dims, (0, 0, 0), (1.0, 0.7, 0.4)) if nodes is None else nodes
    prim = np.zeros(tuple(dims[::-1]) + (5,))
    prim[..., 0], prim[..., 1], prim[..., 2], prim[..., 3], prim[..., 4] = 1.0, u, v, w, 1.0 / 1.4
    blk = cases.Block(tuple(dims), nodes, cases.padded_state(dims, prim))
    bc = cases.bc_records([{"dims": tuple(dims), "faces": faces}])
    kw = dict(gamma=GAM, Rgas=RG, cfl=0.5, limiter="vanleer", riemann="HLLC", n_rk=3, time_accurate=True)
    kw.update(cfgkw)
    return perturb(cases.Case("inlet_outlet", abi.make_cfg(**kw), [blk], bc), amp=amp, seed=seed)


def _inlet(T0=1.04, p0=0.86, alpha=0.1, beta=-0.05, rel=0.8):
    return {"type": 401, "T": T0, "p": p0, "alpha": alpha, "beta": beta, "rel_fac": rel}


def _outlet(p=0.7, rel=1.0):
    return {"type": 406, "p": p, "rel_fac": rel}


def _residual_check(case, tol=1e-12):
    g, o = GpuSolver(case), OracleSolver(case)
    Rg, Ro = g.residual(1), o.residual(1)
    assert rel_l2(Rg, Ro) < tol, rel_l2(Rg, Ro)
    assert np.max(np.abs(interior(Rg) - interior(Ro))) < 1e-13 * flux_scale(case.blocks[0].P)
    g.close()
    o.close()


def _steps_check(case, nsteps=10, tol=1e-12):
    g, o = GpuSolver(case), OracleSolver(case)
    for _ in range(nsteps):
        dg, _, ng = g.step()
        do, _, no = o.step()
        assert abs(dg - do) <= 1e-13 * do
        assert np.allclose(ng, no, rtol=1e-9, atol=1e-12 * np.max(no))
    err = rel_l2(g.get_state(1), o.get_state(1))
    assert err < tol, err
    g.close()
    o.close()


@pytest.mark.parametrize("lim", LIMS)
@pytest.mark.parametrize("rs", ["HLLC", "AUSM+", "HLLE", "SLAU2"])
def test_residual_inlet_outlet(lim, rs):
    case = _box((18, 10, 7), [_inlet(), _outlet(), 300, 300, 400, 400], v=0.05, limiter=lim, riemann=rs)
    _residual_check(case)


@pytest.mark.parametrize("pe", [0.0, 0.3, 0.72, 0.9])
def test_outlet_back_pressures(pe):
    """No back pressure (extrapolation), below critical (p* clip), near and above the cell
    pressure."""
    _residual_check(_box((12, 8, 6), [_inlet(), _outlet(pe), 300, 300, 300, 300]))


def test_supersonic_exit_and_mixed_direction_faces():
    """Large perturbations: some cells of the 401 face flow out and some of the 406 face flow
    in (both take the symmetry flux there), some 406 cells are supersonic."""
    case = _box((14, 9, 6), [_inlet(), _outlet(0.6), 300, 400, 300, 400], u=0.05, amp=0.2, seed=11)
    un_in = interior(case.blocks[0].P)[:, :, 0, 1]
    un_out = interior(case.blocks[0].P)[:, :, -1, 1]
    assert np.any(un_in > 0) and np.any(un_in < 0) and np.any(un_out > 0) and np.any(un_out < 0)
    _residual_check(case)
    case = _box((14, 9, 6), [_inlet(), _outlet(0.6), 300, 400, 300, 400], u=1.3, amp=0.2, seed=12)
    _residual_check(case)


def test_p0_relaxation():
    """p0 far below the cell pressure: the face quadratic has no real root until p0 has been
    relaxed towards the cell pressure (rel_fac 0.3, several iterations)."""
    _residual_check(_box((10, 8, 6), [_inlet(p0=0.01, alpha=1.0, rel=0.3), _outlet(), 300, 300, 300, 300]))


@pytest.mark.parametrize("face", [1, 3, 6])
def test_inlet_on_other_faces_normal_token(face):
    """The inlet on an i, j or k face with the 'normal' direction token; the outlet opposite."""
    faces = [300] * 6
    faces[face - 1] = _inlet(alpha=formats.NORMAL, beta=formats.NORMAL)
    opp = face + 1 if face % 2 else face - 1
    faces[opp - 1] = _outlet()
    vel = [0.0, 0.0, 0.0]
    vel[(face - 1) // 2] = 0.4 if face % 2 else -0.4
    _residual_check(_box((10, 9, 8), faces, u=vel[0], v=vel[1], w=vel[2]))


@pytest.mark.parametrize("rs", ["HLLC", "HLLE"])
def test_curvilinear_inlet_outlet(rs):
    dims = (14, 10, 7)
    case = _box(dims, [_inlet(alpha=0.15, beta=0.05), _outlet(0.68), 300, 300, 300, 300],
                nodes=skewed_nodes(dims), riemann=rs, limiter="vanalbada")
    _residual_check(case)
    _steps_check(case, nsteps=6)
    case = _box(dims, [_inlet(alpha=formats.NORMAL, beta=formats.NORMAL), _outlet(0.68), 300, 300, 300, 300],
                nodes=skewed_nodes(dims), riemann=rs)
    _residual_check(case)


@pytest.mark.parametrize("time_accurate", [True, False])
def test_steps_inlet_outlet(time_accurate):
    case = _box((16, 10, 6), [_inlet(), _outlet(), 300, 300, 400, 400], time_accurate=time_accurate)
    _steps_check(case, nsteps=12)


def test_with_shock_detector_and_residual_smoothing():
    case = _box((16, 10, 6), [_inlet(), _outlet(), 300, 300, 400, 400], amp=0.12, riemann="HLLE", muscl_sd=True,
                irs_beta=0.5, time_accurate=False)
    _steps_check(case, nsteps=8)
    case = _box((16, 10, 6), [_inlet(), _outlet(), 300, 300, 400, 400], amp=0.12, riemann="HLLC+")
    _residual_check(case)


def test_fp32_flux_within_1e5():
    case = _box((16, 10, 6), [_inlet(), _outlet(), 300, 300, 400, 400])
    g64 = GpuSolver(case)
    g32 = GpuSolver(_with(case, precision=abi.PREC_FP32_FLUX))
    assert rel_l2(g32.residual(1), g64.residual(1)) < 1e-5
    g64.close()
    g32.close()


def test_duct_converges_to_isentropic_flow():
    """A constant-area duct between a 401 inlet (T0 = 1, p0 = 1, normal direction) and a 406
    outlet at p_exit = 0.84, symmetry walls: the steady state is uniform with p = p_exit and
    M^2 = ((p0 / p_exit)^((gamma-1)/gamma) - 1) / delta, T0 = T (1 + delta M^2)."""
    dims = (32, 2, 2)
    prim = np.zeros(tuple(dims[::-1]) + (5,))
    prim[..., 0], prim[..., 1], prim[..., 4] = 0.9 / RG, 0.3, 0.9
    blk = cases.Block(dims, cases.cartesian_nodes(dims, (0, 0, 0), (2.0, 0.5, 0.5)), cases.padded_state(dims, prim))
    bc = cases.bc_records([{"dims": dims, "faces": [_inlet(1.0, 1.0, formats.NORMAL, formats.NORMAL, 1.0),
                                                    _outlet(0.84), 300, 300, 300, 300]}])
    cfg = abi.make_cfg(GAM, RG, 0.8, "vanleer", "HLLC", 3, False)
    g = GpuSolver(cases.Case("duct", cfg, [blk], bc))
    _, norms = g.advance(4000)
    P = interior(g.get_state(1))
    g.close()
    assert norms[0] < 1e-11, norms
    M2 = ((1.0 / 0.84) ** (1.0 / EXP) - 1.0) / DELTA
    rho, u, p = P[..., 0], P[..., 1], P[..., 4]
    assert np.allclose(p, 0.84, rtol=1e-8)
    assert np.allclose(u * u / (GAM * p / rho), M2, rtol=1e-7)
    assert np.allclose(p / (rho * RG) * (1.0 + DELTA * M2), 1.0, rtol=1e-8)


def test_refusals():
    """The discrete adjoint does not cover the 401 / 406 fluxes (MOSE_E_ARG); a viscous block
    with them is refused (MOSE_E_BC)."""
    case = _box((8, 8, 6), [_inlet(), _outlet(), 300, 300, 300, 300])
    g = GpuSolver(case)
    with pytest.raises(MoseError) as ei:
        g.residual_vjp([np.ones(abi.padded_shape((8, 8, 6)))])
    assert ei.value.code == abi.E_ARG
    g.close()
    vis = cases.taylor_green(n=8, nbox=1, precision=abi.PREC_FP64)
    blk = vis.blocks[0]
    vis.bc = cases.bc_records([{"dims": blk.dims, "faces": [_inlet(), _outlet(), 0, 0, 0, 0]}])
    with pytest.raises(MoseError) as ei:
        GpuSolver(vis)
    assert ei.value.code == abi.E_BC

"""Discrete adjoint of the residual (hand-written HIP reverse of flux, limiter and
reconstruction) against finite differences.

Checks: directional derivatives <Rbar, dR/dP v> from central differences of the GPU
residual and of the oracle's residual against <Pbar, v>; torch.autograd.gradcheck through
the autograd Function on a small smooth case; every limiter x Riemann solver, boundary types
0/300/400/405, block connections and curvilinear grids.
"""
import numpy as np
import pytest
import torch

from mose_amd import abi, cases
from mose_amd.autograd import residual as torch_residual
from mose_amd.lib import GpuSolver, MoseError
from oracle.oracle import OracleSolver

from helpers import interior, perturb, skewed_nodes

pytestmark = pytest.mark.gpu

LIMS = ["minmod", "vanleer", "vanalbada", "mc", "superbee"]
RSS = ["HLLC", "AUSM+", "HLLE", "SLAU", "SLAU2"]
INFLOW = {"type": 405, "mach": 1.8, "T": 1.05, "p": 0.7, "alpha": 0.1, "beta": -0.05, "rel_fac": 0.8}


def _interior_mask(dims):
    m = np.zeros(abi.padded_shape(dims))
    m[2:-2, 2:-2, 2:-2, :] = 1.0
    return m


def _directional_check(case, solver_cls=GpuSolver, h=1e-7, tol=2e-6, seed=0):
    """<Pbar, v> (GPU adjoint) vs central differences of <Rbar, R(P + t v)> (solver_cls)."""
    rng = np.random.default_rng(seed)
    dims = [b.dims for b in case.blocks]
    P0 = [b.P.copy() for b in case.blocks]
    Rbar = [rng.standard_normal(abi.padded_shape(d)) * _interior_mask(d) for d in dims]
    V = [rng.standard_normal(abi.padded_shape(d)) * _interior_mask(d) * np.abs(P) for d, P in zip(dims, P0)]
    g = GpuSolver(case)
    Pbar = g.residual_vjp(Rbar)
    adj = sum(float(np.sum(pb * v)) for pb, v in zip(Pbar, V))
    s = solver_cls(case)

    def J(t):
        for b, (P, v) in enumerate(zip(P0, V), start=1):
            s.put_state(b, P + t * v)
        return sum(float(np.sum(rb * s.residual(b))) for b, rb in enumerate(Rbar, start=1))

    fd = (J(h) - J(-h)) / (2 * h)
    scale = sum(float(np.sum(np.abs(pb * v))) for pb, v in zip(Pbar, V))
    assert abs(adj - fd) <= tol * scale, (adj, fd, scale)
    g.close()
    return adj, fd


@pytest.mark.parametrize("lim", LIMS)
@pytest.mark.parametrize("rs", RSS)
def test_vjp_periodic(lim, rs):
    case = perturb(cases.vortex(dims=(10, 8, 6), limiter=lim, riemann=rs), seed=1)
    _directional_check(case)


def _box(dims, faces, limiter="vanalbada", riemann="HLLC", mach=1.5, curvi=False, seed=3):
    nodes = skewed_nodes(dims) if curvi else cases.cartesian_nodes(dims, (0, 0, 0), (1.0, 0.7, 0.4))
    prim = np.zeros(tuple(dims[::-1]) + (5,))
    prim[..., 0] = 1.0
    prim[..., 1] = mach
    prim[..., 2] = 0.2
    prim[..., 4] = 1.0 / 1.4
    blk = cases.Block(tuple(dims), nodes, cases.padded_state(dims, prim))
    bc = cases.bc_records([{"dims": tuple(dims), "faces": faces}])
    cfg = abi.make_cfg(1.4, 1.0 / 1.4, 0.5, limiter, riemann, 3, True)
    return perturb(cases.Case("box", cfg, [blk], bc), amp=0.03, seed=seed)


@pytest.mark.parametrize("rs", RSS)
def test_vjp_boundary_types(rs):
    case = _box((9, 7, 5), [INFLOW, 400, 300, 300, 300, 0], riemann=rs)
    _directional_check(case)


def test_vjp_inflow_fallback_and_extrapolation_ghosts():
    inflow = {"type": 405, "mach": 2.0, "T": 1.0, "p": 0.7, "alpha": 0.0, "beta": 0.0, "rel_fac": 1.0}
    case = _box((8, 6, 5), [400, inflow, 300, 400, 0, 300], mach=1.2, limiter="vanleer")
    _directional_check(case)


def test_vjp_multiblock_connections():
    case = perturb(cases.multiblock_chain(nblocks=4, seed=5, lo=5, hi=9, precision=abi.PREC_FP64), seed=4)
    _directional_check(case)


@pytest.mark.parametrize("rs", RSS)
def test_vjp_curvilinear(rs):
    case = _box((9, 7, 5), [INFLOW, 400, 300, 300, 400, 0], riemann=rs, curvi=True)
    _directional_check(case)


def test_vjp_bump_channel_against_oracle_residual():
    """Central differences of the oracle (CPU restatement of the reference residual)."""
    case = perturb(cases.bump_channel(ni=24, nj=12), amp=0.01, seed=2)
    _directional_check(case, solver_cls=OracleSolver)


def test_vjp_against_oracle_residual_3d():
    case = perturb(cases.vortex(dims=(8, 6, 5), limiter="vanalbada", riemann="AUSM+"), seed=7)
    _directional_check(case, solver_cls=OracleSolver)


def test_torch_gradcheck_small_case():
    case = perturb(cases.vortex(dims=(4, 3, 3), limiter="vanalbada", riemann="HLLC"), amp=0.01, seed=9)
    g = GpuSolver(case)
    P = torch.from_numpy(case.blocks[0].P.copy()).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda x: torch_residual(g, x), (P,), eps=1e-6, atol=1e-6, rtol=1e-4)


@pytest.mark.parametrize("pe,mach", [(0.62, 0.5), (0.05, 0.5), (0.0, 0.5), (0.7, 1.5)])
def test_vjp_outflow406(pe, mach):
    """The 406 outlet flux (back pressure p_exit: subsonic, below the critical pressure, none,
    supersonic exit) in the state adjoint, against central differences of the GPU and of the
    oracle residual."""
    outlet = {"type": 406, "p": pe, "rel_fac": 1.0}
    case = _box((9, 7, 5), [400, outlet, 300, 300, 300, 0], mach=mach)
    _directional_check(case)
    _directional_check(case, solver_cls=OracleSolver)


def test_vjp_outflow406_mixed_directions():
    """Cells of the 406 face that flow in take the symmetry flux and its adjoint."""
    outlet = {"type": 406, "p": 0.62, "rel_fac": 1.0}
    case = _box((9, 7, 5), [400, outlet, 300, 400, 300, 0], mach=0.02, seed=11)
    un = interior(case.blocks[0].P)[:, :, -1, 1]
    assert np.any(un > 0) and np.any(un < 0)
    _directional_check(case)


def test_vjp_refuses_inlet401():
    inlet = {"type": 401, "T": 1.04, "p": 0.86, "alpha": 0.1, "beta": -0.05, "rel_fac": 0.8}
    case = _box((8, 6, 5), [inlet, 400, 300, 300, 300, 0], mach=0.4)
    g = GpuSolver(case)
    with pytest.raises(MoseError) as ei:
        g.residual_vjp([np.ones(abi.padded_shape((8, 6, 5)))])
    assert ei.value.code == abi.E_ARG
    g.close()

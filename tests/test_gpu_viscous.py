"""GPU parity of the laminar viscous path (interior Diffusive_Flux, BC_Connection_Visc on
periodic / 1-to-1 connections, Sutherland viscosity, VNN time step) against the oracle.

Bar: FP64 residuals within 1e-12 relative L2; FP64 states after several time-accurate steps
within 1e-12; FP32 flux within 1e-5 of the FP64 path.
"""
import numpy as np
import pytest

from mose_amd import abi, cases
from mose_amd.lib import GpuSolver
from oracle.oracle import OracleSolver

from helpers import flux_scale, interior, perturb, rel_l2

pytestmark = pytest.mark.gpu

VISC = [0.02, 1.0, 0.4, 0.71, 0.3]


def _viscous(case, params=VISC):
    case.meta["viscous"] = [(0, list(params))]
    return case


def _residual_check(case, tol=1e-12):
    g, o = GpuSolver(case), OracleSolver(case)
    for b in range(1, len(case.blocks) + 1):
        Rg, Ro = g.residual(b), o.residual(b)
        err = rel_l2(Rg, Ro)
        assert err < tol, "block %d: rel L2 %.3e" % (b, err)
        assert np.max(np.abs(interior(Rg) - interior(Ro))) < 1e-13 * flux_scale(case.blocks[b - 1].P)
    g.close()


@pytest.mark.parametrize("rs", ["HLLC", "AUSM+", "HLLE", "SLAU", "SLAU2"])
def test_viscous_residual_periodic_box(rs):
    case = _viscous(perturb(cases.vortex(dims=(14, 11, 9), riemann=rs), seed=21))
    _residual_check(case)


def test_viscous_residual_split_blocks():
    case = perturb(cases.vortex(dims=(16, 12, 9), riemann="AUSM+"), seed=22)
    split = _viscous(cases.split_block_case(case, axis=1, parts=3))
    _residual_check(split)


def test_viscous_residual_two_dimensional():
    dims = (20, 14, 1)
    nodes = cases.cartesian_nodes(dims, (0, 0, 0), (1.0, 0.7, 0.1))
    xc = cases.cell_centres(nodes)
    prim = np.zeros(dims[::-1] + (5,))
    prim[..., 0] = 1.0 + 0.1 * np.sin(2 * np.pi * xc[..., 0])
    prim[..., 1] = 0.3 * np.cos(2 * np.pi * xc[..., 1] / 0.7)
    prim[..., 2] = 0.2 * np.sin(2 * np.pi * xc[..., 0])
    prim[..., 4] = 1.0
    blk = cases.Block(dims, nodes, cases.padded_state(dims, prim))
    faces = [{"type": 201, "bs": 1, "fs": 2}, {"type": 201, "bs": 1, "fs": 1},
             {"type": 201, "bs": 1, "fs": 4}, {"type": 201, "bs": 1, "fs": 3}, 0, 0]
    bc = cases.bc_records([{"dims": dims, "faces": faces}])
    case = _viscous(cases.Case("2d", abi.make_cfg(limiter="vanalbada"), [blk], bc))
    _residual_check(perturb(case, seed=5))


def test_viscous_taylor_green_steps():
    case = cases.taylor_green(n=16, nbox=2, precision=abi.PREC_FP64)
    g, o = GpuSolver(case), OracleSolver(case)
    assert abs(g.compute_dt(1) - o.compute_dt(1)) <= 1e-14 * o.compute_dt(1)
    for _ in range(4):
        dg, tg, ng = g.step()
        do, to, no = o.step()
        assert abs(dg - do) <= 1e-13 * do
        assert np.allclose(ng, no, rtol=1e-9, atol=1e-12 * np.max(no))
    for b in (1, 2):
        assert rel_l2(g.get_state(b), o.get_state(b)) < 1e-12


def test_viscous_fp32_flux_against_fp64():
    c64 = cases.taylor_green(n=24, nbox=2, precision=abi.PREC_FP64)
    c32 = cases.taylor_green(n=24, nbox=2, precision=abi.PREC_FP32_FLUX)
    g64, g32 = GpuSolver(c64), GpuSolver(c32)
    g64.advance(10)
    g32.advance(10)
    for b in (1, 2):
        assert rel_l2(g32.get_state(b), g64.get_state(b)) < 1e-5


def test_viscous_rejects_unsupported_boundaries():
    from mose_amd.lib import MoseError
    case = _viscous(cases.sod(n=16))
    with pytest.raises(MoseError):
        GpuSolver(case)

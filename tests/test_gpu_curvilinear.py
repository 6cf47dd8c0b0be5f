"""GPU parity on curvilinear (body-fitted / distorted) grids: the general-metrics kernels
against the oracle, which computes the reference's metrics (Lib_Metrics.f90) on the CPU.

Bar: residuals within 1e-12 relative L2; multi-step states (global and local time
stepping) within 1e-12; free-stream preservation to round-off.
"""
import numpy as np
import pytest

from mose_amd import abi, cases
from mose_amd.lib import GpuSolver
from oracle.oracle import OracleSolver

from helpers import flux_scale, interior, perturb, rel_l2, skewed_nodes

pytestmark = pytest.mark.gpu

LIMS = ["minmod", "vanleer", "vanalbada", "mc", "superbee"]
RSS = ["HLLC", "AUSM+", "HLLE", "SLAU", "SLAU2"]
INFLOW = {"type": 405, "mach": 1.8, "T": 1.05, "p": 0.7, "alpha": 0.1, "beta": -0.05, "rel_fac": 0.8}


def skewed_case(dims, faces, limiter="vanleer", riemann="HLLC", time_accurate=True, mach=1.5, seed=3,
                precision=abi.PREC_FP64):
    nodes = skewed_nodes(dims)
    prim = np.zeros(tuple(dims[::-1]) + (5,))
    prim[..., 0] = 1.0
    prim[..., 1] = mach
    prim[..., 2] = 0.2
    prim[..., 4] = 1.0 / 1.4
    blk = cases.Block(tuple(dims), nodes, cases.padded_state(dims, prim))
    bc = cases.bc_records([{"dims": tuple(dims), "faces": faces}])
    cfg = abi.make_cfg(1.4, 1.0 / 1.4, 0.5, limiter, riemann, 3, time_accurate, precision=precision)
    return perturb(cases.Case("skewed", cfg, [blk], bc), amp=0.03, seed=seed)


def _residual_check(case, tol=1e-12):
    g, o = GpuSolver(case), OracleSolver(case)
    for b in range(1, len(case.blocks) + 1):
        Rg, Ro = g.residual(b), o.residual(b)
        err = rel_l2(Rg, Ro)
        assert err < tol, "block %d: rel L2 %.3e" % (b, err)
        assert np.max(np.abs(interior(Rg) - interior(Ro))) < 1e-13 * flux_scale(case.blocks[b - 1].P)
    g.close()


@pytest.mark.parametrize("lim", LIMS)
@pytest.mark.parametrize("rs", RSS)
def test_residual_skewed_box(lim, rs):
    case = skewed_case((14, 11, 8), [INFLOW, 400, 300, 300, 400, 0], limiter=lim, riemann=rs)
    _residual_check(case)


def test_residual_skewed_two_block_connection():
    case = skewed_case((16, 10, 7), [INFLOW, 400, 300, 300, 400, 400])
    split = cases.split_block_case(case, axis=0, parts=2)
    _residual_check(split)


def test_freestream_preserved_on_skewed_grid():
    dims = (12, 9, 7)
    nodes = skewed_nodes(dims)
    prim = np.zeros(dims[::-1] + (5,))
    prim[..., 0] = 1.2
    prim[..., 1:4] = [0.6, -0.3, 0.2]
    prim[..., 4] = 0.9
    blk = cases.Block(dims, nodes, cases.padded_state(dims, prim))
    bc = cases.bc_records([{"dims": dims, "faces": [400] * 6}])
    g = GpuSolver(cases.Case("fs", abi.make_cfg(limiter="vanleer"), [blk], bc))
    assert np.max(np.abs(interior(g.residual(1)))) < 1e-13


def test_bump_channel_local_dt_steps():
    """Body-fitted 2-D bump channel (405 inflow, 400 outflow, 300 walls), van Albada + HLLC,
    local time stepping: dt field and states after several steps."""
    case = cases.bump_channel(ni=48, nj=24)
    g, o = GpuSolver(case), OracleSolver(case)
    dg, do = g.compute_dt(1), o.compute_dt(1)
    assert abs(dg - do) <= 1e-14 * do
    for _ in range(6):
        _, _, ng = g.step()
        _, _, no = o.step()
        assert np.allclose(ng, no, rtol=1e-9, atol=1e-12 * np.max(no))
    err = rel_l2(g.get_state(1), o.get_state(1))
    assert err < 1e-12, err


def test_skewed_global_dt_steps_fp32_flux():
    c64 = skewed_case((16, 12, 8), [INFLOW, 400, 300, 300, 400, 0])
    c32 = skewed_case((16, 12, 8), [INFLOW, 400, 300, 300, 400, 0], precision=abi.PREC_FP32_FLUX)
    o = OracleSolver(c64)
    g64, g32 = GpuSolver(c64), GpuSolver(c32)
    for _ in range(4):
        o.step()
    g64.advance(4)
    g32.advance(4)
    assert rel_l2(g64.get_state(1), o.get_state(1)) < 1e-12
    assert rel_l2(g32.get_state(1), g64.get_state(1)) < 1e-5

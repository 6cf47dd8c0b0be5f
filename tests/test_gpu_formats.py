"""The reference's own input files (test/1D/Sod79 x1, test/2D/euler/prandtl-meyer) run on
the GPU path against the oracle: residual, steps, and the node / inflow adjoints through the
'normal' inflow direction (angles from the boundary face normals)."""
import os

import numpy as np
import pytest

from mose_amd import abi, formats
from mose_amd.lib import GpuSolver
from oracle.oracle import OracleSolver

from helpers import perturb, rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PM_IC, PM_BC = os.path.join(GOLD, "prandtl_meyer_ic.tec.gz"), os.path.join(GOLD, "prandtl_meyer_bc.txt.gz")
SOD_IC, SOD_BC = os.path.join(GOLD, "sod79_ic_x1.tec"), os.path.join(GOLD, "sod79_bc_x1.txt")


def _pm_case(limiter="vanleer", riemann="HLLC", amp=0.0):
    (b,) = formats.read_tecplot(PM_IC)
    R = float(b["prim"][0, 0, 0, 4] / b["prim"][0, 0, 0, 0])
    cfg = abi.make_cfg(1.4, R, 0.5, limiter, riemann, 3, False)
    case = formats.load_case(PM_IC, PM_BC, cfg)
    return perturb(case, amp=amp, seed=3) if amp else case


@pytest.mark.parametrize("rs", ["HLLC", "AUSM+"])
def test_prandtl_meyer_residual_and_steps(rs):
    case = _pm_case(riemann=rs, amp=0.02)
    g, o = GpuSolver(case), OracleSolver(case)
    assert rel_l2(g.residual(1), o.residual(1)) < 1e-12
    for _ in range(5):
        g.step()
        o.step()
    assert rel_l2(g.get_state(1), o.get_state(1)) < 1e-12
    g.close()


def test_sod79_input_steps():
    cfg = abi.make_cfg(1.4, 1.0 / 1.4, 0.5, "minmod", "HLLC", 3, True)
    case = formats.load_case(SOD_IC, SOD_BC, cfg)
    g, o = GpuSolver(case), OracleSolver(case)
    for _ in range(20):
        _, tg, ng = g.step()
        _, to, no = o.step()
    assert abs(tg - to) <= 1e-13 * to
    assert rel_l2(g.get_state(1), o.get_state(1)) < 1e-12
    g.close()


def test_node_adjoint_with_normal_inflow_direction():
    """The inflow velocity direction follows the face normal: its node adjoint is part of
    the geometry adjoint (central differences of the residual on re-built contexts)."""
    import copy
    # van Albada: the flat walls of this mesh have normals exactly along an axis, where the
    # mirrored velocity slopes are exactly zero and van Leer has only one-sided derivatives
    case = _pm_case(limiter="vanalbada", amp=0.02)
    rng = np.random.default_rng(5)
    d = case.blocks[0].dims
    m = np.zeros(abi.padded_shape(d))
    m[2:-2, 2:-2, 2:-2, :] = 1.0
    Rbar = [rng.standard_normal(abi.padded_shape(d)) * m]
    # the residual of the cells next to the inflow face carries the 'normal' direction
    Rbar[0][2:-2, 2:-2, 2:4, :] *= 50.0
    # nodes of the inflow plane only: the face normals there set the inflow direction.  (Moving
    # all nodes rotates many faces of this Mach 2-3 flow; HLLC's side switch at S* = 0 is a
    # kink, and a rotation of 1e-6 rad shifts u_n by 4e-5, enough for central differences to
    # cross some of those kinks.)
    V = [np.zeros(case.blocks[0].nodes.shape)]
    V[0][:, :, 0, :] = 1e-3 * rng.standard_normal(V[0][:, :, 0, :].shape)
    g = GpuSolver(case)
    _, Xbar, inflow = g.residual_vjp_geometry(Rbar)
    g.close()
    assert inflow[(1, 1)][3] == 0.0 and inflow[(1, 1)][4] == 0.0   # 'normal' angles: no record adjoint

    def J(t):
        c = copy.deepcopy(case)
        c.blocks[0].nodes = np.ascontiguousarray(case.blocks[0].nodes + t * V[0])
        s = GpuSolver(c)
        v = float(np.sum(Rbar[0] * s.residual(1)))
        s.close()
        return v

    h = 1e-5
    fd = (J(h) - J(-h)) / (2 * h)
    adj = float(np.sum(Xbar[0] * V[0]))
    scale = float(np.sum(np.abs(Xbar[0] * V[0])))
    assert abs(adj - fd) <= 1e-6 * scale, (adj, fd, scale)


def test_inflow_face_with_differing_records_is_rejected():
    """One 405 record set per block face: a face whose records differ is refused (MOSE_E_BC)
    rather than run with the first record's values."""
    from mose_amd.lib import MoseError
    case = _pm_case()
    q = np.flatnonzero(case.bc["type"] == 405)[7]
    case.bc["mach"][q] = 2.5
    with pytest.raises(MoseError) as ei:
        GpuSolver(case)
    assert ei.value.code == abi.E_BC

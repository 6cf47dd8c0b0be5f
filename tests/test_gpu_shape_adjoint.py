"""Geometry adjoint of the residual: node adjoints (shape sensitivities through every metric
the residual reads) and inflow-record adjoints, against central differences of the residual
on re-built contexts (GPU) and of the oracle's residual (CPU restatement of the reference).

The residual depends on the nodes through face normals and areas, cell lengths (interior and
the ghost-cell lengths of each bc type: donor copies for 101, mirror copies for 300,
extrapolated node planes for the others) and the symmetry mirror normals; directions are
random, so every one of these paths is exercised.
"""
import copy

import numpy as np
import pytest

from mose_amd import abi, cases
from mose_amd.lib import GpuSolver
from oracle.oracle import OracleSolver

from helpers import perturb, skewed_nodes

pytestmark = pytest.mark.gpu

INFLOW = {"type": 405, "mach": 1.8, "T": 1.05, "p": 0.7, "alpha": 0.1, "beta": -0.05, "rel_fac": 0.8}


def _mask(dims):
    m = np.zeros(abi.padded_shape(dims))
    m[2:-2, 2:-2, 2:-2, :] = 1.0
    return m


def _J(case, Rbar, solver_cls):
    s = solver_cls(case)
    try:
        return sum(float(np.sum(rb * s.residual(b))) for b, rb in enumerate(Rbar, start=1))
    finally:
        if hasattr(s, "close"):
            s.close()


def _with_nodes(case, nodes):
    c2 = copy.deepcopy(case)
    for blk, x in zip(c2.blocks, nodes):
        blk.nodes = np.ascontiguousarray(x)
    return c2


def _shape_check(case, solver_cls=GpuSolver, h=1e-6, tol=1e-6, seed=0, amp=0.01):
    """<Xbar, V> vs central differences of <Rbar, R(X + t V)> with the state fixed."""
    rng = np.random.default_rng(seed)
    dims = [b.dims for b in case.blocks]
    Rbar = [rng.standard_normal(abi.padded_shape(d)) * _mask(d) for d in dims]
    V = [amp * rng.standard_normal(b.nodes.shape) for b in case.blocks]
    g = GpuSolver(case)
    _, Xbar, _ = g.residual_vjp_geometry(Rbar)
    g.close()
    adj = sum(float(np.sum(xb * v)) for xb, v in zip(Xbar, V))
    X0 = [b.nodes for b in case.blocks]
    jp = _J(_with_nodes(case, [x + h * v for x, v in zip(X0, V)]), Rbar, solver_cls)
    jm = _J(_with_nodes(case, [x - h * v for x, v in zip(X0, V)]), Rbar, solver_cls)
    fd = (jp - jm) / (2 * h)
    scale = sum(float(np.sum(np.abs(xb * v))) for xb, v in zip(Xbar, V))
    assert scale > 0.0
    assert abs(adj - fd) <= tol * scale, (adj, fd, scale)
    return adj, fd


def _box(dims, faces, limiter="vanalbada", riemann="HLLC", mach=1.5, seed=3):
    # seeded node jitter: the smooth skewed map alone is mirror-symmetric about x = 1/2, which
    # makes some wall normals exactly orthogonal to an axis; the mirrored velocity slope is
    # then exactly zero, a kink of the van Leer and minmod limiters where only one-sided
    # derivatives exist
    nodes = skewed_nodes(dims) + 0.003 * np.random.default_rng(100 + seed).standard_normal(
        (dims[2] + 1, dims[1] + 1, dims[0] + 1, 3))
    prim = np.zeros(tuple(dims[::-1]) + (5,))
    prim[..., 0] = 1.0
    prim[..., 1] = mach
    prim[..., 2] = 0.2
    prim[..., 4] = 1.0 / 1.4
    blk = cases.Block(tuple(dims), nodes, cases.padded_state(dims, prim))
    bc = cases.bc_records([{"dims": tuple(dims), "faces": faces}])
    cfg = abi.make_cfg(1.4, 1.0 / 1.4, 0.5, limiter, riemann, 3, True)
    return perturb(cases.Case("box", cfg, [blk], bc), amp=0.03, seed=seed)


@pytest.mark.parametrize("rs", ["HLLC", "AUSM+", "HLLE", "SLAU", "SLAU2"])
@pytest.mark.parametrize("lim", ["vanleer", "vanalbada"])
def test_node_adjoint_boundary_types(lim, rs):
    case = _box((9, 7, 5), [INFLOW, 400, 300, 300, 400, 0], limiter=lim, riemann=rs)
    _shape_check(case)


def test_node_adjoint_minmod_symmetry_walls():
    case = _box((8, 6, 5), [400, 400, 300, 300, 300, 300], limiter="minmod", mach=0.6, seed=5)
    _shape_check(case, seed=2)


def test_node_adjoint_connections_and_periodic():
    """Two curvilinear blocks joined by 101 (ghost lengths from the donor), periodic 201
    faces (ghost lengths from extrapolated node planes)."""
    base = _box((10, 6, 5), [{"type": 201, "bs": 1, "fs": 2}, {"type": 201, "bs": 1, "fs": 1}, 300, 400,
                             {"type": 201, "bs": 1, "fs": 6}, {"type": 201, "bs": 1, "fs": 5}],
                mach=0.7, seed=4)
    case = cases.split_block_case(base, axis=0, parts=2)
    _shape_check(case, seed=3)


def test_node_adjoint_bump_against_oracle_residual():
    case = perturb(cases.bump_channel(ni=24, nj=12), amp=0.01, seed=2)
    _shape_check(case, solver_cls=OracleSolver, seed=1)


def test_node_adjoint_bump_heights():
    """Chain rule to the bump control-point heights: dJ/dh = sum Xbar . dX/dh, with the node
    derivative of the parametrisation by central differences of the (linear) grid map."""
    ncontrol = 6
    h0 = np.full(ncontrol, 0.002)
    case = perturb(cases.bump_channel(ni=32, nj=12, ncontrol=ncontrol, heights=h0), amp=0.01, seed=6)
    rng = np.random.default_rng(8)
    Rbar = [rng.standard_normal(abi.padded_shape(b.dims)) * _mask(b.dims) for b in case.blocks]
    g = GpuSolver(case)
    _, Xbar, _ = g.residual_vjp_geometry(Rbar)
    g.close()
    P = case.blocks[0].P
    eps = 1e-6
    for m in (0, 3, 5):
        hp, hm = h0.copy(), h0.copy()
        hp[m] += eps
        hm[m] -= eps
        cp = cases.bump_channel(ni=32, nj=12, ncontrol=ncontrol, heights=hp)
        cm = cases.bump_channel(ni=32, nj=12, ncontrol=ncontrol, heights=hm)
        dX = (cp.blocks[0].nodes - cm.blocks[0].nodes) / (2 * eps)
        adj = float(np.sum(Xbar[0] * dX))
        cp.blocks[0].P = P.copy()
        cm.blocks[0].P = P.copy()
        fd = (_J(cp, Rbar, GpuSolver) - _J(cm, Rbar, GpuSolver)) / (2 * eps)
        scale = float(np.sum(np.abs(Xbar[0] * dX)))
        assert abs(adj - fd) <= 1e-6 * max(scale, 1e-30), (m, adj, fd, scale)


@pytest.mark.parametrize("rs", ["HLLC", "AUSM+"])
def test_inflow_record_adjoint(rs):
    case = _box((9, 7, 5), [INFLOW, 400, 300, 300, 400, 0], riemann=rs)
    rng = np.random.default_rng(11)
    Rbar = [rng.standard_normal(abi.padded_shape(b.dims)) * _mask(b.dims) for b in case.blocks]
    g = GpuSolver(case)
    _, _, inflow = g.residual_vjp_geometry(Rbar)
    g.close()
    grad = inflow[(1, 1)]
    h = 1e-6
    for q, key in enumerate(("mach", "T", "p", "alpha", "beta")):
        cp, cm = copy.deepcopy(case), copy.deepcopy(case)
        sel = (cp.bc["b"] == 1) & (cp.bc["f"] == 1)
        cp.bc[key][sel] += h
        cm.bc[key][sel] -= h
        fd = (_J(cp, Rbar, GpuSolver) - _J(cm, Rbar, GpuSolver)) / (2 * h)
        assert abs(grad[q] - fd) <= 1e-6 * (1.0 + abs(fd)), (key, grad[q], fd)


def test_geometry_adjoint_needs_general_metrics():
    case = perturb(cases.vortex(dims=(6, 5, 4)), seed=1)
    g = GpuSolver(case)
    Rbar = [np.zeros(abi.padded_shape(b.dims)) for b in case.blocks]
    with pytest.raises(Exception):
        g.residual_vjp_geometry(Rbar)
    g.close()


def test_node_adjoint_with_outflow406():
    """Node adjoint of a skewed box with a subsonic 406 outlet (its flux depends on the face
    normal) against central differences of the GPU and oracle residuals."""
    outlet = {"type": 406, "p": 0.62, "rel_fac": 1.0}
    case = _box((9, 7, 5), [400, outlet, 300, 300, 300, 300], mach=0.5)
    _shape_check(case)
    _shape_check(case, solver_cls=OracleSolver)

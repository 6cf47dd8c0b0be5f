"""Host pieces of the bump's steady adjoint gradient (mose_amd/steady_adjoint.py), on the CPU:
the wall-pressure objective's partial derivatives with respect to the primitive state and the
nodes, and the chain rule from node adjoints to control-point heights, each against central
differences of the function it differentiates."""
import numpy as np

from mose_amd import cases, steady_adjoint as sa


def _bump(ni=40, nj=10, nc=6):
    return cases.bump_channel(ni=ni, nj=nj, ncontrol=nc, heights=np.linspace(0.001, 0.003, nc), alpha=0.03)


def test_wall_objective_state_derivative():
    case = _bump()
    rng = np.random.default_rng(1)
    P = case.blocks[0].P * (1.0 + 0.01 * rng.standard_normal(case.blocks[0].P.shape))
    J, dJdP, _ = sa.wall_objective(case, P)
    assert J > 0.0 and np.count_nonzero(dJdP) == np.count_nonzero(dJdP[2, 2, :, 4])
    V = rng.standard_normal(P.shape)
    h = 1e-6
    fd = (sa.wall_objective(case, P + h * V)[0] - sa.wall_objective(case, P - h * V)[0]) / (2 * h)
    assert abs(float(np.sum(dJdP * V)) - fd) <= 1e-9 * abs(fd)


def test_wall_objective_node_derivative():
    case = _bump()
    P = case.blocks[0].P
    _, _, dJdX = sa.wall_objective(case, P)
    rng = np.random.default_rng(2)
    V = 0.01 * rng.standard_normal(case.blocks[0].nodes.shape)
    X0 = case.blocks[0].nodes.copy()
    h = 1e-6
    Js = []
    for s in (1.0, -1.0):
        case.blocks[0].nodes = X0 + s * h * V
        Js.append(sa.wall_objective(case, P)[0])
    case.blocks[0].nodes = X0
    fd = (Js[0] - Js[1]) / (2 * h)
    assert abs(float(np.sum(dJdX * V)) - fd) <= 1e-8 * max(abs(fd), 1e-12)


def test_height_chain_rule_matches_the_grid_map():
    """sum Xbar . dX/dh_m for a random Xbar equals the central difference of sum Xbar . X(h)
    (the nodes are linear in the heights, so the difference is exact to round-off)."""
    case = _bump()
    h0 = case.meta["heights"]
    rng = np.random.default_rng(3)
    Xbar = rng.standard_normal(case.blocks[0].nodes.shape)
    g = sa.height_gradient_of_nodes(case, Xbar)
    eps = 1e-4
    for m in range(len(h0)):
        hp, hm = h0.copy(), h0.copy()
        hp[m] += eps
        hm[m] -= eps
        Xp = cases.bump_channel(ni=40, nj=10, ncontrol=len(h0), heights=hp, alpha=0.03).blocks[0].nodes
        Xm = cases.bump_channel(ni=40, nj=10, ncontrol=len(h0), heights=hm, alpha=0.03).blocks[0].nodes
        fd = float(np.sum(Xbar * (Xp - Xm))) / (2 * eps)
        assert abs(g[m] - fd) <= 1e-9 * max(abs(fd), 1.0), (m, g[m], fd)


def test_wall_faces_are_the_bump():
    case = _bump(ni=40)
    xs = case.meta["xs"]
    ii = sa._wall_faces(case)
    xc = 0.5 * (xs[ii - 1] + xs[ii])
    assert np.all((xc >= 0.0) & (xc <= 1.0)) and len(ii) == 10     # x in [-1.5, 2.5], 40 cells

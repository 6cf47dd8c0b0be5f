"""Pins of the oracle's primitive-variable stage update (integration-variables = prim,
Newstate_Primitive, Lib_Newstate.f90:168-317): its residual conversion is the linearisation
of cons -> prim, so one forward-Euler stage differs from the conservative stage by O(dt^2)
(halving dt quarters the difference), while a planted error in any term of the conversion
leaves an O(dt) difference; a state with zero residual is kept exactly."""
import copy

import numpy as np
import pytest

from mose_amd import abi, cases
from oracle.oracle import OracleSolver

from helpers import interior, perturb


def _with(case, **kw):
    c = copy.deepcopy(case)
    f = {n: getattr(case.cfg, n) for n, _ in abi.MoseCfg._fields_ if n != "reserved"}
    f.update(kw)
    c.cfg = abi.MoseCfg(**f)
    return c


def _one_step_diff(case, cfl):
    out = []
    for prim in (0, 1):
        o = OracleSolver(_with(case, prim_vars=prim, cfl=cfl, n_rk=1))
        o.step()
        out.append(interior(o.get_state(1)))
        o.close()
    return np.max(np.abs(out[1] - out[0]), axis=(0, 1, 2)), out


@pytest.mark.parametrize("time_accurate", [1, 0])
def test_prim_stage_is_second_order_close_to_conservative(time_accurate):
    case = perturb(cases.vortex(dims=(16, 16, 4)), amp=0.05, seed=7)
    case = _with(case, time_accurate=time_accurate)
    d1, (pc, pp) = _one_step_diff(case, 0.4)
    d2, _ = _one_step_diff(case, 0.2)
    P0 = interior(case.blocks[0].P)
    change = np.max(np.abs(pc - P0), axis=(0, 1, 2))
    # the stage moves every variable; density is a primitive and a conserved variable, so its
    # two updates are the same; the others agree to second order in dt
    assert np.all(change > 1e-3)
    assert d1[0] < 1e-14
    ratio = d1[1:] / d2[1:]
    assert np.all((ratio > 3.5) & (ratio < 4.6)), ratio


def test_uniform_state_kept():
    dims = (6, 5, 4)
    prim = np.zeros(dims[::-1] + (5,))
    prim[..., 0], prim[..., 1], prim[..., 2], prim[..., 4] = 1.3, 0.4, -0.2, 0.9
    blk = cases.Block(dims, cases.cartesian_nodes(dims, (0, 0, 0), (1, 1, 1)), cases.padded_state(dims, prim))
    bc = cases.bc_records([{"dims": dims, "faces": [400] * 6}])
    case = cases.Case("uniform", abi.make_cfg(prim_vars=True, irs_beta=0.5, time_accurate=False), [blk], bc)
    o = OracleSolver(case)
    for _ in range(3):
        o.step()
    assert np.allclose(interior(o.get_state(1)), prim, rtol=1e-14, atol=1e-15)


def test_prim_with_irs_agrees_with_cons_with_irs_to_first_order_in_the_amplitude():
    """With IRS the smoothing acts on the primitive residual, which does not commute with the
    cell-wise cons -> prim linearisation, so the two updates differ at first order in dt, by
    a term proportional to the variation of the linearisation over the stencil: on a uniform
    state with a perturbation of amplitude e the difference relative to the update is O(e)
    (halving e halves it), whereas a wrong term in the conversion would leave an O(1)
    relative difference."""
    rel = []
    for amp in (0.02, 0.01):
        case = cases.vortex(dims=(16, 16, 4))
        P = case.blocks[0].P
        P[...] = np.array([1.0, 0.3, 0.2, 0.1, 1.0 / 1.4])
        case = _with(perturb(case, amp=amp, seed=3), irs_beta=0.5, time_accurate=0)
        d, (pc, pp) = _one_step_diff(case, 0.4)
        change = np.max(np.abs(pc - interior(case.blocks[0].P)), axis=(0, 1, 2))
        assert d[0] < 1e-14
        rel.append(d[1:] / change[1:])
    assert np.all(rel[0] < 0.1), rel[0]
    ratio = rel[0] / rel[1]
    assert np.all((ratio > 1.7) & (ratio < 2.3)), ratio

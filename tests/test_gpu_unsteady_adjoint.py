"""Adjoint of the explicit time integration (fixed dt): d KE(t_K) / d mu_ref per block and
d KE(t_K) / d P0 on small viscous Taylor-Green boxes, against central differences of the
same fixed-step march (the Reynolds-number sensitivity of the batched TGV case:
d/dRe = -mu_ref^2 d/dmu_ref)."""
import copy

import numpy as np
import pytest

from mose_amd import abi, cases
from mose_amd.lib import GpuSolver
from mose_amd.unsteady import FixedStepMarch, kinetic_energy, kinetic_energy_ubar

from helpers import perturb

pytestmark = pytest.mark.gpu


def _case(mu=0.04):
    """Two 6^3 boxes with a seeded perturbation: the exact Taylor-Green field has uniform
    density and symmetric velocities, so many limited slopes are exactly zero, where the van
    Leer limiter has only one-sided derivatives."""
    case = cases.taylor_green(n=6, nbox=2, mach=0.3, precision=abi.PREC_FP64)
    case.meta["viscous"] = [(b, [mu * (1 + 0.5 * (b - 1))] + list(p[1:])) for b, p in case.meta["viscous"]]
    return perturb(case, amp=0.02, seed=4)


def _march(case, dt, nsteps, P0):
    g = GpuSolver(case)
    try:
        m = FixedStepMarch(g, case, dt)
        P = m.run(P0, nsteps)
        return kinetic_energy(case, P)
    finally:
        g.close()


def test_kinetic_energy_sensitivities():
    case = _case()
    nsteps = 8
    P0 = [abi.interior(b.P).copy() for b in case.blocks]
    g = GpuSolver(case)
    dt = 0.5 * g.compute_dt()
    m = FixedStepMarch(g, case, dt)
    P = m.run(P0, nsteps)
    ke = kinetic_energy(case, P)
    P0b, mub = m.reverse(kinetic_energy_ubar(case, P, case.cfg.gamma))
    g.close()
    assert ke > 0.0
    # d KE / d mu_ref of each block
    for b in range(1, len(case.blocks) + 1):
        mu = case.meta["viscous"][b - 1][1][0]
        h = 1e-5 * mu

        def shifted(dmu):
            c = copy.deepcopy(case)
            c.meta["viscous"] = [(bb, [p[0] + (dmu if bb == b else 0.0)] + list(p[1:])) for bb, p in c.meta["viscous"]]
            return c

        fd = (_march(shifted(h), dt, nsteps, P0) - _march(shifted(-h), dt, nsteps, P0)) / (2 * h)
        assert mub[b - 1] < 0.0                      # more viscosity, less kinetic energy
        assert abs(mub[b - 1] - fd) <= 1e-6 * abs(fd), (b, mub[b - 1], fd)
    # directional derivative with respect to the initial state
    rng = np.random.default_rng(3)
    V = [0.01 * rng.standard_normal(p.shape) * np.abs(p) for p in P0]
    adj = sum(float(np.sum(pb * v)) for pb, v in zip(P0b, V))
    t = 1e-6
    fd = (_march(case, dt, nsteps, [p + t * v for p, v in zip(P0, V)]) -
          _march(case, dt, nsteps, [p - t * v for p, v in zip(P0, V)])) / (2 * t)
    scale = sum(float(np.sum(np.abs(pb * v))) for pb, v in zip(P0b, V))
    assert abs(adj - fd) <= 1e-6 * scale, (adj, fd, scale)

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


def test_device_march_adjoint_equals_host_tape():
    """mose_unsteady_adjoint_ke (the march, its checkpoints and the reverse sweep on the device)
    against the host-taped sweep above: final kinetic energy, dKE/dmu_ref per block and dKE/dP0."""
    case = _case()
    nsteps = 8
    P0 = [abi.interior(b.P).copy() for b in case.blocks]
    g = GpuSolver(case)
    dt = 0.5 * g.compute_dt()
    ke_d, mub_d, P0b_d = g.march_kinetic_energy(nsteps, dt, stride=3)
    m = FixedStepMarch(g, case, dt)
    P = m.run(P0, nsteps)
    ke_h = kinetic_energy(case, P)
    P0b_h, mub_h = m.reverse(kinetic_energy_ubar(case, P, case.cfg.gamma))
    g.close()
    print("KE device %.15e host %.15e; mu_bar device %s host %s" % (np.sum(ke_d), ke_h, mub_d, mub_h))
    assert abs(np.sum(ke_d) - ke_h) <= 1e-13 * ke_h
    np.testing.assert_allclose(mub_d, mub_h, rtol=1e-11)
    for a, r in zip(P0b_d, P0b_h):
        assert np.linalg.norm(abi.interior(a) - r) <= 1e-11 * np.linalg.norm(r)


def test_device_march_adjoint_checkpoint_stride_invariance():
    """The segments recompute their steps from the checkpoints bit for bit: every stride gives
    the same gradients, and the context's state is left as it was."""
    case = _case()
    g = GpuSolver(case)
    dt = 0.5 * g.compute_dt()
    before = [g.get_state(b) for b in (1, 2)]
    runs = [g.march_kinetic_energy(7, dt, stride=s) for s in (1, 3, 7, 0)]
    after = [g.get_state(b) for b in (1, 2)]
    g.close()
    for ke, mub, P0b in runs[1:]:
        assert np.array_equal(ke, runs[0][0]) and np.array_equal(mub, runs[0][1])
        for a, r in zip(P0b, runs[0][2]):
            assert np.array_equal(a, r)
    for a, r in zip(after, before):
        assert np.array_equal(a, r)


@pytest.mark.parametrize("n,nsteps,stride", [(12, 16, 5), (36, 12, 4)])
def test_device_march_adjoint_reynolds_against_central_differences(n, nsteps, stride):
    """dKE/dmu_ref of each of two n^3 boxes over nsteps steps (checkpoint stride) against central
    differences of the device march; d/dRe = -mu_ref^2 d/dmu_ref.  (36^3: several patches of
    the viscous adjoint's interior-face kernel and two k chunks.)"""
    case = cases.taylor_green(n=n, nbox=2, mach=0.2, precision=abi.PREC_FP64)
    case.meta["viscous"] = [(b, [0.02 * b] + list(p[1:])) for b, p in case.meta["viscous"]]
    case = perturb(case, amp=0.01, seed=7)
    g = GpuSolver(case)
    dt = 0.5 * g.compute_dt()
    ke, mub, _ = g.march_kinetic_energy(nsteps, dt, stride=stride)
    g.close()
    for b in (1, 2):
        mu = case.meta["viscous"][b - 1][1][0]
        h = 1e-6 * mu
        kes = []
        for sgn in (1.0, -1.0):
            c = copy.deepcopy(case)
            c.meta["viscous"] = [(bb, [p[0] + (sgn * h if bb == b else 0.0)] + list(p[1:])) for bb, p in c.meta["viscous"]]
            gg = GpuSolver(c)
            kes.append(float(np.sum(gg.march_kinetic_energy(nsteps, dt, adjoint=False))))
            gg.close()
        fd = (kes[0] - kes[1]) / (2 * h)
        print("box %d: dKE/dmu_ref adjoint %.10e fd %.10e; dKE/dRe %.6e" % (b, mub[b - 1], fd, -mu * mu * mub[b - 1]))
        assert mub[b - 1] < 0.0
        assert abs(mub[b - 1] - fd) <= 1e-6 * abs(fd)

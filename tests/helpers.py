"""Shared helpers of the parity tests."""
import numpy as np

from mose_amd import abi


def interior(a):
    return abi.interior(a)


def rel_l2(a, b):
    """||a - b||_2 / ||b||_2 over every interior cell and variable."""
    a = interior(a).ravel()
    b = interior(b).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def max_rel(a, b, scale):
    """max |a - b| / scale over interior cells and variables."""
    return float(np.max(np.abs(interior(a) - interior(b))) / scale)


def flux_scale(P):
    """A face-flux magnitude of the state: max over cells of rho |u|^2 + p, rho a."""
    Pi = interior(P)
    rho, p = Pi[..., 0], Pi[..., 4]
    q2 = Pi[..., 1] ** 2 + Pi[..., 2] ** 2 + Pi[..., 3] ** 2
    return float(np.max(rho * q2 + p))


def perturb(case, amp=0.02, seed=0):
    """Add a seeded smooth-random perturbation to every block's interior primitive state so
    that the residual is non-trivial in every direction and variable."""
    rng = np.random.default_rng(seed)
    for blk in case.blocks:
        Pi = blk.P[2:-2, 2:-2, 2:-2, :]
        noise = rng.standard_normal(Pi.shape)
        Pi[..., 0] *= 1.0 + amp * noise[..., 0]
        Pi[..., 1:4] += amp * noise[..., 1:4]
        Pi[..., 4] *= 1.0 + amp * noise[..., 4]
    return case


def skewed_nodes(dims, amp=0.08, hi=(1.0, 0.8, 0.6)):
    """Smoothly distorted (curvilinear) grid of a box: node (3, 0:ni, 0:nj, 0:nk) array in
    numpy (k, j, i, 3) order."""
    from mose_amd import cases
    nodes = cases.cartesian_nodes(dims, (0, 0, 0), hi)
    X, Y, Z = nodes[..., 0].copy(), nodes[..., 1].copy(), nodes[..., 2].copy()
    nodes[..., 0] = X + amp * np.sin(2 * np.pi * Y) * np.sin(np.pi * Z)
    nodes[..., 1] = Y + amp * np.sin(2 * np.pi * X) * np.cos(np.pi * Z)
    nodes[..., 2] = Z + 0.5 * amp * np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)
    return np.ascontiguousarray(nodes)

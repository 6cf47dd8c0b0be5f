"""The Fortran host driver (ISO_C_BINDING over the C ABI) on the GPU: Sod to t = 0.2.

The program mose_amd/fortran/mose_run builds the Sod79 case itself, advances it with
Explicit_Step and writes x, rho, u, p, e per cell.  Its profile must equal the Python
driver's run of the same case through the same library bit for bit, and pass the
reference's probe rule against test/1D/Sod79/exact.txt.
"""
import os
import subprocess

import numpy as np
import pytest

from mose_amd import cases
from mose_amd.abi import interior
from mose_amd.lib import GpuSolver
from oracle import exact_riemann

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "mose_amd", "fortran", "mose_run")


@pytest.mark.gpu
def test_fortran_driver_sod(tmp_path):
    assert os.path.exists(EXE), "Fortran driver not built: run __graft_entry__.build()"
    out = tmp_path / "sod.txt"
    r = subprocess.run([EXE, str(out), "1000"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    prof = np.loadtxt(out)
    assert prof.shape == (1000, 5)
    t_f = float(r.stdout.split("time =")[1].split()[0])

    case = cases.sod(n=1000)
    g = GpuSolver(case)
    t = 0.0
    while t < 0.2:
        _, t, _ = g.step()
    P = interior(g.get_state(1))[0, 0]
    g.close()
    assert t_f == t
    np.testing.assert_array_equal(prof[:, 1], P[:, 0])
    np.testing.assert_array_equal(prof[:, 2], P[:, 1])
    np.testing.assert_array_equal(prof[:, 3], P[:, 4])

    table = np.loadtxt(os.path.join(HERE, "golden", "sod79_exact.txt"))
    fails = exact_riemann.probe_check(prof[:, 0], prof[:, 1], prof[:, 2], prof[:, 3], prof[:, 4],
                                      table, threshold=1.0)
    assert not fails, fails

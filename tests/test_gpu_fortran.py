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


CASE_EXE = os.path.join(os.path.dirname(HERE), "mose_amd", "fortran", "mose_case")
GOLD = os.path.join(HERE, "golden")


def _unpack(src, dst):
    import gzip
    opener = gzip.open if src.endswith(".gz") else open
    with opener(src, "rt") as f, open(dst, "w") as g:
        g.write(f.read())
    return str(dst)


def _run_case(tmp_path, name, ini, extra=()):
    ic = _unpack(os.path.join(GOLD, name + "_ic" + (".tec.gz" if name != "sod79" else "_x1.tec")), tmp_path / "ic.tec")
    bc = _unpack(os.path.join(GOLD, name + "_bc" + (".txt.gz" if name != "sod79" else "_x1.txt")), tmp_path / "bc.txt")
    (tmp_path / "input.ini").write_text(ini)
    out = tmp_path / "field.tec"
    assert os.path.exists(CASE_EXE), "Fortran case driver not built: run __graft_entry__.build()"
    r = subprocess.run([CASE_EXE, str(tmp_path / "input.ini"), ic, bc, str(out)] + list(extra), capture_output=True,
                       text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    return ic, bc, out, r.stdout


@pytest.mark.gpu
def test_fortran_case_driver_prandtl_meyer(tmp_path):
    """mose_case runs the reference's Prandtl-Meyer INPUT files with the numerics of its
    input.ini (HLLE + van Leer, RK3, local dt, CFL 1.0, irs-beta 0.5; 300 iterations here): its
    output field.tec equals the Python driver's run of the same case through the library."""
    from mose_amd import abi, formats
    ini = """[MOSE-Parameters]
iter-threshold = 300
[MOSE-Numerics]
cfl = 1.0
time-scheme = RK3
time-accurate = false
irs-beta = 0.5
space-reconstruction = MUSCL
flux-limiter = vanleer
riemann-solver = HLLE ; the shipped solver
integration-variables = cons
"""
    # the input starts at T = 1 in its gas: R = p / rho of any cell, passed with gamma (its
    # 405 inflow records need the gas constant)
    (b0,) = formats.read_tecplot(os.path.join(GOLD, "prandtl_meyer_ic.tec.gz"))
    R = float(b0["prim"][0, 0, 0, 4] / b0["prim"][0, 0, 0, 0])
    ic, bc, out, _ = _run_case(tmp_path, "prandtl_meyer", ini, extra=("1.4", repr(R)))
    (fb,) = formats.read_tecplot(out)
    cfg = abi.make_cfg(1.4, R, 1.0, "vanleer", "HLLE", 3, False, irs_beta=0.5)
    g = GpuSolver(formats.load_case(ic, bc, cfg))
    g.advance(300)
    P = interior(g.get_state(1))
    g.close()
    assert np.array_equal(fb["prim"], P)
    assert np.array_equal(fb["nodes"], b0["nodes"])


@pytest.mark.gpu
def test_fortran_case_driver_sod79_shipped(tmp_path):
    """The Sod79 input with its shipped numerics (MC + SLAU, RK3, CFL 0.9, time-accurate to
    t = 0.2) through mose_case (gamma and R = 1/1.4 given: the input starts at T = 1.4) equals
    the Python driver's run."""
    from mose_amd import abi, formats
    ini = """[MOSE-Parameters]
time-threshold = 0.2
[MOSE-Numerics]
cfl = 0.9
time-scheme = RK3
time-accurate = true
space-reconstruction = MUSCL
flux-limiter = mc
riemann-solver = SLAU
"""
    ic, bc, out, stdout = _run_case(tmp_path, "sod79", ini, extra=("1.4", repr(1.0 / 1.4)))
    (fb,) = formats.read_tecplot(out)
    cfg = abi.make_cfg(1.4, 1.0 / 1.4, 0.9, "mc", "SLAU", 3, True)
    g = GpuSolver(formats.load_case(ic, bc, cfg))
    t = 0.0
    while t < 0.2:
        _, t, _ = g.step()
    P = interior(g.get_state(1))
    g.close()
    assert np.array_equal(fb["prim"], P)
    assert abs(float(stdout.split("time =")[1].split()[0]) - t) <= 1e-15

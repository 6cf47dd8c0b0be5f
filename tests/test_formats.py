"""The reference's on-disk inputs (Tecplot ASCII ic.tec, bc.txt) read, written and run by
the oracle.  Fixtures: the reference's own input files of test/1D/Sod79 (x1) and
test/2D/euler/prandtl-meyer (tests/golden, copied by make_format_fixtures.py)."""
import os

import numpy as np
import pytest

from mose_amd import abi, cases, formats
from oracle.oracle import OracleSolver

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SOD_IC, SOD_BC = os.path.join(GOLD, "sod79_ic_x1.tec"), os.path.join(GOLD, "sod79_bc_x1.txt")
PM_IC, PM_BC = os.path.join(GOLD, "prandtl_meyer_ic.tec.gz"), os.path.join(GOLD, "prandtl_meyer_bc.txt.gz")


def test_sod79_input_files():
    (b,) = formats.read_tecplot(SOD_IC)
    assert b["dims"] == (100, 1, 1)
    x = b["nodes"][0, 0, :, 0]
    np.testing.assert_allclose(x, np.linspace(0.0, 1.0, 101), atol=1e-12)
    P = b["prim"][0, 0]
    np.testing.assert_allclose(P[:50, 4], 1.0)
    np.testing.assert_allclose(P[50:, 4], 0.1)
    assert np.all(P[:, 1:4] == 0.0)
    bc = formats.read_bc(SOD_BC, [b["dims"]])
    assert len(bc) == 402 == cases.expected_bc_count([b["dims"]])
    for f in range(1, 7):      # the 1-D pattern: x and y faces extrapolate, z faces no flux
        assert set(bc["type"][bc["f"] == f]) == ({400} if f <= 4 else {0})


def test_prandtl_meyer_input_files_and_normal_token():
    (b,) = formats.read_tecplot(PM_IC)
    assert b["dims"] == (50, 50, 1)
    bc = formats.read_bc(PM_BC, [b["dims"]])
    inflow = bc[bc["type"] == 405]
    assert len(inflow) == 50 and np.all(inflow["f"] == 1)
    assert np.all(inflow["alpha"] == formats.NORMAL) and np.all(inflow["beta"] == formats.NORMAL)
    np.testing.assert_allclose(inflow["mach"], 2.0)
    # the initial state is the inflow state: T = p / (rho R) = 1 with the Air gas constant
    R = b["prim"][..., 4] / b["prim"][..., 0]
    np.testing.assert_allclose(R, R.flat[0], rtol=1e-12)
    assert 286.0 < R.flat[0] < 288.0


def test_round_trip(tmp_path):
    blocks = formats.read_tecplot(PM_IC)
    bc = formats.read_bc(PM_BC)
    formats.write_tecplot(tmp_path / "ic.tec", [(b["nodes"], b["prim"]) for b in blocks], time=2.5)
    formats.write_bc(tmp_path / "bc.txt", bc)
    again = formats.read_tecplot(tmp_path / "ic.tec")
    assert again[0]["time"] == 2.5
    np.testing.assert_array_equal(again[0]["nodes"], blocks[0]["nodes"])
    np.testing.assert_array_equal(again[0]["prim"], blocks[0]["prim"])
    np.testing.assert_array_equal(formats.read_bc(tmp_path / "bc.txt"), bc)


def test_multiblock_round_trip(tmp_path):
    c = cases.multiblock_chain(nblocks=3, seed=4, lo=5, hi=8)
    formats.write_tecplot(tmp_path / "ic.tec", [(b.nodes, abi.interior(b.P)) for b in c.blocks])
    formats.write_bc(tmp_path / "bc.txt", c.bc)
    back = formats.load_case(tmp_path / "ic.tec", tmp_path / "bc.txt", c.cfg)
    assert [tuple(b.dims) for b in back.blocks] == [tuple(int(x) for x in b.dims) for b in c.blocks]
    for a, b in zip(back.blocks, c.blocks):
        np.testing.assert_array_equal(abi.interior(a.P), abi.interior(b.P))
    np.testing.assert_array_equal(back.bc, c.bc)


def test_bc_count_and_type_errors(tmp_path):
    lines = open(SOD_BC).read().split("\n")
    (tmp_path / "short.txt").write_text("\n".join(lines[:-3]) + "\n")
    with pytest.raises(ValueError, match="Boundary conditions number"):
        formats.read_bc(tmp_path / "short.txt", [(100, 1, 1)])
    (tmp_path / "wall.txt").write_text("       1       1       1       1       3     302\n  1.0, 0.0, 0.0\n")
    with pytest.raises(ValueError, match="type 302"):
        formats.read_bc(tmp_path / "wall.txt")


def test_oracle_runs_the_prandtl_meyer_input():
    """The reference's Prandtl-Meyer input (405 inflow with the 'normal' direction, 400,
    300 and 0 faces, body-fitted expansion corner) through the oracle."""
    (b,) = formats.read_tecplot(PM_IC)
    R = float(b["prim"][0, 0, 0, 4] / b["prim"][0, 0, 0, 0])
    cfg = abi.make_cfg(1.4, R, 0.5, "vanleer", "HLLC", 3, False)
    case = formats.load_case(PM_IC, PM_BC, cfg)
    o = OracleSolver(case)
    res = o.residual(1)
    assert np.all(np.isfinite(res))
    for _ in range(3):
        o.step()
    P = abi.interior(o.get_state(1))
    assert np.all(np.isfinite(P)) and np.all(P[..., 4] > 0)

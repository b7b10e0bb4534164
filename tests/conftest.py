import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU checks")


def load_golden(name):
    return np.load(GOLDEN / name, allow_pickle=False)


@pytest.fixture(scope="session")
def ref_tables():
    d = load_golden("fn_crek.npz")
    return d["xr"], d["xz"], (d["apd1x"], d["apd1z"], d["apd2x"], d["apd2z"])


def case_geometry(d):
    keys = ("xg", "yg", "zg", "xn", "yn", "zn", "aire", "tdis")
    geo = dict(zip(keys, d["geo"].T))
    geo["zer"] = float(d["zer"])
    return geo


def write_synthetic_case(work, name):
    """Inputs of the synthetic golden cases (the reference preProcessor and
    Solver read them), with the exact mesh they were made from."""
    from nemoh_amd import io as nio
    from nemoh_amd import mesh as nmesh
    (work / "mesh").mkdir(parents=True)
    (work / "results").mkdir()
    (work / "ID.dat").write_text("1\n.\n")
    (work / "input.txt").write_text("--- Calculation parameters ---\n0\t! Indiq_solver\n")
    if name == "cyl512":
        m = nmesh.cylinder(10.0, 5.0, 32, 8, 8, half=True)
        m.write_dat(work / "cyl.dat")
        body = nio.BodySpec("cyl.dat", m.npoints, m.npanels, nio.six_dof(), nio.six_dof())
        nio.write_nemoh_cal(work / "Nemoh.cal", 1025.0, 9.81, 0.0, [body], 20, 0.1, 2.0, 1, 0.0, 0.0)
        depth = 0.0
    elif name == "barge4000":
        m = nmesh.box(100.0, 20.0, 10.0, 90, 20, 10)
        m.write_dat(work / "barge.dat")
        body = nio.BodySpec("barge.dat", m.npoints, m.npanels, nio.six_dof(), nio.six_dof())
        nio.write_nemoh_cal(work / "Nemoh.cal", 1025.0, 9.81, 50.0, [body], 3, 0.2, 2.0, 3, 0.0, 90.0)
        depth = 50.0
    else:
        m = nmesh.box(100.0, 20.0, 10.0, 18, 4, 2)
        m.write_dat(work / "barge.dat")
        body = nio.BodySpec("barge.dat", m.npoints, m.npanels, nio.six_dof(), nio.six_dof())
        nio.write_nemoh_cal(work / "Nemoh.cal", 1025.0, 9.81, 50.0, [body], 3, 0.2, 2.0, 3, 0.0, 90.0)
        depth = 50.0
    return m, depth

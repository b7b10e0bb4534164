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

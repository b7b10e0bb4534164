"""Outputs computed from the sources on the device: the Kochin function.

The device's Kochin function (COMPUTE_KOCHIN.f90:25-153) of the sources of
a real solve equals the oracle's on the same sources (the oracle is pinned
to the reference build's COMPUTE_KOCHIN), for a y-symmetric case in
infinite depth and a finite-depth case with an off-origin reference point,
and per problem after an all-problems solve.
"""
import numpy as np
import pytest

from conftest import load_golden
from nemoh_amd import core
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _tables():
    t = load_golden("fn_crek.npz")
    return t["xr"], t["xz"], (t["apd1x"], t["apd1z"], t["apd2x"], t["apd2z"])


@pytest.mark.parametrize("name,xeff,yeff", [("Cylinder", 0.0, 0.0), ("barge_fd", 4.0, -3.0)])
def test_kochin_matches_oracle(name, xeff, yeff):
    d = load_golden(f"case_{name}.npz")
    isym, depth, g = int(d["isym"]), float(d["depth"]), float(d["g"])
    c = core.Core(d["vertices"], d["panels"], isym, depth, float(d["rho"]), g, tables=_tables())
    oc = po.Case(d["vertices"], d["panels"], isym, depth=depth)
    thetas = np.linspace(0.0, 2.0 * np.pi, 37)
    omegas = np.unique(d["omega"])
    for w in omegas[[0, -1]]:
        period = 2 * np.pi / w
        cols = np.where(d["omega"] == w)[0]
        f = c.freq_setup(period)
        kwave = f.amh / depth if f.finite_depth else w * w / g
        per_problem, refs = [], []
        for j in cols:
            zigb, zigs, _, _ = c.solve(period, f, d["nvel"][:, j])
            per_problem.append(c.kochin(thetas, xeff, yeff)[:, 0])
            refs.append(po.kochin(oc, kwave, xeff, yeff, thetas, zigb, zigs if isym else None))
        # normwise over the frequency's problems: the cylinder's yaw about its
        # own axis moves no water, so its sources, and Kochin function, are
        # rounding noise on both sides
        scale = max(np.abs(r).max() for r in refs)
        err = max(np.abs(h - r).max() for h, r in zip(per_problem, refs)) / scale
        # (CEP(l+1) - CEP(l)) / ZC cancels when k x (edge length) is small,
        # amplifying the ~1e-16 differences between the device's and the C
        # library's exp/sin/cos by about 1 / (k x edge); for a radiating mode
        # at small k the panel sum cancels as well
        tol = 1e-12 if kwave > 0.05 else 1e-9
        print(f"{name} w={w:.3f} k={kwave:.4f}: device vs oracle {err:.2e}")
        assert err < tol, (name, w, err)
        # all problems of the frequency in one solve: same functions per column
        c.solve_frequency(period, d["nvel"][:, cols])
        hall = c.kochin(thetas, xeff, yeff)
        assert hall.shape == (thetas.size, cols.size)
        ref = np.stack(per_problem, axis=1)
        assert np.abs(hall - ref).max() <= 1e-11 * np.abs(ref).max()
    c.close()


def test_kochin_needs_a_solve():
    d = load_golden("case_Cylinder.npz")
    c = core.Core(d["vertices"], d["panels"], int(d["isym"]), 0.0, tables=_tables())
    with pytest.raises(core.NemohError) as e:
        c.kochin(np.array([0.0]), nrhs=1)
    assert e.value.code == core.ERR_ARG
    c.close()

"""Outputs computed from the sources on the device: the Kochin function
and the potential at field points.

The device's Kochin function (COMPUTE_KOCHIN.f90:25-153) and field
potential (COMPUTE_POTENTIAL_DOMAIN.f90:24-82) of the sources of a real
solve equal the oracle's on the same sources (the oracle is pinned to the
reference build's routines), for a y-symmetric case in infinite depth and a
barge at h = 50 m (finite depth, and an infinite-depth frequency at finite
depth), and per problem after an all-problems solve.
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


def _field_points(vertices):
    """Free-surface points (z = 0) on a ring and a line beyond the hull, as
    the free-surface mesh gives them, and submerged points near the hull."""
    v = np.asarray(vertices)
    rx, ry, zmin = np.abs(v[:, 0]).max(), np.abs(v[:, 1]).max(), v[:, 2].min()
    t = np.linspace(0.0, 2.0 * np.pi, 24, endpoint=False)
    ring = np.stack([1.3 * rx * np.cos(t), 1.3 * max(ry, 0.3 * rx) * np.sin(t), 0.0 * t], axis=1)
    line = np.stack([np.linspace(-4 * rx, 4 * rx, 16), np.full(16, 2.5 * max(ry, 1.0)), np.zeros(16)], axis=1)
    sub = np.array([[1.05 * rx, 0.1 * ry, 0.5 * zmin], [0.0, 1.1 * ry + 0.5, 0.3 * zmin],
                    [-2.0 * rx, -ry, 1.2 * zmin], [0.3 * rx, 0.2 * ry, 1.05 * zmin]])
    return np.concatenate([ring, line, sub])


@pytest.mark.parametrize("name", ["Cylinder", "barge_fd"])
def test_potential_matches_oracle(name):
    d = load_golden(f"case_{name}.npz")
    isym, depth, g = int(d["isym"]), float(d["depth"]), float(d["g"])
    c = core.Core(d["vertices"], d["panels"], isym, depth, float(d["rho"]), g, tables=_tables())
    oc = po.Case(d["vertices"], d["panels"], isym, depth=depth)
    pts = _field_points(d["vertices"])
    for w in np.unique(d["omega"]):
        period = 2 * np.pi / w
        cols = np.where(d["omega"] == w)[0]
        f = c.freq_setup(period)
        per_problem, refs = [], []
        for j in cols:
            zigb, zigs, _, _ = c.solve(period, f, d["nvel"][:, j])
            print(f"{name} w={w:.3f} problem {j}: device potential at {len(pts)} points", flush=True)
            per_problem.append(c.potential(pts)[:, 0])
            refs.append(po.potential(oc, period, pts, zigb, zigs if isym else None))
        # normwise over the frequency's problems (a mode that moves no water
        # has rounding-noise sources on both sides)
        scale = max(np.abs(r).max() for r in refs)
        err = max(np.abs(p - r).max() for p, r in zip(per_problem, refs)) / scale
        print(f"{name} w={w:.3f} regime {'FD' if f.finite_depth else 'INFD'}: device vs oracle {err:.2e}")
        assert err < 1e-12, (name, w, err)
        c.solve_frequency(period, d["nvel"][:, cols])
        pall = c.potential(pts)
        assert pall.shape == (len(pts), cols.size)
        ref = np.stack(per_problem, axis=1)
        assert np.abs(pall - ref).max() <= 1e-11 * np.abs(ref).max()
    c.close()


def test_potential_at_collocation_points_is_body_potential():
    """Infinite depth: at the collocation points the device's field
    potential is the body potential ZPB of its own solve, to rounding."""
    d = load_golden("case_Cylinder.npz")
    isym = int(d["isym"])
    c = core.Core(d["vertices"], d["panels"], isym, 0.0, float(d["rho"]), float(d["g"]), tables=_tables())
    oc = po.Case(d["vertices"], d["panels"], isym, depth=0.0)
    w = float(d["omega"][-1])
    f = c.freq_setup(2 * np.pi / w)
    _, _, zpb, _ = c.solve(2 * np.pi / w, f, d["nvel"][:, -1])
    geo = oc.geo
    phi = c.potential(np.stack([geo["xg"], geo["yg"], geo["zg"]], axis=1))[:, 0]
    assert np.abs(phi - zpb).max() <= 1e-12 * np.abs(zpb).max()
    c.close()


def test_potential_errors():
    d = load_golden("case_Cylinder.npz")
    c = core.Core(d["vertices"], d["panels"], int(d["isym"]), 0.0, tables=_tables())
    with pytest.raises(core.NemohError) as e:
        c.potential(np.zeros((1, 3)), nrhs=1)
    assert e.value.code == core.ERR_ARG  # no solve yet
    w = float(d["omega"][-1])
    f = c.freq_setup(2 * np.pi / w)
    c.solve(2 * np.pi / w, f, d["nvel"][:, 0])
    with pytest.raises(core.NemohError) as e:  # above every panel: AKZ > -1.5e-6
        c.potential(np.array([[0.0, 0.0, 0.0], [20.0, 0.0, 5.0]]))
    assert e.value.code == core.ERR_ARG and "AKZ" in str(e.value)
    assert c.potential(np.array([[20.0, 0.0, 0.0]])).shape == (1, 1)  # the handle still works
    c.close()

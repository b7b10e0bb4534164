"""Right-hand sides and force weights against the reference preProcessor
(CPU).

The synthetic golden cases were made by the reference preProcessor from
six-DOF bodies about the origin: cyl512 (half cylinder, infinite depth,
heading 0) and barge_fd (h = 50 m, headings 0/45/90, the last frequency on
the preProcessor's deep-water branch).  The fixture mesh is the
preProcessor's L12.dat, written with about seven significant digits, so
normals and centres agree with the ones it used to about 1e-6 of the body
scale: that is the tolerance (2e-6 of each group's largest value).
"""
import numpy as np
import pytest

from conftest import load_golden
from nemoh_amd import bodycond as bcm
from nemoh_amd import mesh as nmesh

TOL = 2e-6


@pytest.mark.parametrize("name,betas", [("cyl512", [0.0]), ("barge_fd", [0.0, 45.0, 90.0])])
def test_problems_and_weights_match_preprocessor(name, betas):
    d = load_golden(f"case_{name}.npz")
    m = nmesh.Mesh(d["vertices"], d["panels"], int(d["isym"]))
    modes = bcm.six_dof()
    om, kind, nvel = bcm.problems(m, np.unique(d["omega"]), betas, modes, float(d["depth"]), float(d["g"]),
                                  float(d["rho"]))
    assert np.array_equal(om, d["omega"]) and np.array_equal(kind, d["kind"])
    ref = d["nvel"]
    rad = kind == -1.0
    err_rad = np.abs(nvel - ref)[:, rad].max() / np.abs(ref[:, rad]).max()
    err_dif = (np.abs(nvel - ref)[:, ~rad].max(axis=0) / np.abs(ref[:, ~rad]).max(axis=0)).max()
    print(f"{name}: radiation {err_rad:.2e}, diffraction {err_dif:.2e}")
    assert err_rad < TOL and err_dif < TOL
    nds = bcm.integration(m, modes)
    assert np.abs(nds - d["nds"]).max() < TOL * np.abs(d["nds"]).max()


def test_diffraction_closed_forms():
    """Deep water, heading 0, one horizontal panel with normal -z: the
    incident velocity has v_z = -i g k / w e^{kz + ikx}, so NVEL = -v.n =
    v_z, and the incident pressure is rho g e^{kz + ikx}; with k h > 20 the
    finite-depth factors are the deep-water ones."""
    v = np.array([[0, 0, -2.0], [1, 0, -2.0], [1, 1, -2.0], [0, 1, -2.0]])
    m = nmesh.Mesh(v, np.array([[0, 3, 2, 1]], np.int32), 0)
    xm, n, _ = bcm.panel_properties(m)
    assert np.allclose(n[0], [0, 0, -1])
    w, g = 1.3, 9.81
    k = w * w / g
    nv, pr = bcm.diffraction(m, w, 0.0, 0.0, g, 1025.0)
    e = np.exp(k * xm[0, 2]) * np.exp(1j * k * xm[0, 0])
    assert np.allclose(nv[0], -(-1j * g / w * k * e) * -1.0, rtol=1e-14)
    assert np.allclose(pr[0], 1025.0 * g * e, rtol=1e-14)
    nv_h, _ = bcm.diffraction(m, w, 0.0, 1000.0, g, 1025.0)  # k h = 172 > 20
    assert np.allclose(nv_h, nv, rtol=1e-14)

"""Parametric meshes of the benchmark configurations (CPU)."""
import numpy as np

from nemoh_amd import core
from nemoh_amd import mesh as nmesh


def _geometry(m):
    return core.prepare_mesh(m.vertices, m.panels, m.isym)


def test_closed_box_volume_and_outward_normals():
    """Divergence theorem on a closed surface: sum(x n_x A) = sum(y n_y A) =
    sum(z n_z A) = volume, positive only with outward normals."""
    m = nmesh.closed_box(38.0, 8.0, 6.0, 19, 4, 3, center=(1.0, -25.0, -17.0))
    g = _geometry(m)
    for c, n in (("xg", "xn"), ("yg", "yn"), ("zg", "zn")):
        vol = np.sum(g[c] * g[n] * g["aire"])
        assert abs(vol - 38.0 * 8.0 * 6.0) < 1e-9 * 38.0 * 8.0 * 6.0


def test_box_and_cylinder_normals_point_outward():
    barge = _geometry(nmesh.box(100.0, 20.0, 10.0, 18, 4, 2))
    # open at z = 0: the x and y components still close
    assert abs(np.sum(barge["xg"] * barge["xn"] * barge["aire"]) - 100 * 20 * 10) < 1e-6
    cyl = _geometry(nmesh.cylinder(10.0, 5.0, 32, 8, 8, half=False))
    vol = np.sum(cyl["xg"] * cyl["xn"] * cyl["aire"]) + np.sum(cyl["yg"] * cyl["yn"] * cyl["aire"])
    # 2 x the volume of the inscribed 32-gon prism
    prism = 0.5 * 32 * 10.0 ** 2 * np.sin(2 * np.pi / 32) * 5.0
    assert abs(vol - 2 * prism) < 1e-9 * prism


def test_semisub_configuration():
    """Four columns r = 6 m at (+-25, +-25), draft 20 m, pontoons 8 x 6 m at
    the keel: 12,048 panels, all wetted, each pontoon a closed box."""
    m = nmesh.semisub()
    assert m.npanels == 12048 and m.isym == 0
    g = _geometry(m)
    assert g["zg"].max() < 0 and g["zg"].min() > -20.0 - 1e-9
    cols, pon = slice(0, 4 * 1824), slice(4 * 1824, None)
    for sl, vol in ((pon, 4 * 38.0 * 8.0 * 6.0),):
        v = np.sum(g["zg"][sl] * g["zn"][sl] * g["aire"][sl])
        assert abs(v - vol) < 1e-9 * vol
    r = np.hypot(np.abs(g["xg"][cols]) - 25.0, np.abs(g["yg"][cols]) - 25.0)
    assert r.max() <= 6.0 + 1e-9


def test_spheroid_wetted_part_and_volume():
    """Prolate spheroid 40 x 10 m cut at z = 0 for drafts below, at and
    above the axis: no vertex above the free surface, and sum(z n_z A)
    within 3% below the exact submerged volume (the inscribed polyhedron
    is slightly smaller), the integral of the circular-segment areas."""
    import math
    from scipy.integrate import quad
    a, b = 20.0, 5.0
    for df in (0.2, 0.5, 0.8):
        m = nmesh.spheroid(2 * a, 2 * b, df, 24, 16)
        assert m.vertices[:, 2].max() <= 0.0
        g = _geometry(m)
        vol = np.sum(g["zg"] * g["zn"] * g["aire"])
        d = b - df * 2 * b  # height of the axis above the free surface

        def seg(x):
            r = b * math.sqrt(max(0.0, 1 - (x / a) ** 2))
            if d >= r:
                return 0.0
            if d <= -r:
                return math.pi * r * r
            return r * r * math.acos(d / r) - d * math.sqrt(r * r - d * d)

        exact = quad(seg, -a, a, limit=200)[0]
        assert 0.97 * exact < vol < exact, (df, vol, exact)

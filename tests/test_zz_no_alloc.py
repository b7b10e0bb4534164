"""nemoh_core_kochin_noalloc (the Kochin function through a fixed device
array, no device allocation) against nemoh_core_kochin on the same
handle: identical values, two calls in a row with the second smaller, on two
handles sharing the array in turn, and the size limit as an error that
leaves the handle working."""
import numpy as np
import pytest

from conftest import load_golden
from nemoh_amd import core

pytestmark = pytest.mark.gpu


def _tables():
    t = load_golden("fn_crek.npz")
    return t["xr"], t["xz"], (t["apd1x"], t["apd1z"], t["apd2x"], t["apd2z"])


def _solved(name, prec):
    d = load_golden(f"case_{name}.npz")
    c = core.Core(d["vertices"], d["panels"], int(d["isym"]), float(d["depth"]), float(d["rho"]), float(d["g"]),
                  tables=_tables())
    c.set_precision(prec)
    w = d["omega"][0]
    cols = np.where(d["omega"] == w)[0]
    c.solve_frequency(2 * np.pi / w, d["nvel"][:, cols], nds=d["nds"], want_pressure=False)
    return c, cols.size


@pytest.mark.parametrize("name,prec", [("Cylinder", core.FP64), ("barge_fd", core.MIXED)])
def test_kochin_noalloc_equals_kochin(name, prec):
    c, nrhs = _solved(name, prec)
    th = np.linspace(0.0, 2.0 * np.pi, 181)
    for angles, xe, ye in ((th, 0.0, 0.0), (th[5:42], 4.0, -3.0)):  # the second call is smaller
        ref = c.kochin(angles, xe, ye)
        got = c.kochin_noalloc(angles, xe, ye)
        assert got.shape == (angles.size, nrhs)
        assert np.all(np.isfinite(got)) and np.abs(ref).max() > 0
        assert np.array_equal(got, ref)
    c.close()


def test_kochin_noalloc_two_handles_and_limit():
    a, _ = _solved("Cylinder", core.FP64)
    b, _ = _solved("barge_fd", core.FP64)
    th = np.linspace(0.0, np.pi, 64)
    ra, rb = a.kochin(th), b.kochin(th)
    assert np.array_equal(a.kochin_noalloc(th), ra)
    assert np.array_equal(b.kochin_noalloc(th), rb)
    assert np.array_equal(a.kochin_noalloc(th[:7]), ra[:7])
    with pytest.raises(core.NemohError) as e:
        a.kochin_noalloc(np.zeros(4097))
    assert e.value.code == core.ERR_ARG
    assert np.array_equal(a.kochin_noalloc(th), ra)
    a.close()
    b.close()

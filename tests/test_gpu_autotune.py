"""The factorisation-form autotuner (nemoh_core_autotune) and the forms it
chooses between: every form's solve matches the reference Solver, the
tile-band height leaves the factors bit for bit unchanged, and a tuned
handle (and any handle of the same size class that opts in) factorises in
the form the tuner recorded."""
import numpy as np
import pytest

from conftest import load_golden
from nemoh_amd import core
from nemoh_amd import io as nio
from nemoh_amd import mesh as nmesh
from nemoh_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _tables():
    t = load_golden("fn_crek.npz")
    return t["xr"], t["xz"], (t["apd1x"], t["apd1z"], t["apd2x"], t["apd2z"])


def _case(name):
    d = load_golden(f"case_{name}.npz")
    m = nmesh.Mesh(d["vertices"], d["panels"], int(d["isym"]))
    return d, m, nio.table_to_forces(d["forces_table"], d["omega"], d["kind"])


def _normwise(a, ref):
    return float((np.abs(a - ref).max(axis=0) / np.abs(ref).max(axis=0)).max())


@pytest.mark.parametrize("prec,tol", [(core.FP64, 1e-10), (core.MIXED, 1e-6)])
@pytest.mark.parametrize("name", ["barge_fd", "cyl512"])
def test_autotune_then_solve_matches_reference(name, prec, tol):
    d, m, ref = _case(name)
    s = Solver(m, float(d["depth"]), float(d["rho"]), float(d["g"]), precision=prec, tables=_tables())
    entries, k = s.core.autotune(reps=2)
    print(name, prec, [(r, b, round(ms, 3)) for r, b, ms in entries], "chosen", entries[k][:2], flush=True)
    assert len(entries) == core.TUNE_CANDIDATES and all(ms > 0 for _, _, ms in entries)
    assert {(r, b) for r, b, _ in entries} == {(r, b) for r in (64, 128) for b in (16, 32, 64)}
    assert k == int(np.argmin([ms for _, _, ms in entries]))
    f = s.solve_problems(d["omega"], d["nvel"], d["nds"])
    st = s.core.stats()
    assert (st.lu_rank, st.lu_band) == entries[k][:2]
    s.close()
    err = _normwise(f, ref)
    print(f"{name}: tuned form {entries[k][:2]}, precision {prec}: max normwise rel err {err:.3e}")
    assert err < tol
    # a handle of the same size class that opts in uses the recorded form;
    # one that does not keeps its own
    for opt, want in ((1, entries[k][:2]), (0, (128, 16))):
        s2 = Solver(m, float(d["depth"]), float(d["rho"]), float(d["g"]), precision=prec, tables=_tables())
        s2.core.set_option(core.OPT_LU_TUNED, opt)
        s2.solve_problems(d["omega"][:1], d["nvel"][:, :1], d["nds"])
        st = s2.core.stats()
        s2.close()
        assert (st.lu_rank, st.lu_band) == want, opt


@pytest.mark.parametrize("prec", [core.FP64, core.MIXED])
def test_lu_band_leaves_results_unchanged(prec):
    d, m, ref = _case("barge4000")
    got = []
    for band in (16, 32, 64):
        s = Solver(m, float(d["depth"]), float(d["rho"]), float(d["g"]), precision=prec, tables=_tables())
        s.core.set_option(core.OPT_LU_BAND, band)
        got.append(s.solve_problems(d["omega"][:18], d["nvel"][:, :18], d["nds"]))  # two frequencies
        st = s.core.stats()
        s.close()
        assert st.lu_band == band
    assert np.array_equal(got[0], got[1]) and np.array_equal(got[0], got[2])


def test_lu_band_option_is_checked():
    d, m, _ = _case("barge_fd")
    c = core.Core(d["vertices"], d["panels"], 0, float(d["depth"]), tables=_tables())
    with pytest.raises(core.NemohError):
        c.set_option(core.OPT_LU_BAND, 24)
    c.close()

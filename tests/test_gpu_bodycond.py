"""Body conditions, force weights and Froude-Krylov forces generated on the
device, against the reference preProcessor run on the same exact mesh.

The reference preProcessor (the FP64 build in oracle/_ref) writes
Normalvelocities.dat, mesh/Integration.dat and results/FKForce.dat in full
precision (list-directed output).  The device reproduces each to rounding
(1e-12 of each column's scale: the transcendental functions are the
device's own).  Then nemoh_core_solve_waves, with no per-frequency host
input, gives the forces that nemoh_core_solve_frequency gives from the
preProcessor's files (FP64 within 1e-10, mixed within 1e-6).
"""
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden, write_synthetic_case
from nemoh_amd import bodycond as bcm
from nemoh_amd import core
from nemoh_amd import io as nio

pytestmark = pytest.mark.gpu
PRE = ROOT / "oracle" / "_ref" / "preproc"


def _tables():
    t = load_golden("fn_crek.npz")
    return t["xr"], t["xz"], (t["apd1x"], t["apd1z"], t["apd2x"], t["apd2z"])


def _normwise(a, ref, axis=0):
    scale = np.abs(ref).max(axis=axis)
    return float((np.abs(a - ref).max(axis=axis) / np.where(scale > 0, scale, 1.0)).max())


@pytest.fixture(scope="module", params=["cyl512", "barge_fd"])
def case(request, tmp_path_factory):
    if not PRE.exists():
        pytest.skip("reference preProcessor not built (make -f oracle/Makefile.ref)")
    name = request.param
    work = tmp_path_factory.mktemp(name)
    m, depth = write_synthetic_case(work, name)
    r = subprocess.run([str(PRE)], cwd=work, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-1000:] + r.stderr[-1000:]
    ntot = m.npanels * 2 ** m.isym
    bc = nio.read_normal_velocities(work / "Normalvelocities.dat", ntot)
    nds = nio.read_integration(work / "mesh" / "Integration.dat", ntot)
    nprob = bc.omega.size
    fk = nio.table_to_forces(nio.read_forces_dat(work / "results" / "FKForce.dat", nds.shape[0], nprob), bc.omega,
                             bc.kind)
    c = core.Core(m.vertices, m.panels, m.isym, depth, 1025.0, 9.81, tables=_tables())
    c.set_dofs(bcm.six_dof())
    yield name, c, bc, nds, fk
    c.close()


def _by_frequency(bc):
    """(omega, column indices, headings in radians) in file order."""
    out, j = [], 0
    while j < bc.omega.size:
        k = j
        while k + 1 < bc.omega.size and bc.omega[k + 1] == bc.omega[j]:
            k += 1
        cols = np.arange(j, k + 1)
        out.append((bc.omega[j], cols, bc.kind[cols][bc.kind[cols] != -1.0]))
        j = k + 1
    return out


def test_body_conditions_match_preprocessor(case):
    name, c, bc, nds, fk = case
    worst = {"nvel": 0.0, "nds": 0.0, "fk": 0.0}
    for w, cols, betas in _by_frequency(bc):
        nvel, pinc, nds_dev = c.body_conditions(w, betas)
        worst["nvel"] = max(worst["nvel"], _normwise(nvel, bc.nvel[:, cols]))
        worst["nds"] = max(worst["nds"], _normwise(nds_dev, nds, axis=1))
        fk_dev = -(nds @ pinc)
        worst["fk"] = max(worst["fk"], _normwise(fk_dev, fk[:, cols[:betas.size]]))
    print(f"{name}: device vs reference preProcessor: {worst}")
    assert worst["nvel"] < 1e-12 and worst["nds"] < 1e-12 and worst["fk"] < 1e-12


@pytest.mark.parametrize("prec", [core.FP64, core.MIXED])
def test_solve_waves_matches_file_driven_solve(case, prec):
    name, c, bc, nds, fk = case
    c.set_precision(prec)
    err, err_fk = 0.0, 0.0
    for w, cols, betas in _by_frequency(bc):
        forces, fk_dev = c.solve_waves(w, betas)
        _, ref = c.solve_frequency(2 * np.pi / w, bc.nvel[:, cols], nds=nds, want_pressure=False)
        err = max(err, _normwise(forces, ref))
        if betas.size:
            err_fk = max(err_fk, _normwise(fk_dev, fk[:, cols[:betas.size]]))
    c.set_precision(core.FP64)
    print(f"{name} precision {prec}: solve_waves vs file-driven solve {err:.3e}, FK vs reference {err_fk:.3e}")
    assert err < (1e-10 if prec == core.FP64 else 1e-6) and err_fk < 1e-12


def test_outputs_after_solve_waves(case):
    """The Kochin function and the field potential read the sources of the
    last solve: after nemoh_core_solve_waves (problems generated on the
    device) they equal those after the file-driven solve of the same
    problems."""
    name, c, bc, nds, fk = case
    c.set_precision(core.FP64)
    w, cols, betas = _by_frequency(bc)[0]
    v = np.asarray(c.geo["xg"]), np.asarray(c.geo["yg"])
    r = 3.0 * max(np.abs(v[0]).max(), np.abs(v[1]).max()) + 1.0
    t = np.linspace(0.0, 2.0 * np.pi, 16, endpoint=False)
    pts = np.concatenate([np.stack([r * np.cos(t), r * np.sin(t), np.full(16, z)], axis=1) for z in (0.0, -1.0)])
    th = np.linspace(0.0, 2.0 * np.pi, 13)
    c.solve_waves(w, betas)
    h1, p1 = c.kochin(th), c.potential(pts)
    c.solve_frequency(2 * np.pi / w, bc.nvel[:, cols], nds=nds, want_pressure=False)
    h2, p2 = c.kochin(th), c.potential(pts)
    assert h1.shape == h2.shape == (th.size, cols.size) and p1.shape == p2.shape == (pts.shape[0], cols.size)
    assert _normwise(h1, h2) < 1e-10 and _normwise(p1, p2) < 1e-10


def test_set_dofs_rejects_bad_cases(case):
    name, c, bc, nds, fk = case
    bad = bcm.Mode(3, (1, 0, 0))
    with pytest.raises(core.NemohError) as e:
        c.set_dofs([bad])
    assert e.value.code == core.ERR_ARG
    with pytest.raises(core.NemohError):
        c.body_conditions(1.0, [0.0])  # the failed set-up leaves none
    c.set_dofs(bcm.six_dof())

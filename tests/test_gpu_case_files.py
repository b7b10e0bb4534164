"""Whole cases through the file-format layer on the GPU, against the files
the reference preProcessor and Solver wrote for the same inputs (fixtures
files_<case>.npz of four reference Verification cases).

* preProcessor + Solver + postProcessor stages (nemoh_case_run): the
  host-side mesh files
  record for record; Normalvelocities.dat, Integration.dat and FKForce.dat
  within 1e-12 of each problem's scale; Forces.dat within 1e-10 (FP64) and
  1e-6 (mixed) normwise per problem; for the case that asks for them
  (Postprocessing: 181 Kochin angles, a 50 x 50 free-surface grid, pressure
  output) the Kochin files within 1e-10 of the largest over the problems,
  the pressure files within 1e-10 and
  the free-surface elevation within the 7 digits its file keeps.
* the Solver stage alone on the reference preProcessor's own files: the
  reference's Forces.dat within 1e-10.
* the command-line runner on an ID.dat case.
"""
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from nemoh_amd import cases
from nemoh_amd import io as nio
from test_case_files import CASES, assert_same_records, fixture_files

pytestmark = pytest.mark.gpu


def _stage(tmp_path, name, with_outputs=False):
    ins, outs = fixture_files(name)
    work = tmp_path / name
    work.mkdir()
    for f, b in ins.items():
        (work / f).write_bytes(b)
    if with_outputs:  # the reference preProcessor's files
        for rel, b in outs.items():
            if rel.startswith("mesh/") or rel == "Normalvelocities.dat":
                (work / rel).parent.mkdir(parents=True, exist_ok=True)
                (work / rel).write_bytes(b)
    return work, outs


def _counts(outs):
    nw, nbeta, nrad, nint, ntheta = map(int, outs["results/index.dat"].decode().split()[:5])
    l10 = outs["mesh/L10.dat"].decode().split("\n")[1].split()
    isym, nfa = int(l10[0]), int(l10[2])
    return nw, nbeta, nrad, nint, ntheta, nfa * 2 ** isym


def _normwise(a, b, axis=0):
    scale = np.abs(b).max(axis=axis)
    return float((np.abs(a - b).max(axis=axis) / np.where(scale > 0, scale, 1.0)).max())


def _forces(path_or_bytes, tmp, nint, nprob, omega, kind):
    if isinstance(path_or_bytes, bytes):
        p = tmp / "ref_forces.dat"
        p.write_bytes(path_or_bytes)
        path_or_bytes = p
    return nio.table_to_forces(nio.read_forces_dat(path_or_bytes, nint, nprob), omega, kind)


def _check_solver_outputs(work, outs, tmp_path, name, tol):
    nw, nbeta, nrad, nint, ntheta, ntot = _counts(outs)
    nprob = nw * (nbeta + nrad)
    bc = nio.read_normal_velocities(work / "Normalvelocities.dat", ntot)
    got = _forces(work / "results" / "Forces.dat", tmp_path, nint, nprob, bc.omega, bc.kind)
    ref = _forces(outs["results/Forces.dat"], tmp_path, nint, nprob, bc.omega, bc.kind)
    err = _normwise(got, ref)
    print(f"{name}: Forces.dat vs the reference's {err:.2e}", flush=True)
    assert err < tol, (name, err)
    return err


@pytest.mark.parametrize("name", CASES)
def test_case_pipeline_matches_reference_files(tmp_path, name):
    work, outs = _stage(tmp_path, name)
    print(f"{name}: preProcessor + Solver stages on the device", flush=True)
    cases.run_case(work)
    nw, nbeta, nrad, nint, ntheta, ntot = _counts(outs)
    for rel in ("mesh/L10.dat", "mesh/L12.dat", "mesh/Mesh.tec", "mesh/Kochin.dat", "mesh/Freesurface.dat",
                "results/index.dat"):
        assert_same_records((work / rel).read_text(), outs[rel].decode(), (name, rel))
    # body conditions and force weights from the device
    bc = nio.read_normal_velocities(work / "Normalvelocities.dat", ntot)
    (tmp_path / "ref_nv.dat").write_bytes(outs["Normalvelocities.dat"])
    bref = nio.read_normal_velocities(tmp_path / "ref_nv.dat", ntot)
    assert np.array_equal(bc.omega, bref.omega) and np.array_equal(bc.kind, bref.kind)
    assert _normwise(bc.nvel, bref.nvel) < 1e-12
    (tmp_path / "ref_int.dat").write_bytes(outs["mesh/Integration.dat"])
    nds = nio.read_integration(work / "mesh" / "Integration.dat", ntot)
    assert _normwise(nds, nio.read_integration(tmp_path / "ref_int.dat", ntot), axis=1) < 1e-12
    fk = _forces(work / "results" / "FKForce.dat", tmp_path, nint, bc.omega.size, bc.omega, bc.kind)
    fk_ref = _forces(outs["results/FKForce.dat"], tmp_path, nint, bc.omega.size, bc.omega, bc.kind)
    diff = bc.kind != -1.0
    assert _normwise(fk[:, diff], fk_ref[:, diff]) < 1e-12
    # FKForce.tec: the same records; its 7-digit (|F|, arg F) pairs compared
    # as complex values against the record's largest (the phases of the
    # components that vanish by symmetry are rounding noise on both sides)
    tg, tr = (work / "results" / "FKForce.tec").read_text().split("\n"), outs["results/FKForce.tec"].decode().split("\n")
    assert len(tg) == len(tr)
    for a, b in zip(tg, tr):
        if a != b:
            va, vb = np.array(a.split(), float), np.array(b.split(), float)
            assert va.size == vb.size and va[0] == vb[0], (a, b)
            fa, fb = va[1::2] * np.exp(1j * va[2::2]), vb[1::2] * np.exp(1j * vb[2::2])
            assert np.abs(fa - fb).max() <= 2e-6 * np.abs(fb).max(), (a, b)
    _check_solver_outputs(work, outs, tmp_path, name, 1e-10)
    # the optional outputs; Kochin functions normwise over the problems (the
    # yaw of the cube radiates 1000 times less than the other modes: its
    # function is cancellation on both sides)
    kochin = {}
    for rel, b in outs.items():
        m = re.match(r"results/(Kochin|freesurface|pressure)\.\s*(\d+)\.dat", rel)
        if not m:
            continue
        got = (work / rel).read_text()
        ref = b.decode()
        if m.group(1) == "Kochin":
            g, r = np.array(got.split(), float).reshape(-1, 3), np.array(ref.split(), float).reshape(-1, 3)
            assert np.array_equal(g[:, 0], r[:, 0])
            kochin[rel] = (g[:, 1] * np.exp(1j * g[:, 2]), r[:, 1] * np.exp(1j * r[:, 2]))
        elif m.group(1) == "pressure":
            lg, lr = got.split("\n"), ref.split("\n")
            assert lg[:2] == lr[:2] and len(lg) == len(lr)
            data = [i for i, l in enumerate(lr) if "." in l and i >= 2]
            g = np.array([lg[i].split() for i in data], float)
            r = np.array([lr[i].split() for i in data], float)
            assert np.array_equal(g[:, :3], r[:, :3])
            pg, pr = g[:, 3] * np.exp(1j * g[:, 4]), r[:, 3] * np.exp(1j * r[:, 4])
            assert np.abs(pg - pr).max() <= 1e-10 * np.abs(pr).max(), rel
            assert [lg[i] for i in range(2, len(lg)) if i not in set(data)] == \
                [lr[i] for i in range(2, len(lr)) if i not in set(data)]
        else:
            lg, lr = got.split("\n"), ref.split("\n")
            assert lg[:2] == lr[:2] and len(lg) == len(lr)
            npts = int(re.search(r"N=\s*(\d+)", lr[1]).group(1))
            g = np.array([l.split() for l in lg[2:2 + npts]], float)
            r = np.array([l.split() for l in lr[2:2 + npts]], float)
            assert np.array_equal(g[:, :2], r[:, :2])
            eg, er = g[:, 4] + 1j * g[:, 5], r[:, 4] + 1j * r[:, 5]
            err = np.abs(eg - er).max() / np.abs(er).max()
            print(f"{name} {rel}: elevation vs the reference's {err:.2e}")
            assert err < 2e-6, rel
            assert lg[2 + npts:] == lr[2 + npts:]
    # the postProcessor stage ran on these files (its arithmetic is checked
    # record for record on the reference's own files without a GPU): the
    # radiation coefficients agree to the 7 digits they keep
    rg = (work / "results" / "RadiationCoefficients.tec").read_text().split("\n")
    rr = outs["results/RadiationCoefficients.tec"].decode().split("\n")
    assert len(rg) == len(rr)
    for a, b in zip(rg, rr):
        if a != b:
            va, vb = np.array(a.split(), float), np.array(b.split(), float)
            assert va.size == vb.size and np.abs(va - vb).max() <= 2e-6 * np.abs(vb).max(), (a, b)
    for rel in ("results/DiffractionForce.tec", "results/ExcitationForce.tec", "results/CA.dat", "results/CM.dat",
                "results/Fe.dat"):
        assert (work / rel).exists(), rel
    if kochin:
        scale = max(np.abs(hr).max() for _, hr in kochin.values())
        err = max(np.abs(hg - hr).max() for hg, hr in kochin.values()) / scale
        print(f"{name}: {len(kochin)} Kochin files vs the reference's {err:.2e}")
        assert err < 1e-10


@pytest.mark.parametrize("name", ["Cylinder", "2Bodies"])
def test_solver_stage_on_reference_preprocessor_files(tmp_path, name):
    work, outs = _stage(tmp_path, name, with_outputs=True)
    print(f"{name}: Solver stage on the reference preProcessor's files", flush=True)
    cases.run_case(work, stages=cases.SOLVE)
    _check_solver_outputs(work, outs, tmp_path, name, 1e-10)


def test_case_mixed_precision(tmp_path):
    work, outs = _stage(tmp_path, "NonSymmetrical")
    print("NonSymmetrical: mixed precision", flush=True)
    cases.run_case(work, precision=cases.MIXED)
    _check_solver_outputs(work, outs, tmp_path, "NonSymmetrical", 1e-6)


def test_command_line_runner(tmp_path):
    work, outs = _stage(tmp_path, "Cylinder")
    (work / "ID.dat").write_text("1\n.\n")
    exe = ROOT / "nemoh_amd" / "nemoh_run"
    print("Cylinder: nemoh_run from the case directory", flush=True)
    r = subprocess.run([str(exe)], cwd=work, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _check_solver_outputs(work, outs, tmp_path, "Cylinder", 1e-10)
    bad = subprocess.run([str(exe), str(work / "missing")], capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "Nemoh.cal" in bad.stderr

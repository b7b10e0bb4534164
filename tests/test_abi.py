"""The C-ABI library: loads, exports every declared symbol, and its host
routines (no GPU involved) match the FP64 reference build bit for bit."""
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT, case_geometry, load_golden
from nemoh_amd import core


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "nemoh_core.h").read_text()
    declared = set(re.findall(r"\b(nemoh_(?:core|server|batch)_\w+)\s*\(", header))
    L = core.load()
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert declared == set(core.EXPORTS)
    case_header = (ROOT / "include" / "nemoh_case.h").read_text()
    declared = set(re.findall(r"\b(nemoh_(?:case|format)_\w+)\s*\(", case_header))
    assert declared == {"nemoh_case_run", "nemoh_case_write_mesh_files", "nemoh_format_list", "nemoh_format_e"}
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing


def test_prony_fit_matches_reference_bitwise():
    d = load_golden("fn_lisc.npz")
    for k in range(d["akh"].size):
        n, lam, ar = core.prony(d["akh"][k], d["amh"][k])
        assert n == d["nexp"][k]
        assert np.array_equal(lam, d["ambda"][k, :n]) and np.array_equal(ar, d["ar"][k, :n])


def test_wavenumber_matches_reference_bitwise():
    d = load_golden("fn_x0.npz")
    assert all(core.wavenumber(a) == b for a, b in zip(d["ak"], d["x0"]))


def test_theta_rule_matches_reference():
    d = load_golden("fn_cint.npz")
    for k, akk in enumerate(d["akk"]):
        n, qq, cq = core.theta_rule(akk)
        assert n == d["nq"][k]
        assert np.array_equal(qq, d["qq"][k, :n]) and np.array_equal(cq, d["cq"][k, :n])


@pytest.mark.parametrize("name", ["pairs_lidcyl.npz", "pairs_barge.npz"])
def test_prepare_mesh_matches_reference_bitwise(name):
    d = load_golden(name)
    geo = core.prepare_mesh(d["vertices"], d["panels"], int(d["isym"]))
    ref = case_geometry(d)
    for k in ref:
        assert np.array_equal(np.asarray(geo[k]), np.asarray(ref[k])), k


def test_degenerate_panel_is_an_error():
    v = np.array([[0, 0, -1], [1, 0, -1], [2, 0, -1], [3, 0, -1]], float)
    with pytest.raises(core.NemohError) as e:
        core.prepare_mesh(v, np.array([[0, 1, 2, 3]], np.int32), 0)
    assert e.value.code == core.ERR_PANEL


def _has_gpu():
    import torch
    return torch.cuda.device_count() > 0


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_create_without_device_fails_loudly():
    d = load_golden("pairs_barge.npz")
    t = load_golden("fn_crek.npz")
    with pytest.raises(core.NemohError) as e:
        core.Core(d["vertices"], d["panels"], 0, 0.0, geometry=case_geometry(d),
                  tables=(t["xr"], t["xz"], (t["apd1x"], t["apd1z"], t["apd2x"], t["apd2z"])))
    assert e.value.code in (core.ERR_HIP, core.ERR_ARG)


def test_incident_wavenumber_matches_reference_bitwise():
    """The preProcessor's own wave number (Common/Environment.f90:45-84),
    deep-water branch and bisection, against the reference build."""
    d = load_golden("fn_wavenumber.npz")
    got = [core.incident_wavenumber(w, h, g) for w, h, g in zip(d["w"], d["depth"], d["g"])]
    assert np.array_equal(np.array(got), d["k"])

"""The reference's own Fortran host calling the core through ISO_C_BINDING.

oracle/_ref/solver_core is the reference Solver built from its unmodified
sources with the three drop-in edits of oracle/drop_in_edit.py and the
binding modules of nemoh_amd/fortran/, linked against
nemoh_amd/libnemoh_core.so.  On the GPU it solves the synthetic cases from
the inputs the reference preProcessor writes, and its Forces.dat must match
the one the unmodified reference Solver wrote (the golden fixture): FP64
within 1e-10, mixed within 1e-6, normwise per problem.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden, write_synthetic_case
from nemoh_amd import io as nio

REF = ROOT / "oracle" / "_ref"
FLANG = "/opt/rocm/lib/llvm/bin/flang"


def test_binding_module_compiles_and_names_exports(tmp_path):
    """The ISO_C_BINDING module compiles with the host's FP64 flags, and
    every C name it binds is declared in include/nemoh_core.h."""
    if not os.path.exists(FLANG):
        pytest.skip("flang not available")
    src = ROOT / "nemoh_amd" / "fortran" / "nemoh_core_bind.f90"
    r = subprocess.run([FLANG, "-fdefault-real-8", "-fdefault-double-8", "-module-dir", str(tmp_path), "-c",
                        str(src), "-o", str(tmp_path / "bind.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    bound = set(re.findall(r"NAME='(\w+)'", src.read_text()))
    header = (ROOT / "include" / "nemoh_core.h").read_text()
    assert bound and all(re.search(rf"\b{n}\s*\(", header) for n in bound)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cyl512", "barge_fd", "barge4000"])
@pytest.mark.parametrize("precision", ["FP64", "MIXED"])
def test_reference_solver_through_core(tmp_path, name, precision):
    exe, pre = REF / "solver_core", REF / "preproc"
    if not (exe.exists() and pre.exists()):
        pytest.skip("reference host not built (make -f oracle/Makefile.ref)")
    work = tmp_path / name
    write_synthetic_case(work, name)
    r = subprocess.run([str(pre)], cwd=work, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-1000:] + r.stderr[-1000:]
    env = dict(os.environ, NEMOH_CORE_PRECISION=precision)
    print(f"{name}: reference Solver through the core, {precision}", flush=True)
    r = subprocess.run([str(exe)], cwd=work, capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0 and "nemoh_core error" not in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
    d = load_golden(f"case_{name}.npz")
    nint, nprob = d["nds"].shape[0], d["omega"].size
    got = nio.table_to_forces(nio.read_forces_dat(work / "results" / "Forces.dat", nint, nprob), d["omega"],
                              d["kind"])
    ref = nio.table_to_forces(d["forces_table"], d["omega"], d["kind"])
    err = float((np.abs(got - ref).max(axis=0) / np.abs(ref).max(axis=0)).max())
    print(f"{name} {precision}: max normwise rel err vs the unmodified reference {err:.3e}")
    assert err < (1e-10 if precision == "FP64" else 1e-6)
    shutil.rmtree(work)

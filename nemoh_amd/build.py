"""Builds libnemoh_core.so in-tree with hipcc for gfx950.

The library is plain HIP + C++ behind the C ABI of include/nemoh_core.h;
Python reaches it through ctypes (nemoh_amd.core).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libnemoh_core.so"
SOURCES = ["core.hip", "kernels.hip", "lu.hip", "bodycond.hip", "potential.hip", "host_math.cpp", "server.cpp",
           "fortran_io.cpp", "case_io.cpp", "postproc.cpp", "run_case.cpp"]
EXE = PKG / "nemoh_run"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NEMOH_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
# hipcc's host pass warns that it does not know this device feature
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose: bool = False) -> Path:
    objs = []
    out_dir = PKG / "_build"
    out_dir.mkdir(exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = out_dir / (src + ".o")
        cmd = [HIPCC, *COMMON, *DEVICE, "-c", str(CSRC / src), "-o", str(obj)]
        if src.endswith(".cpp"):
            cmd = [HIPCC, *COMMON, "-pthread", "-x", "c++", "-c", str(CSRC / src), "-o", str(obj)]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    for src, p in procs:
        out, _ = p.communicate()
        msg = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and "unknown target feature" not in l)
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{msg}")
        if verbose and msg.strip():
            print(msg)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *DEVICE, "-shared", "-pthread", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the command-line runner over the library (rpath: beside it)
    tmp = EXE.with_suffix(".tmp")
    cmd = [HIPCC, *COMMON, "-x", "c++", str(CSRC / "nemoh_run.cpp"), "-x", "none", "-o", str(tmp), f"-L{PKG}",
           "-lnemoh_core", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"nemoh_run link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, EXE)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

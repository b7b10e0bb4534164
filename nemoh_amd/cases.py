"""Whole-case runs over the reference's files (include/nemoh_case.h).

``run_case(dir)`` does what the reference's preProcessor, Solver and
postProcessor executables do for a case directory (Nemoh.cal + body meshes), with the
per-frequency work on the GPU, and writes the same files in the same text
formats; ``nemoh_amd/nemoh_run`` is the same as a command.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

from .core import NemohError, load

PRE, SOLVE, POST = 1, 2, 4
FP64, MIXED = 0, 1


def _lib():
    L = load()
    L.nemoh_case_run.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_char_p, C.c_int]
    L.nemoh_case_write_mesh_files.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    L.nemoh_format_list.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_char_p, C.c_size_t]
    L.nemoh_format_e.argtypes = [C.c_double, C.c_int, C.c_int, C.c_char_p, C.c_size_t]
    return L


def run_case(case_dir, stages: int = PRE | SOLVE | POST, precision: int = FP64) -> None:
    """Runs the preProcessor, Solver and/or postProcessor stage on a case
    directory (the postProcessor stage needs no GPU)."""
    msg = C.create_string_buffer(1024)
    r = _lib().nemoh_case_run(str(Path(case_dir)).encode(), int(stages), int(precision), msg, len(msg))
    if r:
        raise NemohError(r, msg.value.decode(errors="replace"))


def write_mesh_files(case_dir) -> None:
    """The host-only preProcessor files (mesh/L10.dat, L12.dat, Mesh.tec,
    Kochin.dat, Freesurface.dat, results/index.dat); no device needed."""
    msg = C.create_string_buffer(1024)
    r = _lib().nemoh_case_write_mesh_files(str(Path(case_dir)).encode(), msg, len(msg))
    if r:
        raise NemohError(r, msg.value.decode(errors="replace"))


def format_list(values) -> str:
    """One list-directed WRITE of reals, as the reference's runtime writes it."""
    v = [float(x) for x in values]
    arr = (C.c_double * max(1, len(v)))(*v)
    cap = 64 + 32 * len(v)
    out = C.create_string_buffer(cap)
    n = _lib().nemoh_format_list(len(v), arr, out, cap)
    if n < 0:
        raise ValueError("formatting failed")
    return out.value.decode()


def format_e(value: float, w: int, d: int) -> str:
    out = C.create_string_buffer(64)
    n = _lib().nemoh_format_e(float(value), int(w), int(d), out, 64)
    if n < 0:
        raise ValueError("formatting failed")
    return out.value.decode()

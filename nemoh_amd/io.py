"""Readers and writers for the files at the Solver boundary.

Formats follow the reference exactly, so a case prepared by the reference
preProcessor can be solved here and the results post-processed by the
reference postProcessor:

* ``Nemoh.cal`` -- fixed line positions (``Solver/Core/INITIALIZATION.f90:
  44-50``, ``preProcessor/Main.f90:93-150``).
* ``Normalvelocities.dat`` -- ``Nproblems``, five per-problem records (omega,
  type = heading in rad or -1 for radiation, three switches), then one row
  of (Re, Im) pairs per panel (``Solver/Core/Bodyconditions.f90:65-89``).
* ``mesh/Integration.dat`` -- ``Nint`` then ``Nint`` rows of ``N*2^isym``
  reals (``Solver/NEMOH.f90:88-94``).
* ``results/Forces.dat`` -- ``Nint`` rows of ``2*Nproblems`` values:
  ``|F|, arg F`` for diffraction, ``Im F / omega, -Re F`` for radiation
  (``Solver/NEMOH.f90:152-166``).

All readers are token-stream based, as Fortran list-directed input is.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np


def _tokens(path):
    return Path(path).read_text().split()


@dataclass
class BodyConditions:
    omega: np.ndarray          # (nprob,)
    kind: np.ndarray           # (nprob,) heading in rad, or -1 for radiation
    switch_potential: np.ndarray
    switch_fs: np.ndarray
    switch_kochin: np.ndarray
    nvel: np.ndarray           # (ntot, nprob) complex128


def read_normal_velocities(path, ntot: int) -> BodyConditions:
    t = _tokens(path)
    n = int(t[0])
    pos = 1

    def take(k, conv=float):
        nonlocal pos
        out = [conv(v) for v in t[pos:pos + k]]
        pos += k
        return np.asarray(out)

    omega = take(n)
    kind = take(n)
    sp = take(n, lambda v: int(float(v)))
    sf = take(n, lambda v: int(float(v)))
    sk = take(n, lambda v: int(float(v)))
    vals = take(ntot * 2 * n).reshape(ntot, n, 2)
    return BodyConditions(omega, kind, sp, sf, sk, vals[..., 0] + 1j * vals[..., 1])


def write_normal_velocities(path, bc: BodyConditions) -> None:
    n = bc.omega.size
    lines = [f"{n}"]
    for arr in (bc.omega, bc.kind):
        lines.append(" ".join(f"{v:.17e}" for v in arr))
    for arr in (bc.switch_potential, bc.switch_fs, bc.switch_kochin):
        lines.append(" ".join(str(int(v)) for v in arr))
    for row in bc.nvel:
        lines.append(" ".join(f"{v.real:.17e} {v.imag:.17e}" for v in row))
    Path(path).write_text("\n".join(lines) + "\n")


def read_integration(path, ntot: int) -> np.ndarray:
    t = _tokens(path)
    nint = int(t[0])
    return np.asarray([float(v) for v in t[1:1 + nint * ntot]]).reshape(nint, ntot)


def write_integration(path, nds: np.ndarray) -> None:
    lines = [f"{nds.shape[0]}"] + [" ".join(f"{v:.17e}" for v in row) for row in nds]
    Path(path).write_text("\n".join(lines) + "\n")


def forces_to_table(force: np.ndarray, omega: np.ndarray, kind: np.ndarray) -> np.ndarray:
    """Complex forces (nint, nprob) -> the Forces.dat table (nint, 2*nprob)
    with the reference output conventions (NEMOH.f90:154-163)."""
    nint, nprob = force.shape
    out = np.empty((nint, 2 * nprob))
    for j in range(nprob):
        if kind[j] != -1.0:
            out[:, 2 * j] = np.abs(force[:, j])
            out[:, 2 * j + 1] = np.arctan2(force[:, j].imag, force[:, j].real)
        else:
            out[:, 2 * j] = force[:, j].imag / omega[j]
            out[:, 2 * j + 1] = -force[:, j].real
    return out


def table_to_forces(table: np.ndarray, omega: np.ndarray, kind: np.ndarray) -> np.ndarray:
    """Inverse of :func:`forces_to_table` (complex values avoid +-pi wrap)."""
    nint = table.shape[0]
    nprob = omega.size
    f = np.empty((nint, nprob), np.complex128)
    for j in range(nprob):
        a, b = table[:, 2 * j], table[:, 2 * j + 1]
        if kind[j] != -1.0:
            f[:, j] = a * np.exp(1j * b)
        else:
            f[:, j] = -b + 1j * a * omega[j]
    return f


def read_forces_dat(path, nint: int, nprob: int) -> np.ndarray:
    t = _tokens(path)
    return np.asarray([float(v) for v in t[:nint * 2 * nprob]]).reshape(nint, 2 * nprob)


def write_forces_dat(path, table: np.ndarray) -> None:
    Path(path).write_text("\n".join(" ".join(f"{v:.17e}" for v in row) for row in table) + "\n")


@dataclass
class BodySpec:
    mesh_file: str
    npoints: int
    npanels: int
    dofs: list = field(default_factory=list)    # [(type, dx, dy, dz, cx, cy, cz)]
    forces: list = field(default_factory=list)  # same layout


def six_dof(cx=0.0, cy=0.0, cz=0.0):
    """Surge, sway, heave, then roll, pitch, yaw about (cx, cy, cz)."""
    return [(1, 1., 0., 0., 0., 0., 0.), (1, 0., 1., 0., 0., 0., 0.), (1, 0., 0., 1., 0., 0., 0.),
            (2, 1., 0., 0., cx, cy, cz), (2, 0., 1., 0., cx, cy, cz), (2, 0., 0., 1., cx, cy, cz)]


def write_nemoh_cal(path, rho, g, depth, bodies, nw, wmin, wmax, nbeta, bmin, bmax,
                    kochin=(0, 0.0, 180.0), fs=(0, 2, 1000.0, 2.0), show_pressure=0) -> None:
    """Write a Nemoh.cal in the reference's fixed line layout
    (Verification/Cylinder/Nemoh.cal)."""
    L = ["--- Environment ---",
         f"{rho}\t\t! RHO", f"{g}\t\t! G", f"{depth}\t\t! DEPTH", "0.\t0.\t\t! XEFF YEFF",
         "--- Description of floating bodies ---", f"{len(bodies)}\t\t! Number of bodies"]
    for k, b in enumerate(bodies, start=1):
        L.append(f"--- Body {k} ---")
        L.append(f"{b.mesh_file}\t\t! Name of mesh file")
        L.append(f"{b.npoints}\t{b.npanels}\t\t! Number of points and number of panels")
        L.append(f"{len(b.dofs)}\t\t! Number of degrees of freedom")
        for d in b.dofs:
            L.append(f"{d[0]} " + " ".join(f"{v!r}" for v in d[1:]))
        L.append(f"{len(b.forces)}\t\t! Number of resulting generalised forces")
        for d in b.forces:
            L.append(f"{d[0]} " + " ".join(f"{v!r}" for v in d[1:]))
        L.append("0\t\t! Number of lines of additional information")
    L.append("--- Load cases to be solved ---")
    L.append(f"{nw}\t{wmin!r}\t{wmax!r}\t\t! Number of wave frequencies, Min, and Max (rad/s)")
    L.append(f"{nbeta}\t{bmin!r}\t{bmax!r}\t\t! Number of wave directions, Min and Max (degrees)")
    L.append("--- Post processing ---")
    L.append("0\t0.1\t10.\t\t! IRF")
    L.append(f"{show_pressure}\t\t! Show pressure")
    L.append(f"{kochin[0]}\t{kochin[1]!r}\t{kochin[2]!r}\t\t! Kochin function")
    L.append(f"{fs[0]}\t{fs[1]}\t{fs[2]!r}\t{fs[3]!r}\t! Free surface elevation")
    Path(path).write_text("\n".join(L) + "\n")

"""The Solver stage on the MI355X core, mirroring the reference's interface.

``Solver.solve_bvp(period, nvel)`` is SOLVE_BVP (Solver/Core/SOLVE_BEM.f90:
28-96) for one problem: normal velocities in, pressure out.
``Solver.solve_problems`` runs the problem loop of the Solver program
(Solver/NEMOH.f90:129-143) -- problems are grouped by frequency in file
order, all right-hand sides of a frequency go through one factorisation --
and integrates forces (NEMOH.f90:137-141).  ``run_case`` reads a case
directory prepared by the reference preProcessor and writes
results/Forces.dat in the reference's format (NEMOH.f90:152-166).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from . import io as nio
from . import mesh as nmesh
from .core import FP64, MIXED, Core

MAX_RHS = 16  # right-hand sides per device pass


class Solver:
    def __init__(self, mesh: nmesh.Mesh, depth: float, rho: float = 1000.0, g: float = 9.81,
                 precision: int = FP64, tables=None, geometry=None, tol: float = 0.0):
        self.mesh = mesh
        self.core = Core(mesh.vertices, mesh.panels, mesh.isym, depth, rho, g, geometry=geometry, tables=tables)
        self.core.set_precision(precision, tol)
        self.ntot = self.core.ntot

    def set_precision(self, precision: int, tol: float = 0.0):
        self.core.set_precision(precision, tol)

    def solve_bvp(self, period: float, nvel) -> np.ndarray:
        pres, _ = self.core.solve_frequency(period, np.asarray(nvel).reshape(-1, 1))
        return pres[:, 0]

    def solve_problems(self, omega, nvel, nds) -> np.ndarray:
        """Forces (nint, nprob) for problems with frequencies omega (nprob,)
        and normal velocities nvel (ntot, nprob), in the given order."""
        omega = np.asarray(omega, dtype=np.float64)
        nprob = omega.size
        forces = np.zeros((nds.shape[0], nprob), np.complex128)
        j = 0
        while j < nprob:
            k = j
            while k + 1 < nprob and omega[k + 1] == omega[j] and k + 1 - j < MAX_RHS:
                k += 1
            period = 2.0 * np.pi / omega[j]
            _, f = self.core.solve_frequency(period, nvel[:, j:k + 1], nds=nds, want_pressure=False)
            forces[:, j:k + 1] = f
            j = k + 1
        return forces

    def close(self):
        self.core.close()


def read_environment(case_dir) -> tuple:
    """RHO, G, DEPTH from Nemoh.cal lines 2-4 (INITIALIZATION.f90:44-50)."""
    lines = Path(case_dir, "Nemoh.cal").read_text(errors="replace").split("\n")
    return tuple(float(lines[k].split()[0]) for k in (1, 2, 3))


def run_case(case_dir, precision: int = FP64) -> np.ndarray:
    """Solve a case directory prepared by the reference preProcessor and
    write results/Forces.dat; returns the complex forces (nint, nprob)."""
    case = Path(case_dir)
    rho, g, depth = read_environment(case)
    m = nmesh.read_dat(case / "mesh" / "L12.dat")
    ntot = m.npanels * 2 ** m.isym
    bc = nio.read_normal_velocities(case / "Normalvelocities.dat", ntot)
    nds = nio.read_integration(case / "mesh" / "Integration.dat", ntot)
    s = Solver(m, depth, rho, g, precision=precision)
    try:
        forces = s.solve_problems(bc.omega, bc.nvel, nds)
    finally:
        s.close()
    (case / "results").mkdir(exist_ok=True)
    nio.write_forces_dat(case / "results" / "Forces.dat", nio.forces_to_table(forces, bc.omega, bc.kind))
    return forces

"""Right-hand sides and force-integration weights (the preProcessor's job).

Restates the reference preProcessor so that standalone, batched and
serving runs need no Fortran: panel centres / normals / areas from two
triangles (preProcessor/Mesh.f90:207-233), radiation conditions
n . (translation | rotation x r) (preProcessor/BodyConditions.f90:27-88),
diffraction conditions -grad(phi_incident) . n (BodyConditions.f90:90-137 with
Compute_Wave, Common/Environment.f90:86-99), the preProcessor's own wave
number (Common/Environment.f90:45-84, not the Solver's X0), force weights
NDS = (n . mode) A (preProcessor/Integration.f90:27-86), and the problem
order per frequency: headings first, then radiation modes
(preProcessor/Main.f90:223-243).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mesh import Mesh


@dataclass
class Mode:
    kind: int          # 1 translation, 2 rotation
    direction: tuple
    axis: tuple = (0.0, 0.0, 0.0)
    body: int = 1


def six_dof(axis=(0.0, 0.0, 0.0), body: int = 1):
    return [Mode(1, (1, 0, 0), axis, body), Mode(1, (0, 1, 0), axis, body), Mode(1, (0, 0, 1), axis, body),
            Mode(2, (1, 0, 0), axis, body), Mode(2, (0, 1, 0), axis, body), Mode(2, (0, 0, 1), axis, body)]


def panel_properties(mesh: Mesh):
    """Centres XM (n,3), unit normals N (n,3), areas A (n,) as the
    preProcessor computes them (two triangles P1P2P4 and P3P4P2)."""
    P = mesh.vertices[mesh.panels]
    w1 = np.cross(P[:, 1] - P[:, 0], P[:, 3] - P[:, 1])
    w2 = np.cross(P[:, 3] - P[:, 2], P[:, 1] - P[:, 2])
    a1 = 0.5 * np.linalg.norm(w1, axis=1)
    a2 = 0.5 * np.linalg.norm(w2, axis=1)
    A = a1 + a2
    if np.any(A < 1.0e-7):
        i = int(np.argmax(A < 1.0e-7))
        raise ValueError(f"Error: surface of panel {i + 1} is too small ({A[i]:.7e})")
    xm = (1.0 / 3 * (P[:, 0] + P[:, 1] + P[:, 3]) * (a1 / A)[:, None] +
          1.0 / 3 * (P[:, 1] + P[:, 2] + P[:, 3]) * (a2 / A)[:, None])
    u = w1 + w2
    n = u / np.linalg.norm(u, axis=1)[:, None]
    return xm, n, A


def wavenumber(w: float, depth: float, g: float) -> float:
    """Bisection of k h tanh(k h) = w^2 h / g to 1e-6 (preProcessor's own)."""
    k = w * w / g
    x0 = k * depth
    if not (0.0 < x0 <= 20.0):
        return k
    xg, xd = 0.0, x0
    while (x0 - xg * np.tanh(xg)) * (x0 - xd * np.tanh(xd)) > 0.0:
        xg, xd = xd, 2.0 * xd
    xc = 0.5 * (xd + xg)
    while abs(xd - xg) / abs(xc) >= 1.0e-6:
        xc = 0.5 * (xd + xg)
        if (x0 - xg * np.tanh(xg)) * (x0 - xc * np.tanh(xc)) > 0.0:
            xg = xc
        else:
            xd = xc
    return xc / depth


def _cih_sih(k, z, h):
    if 0.0 < k * h <= 20.0:
        return np.cosh(k * (z + h)) / np.cosh(k * h), np.sinh(k * (z + h)) / np.cosh(k * h)
    e = np.exp(k * z)
    return e, e


def _mode_velocity(m: Mode, xm, sign_y):
    d = np.asarray(m.direction, dtype=np.float64)
    if m.kind == 1:
        return np.broadcast_to(d, xm.shape)
    c = np.asarray(m.axis, dtype=np.float64)
    x, y, z = xm[:, 0], sign_y * xm[:, 1], xm[:, 2]
    return np.stack([d[1] * (z - c[2]) - d[2] * (y - c[1]),
                     d[2] * (x - c[0]) - d[0] * (z - c[2]),
                     d[0] * (y - c[1]) - d[1] * (x - c[0])], axis=1)


def radiation(mesh: Mesh, m: Mode, body_of_panel=None) -> np.ndarray:
    xm, n, _ = panel_properties(mesh)
    on = np.ones(mesh.npanels, bool) if body_of_panel is None else (body_of_panel == m.body)
    out = [np.where(on, (n * _mode_velocity(m, xm, 1.0)).sum(1), 0.0)]
    if mesh.isym:
        nm = n * np.array([1.0, -1.0, 1.0])
        out.append(np.where(on, (nm * _mode_velocity(m, xm, -1.0)).sum(1), 0.0))
    return np.concatenate(out).astype(np.complex128)


def nds(mesh: Mesh, m: Mode, body_of_panel=None) -> np.ndarray:
    xm, n, A = panel_properties(mesh)
    on = np.ones(mesh.npanels, bool) if body_of_panel is None else (body_of_panel == m.body)
    out = [np.where(on, (n * _mode_velocity(m, xm, 1.0)).sum(1) * A, 0.0)]
    if mesh.isym:
        nm = n * np.array([1.0, -1.0, 1.0])
        out.append(np.where(on, (nm * _mode_velocity(m, xm, -1.0)).sum(1) * A, 0.0))
    return np.concatenate(out)


def diffraction(mesh: Mesh, w: float, beta: float, depth: float, g: float, rho: float = 1025.0,
                xeff: float = 0.0, yeff: float = 0.0):
    """(NVEL, incident pressure) of a regular wave of heading beta (rad)."""
    xm, n, _ = panel_properties(mesh)
    k = wavenumber(w, depth, g)
    rows = [(xm[:, 0], xm[:, 1], n * np.array([1.0, 1.0, 1.0]))]
    if mesh.isym:
        rows.append((xm[:, 0], -xm[:, 1], n * np.array([1.0, -1.0, 1.0])))
    nv, pr = [], []
    for x, y, nn in rows:
        wbar = (x - xeff) * np.cos(beta) + (y - yeff) * np.sin(beta)
        cih, sih = _cih_sih(k, xm[:, 2], depth)
        e = np.exp(1j * k * wbar)
        vx = g / w * k * np.cos(beta) * cih * e
        vy = g / w * k * np.sin(beta) * cih * e
        vz = -1j * g / w * k * sih * e
        nv.append(-(vx * nn[:, 0] + vy * nn[:, 1] + vz * nn[:, 2]))
        pr.append(rho * g * cih * e)
    return np.concatenate(nv), np.concatenate(pr)


def problems(mesh: Mesh, omegas, betas_deg, modes, depth: float, g: float, rho: float = 1025.0,
             body_of_panel=None):
    """All problems in the reference order: for each omega, one diffraction
    problem per heading then one radiation problem per mode (a mode moves
    only the panels of its body when ``body_of_panel`` is given).  Returns
    (omega (nprob,), kind (nprob,), nvel (ntot, nprob))."""
    om, kind, cols = [], [], []
    for w in omegas:
        for b in betas_deg:
            nv, _ = diffraction(mesh, w, np.deg2rad(b), depth, g, rho)
            om.append(w)
            kind.append(np.deg2rad(b))
            cols.append(nv)
        for m in modes:
            om.append(w)
            kind.append(-1.0)
            cols.append(radiation(mesh, m, body_of_panel))
    return np.array(om), np.array(kind), np.stack(cols, axis=1)


def integration(mesh: Mesh, modes, body_of_panel=None) -> np.ndarray:
    return np.stack([nds(mesh, m, body_of_panel) for m in modes])

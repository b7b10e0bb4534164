"""Panel meshes: the Nemoh mesh file format and parametric generators.

A mesh is a set of vertices and quadrilateral panels (triangles repeat a
vertex), with ``isym = 1`` when only the y >= 0 (or y <= 0) half is stored
and the other half is implied by symmetry about the xOz plane.

File format (reference ``Common/Mesh.f90:81-132``, ``preProcessor/Mesh.f90:
237-247``): a header ``2 isym``, then ``id x y z`` per vertex closed by
``0 0. 0. 0.``, then ``p1 p2 p3 p4`` (1-based) per panel closed by
``0 0 0 0``.

Generators produce outward normals (pointing into the fluid) under the
reference's normal definition ``(P3 - P1) x (P4 - P2)``
(``Solver/Core/PREPARE_MESH.f90:102-123``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np


@dataclass
class Mesh:
    vertices: np.ndarray  # (np, 3) float64
    panels: np.ndarray    # (nfa, 4) int32, 0-based
    isym: int = 0

    @property
    def npoints(self) -> int:
        return int(self.vertices.shape[0])

    @property
    def npanels(self) -> int:
        return int(self.panels.shape[0])

    def normals(self) -> np.ndarray:
        p = self.vertices[self.panels]
        n = np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 1])
        return n / np.linalg.norm(n, axis=1)[:, None]

    def centers(self) -> np.ndarray:
        return self.vertices[self.panels].mean(axis=1)

    def write_dat(self, path) -> None:
        """Write the Nemoh ``.dat`` mesh format (full precision)."""
        lines = [f"{2:>21d}{self.isym:>11d}"]
        for i, (x, y, z) in enumerate(self.vertices, start=1):
            lines.append(f"{i:>14d} {x:>23.16e} {y:>23.16e} {z:>23.16e}")
        lines.append(f"{0:>14d} {0.0:>23.16e} {0.0:>23.16e} {0.0:>23.16e}")
        for p in self.panels + 1:
            lines.append(" ".join(f"{int(v):>13d}" for v in p))
        lines.append(" ".join(f"{0:>13d}" for _ in range(4)))
        Path(path).write_text("\n".join(lines) + "\n")


def read_dat(path) -> Mesh:
    """Read a Nemoh ``.dat`` (or ``mesh/L12.dat``) mesh file."""
    tok = Path(path).read_text().split("\n")
    head = tok[0].split()
    if int(head[0]) != 2:
        raise ValueError(f"{path}: the mesh file format is not correct")
    isym = int(head[1])
    verts, pans, i = [], [], 1
    while True:
        t = tok[i].split()
        i += 1
        if int(t[0]) == 0:
            break
        verts.append([float(v) for v in t[1:4]])
    while True:
        t = tok[i].split()
        i += 1
        if not t or int(t[0]) == 0:
            break
        pans.append([int(v) - 1 for v in t[:4]])
    return Mesh(np.asarray(verts, dtype=np.float64), np.asarray(pans, dtype=np.int32), isym)


class _VertexPool:
    """Deduplicates vertices generated on structured grids."""

    def __init__(self, scale: float):
        self._index: dict = {}
        self._pts: list = []
        self._q = 1e-9 * max(scale, 1.0)

    def add(self, p) -> int:
        key = tuple(int(round(c / self._q)) for c in p)
        idx = self._index.get(key)
        if idx is None:
            idx = len(self._pts)
            self._index[key] = idx
            self._pts.append([float(c) for c in p])
        return idx

    def array(self) -> np.ndarray:
        return np.asarray(self._pts, dtype=np.float64)


def _orient(mesh_pts: np.ndarray, quads: list, outward) -> np.ndarray:
    """Reverse any quad whose reference normal opposes ``outward(center)``."""
    q = np.asarray(quads, dtype=np.int32)
    p = mesh_pts[q]
    n = np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 1])
    c = p.mean(axis=1)
    o = np.asarray([outward(ci) for ci in c])
    flip = (n * o).sum(axis=1) < 0
    q[flip] = q[flip][:, [0, 3, 2, 1]]
    return q


def cylinder(radius: float, draft: float, n_theta: int, n_z: int, n_r: int,
             half: bool = True, x0: float = 0.0, y0: float = 0.0) -> Mesh:
    """Vertical truncated cylinder: side ``n_theta x n_z`` panels, bottom
    ``n_r`` rings of ``n_theta`` sectors with degenerate quads at the centre.
    ``half`` stores theta in [0, pi] (y >= 0) with ``isym = 1``."""
    pool = _VertexPool(max(radius, draft))
    span = math.pi if half else 2.0 * math.pi
    th = [span * k / n_theta for k in range(n_theta + 1)]
    if not half:
        th[-1] = 0.0
    zs = [-draft + draft * k / n_z for k in range(n_z + 1)]
    rs = [radius * k / n_r for k in range(n_r + 1)]

    def pt(r, t, z):
        return (x0 + r * math.cos(t), y0 + r * math.sin(t) if r else y0, z)

    quads = []
    for a in range(n_theta):
        for b in range(n_z):
            quads.append([pool.add(pt(radius, th[a], zs[b])), pool.add(pt(radius, th[a + 1], zs[b])),
                          pool.add(pt(radius, th[a + 1], zs[b + 1])), pool.add(pt(radius, th[a], zs[b + 1]))])
    for k in range(n_r):
        for a in range(n_theta):
            if k == 0:
                c = pool.add(pt(0.0, 0.0, -draft))
                quads.append([c, pool.add(pt(rs[1], th[a], -draft)),
                              pool.add(pt(rs[1], th[a + 1], -draft)), c])
            else:
                quads.append([pool.add(pt(rs[k], th[a], -draft)), pool.add(pt(rs[k + 1], th[a], -draft)),
                              pool.add(pt(rs[k + 1], th[a + 1], -draft)), pool.add(pt(rs[k], th[a + 1], -draft))])
    pts = pool.array()

    def outward(c):
        if abs(c[2] + draft) < 1e-9 * max(1.0, draft):
            return (0.0, 0.0, -1.0)
        return (c[0] - x0, c[1] - y0, 0.0)

    return Mesh(pts, _orient(pts, quads, outward), 1 if half else 0)


def box(length: float, beam: float, draft: float, nx: int, ny: int, nz: int,
        half: bool = False) -> Mesh:
    """Rectangular barge centred on the origin: bottom ``nx x ny``, sides
    ``nz`` panels high.  ``half`` keeps y >= 0 with ``isym = 1``."""
    pool = _VertexPool(max(length, beam, draft))
    xs = np.linspace(-length / 2, length / 2, nx + 1)
    y_lo = 0.0 if half else -beam / 2
    ys = np.linspace(y_lo, beam / 2, ny + 1)
    zs = np.linspace(-draft, 0.0, nz + 1)
    quads = []
    for i in range(nx):
        for j in range(ny):
            quads.append([pool.add((xs[i], ys[j], -draft)), pool.add((xs[i + 1], ys[j], -draft)),
                          pool.add((xs[i + 1], ys[j + 1], -draft)), pool.add((xs[i], ys[j + 1], -draft))])
    for k in range(nz):
        for i in range(nx):
            for yv in ([beam / 2] if half else [-beam / 2, beam / 2]):
                quads.append([pool.add((xs[i], yv, zs[k])), pool.add((xs[i + 1], yv, zs[k])),
                              pool.add((xs[i + 1], yv, zs[k + 1])), pool.add((xs[i], yv, zs[k + 1]))])
        for j in range(ny):
            for xv in (-length / 2, length / 2):
                quads.append([pool.add((xv, ys[j], zs[k])), pool.add((xv, ys[j + 1], zs[k])),
                              pool.add((xv, ys[j + 1], zs[k + 1])), pool.add((xv, ys[j], zs[k + 1]))])
    pts = pool.array()
    tol = 1e-9 * max(length, beam, draft)

    def outward(c):
        if abs(c[2] + draft) < tol:
            return (0.0, 0.0, -1.0)
        if abs(abs(c[0]) - length / 2) < tol:
            return (math.copysign(1.0, c[0]), 0.0, 0.0)
        return (0.0, math.copysign(1.0, c[1]), 0.0)

    return Mesh(pts, _orient(pts, quads, outward), 1 if half else 0)


def spheroid(length: float, diameter: float, draft_frac: float, n_u: int, n_v: int,
             half: bool = False) -> Mesh:
    """Prolate spheroid with axis along x, cut at the free surface: the
    wetted part below z = 0, with the centre at depth so that the wetted
    draft is ``draft_frac * diameter``.  Each cross-section x = a u is a
    circle of radius r(u) about (0, zc); its wetted arc is where
    zc + r cos(phi) <= 0, i.e. phi in [phi0(u), 2 pi - phi0(u)] with
    cos(phi0) = -zc / r (the whole circle when -zc >= r).  Sections that are
    dry (r <= zc, above-axis cuts) are left out, so the mesh closes at a
    waterline point where a section dries out and at a tip otherwise."""
    a, b = length / 2, diameter / 2
    zc = b - draft_frac * diameter
    pool = _VertexPool(max(a, b))
    u_max = math.sqrt(max(0.0, 1.0 - (zc / b) ** 2)) if zc > 0 else 1.0
    us = np.linspace(-u_max, u_max, n_u + 1)
    ts = np.linspace(0.0, 1.0, n_v + 1)

    def pt(u, t):
        r = b * math.sqrt(max(0.0, 1.0 - u * u))
        c = -zc / r if r > 0 else (1.0 if zc < 0 else -1.0)
        phi0 = math.acos(max(-1.0, min(1.0, c)))
        phi = phi0 + t * ((math.pi if half else 2 * math.pi - phi0) - phi0)
        z = 0.0 if (r > 0 and -1.0 < c < 1.0 and t in (0.0, 1.0) and not (half and t == 1.0)) \
            else zc + r * math.cos(phi)
        return (a * u, r * math.sin(phi), min(z, 0.0))

    quads = []
    for i in range(n_u):
        for j in range(n_v):
            q = [pool.add(pt(us[i], ts[j])), pool.add(pt(us[i + 1], ts[j])),
                 pool.add(pt(us[i + 1], ts[j + 1])), pool.add(pt(us[i], ts[j + 1]))]
            if len(set(q)) < 3:
                continue
            if len(set(q)) == 3:  # degenerate at a tip: repeat the shared vertex
                seen = []
                for v in q:
                    if v not in seen:
                        seen.append(v)
                q = [seen[0], seen[1], seen[2], seen[0]]
            quads.append(q)
    pts = pool.array()

    def outward(c):
        return (c[0] / (a * a), c[1] / (b * b), (c[2] - zc) / (b * b))

    return Mesh(pts, _orient(pts, quads, outward), 1 if half else 0)


def merge(meshes) -> Mesh:
    """Concatenate bodies into one mesh (same symmetry flag required)."""
    isym = meshes[0].isym
    verts, pans, off = [], [], 0
    for m in meshes:
        if m.isym != isym:
            raise ValueError("there is an inconsistency in the mesh files regarding the xOz symmetries")
        verts.append(m.vertices)
        pans.append(m.panels + off)
        off += m.npoints
    return Mesh(np.concatenate(verts), np.concatenate(pans).astype(np.int32), isym)


def translate(m: Mesh, dx=0.0, dy=0.0, dz=0.0) -> Mesh:
    return Mesh(m.vertices + np.array([dx, dy, dz]), m.panels.copy(), m.isym)


def closed_box(length: float, width: float, height: float, nx: int, ny: int, nz: int,
               center=(0.0, 0.0, 0.0)) -> Mesh:
    """Fully submerged closed box (all six faces), axes along x/y/z, outward
    normals; ``nx x ny`` panels on top and bottom, ``nz`` panels high."""
    cx, cy, cz = center
    pool = _VertexPool(max(length, width, height))
    xs = np.linspace(cx - length / 2, cx + length / 2, nx + 1)
    ys = np.linspace(cy - width / 2, cy + width / 2, ny + 1)
    zs = np.linspace(cz - height / 2, cz + height / 2, nz + 1)
    quads = []
    for zv in (zs[0], zs[-1]):
        for i in range(nx):
            for j in range(ny):
                quads.append([pool.add((xs[i], ys[j], zv)), pool.add((xs[i + 1], ys[j], zv)),
                              pool.add((xs[i + 1], ys[j + 1], zv)), pool.add((xs[i], ys[j + 1], zv))])
    for k in range(nz):
        for i in range(nx):
            for yv in (ys[0], ys[-1]):
                quads.append([pool.add((xs[i], yv, zs[k])), pool.add((xs[i + 1], yv, zs[k])),
                              pool.add((xs[i + 1], yv, zs[k + 1])), pool.add((xs[i], yv, zs[k + 1]))])
        for j in range(ny):
            for xv in (xs[0], xs[-1]):
                quads.append([pool.add((xv, ys[j], zs[k])), pool.add((xv, ys[j + 1], zs[k])),
                              pool.add((xv, ys[j + 1], zs[k + 1])), pool.add((xv, ys[j], zs[k + 1]))])
    pts = pool.array()
    tol = 1e-9 * max(length, width, height)

    def outward(c):
        if abs(abs(c[2] - cz) - height / 2) < tol:
            return (0.0, 0.0, math.copysign(1.0, c[2] - cz))
        if abs(abs(c[0] - cx) - length / 2) < tol:
            return (math.copysign(1.0, c[0] - cx), 0.0, 0.0)
        return (0.0, math.copysign(1.0, c[1] - cy), 0.0)

    return Mesh(pts, _orient(pts, quads, outward), 0)


def semisub(col_radius: float = 6.0, spacing: float = 50.0, pontoon_w: float = 8.0, pontoon_h: float = 6.0,
            draft: float = 20.0, n_theta: int = 48, n_z: int = 30, n_r: int = 8, nx: int = 39, ny: int = 8,
            nz: int = 6) -> Mesh:
    """Four-column semi-submersible (no symmetry): vertical columns of radius
    ``col_radius`` at (+-s/2, +-s/2), s = ``spacing``, and closed box
    pontoons ``pontoon_w`` wide x ``pontoon_h`` high at the keel joining
    neighbouring columns between their walls.  The default is 4 x 1824
    column panels + 4 x 1188 pontoon panels = 12,048 panels.  The pontoon
    ends touch the column walls without a shared seam (a common
    simplification for synthetic meshes)."""
    h = spacing / 2
    cols = [cylinder(col_radius, draft, n_theta, n_z, n_r, half=False, x0=sx * h, y0=sy * h)
            for sx in (-1, 1) for sy in (-1, 1)]
    length = spacing - 2 * col_radius
    zc = -draft + pontoon_h / 2
    pons = []
    for sy in (-1, 1):  # along x at y = +-s/2
        pons.append(closed_box(length, pontoon_w, pontoon_h, nx, ny, nz, center=(0.0, sy * h, zc)))
    for sx in (-1, 1):  # along y at x = +-s/2
        pons.append(closed_box(pontoon_w, length, pontoon_h, ny, nx, nz, center=(sx * h, 0.0, zc)))
    return merge(cols + pons)

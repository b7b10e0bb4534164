"""Request loop on one GPU: many independent solves of different sizes.

``Server`` wraps the nemoh_server_* C ABI (the loop itself is native:
worker threads, each owning a solver handle and HIP stream, take one
(request, frequency) at a time first come first served while the device
memory that the workers' handles hold fits the budget).
``two_cylinder_request`` draws the serving workload: two vertical
cylinders of one random radius r ~ U(2, 15) m at a spacing U(3, 10) r,
N_total log-uniform in [200, 16000] panels, 12 degrees of freedom (six
per body), 1-4 headings and 1-64 frequencies, infinite depth.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import bodycond as bcm
from . import core
from . import mesh as nmesh


@dataclass
class Request:
    """One solve request: mesh, frequencies and the problems at each."""
    mesh: nmesh.Mesh
    depth: float
    omegas: np.ndarray      # (nfreq,) rad/s
    nvel: np.ndarray        # (nfreq, ntot, nrhs) complex normal velocities
    nds: np.ndarray         # (nint, ntot) force weights
    rho: float = 1025.0
    g: float = 9.81

    @property
    def npanels(self) -> int:
        return self.mesh.npanels * 2 ** self.mesh.isym

    @property
    def nfreq(self) -> int:
        return self.omegas.size


@dataclass
class WaveRequest:
    """A request whose problems are generated on the device: diffraction at
    each heading then radiation for each degree of freedom, at every
    frequency; forces on the force degrees of freedom (default: the
    radiation ones) and the Froude-Krylov forces."""
    mesh: nmesh.Mesh
    depth: float
    omegas: np.ndarray      # (nfreq,) rad/s
    betas: np.ndarray       # (nbeta,) headings, radians
    radiation: list         # bodycond.Mode
    forces: list = None     # bodycond.Mode (None: the radiation ones)
    body_of_panel: np.ndarray = None  # (npanels,) 1-based body of each panel
    rho: float = 1025.0
    g: float = 9.81
    xeff: float = 0.0
    yeff: float = 0.0

    @property
    def npanels(self) -> int:
        return self.mesh.npanels * 2 ** self.mesh.isym

    @property
    def nfreq(self) -> int:
        return self.omegas.size


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Server:
    def __init__(self, workers: int = 4, budget_bytes: float | None = None, tables=None, batch: int = 1):
        """batch: most frequencies of one wave request a worker solves at
        once, as one batch of handles (nemoh_server_set_batch; 1, the
        default, takes one frequency at a time: on the batched-bodies and
        serving workloads, whose requests each bring a new mesh, re-binding
        and re-tabling several handles per request cost more than the
        batches gained)."""
        L = core.load()
        if tables is None:
            tables = core.build_tables()
        xr, xz, (a1x, a1z, a2x, a2z) = tables
        self._tabs = [_f64(xr), _f64(xz)] + [np.asfortranarray(np.asarray(a, dtype=np.float64)).ravel(order="F")
                                             for a in (a1x, a1z, a2x, a2z)]
        if budget_bytes is None:
            budget_bytes = 32.0 * 2 ** 30
        h = C.c_void_p()
        r = L.nemoh_server_create(C.byref(h), int(workers), float(budget_bytes), *[core._p(a) for a in self._tabs])
        if r:
            raise core.NemohError(r, "nemoh_server_create failed (is a HIP device present?)")
        self.h = h
        r = L.nemoh_server_set_batch(h, int(batch))
        if r:
            L.nemoh_server_destroy(h)
            self.h = None
            raise core.NemohError(r, "nemoh_server_set_batch: batch out of range")
        self._shapes = {}

    def submit(self, req, precision: int = core.MIXED) -> int:
        if isinstance(req, WaveRequest):
            return self._submit_waves(req, precision)
        m = req.mesh
        v = _f64(m.vertices)
        pc = [np.ascontiguousarray(m.panels[:, k] + 1, dtype=np.int32) for k in range(4)]
        nfreq, ntot, nrhs = req.nvel.shape
        if ntot != req.npanels or req.nds.shape[1] != ntot or req.omegas.size != nfreq:
            raise ValueError("request arrays do not match the mesh")
        periods = _f64(2.0 * np.pi / np.asarray(req.omegas, dtype=np.float64))
        # per frequency (ntot, nrhs) column-major, interleaved complex
        nv = np.ascontiguousarray(np.asarray(req.nvel, dtype=np.complex128).transpose(0, 2, 1))
        nds = _f64(np.asarray(req.nds).T)  # (nint, ntot) column-major
        nint = req.nds.shape[0]
        rid = C.c_longlong()
        r = core.load().nemoh_server_submit(
            self.h, v.shape[0], m.npanels, int(m.isym), core._p(_f64(v[:, 0])), core._p(_f64(v[:, 1])),
            core._p(_f64(v[:, 2])), *[core._ip_(a) for a in pc], float(req.depth), float(req.rho), float(req.g),
            int(precision), nfreq, core._p(periods), nrhs, nv.ctypes.data_as(core._dp), nint, core._p(nds),
            C.byref(rid))
        if r:
            raise core.NemohError(r, "nemoh_server_submit: bad request")
        self._shapes[rid.value] = (nfreq, nrhs, nint, 0)
        return rid.value

    def _submit_waves(self, req: WaveRequest, precision: int) -> int:
        m = req.mesh
        v = _f64(m.vertices)
        pc = [np.ascontiguousarray(m.panels[:, k] + 1, dtype=np.int32) for k in range(4)]
        om = _f64(req.omegas)
        be = _f64(np.atleast_1d(req.betas)) if len(req.betas) else np.zeros(1)
        nbeta = len(req.betas)
        rad, nrad = core._dofs(req.radiation)
        frc, nforce = core._dofs(req.radiation if req.forces is None else req.forces)
        body = None if req.body_of_panel is None else np.ascontiguousarray(req.body_of_panel, dtype=np.int32)
        rid = C.c_longlong()
        r = core.load().nemoh_server_submit_waves(
            self.h, v.shape[0], m.npanels, int(m.isym), core._p(_f64(v[:, 0])), core._p(_f64(v[:, 1])),
            core._p(_f64(v[:, 2])), *[core._ip_(a) for a in pc], float(req.depth), float(req.rho), float(req.g),
            int(precision), om.size, core._p(om), nbeta, core._p(be), None if body is None else core._ip_(body),
            nrad, rad, nforce, frc, float(req.xeff), float(req.yeff), C.byref(rid))
        if r:
            raise core.NemohError(r, "nemoh_server_submit_waves: bad request")
        self._shapes[rid.value] = (om.size, nbeta + nrad, nforce, nbeta)
        return rid.value

    def wait(self, rid: int, with_fk: bool = False):
        """Forces (nfreq, nint, nrhs) complex and the request's timings (ms):
        latency (submit to done), queue (submit to first start), service
        (first start to done).  with_fk: also the Froude-Krylov forces
        (nfreq, nint, nbeta) of a wave request, as the middle element."""
        nfreq, nrhs, nint, nbeta = self._shapes.pop(rid)
        out = np.zeros((nfreq, nrhs, nint), np.complex128)
        fk = np.zeros((nfreq, max(nbeta, 1), nint), np.complex128)
        t = np.zeros(3)
        r = core.load().nemoh_server_wait(self.h, C.c_longlong(rid), out.ctypes.data_as(core._dp),
                                          fk.ctypes.data_as(core._dp), core._p(t))
        if r:
            raise core.NemohError(r, core.load().nemoh_server_last_error(self.h).decode())
        timing = dict(zip(("latency", "queue", "service"), t.tolist()))
        if with_fk:
            return out.transpose(0, 2, 1), fk[:, :nbeta].transpose(0, 2, 1), timing
        return out.transpose(0, 2, 1), timing

    def memory(self):
        """(bytes the workers' handles hold now, largest total charged against
        the budget while other frequencies ran)."""
        held, peak = C.c_double(), C.c_double()
        r = core.load().nemoh_server_memory(self.h, C.byref(held), C.byref(peak))
        if r:
            raise core.NemohError(r, "nemoh_server_memory failed")
        return held.value, peak.value

    def close(self):
        if getattr(self, "h", None):
            core.load().nemoh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _cylinder_for(npan: int, radius: float, draft: float, x0: float) -> nmesh.Mesh:
    """Full vertical cylinder of about npan panels, side panels near square:
    npan = n_theta (n_z + n_r) with n_z ~ n_theta draft / (2 pi r) and
    n_r ~ n_theta / (2 pi)."""
    f = (draft / radius + 1.0) / (2.0 * math.pi)
    n_theta = max(8, 4 * round(math.sqrt(npan / f) / 4))
    n_r = max(1, round(n_theta / (2.0 * math.pi)))
    n_z = max(1, round(npan / n_theta) - n_r)
    return nmesh.cylinder(radius, draft, n_theta, n_z, n_r, half=False, x0=x0)


def two_cylinder_parts(rng: np.random.Generator, n_range=(200, 16000), max_freq: int = 64,
                       max_headings: int = 4):
    """The serving workload's draw: (mesh, body_of_panel, 12 modes, omegas,
    headings in radians)."""
    r = rng.uniform(2.0, 15.0)
    spacing = rng.uniform(3.0, 10.0) * r
    n_total = math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))
    nh = int(rng.integers(1, max_headings + 1))
    nf = int(rng.integers(1, max_freq + 1))
    draft = r
    a = _cylinder_for(int(n_total / 2), r, draft, -spacing / 2)
    b = _cylinder_for(int(n_total / 2), r, draft, spacing / 2)
    m = nmesh.merge([a, b])
    body = np.concatenate([np.ones(a.npanels, int), np.full(b.npanels, 2)])
    modes = bcm.six_dof(axis=(-spacing / 2, 0.0, 0.0), body=1) + bcm.six_dof(axis=(spacing / 2, 0.0, 0.0), body=2)
    omegas = np.linspace(0.3, 2.0, nf) if nf > 1 else np.array([0.8])
    betas = np.deg2rad([180.0 * k / nh for k in range(nh)])
    return m, body, modes, omegas, betas


def _arrays(m, body, modes, omegas, betas, depth, g, rho) -> Request:
    rad = np.stack([bcm.radiation(m, md, body) for md in modes], axis=1)
    nh = len(betas)
    nvel = np.empty((omegas.size, m.npanels * 2 ** m.isym, nh + len(modes)), np.complex128)
    for i, w in enumerate(omegas):
        for j, b in enumerate(betas):
            nvel[i, :, j] = bcm.diffraction(m, w, b, depth, g, rho)[0]
        nvel[i, :, nh:] = rad
    return Request(m, depth, omegas, nvel, bcm.integration(m, modes, body), rho, g)


def two_cylinder_request(rng: np.random.Generator, n_range=(200, 16000), max_freq: int = 64,
                         max_headings: int = 4, g: float = 9.81, rho: float = 1025.0) -> Request:
    """The serving workload with its problems as host arrays."""
    m, body, modes, omegas, betas = two_cylinder_parts(rng, n_range, max_freq, max_headings)
    return _arrays(m, body, modes, omegas, betas, 0.0, g, rho)


def two_cylinder_waves(rng: np.random.Generator, n_range=(200, 16000), max_freq: int = 64,
                       max_headings: int = 4, g: float = 9.81, rho: float = 1025.0) -> WaveRequest:
    """The serving workload with its problems generated on the device."""
    m, body, modes, omegas, betas = two_cylinder_parts(rng, n_range, max_freq, max_headings)
    return WaveRequest(m, 0.0, omegas, betas, modes, None, body, rho, g)


def random_body_parts(rng: np.random.Generator, n_range=(300, 6000), nfreq: int = 8):
    """One body of the batched workload: a cylinder, box or prolate spheroid
    (equally likely) with dimensions U(2, 50) m and a draft of U(0.2, 0.8)
    of its height, meshed to about N panels (N log-uniform in n_range);
    6 DOF about the waterplane centre + heading 0 (7 right-hand sides),
    nfreq frequencies uniform in 0.2-2.0 rad/s, infinite depth.  Returns
    (mesh, modes, omegas, headings in radians)."""
    kind = int(rng.integers(0, 3))
    length, width, height = rng.uniform(2.0, 50.0, 3)
    draft = rng.uniform(0.2, 0.8) * height
    target = math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))
    if kind == 0:  # vertical cylinder of diameter `width`
        r = width / 2  # n_theta (n_z + n_r) panels: every bottom ring has n_theta sectors
        h = math.sqrt(2 * math.pi * r * (draft + r) / target)
        m = nmesh.cylinder(r, draft, max(8, 4 * round(2 * math.pi * r / h / 4)), max(1, round(draft / h)),
                           max(1, round(r / h)), half=False)
    elif kind == 1:  # box length x width x draft
        h = math.sqrt((length * width + 2 * draft * (length + width)) / target)
        m = nmesh.box(length, width, draft, max(1, round(length / h)), max(1, round(width / h)),
                      max(1, round(draft / h)))
    else:  # prolate spheroid, axis along x, diameter min(width, length)
        d = min(width, length)
        n_u = max(4, round(math.sqrt(target * length / (length + d))))
        m = nmesh.spheroid(max(length, d * 1.01), d, draft / height, n_u, max(4, round(target / n_u)))
    return m, bcm.six_dof(), np.linspace(0.2, 2.0, nfreq), np.zeros(1)


def random_body_request(rng: np.random.Generator, n_range=(300, 6000), nfreq: int = 8, g: float = 9.81,
                        rho: float = 1025.0) -> Request:
    """The batched-bodies workload with its problems as host arrays."""
    m, modes, omegas, betas = random_body_parts(rng, n_range, nfreq)
    return _arrays(m, None, modes, omegas, betas, 0.0, g, rho)


def random_body_waves(rng: np.random.Generator, n_range=(300, 6000), nfreq: int = 8, g: float = 9.81,
                      rho: float = 1025.0) -> WaveRequest:
    """The batched-bodies workload with its problems generated on the device."""
    m, modes, omegas, betas = random_body_parts(rng, n_range, nfreq)
    return WaveRequest(m, 0.0, omegas, betas, modes, None, None, rho, g)

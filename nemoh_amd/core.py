"""ctypes bindings to libnemoh_core.so (the C ABI of include/nemoh_core.h).

The HIP library is the product; this module only marshals numpy arrays.
Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libnemoh_core.so"

OK, ERR_ARG, ERR_HIP, ERR_PIVOT, ERR_OMEGA, ERR_FIT, ERR_PANEL, ERR_CONVERGENCE, ERR_WAVENUMBER = range(9)
FP64, MIXED = 0, 1
OPT_STEPWISE_SOLVE = 1
OPT_FP32_ASSEMBLY = 2
OPT_SOLVE_SPIN_LIMIT = 3
OPT_GRAPHS = 4
OPT_LU_RANK = 5
OPT_LU_LOOKAHEAD = 6
OPT_LU_FUSED_PANEL = 7
OPT_PAIRED_ASSEMBLY = 8
OPT_LU_BAND = 9
OPT_LU_TUNED = 10
TUNE_CANDIDATES = 6

EXPORTS = (
    "nemoh_core_create", "nemoh_core_solve", "nemoh_core_solve_multi", "nemoh_core_solve_frequency",
    "nemoh_core_freq_setup", "nemoh_core_set_precision", "nemoh_core_set_option", "nemoh_core_get_stats", "nemoh_core_enable_timing",
    "nemoh_core_last_error", "nemoh_core_destroy", "nemoh_core_prepare_mesh", "nemoh_core_build_tables",
    "nemoh_core_wavenumber", "nemoh_core_theta_rule", "nemoh_core_prony", "nemoh_core_pairs",
    "nemoh_core_mfma_probe", "nemoh_core_get_extent", "nemoh_core_set_dofs", "nemoh_core_body_conditions",
    "nemoh_core_solve_waves", "nemoh_core_incident_wavenumber", "nemoh_core_kochin", "nemoh_core_kochin_noalloc", "nemoh_core_potential", "nemoh_server_create", "nemoh_server_submit",
    "nemoh_server_submit_waves", "nemoh_server_wait", "nemoh_server_memory", "nemoh_server_set_batch",
    "nemoh_server_last_error",
    "nemoh_server_destroy", "nemoh_batch_create", "nemoh_batch_solve_waves", "nemoh_batch_last_error",
    "nemoh_batch_destroy", "nemoh_core_autotune", "nemoh_batch_solve_waves_multi", "nemoh_batch_set_timing",
    "nemoh_batch_update_time",
)


class NemohError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"nemoh_core error {code}: {msg}")
        self.code = code


class FreqSetup(C.Structure):
    _fields_ = [("finite_depth", C.c_int), ("amh", C.c_double), ("nexp", C.c_int),
                ("ambda", C.c_double * 31), ("ar", C.c_double * 31), ("nq", C.c_int),
                ("qq", C.c_double * 101), ("cq", C.c_double * 101)]


class Dof(C.Structure):
    """nemoh_dof: body (1-based), kind 1 translation / 2 rotation, dir, axis."""
    _fields_ = [("body", C.c_int), ("kind", C.c_int), ("dir", C.c_double * 3), ("axis", C.c_double * 3)]


def _dofs(modes):
    """bodycond.Mode list -> nemoh_dof array (None when empty)."""
    if not modes:
        return None, 0
    arr = (Dof * len(modes))()
    for a, m in zip(arr, modes):
        a.body, a.kind = int(m.body), int(m.kind)
        a.dir[:] = [float(v) for v in m.direction]
        a.axis[:] = [float(v) for v in m.axis]
    return arr, len(modes)


class Stats(C.Structure):
    _fields_ = [("ms_assembly", C.c_double), ("ms_factor", C.c_double), ("ms_solve", C.c_double),
                ("ms_total", C.c_double), ("refine_iterations", C.c_int), ("refine_residual", C.c_double),
                ("cache_hit", C.c_int), ("ms_lu_update", C.c_double), ("lu_update_launches", C.c_int),
                ("lu_update_flops", C.c_double), ("lu_precision", C.c_int),
                ("solve_fallbacks", C.c_int), ("graph_replays", C.c_int), ("lu_rank", C.c_int),
                ("lu_band", C.c_int)]


class TuneEntry(C.Structure):
    _fields_ = [("rank", C.c_int), ("band", C.c_int), ("ms", C.c_double)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None


def load():
    """Load the built library (build it with nemoh_amd.build / build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m nemoh_amd.build`")
    L = C.CDLL(str(LIB_PATH))
    L.nemoh_core_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int] + [_dp] * 3 + [_ip] * 4 + \
        [_dp] * 8 + [C.c_double] * 4 + [_dp] * 6
    L.nemoh_core_solve.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp,
                                   _dp, _dp, _dp, _dp]
    L.nemoh_core_solve_multi.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _dp, _dp, C.c_int, _dp, _dp,
                                         C.c_int, _dp, _dp, C.c_int, _dp, _dp]
    L.nemoh_core_solve_frequency.argtypes = [C.c_void_p, C.c_double, C.c_int, _dp, _dp, C.c_int, _dp, _dp]
    L.nemoh_core_freq_setup.argtypes = [C.c_void_p, C.c_double, C.POINTER(FreqSetup)]
    L.nemoh_core_set_precision.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
    L.nemoh_core_set_option.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.nemoh_core_get_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
    L.nemoh_core_enable_timing.argtypes = [C.c_void_p, C.c_int]
    L.nemoh_core_last_error.argtypes = [C.c_void_p]
    L.nemoh_core_last_error.restype = C.c_char_p
    L.nemoh_core_destroy.argtypes = [C.c_void_p]
    L.nemoh_core_destroy.restype = None
    L.nemoh_core_prepare_mesh.argtypes = [C.c_int] * 3 + [_dp] * 3 + [_ip] * 4 + [_dp] * 9
    L.nemoh_core_build_tables.argtypes = [_dp] * 6
    L.nemoh_core_wavenumber.argtypes = [C.c_double]
    L.nemoh_core_wavenumber.restype = C.c_double
    L.nemoh_core_theta_rule.argtypes = [C.c_double, _dp, _dp]
    L.nemoh_core_prony.argtypes = [C.c_double, C.c_double, _ip, _dp, _dp]
    L.nemoh_core_pairs.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp]
    L.nemoh_core_mfma_probe.argtypes = [C.c_int, _dp]
    L.nemoh_core_get_extent.argtypes = [C.c_void_p, _dp]
    L.nemoh_core_set_dofs.argtypes = [C.c_void_p, _ip, C.c_int, C.POINTER(Dof), C.c_int, C.POINTER(Dof), C.c_double,
                                      C.c_double]
    L.nemoh_core_body_conditions.argtypes = [C.c_void_p, C.c_double, C.c_int, _dp, _dp, _dp, _dp]
    L.nemoh_core_solve_waves.argtypes = [C.c_void_p, C.c_double, C.c_int, _dp, _dp, _dp]
    L.nemoh_core_kochin.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _dp, _dp]
    L.nemoh_core_kochin_noalloc.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _dp, _dp]
    L.nemoh_core_potential.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    L.nemoh_core_incident_wavenumber.argtypes = [C.c_double, C.c_double, C.c_double]
    L.nemoh_core_incident_wavenumber.restype = C.c_double
    L.nemoh_server_memory.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.nemoh_server_set_batch.argtypes = [C.c_void_p, C.c_int]
    L.nemoh_server_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_double] + [_dp] * 6
    L.nemoh_server_submit.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int] + [_dp] * 3 + [_ip] * 4 + \
        [C.c_double] * 3 + [C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, C.POINTER(C.c_longlong)]
    L.nemoh_server_submit_waves.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int] + [_dp] * 3 + [_ip] * 4 + \
        [C.c_double] * 3 + [C.c_int, C.c_int, _dp, C.c_int, _dp, _ip, C.c_int, C.POINTER(Dof), C.c_int,
                            C.POINTER(Dof), C.c_double, C.c_double, C.POINTER(C.c_longlong)]
    L.nemoh_server_wait.argtypes = [C.c_void_p, C.c_longlong, _dp, _dp, _dp]
    L.nemoh_server_last_error.argtypes = [C.c_void_p]
    L.nemoh_server_last_error.restype = C.c_char_p
    L.nemoh_server_destroy.argtypes = [C.c_void_p]
    L.nemoh_server_destroy.restype = None
    L.nemoh_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p), C.c_int, C.c_int]
    L.nemoh_batch_solve_waves.argtypes = [C.c_void_p, C.c_double, C.c_int, _dp, _dp, _dp]
    L.nemoh_batch_solve_waves_multi.argtypes = [C.c_void_p, _dp, C.c_int, _dp, _dp, _dp]
    L.nemoh_batch_set_timing.argtypes = [C.c_void_p, C.c_int]
    L.nemoh_batch_update_time.argtypes = [C.c_void_p, _dp, _dp, _dp, C.POINTER(C.c_int)]
    L.nemoh_batch_last_error.argtypes = [C.c_void_p]
    L.nemoh_batch_last_error.restype = C.c_char_p
    L.nemoh_core_autotune.argtypes = [C.c_void_p, C.c_int, C.POINTER(TuneEntry), C.POINTER(C.c_int)]
    L.nemoh_batch_destroy.argtypes = [C.c_void_p]
    L.nemoh_batch_destroy.restype = None
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _ip_(a):
    return a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# -- host routines of the reference (no GPU needed) -------------------------
def wavenumber(ak: float) -> float:
    """X0: root of x tanh x = ak (COMPUTE_GREEN_FD.f90:1635)."""
    return load().nemoh_core_wavenumber(float(ak))


def incident_wavenumber(w: float, depth: float, g: float) -> float:
    """The preProcessor's wave number (Common/Environment.f90:45-84)."""
    return load().nemoh_core_incident_wavenumber(float(w), float(depth), float(g))


def theta_rule(akk: float):
    qq, cq = np.zeros(101), np.zeros(101)
    n = load().nemoh_core_theta_rule(float(akk), _p(qq), _p(cq))
    return n, qq[:n].copy(), cq[:n].copy()


def prony(akh: float, amh: float):
    """LISC: exponents and amplitudes of the finite-depth kernel fit."""
    n = C.c_int()
    ambda, ar = np.zeros(31), np.zeros(31)
    r = load().nemoh_core_prony(float(akh), float(amh), C.byref(n), _p(ambda), _p(ar))
    if r:
        raise NemohError(r, "exponential fit of the finite-depth kernel failed")
    return n.value, ambda[:n.value].copy(), ar[:n.value].copy()


def prepare_mesh(vertices, panels, isym: int) -> dict:
    """GOMG: collocation points, normals, areas, TDIS and ZER (0-based panels)."""
    v = _f64(vertices)
    p = np.ascontiguousarray(panels, dtype=np.int32) + 1
    nfa = p.shape[0]
    outs = [np.zeros(nfa) for _ in range(8)]
    zer = C.c_double()
    cols = [np.ascontiguousarray(p[:, k]) for k in range(4)]
    r = load().nemoh_core_prepare_mesh(v.shape[0], nfa, int(isym), _p(_f64(v[:, 0])), _p(_f64(v[:, 1])),
                                       _p(_f64(v[:, 2])), *[_ip_(c) for c in cols], *[_p(o) for o in outs],
                                       C.byref(zer))
    if r:
        raise NemohError(r, "Error: area of a panel is too small")
    geo = dict(zip(("xg", "yg", "zg", "xn", "yn", "zn", "aire", "tdis"), outs))
    geo["zer"] = zer.value
    return geo


def build_tables():
    """CREK on the GPU: xr(328), xz(46) and the four (328, 46) tables."""
    xr, xz = np.zeros(328), np.zeros(46)
    t = [np.zeros(328 * 46) for _ in range(4)]
    r = load().nemoh_core_build_tables(_p(xr), _p(xz), *[_p(a) for a in t])
    if r:
        raise NemohError(r, "table build failed")
    return xr, xz, tuple(a.reshape(46, 328).T.copy() for a in t)


def mfma_probe(fp32: bool = False) -> np.ndarray:
    out = np.zeros(256)
    r = load().nemoh_core_mfma_probe(int(fp32), _p(out))
    if r:
        raise NemohError(r, "mfma probe failed")
    return out.reshape(16, 16)


# -- the solver handle -------------------------------------------------------
class Core:
    """A device-resident body: the handle of nemoh_core_create."""

    def __init__(self, vertices, panels, isym, depth, rho=1025.0, g=9.81, geometry=None, tables=None):
        L = load()
        v = _f64(vertices)
        p = np.ascontiguousarray(panels, dtype=np.int32)
        self.nfa, self.np_, self.isym = p.shape[0], v.shape[0], int(isym)
        self.ntot = self.nfa * 2 ** self.isym
        self.depth, self.rho, self.g = float(depth), float(rho), float(g)
        geo = geometry if geometry is not None else prepare_mesh(v, p, isym)
        if tables is None:
            tables = build_tables()
        xr, xz, (a1x, a1z, a2x, a2z) = tables
        self.geo, self.tables = geo, tables
        keep = [_f64(v[:, 0]), _f64(v[:, 1]), _f64(v[:, 2])]
        pc = [np.ascontiguousarray(p[:, k] + 1, dtype=np.int32) for k in range(4)]
        g8 = [_f64(geo[k]) for k in ("xg", "yg", "zg", "xn", "yn", "zn", "aire", "tdis")]
        tf = [_f64(xr), _f64(xz)] + [np.asfortranarray(np.asarray(a, dtype=np.float64)).ravel(order="F")
                                     for a in (a1x, a1z, a2x, a2z)]
        h = C.c_void_p()
        r = L.nemoh_core_create(C.byref(h), self.np_, self.nfa, self.isym, *[_p(a) for a in keep],
                                *[_ip_(a) for a in pc], *[_p(a) for a in g8], float(geo["zer"]), self.g, self.rho,
                                self.depth, *[_p(a) for a in tf])
        if r:
            raise NemohError(r, "nemoh_core_create failed (is a HIP device present?)")
        self.h = h
        self._last_nrhs = 0

    def close(self):
        if getattr(self, "h", None):
            load().nemoh_core_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, r):
        if r:
            raise NemohError(r, load().nemoh_core_last_error(self.h).decode())

    def autotune(self, reps: int = 3):
        """Times the factorisation forms (update rank x tile band) at this
        handle's size and precision; returns ([(rank, band, ms)], chosen
        index).  The choice goes to the process's table and this handle uses
        it from now on."""
        e = (TuneEntry * TUNE_CANDIDATES)()
        k = C.c_int()
        self._check(load().nemoh_core_autotune(self.h, int(reps), e, C.byref(k)))
        return [(x.rank, x.band, x.ms) for x in e], k.value

    def set_precision(self, mode: int, tol: float = 0.0, max_iter: int = 0):
        self._check(load().nemoh_core_set_precision(self.h, int(mode), float(tol), int(max_iter)))

    def set_dofs(self, radiation, forces=None, body_of_panel=None, xeff: float = 0.0, yeff: float = 0.0):
        """Degrees of freedom (bodycond.Mode lists) of the radiation problems
        and of the force integration (default: the same), the body of each
        panel (1-based, default all 1) and the wave reference point."""
        self._nrad = len(radiation)
        forces = radiation if forces is None else forces
        self._nforce = len(forces)
        r, nr = _dofs(radiation)
        f, nf = _dofs(forces)
        body = None if body_of_panel is None else np.ascontiguousarray(body_of_panel, dtype=np.int32)
        self._check(load().nemoh_core_set_dofs(self.h, None if body is None else _ip_(body), nr, r, nf, f,
                                               float(xeff), float(yeff)))

    def body_conditions(self, omega: float, betas):
        """(nvel (ntot, nbeta + nrad), pinc (ntot, nbeta), nds (nforce, ntot))
        generated on the device; betas in radians."""
        b = _f64(np.atleast_1d(betas)) if len(betas) else np.zeros(0)
        nb = b.size
        nvel = np.zeros((self.ntot, nb + self._nrad), np.complex128, order="F")
        pinc = np.zeros((self.ntot, nb), np.complex128, order="F")
        nds = np.zeros((self._nforce, self.ntot), order="F")
        self._check(load().nemoh_core_body_conditions(self.h, float(omega), nb, _p(b) if nb else None,
                                                      nvel.ctypes.data_as(_dp), pinc.ctypes.data_as(_dp),
                                                      nds.ctypes.data_as(_dp)))
        return nvel, pinc, nds

    def solve_waves(self, omega: float, betas):
        """Forces (nforce, nbeta + nrad) and Froude-Krylov forces (nforce,
        nbeta) with the problems generated on the device."""
        b = _f64(np.atleast_1d(betas)) if len(betas) else np.zeros(0)
        nb = b.size
        forces = np.zeros((self._nforce, nb + self._nrad), np.complex128, order="F")
        fk = np.zeros((self._nforce, max(nb, 1)), np.complex128, order="F")
        self._check(load().nemoh_core_solve_waves(self.h, float(omega), nb, _p(b) if nb else None,
                                                  forces.ctypes.data_as(_dp), fk.ctypes.data_as(_dp)))
        self._last_nrhs = nb + self._nrad
        return forces, fk[:, :nb]

    def kochin(self, thetas, xeff: float = 0.0, yeff: float = 0.0, nrhs: int = None) -> np.ndarray:
        """Kochin function (ntheta, nrhs) of the last solve's problems."""
        th = _f64(np.atleast_1d(thetas))
        n = nrhs if nrhs is not None else self._last_nrhs
        out = np.zeros((th.size, n), np.complex128, order="F")
        self._check(load().nemoh_core_kochin(self.h, float(xeff), float(yeff), th.size, _p(th),
                                             out.ctypes.data_as(_dp)))
        return out

    def kochin_noalloc(self, thetas, xeff: float = 0.0, yeff: float = 0.0, nrhs: int = None) -> np.ndarray:
        """As kochin, through the entry point that allocates no device memory
        (at most 4096 angles)."""
        th = _f64(np.atleast_1d(thetas))
        n = nrhs if nrhs is not None else self._last_nrhs
        out = np.zeros((th.size, n), np.complex128, order="F")
        self._check(load().nemoh_core_kochin_noalloc(self.h, float(xeff), float(yeff), th.size, _p(th),
                                                     out.ctypes.data_as(_dp)))
        return out

    def potential(self, points, nrhs: int = None) -> np.ndarray:
        """Potential (npoints, nrhs) of the last solve's sources at field
        points (npoints, 3); the free-surface elevation at z = 0 is
        1j * w / g times it."""
        pts = _f64(np.asarray(points, dtype=np.float64).reshape(-1, 3))
        n = nrhs if nrhs is not None else self._last_nrhs
        out = np.zeros((pts.shape[0], n), np.complex128, order="F")
        self._check(load().nemoh_core_potential(self.h, pts.shape[0], _p(pts), out.ctypes.data_as(_dp)))
        return out

    def extent(self) -> float:
        """GM: the largest horizontal vertex distance (to the y-images too
        when the mesh is symmetric), as the handle computed it."""
        gm = C.c_double()
        self._check(load().nemoh_core_get_extent(self.h, C.byref(gm)))
        return gm.value

    def set_option(self, option: int, value: int):
        self._check(load().nemoh_core_set_option(self.h, int(option), int(value)))

    def enable_timing(self, on=True):
        self._check(load().nemoh_core_enable_timing(self.h, int(on)))

    def stats(self) -> Stats:
        s = Stats()
        self._check(load().nemoh_core_get_stats(self.h, C.byref(s)))
        return s

    def freq_setup(self, period: float) -> FreqSetup:
        f = FreqSetup()
        self._check(load().nemoh_core_freq_setup(self.h, float(period), C.byref(f)))
        return f

    @staticmethod
    def _fargs(f):
        """(amh, nexp, ambda, ar, nq, qq, cq) from a FreqSetup or a dict."""
        if isinstance(f, FreqSetup):
            return (f.amh, f.nexp, np.array(f.ambda[:]), np.array(f.ar[:]), f.nq, np.array(f.qq[:]),
                    np.array(f.cq[:]))
        return (float(f.get("amh", 0.0)), int(f.get("nexp", 0)), _f64(np.resize(f.get("ambda", [0.0]), 31)),
                _f64(np.resize(f.get("ar", [0.0]), 31)), int(f.get("nq", 0)),
                _f64(np.resize(f.get("qq", [0.0]), 101)), _f64(np.resize(f.get("cq", [0.0]), 101)))

    def solve(self, period, freq, nvel):
        """One problem (drop-in of SOLVE_POTENTIAL_*_DIRECT): ZIGB, ZIGS, ZPB, ZPS."""
        amh, nexp, ambda, ar, nq, qq, cq = self._fargs(freq)
        nv = np.ascontiguousarray(nvel, dtype=np.complex128)
        outs = [np.zeros(self.nfa, np.complex128) for _ in range(4)]
        self._check(load().nemoh_core_solve(self.h, float(period), amh, nexp, _p(ambda), _p(ar), nq, _p(qq), _p(cq),
                                            nv.ctypes.data_as(_dp), *[o.ctypes.data_as(_dp) for o in outs]))
        self._last_nrhs = 1
        return outs

    def solve_multi(self, period, freq, nvel, nds=None, want_pressure=True):
        """All problems of one frequency: nvel (ntot, nrhs) -> pressure (ntot,
        nrhs) and forces (nint, nrhs)."""
        amh, nexp, ambda, ar, nq, qq, cq = self._fargs(freq)
        nv = np.asfortranarray(nvel, dtype=np.complex128)
        nrhs = nv.shape[1]
        pres = np.zeros((self.ntot, nrhs), np.complex128, order="F") if want_pressure else None
        nint = 0 if nds is None else nds.shape[0]
        nd = None if nds is None else np.asfortranarray(nds, dtype=np.float64)
        forces = None if nds is None else np.zeros((nint, nrhs), np.complex128, order="F")
        self._check(load().nemoh_core_solve_multi(
            self.h, float(period), amh, nexp, _p(ambda), _p(ar), nq, _p(qq), _p(cq), nrhs, nv.ctypes.data_as(_dp),
            None if pres is None else pres.ctypes.data_as(_dp), nint, None if nd is None else nd.ctypes.data_as(_dp),
            None if forces is None else forces.ctypes.data_as(_dp)))
        self._last_nrhs = nrhs
        return pres, forces

    def solve_frequency(self, period, nvel, nds=None, want_pressure=True):
        """Standalone: frequency set-up by the core, then solve_multi."""
        nv = np.asfortranarray(nvel, dtype=np.complex128)
        nrhs = nv.shape[1]
        pres = np.zeros((self.ntot, nrhs), np.complex128, order="F") if want_pressure else None
        nint = 0 if nds is None else nds.shape[0]
        nd = None if nds is None else np.asfortranarray(nds, dtype=np.float64)
        forces = None if nds is None else np.zeros((nint, nrhs), np.complex128, order="F")
        self._check(load().nemoh_core_solve_frequency(
            self.h, float(period), nrhs, nv.ctypes.data_as(_dp), None if pres is None else pres.ctypes.data_as(_dp),
            nint, None if nd is None else nd.ctypes.data_as(_dp),
            None if forces is None else forces.ctypes.data_as(_dp)))
        self._last_nrhs = nrhs
        return pres, forces

    def pairs(self, period, freq):
        amh, nexp, ambda, ar, nq, qq, cq = self._fargs(freq)
        out = np.zeros(self.nfa * self.nfa * 16)
        self._check(load().nemoh_core_pairs(self.h, float(period), amh, nexp, _p(ambda), _p(ar), nq, _p(qq), _p(cq),
                                            _p(out)))
        return out.reshape(self.nfa, self.nfa, 16)


class Batch:
    """Many bodies solved together (nemoh_batch_*): each heavy step of a
    frequency is one launch over every body.  ``cores``: Core handles with
    their degrees of freedom set; they must stay open while the batch
    exists."""

    def __init__(self, cores, precision: int = FP64, max_rhs: int = 16):
        L = load()
        self.cores = list(cores)
        arr = (C.c_void_p * len(self.cores))(*[c.h for c in self.cores])
        h = C.c_void_p()
        r = L.nemoh_batch_create(C.byref(h), len(self.cores), arr, int(precision), int(max_rhs))
        if r:
            raise NemohError(r, L.nemoh_batch_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().nemoh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_timing(self, on: bool):
        """Events around the grouped rank-128 update launches of the calls
        that follow (sums reset)."""
        load().nemoh_batch_set_timing(self.h, int(bool(on)))

    def update_time(self):
        """(milliseconds, algorithmic flops, algorithmic bytes, launches) of
        the timed update launches since set_timing(True)."""
        ms, fl, by, n = C.c_double(), C.c_double(), C.c_double(), C.c_int()
        load().nemoh_batch_update_time(self.h, C.byref(ms), C.byref(fl), C.byref(by), C.byref(n))
        return ms.value, fl.value, by.value, n.value

    def solve_waves(self, omega, betas):
        """Per body: forces (nforce, nbeta + nrad) and Froude-Krylov forces
        (nforce, nbeta), as Core.solve_waves gives them.  omega: one
        frequency for every body, or one per body."""
        om = np.atleast_1d(np.asarray(omega, dtype=np.float64))
        per_body = om.size > 1
        if per_body and om.size != len(self.cores):
            raise ValueError("one frequency, or one per body")
        b = _f64(np.atleast_1d(betas)) if len(betas) else np.zeros(0)
        nb = b.size
        shapes = [(c._nforce, nb + c._nrad) for c in self.cores]
        total = sum(a * q for a, q in shapes)
        ktotal = sum(c._nforce * nb for c in self.cores)
        f = np.zeros(max(total, 1), np.complex128)
        k = np.zeros(max(ktotal, 1), np.complex128)
        if per_body:
            om = _f64(om)
            r = load().nemoh_batch_solve_waves_multi(self.h, _p(om), nb, _p(b) if nb else None,
                                                     f.ctypes.data_as(_dp), k.ctypes.data_as(_dp))
        else:
            r = load().nemoh_batch_solve_waves(self.h, float(om[0]), nb, _p(b) if nb else None,
                                               f.ctypes.data_as(_dp), k.ctypes.data_as(_dp))
        if r:
            raise NemohError(r, load().nemoh_batch_last_error(self.h).decode())
        out, fo, ko = [], 0, 0
        for c, (a, q) in zip(self.cores, shapes):
            out.append((f[fo:fo + a * q].reshape((a, q), order="F"), k[ko:ko + a * nb].reshape((a, nb), order="F")))
            fo += a * q
            ko += a * nb
            c._last_nrhs = q
        return out

"""ctypes bindings to the C restatement of the NEMOH hot path (the oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's CPU-baseline leg, as the checker.  The product
(``nemoh_amd``) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import math
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
NR, NZ, PAIR_N = 328, 46, 16

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")


class _Ctx(C.Structure):
    _fields_ = [("np", C.c_int), ("nfa", C.c_int), ("nsym", C.c_int)] + \
        [(n, C.c_void_p) for n in ("x", "y", "z", "m1", "m2", "m3", "m4", "xg", "yg", "zg",
                                   "xn", "yn", "zn", "aire", "tdis")] + \
        [(n, C.c_double) for n in ("zer", "depth", "g", "rho")] + \
        [(n, C.c_void_p) for n in ("xr", "xz", "apd1x", "apd1z", "apd2x", "apd2z")]


class Freq(C.Structure):
    _fields_ = [("t", C.c_double), ("amh", C.c_double), ("nexp", C.c_int),
                ("ambda", C.c_double * 31), ("ar", C.c_double * 31), ("nq", C.c_int),
                ("qq", C.c_double * 251), ("cq", C.c_double * 251)]


def build() -> None:
    subprocess.run(["make", "-s", "-f", "oracle/Makefile"], cwd=HERE.parent, check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.ora_crek.argtypes = [_dp] * 6
        L.ora_crek.restype = None
        L.ora_cint.argtypes = [C.c_double, C.POINTER(C.c_int), _dp, _dp]
        L.ora_x0.argtypes = [C.c_double]
        L.ora_x0.restype = C.c_double
        L.ora_lisc.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_int), _dp, _dp]
        L.ora_prepare_mesh.argtypes = [C.c_int] * 3 + [_dp] * 3 + [_ip] * 4 + [_dp] * 8 + [C.POINTER(C.c_double)]
        L.ora_freq_setup.argtypes = [C.POINTER(_Ctx), C.c_double, C.POINTER(C.c_int), C.POINTER(Freq), C.POINTER(C.c_int)]
        L.ora_pairs_all.argtypes = [C.POINTER(_Ctx), C.POINTER(Freq), C.c_int, _dp]
        L.ora_pair_infd.argtypes = [C.POINTER(_Ctx), C.POINTER(Freq), C.c_int, C.c_int, _dp]
        L.ora_pair_fd.argtypes = [C.POINTER(_Ctx), C.POINTER(Freq), C.c_int, C.c_int, _dp]
        L.ora_session_new.argtypes = [C.POINTER(_Ctx)]
        L.ora_session_new.restype = C.c_void_p
        L.ora_session_free.argtypes = [C.c_void_p]
        L.ora_session_undefined_pairs.argtypes = [C.c_void_p]
        L.ora_solve_bvp.argtypes = [C.c_void_p, C.c_double, _cp, _cp, _cp, _cp, _cp, _cp]
        L.ora_gaussz.argtypes = [_cp, C.c_int, C.c_int, C.c_int]
        L.ora_forces.argtypes = [C.c_int, C.c_int, _dp, _cp, _cp]
        L.ora_forces.restype = None
        L.ora_pairs_rows.argtypes = [C.POINTER(_Ctx), C.POINTER(Freq), C.c_int, C.c_int, C.c_int]
        L.ora_pairs_rows.restype = C.c_double
        L.ora_gg_vec.argtypes = [C.c_int, _cp, _cp, _cp]
        L.ora_gg_vec.restype = None
        _lib = L
    return _lib


def gg(z, cex=None) -> np.ndarray:
    z = np.ascontiguousarray(z, dtype=np.complex128).ravel()
    cex = np.exp(z) if cex is None else np.ascontiguousarray(cex, dtype=np.complex128).ravel()
    out = np.empty_like(z)
    lib().ora_gg_vec(z.size, z, cex, out)
    return out


def crek():
    xr, xz = np.zeros(NR), np.zeros(NZ)
    t = [np.zeros(NR * NZ) for _ in range(4)]
    lib().ora_crek(xr, xz, *t)
    return xr, xz, [a.reshape(NZ, NR).T.copy() for a in t]  # (328, 46) each


def cint(akk: float):
    qq, cq = np.zeros(251), np.zeros(251)
    n = C.c_int()
    lib().ora_cint(float(akk), C.byref(n), qq, cq)
    return n.value, qq[:n.value].copy(), cq[:n.value].copy()


def x0(ak: float) -> float:
    return lib().ora_x0(float(ak))


def lisc(akh: float, amh: float):
    ambda, ar = np.zeros(31), np.zeros(31)
    n = C.c_int()
    err = lib().ora_lisc(float(akh), float(amh), C.byref(n), ambda, ar)
    return err, n.value, ambda[:n.value].copy(), ar[:n.value].copy()


def ff(x: float, ak: float, am: float) -> float:
    """FF (COMPUTE_GREEN_FD.f90:1612): the regular part of the FD kernel."""
    L = lib()
    L.ora_ff_pub.argtypes = [C.c_double] * 3
    L.ora_ff_pub.restype = C.c_double
    return L.ora_ff_pub(float(x), float(ak), float(am))


def housrs(a, b):
    """HOUSRS (COMPUTE_GREEN_FD.f90:1454): least-squares solution of
    a (nl x ncc) x = b by Householder reflections."""
    a = np.asarray(a, np.float64)
    nl, ncc = a.shape
    s = np.zeros((32, 120))  # column-major (120, 32)
    s[:ncc, :nl] = a.T
    s[ncc, :nl] = b
    flat = np.ascontiguousarray(s.ravel())
    L = lib()
    L.ora_housrs_pub.argtypes = [_dp, C.c_int, C.c_int]
    if L.ora_housrs_pub(flat, nl, ncc):
        raise ValueError("HOUSRS failed")
    return flat.reshape(32, 120)[ncc, :ncc].copy()


def sprbm(c):
    """SPRBM (COMPUTE_GREEN_FD.f90:1182): roots of sum_i c[i] x^i (c[0] the
    constant term) by Bairstow factors; returns (ier, roots re, roots im)."""
    c = np.ascontiguousarray(c, np.float64).copy()
    ic = c.size
    rr, rc = np.zeros(ic), np.zeros(ic)
    L = lib()
    L.ora_sprbm_pub.argtypes = [_dp, C.c_int, _dp, _dp]
    ier = L.ora_sprbm_pub(c, ic, rr, rc)
    return ier, rr, rc


def kochin(case, kwave: float, xeff: float, yeff: float, thetas, zigb, zigs=None) -> np.ndarray:
    """COMPUTE_KOCHIN at each angle for sources zigb, zigs (NFA each)."""
    L = lib()
    L.ora_kochin_vec.argtypes = [C.POINTER(_Ctx), C.c_double, C.c_double, C.c_double, C.c_int, _dp, _cp, _cp,
                                 _cp]
    th = np.ascontiguousarray(np.atleast_1d(thetas), dtype=np.float64)
    zb = np.ascontiguousarray(zigb, dtype=np.complex128)
    zs = np.ascontiguousarray(np.zeros_like(zb) if zigs is None else zigs, dtype=np.complex128)
    out = np.zeros(th.size, np.complex128)
    L.ora_kochin_vec(C.byref(case.ctx), float(kwave), float(xeff), float(yeff), th.size, th, zb, zs, out)
    return out


def potential_setup(case, period: float):
    """(AM0, AMH, NEXP, AMBDA, AR) that the Solver hands COMPUTE_POTENTIAL_DOMAIN
    for this period (SOLVE_BEM.f90:66-89, SOLVE_BEM_FD_DIRECT.f90:53-100):
    AM0 = w^2/g and AMH = AM0 Depth in infinite depth (Depth = 0 or
    w^2 Depth / g >= 20), else AMH = X0(w^2 Depth / g), AM0 = AMH / Depth and
    LISC's expansion of AMH tanh(AMH).  LISC is ill-conditioned in its
    first argument (one ulp of tanh can change NEXP), so the whole set-up
    comes from the C restatement with the C library's tanh, as the
    reference's does, never from numpy."""
    f, fd, _ = case.freq(period)
    if fd:
        return f.amh / case.depth, f.amh, f.nexp, np.array(f.ambda[:]), np.array(f.ar[:])
    w = 2 * (4 * math.atan(1.0)) / period
    am0 = w * w / case.g
    return am0, am0 * case.depth, 0, np.zeros(31), np.zeros(31)


def potential(case, period: float, points, zigb, zigs=None) -> np.ndarray:
    """COMPUTE_POTENTIAL_DOMAIN at each point (n, 3) for sources zigb, zigs,
    with the period's set-up from potential_setup."""
    am0, amh, nexp, lam, a = potential_setup(case, period)
    L = lib()
    L.ora_potential_vec.argtypes = [C.POINTER(_Ctx), C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp,
                                    C.c_int, _dp, _cp, _cp, _cp]
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    zb = np.ascontiguousarray(zigb, dtype=np.complex128)
    zs = np.ascontiguousarray(np.zeros_like(zb) if zigs is None else zigs, dtype=np.complex128)
    out = np.zeros(pts.shape[0], np.complex128)
    r = L.ora_potential_vec(C.byref(case.ctx), float(period), float(am0), float(amh), int(nexp), lam, a,
                            pts.shape[0], pts.ravel(), zb, zs, out)
    if r:
        raise ValueError("field point too close to the free surface (AKZ > -1.5e-6)")
    return out


def prepare_mesh(vertices, panels, isym):
    v = np.ascontiguousarray(vertices, dtype=np.float64)
    p = np.ascontiguousarray(panels, dtype=np.int32)
    nfa = p.shape[0]
    outs = [np.zeros(nfa) for _ in range(8)]
    zer = C.c_double()
    err = lib().ora_prepare_mesh(v.shape[0], nfa, int(isym), v[:, 0].copy(), v[:, 1].copy(),
                                 v[:, 2].copy(), p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(),
                                 p[:, 3].copy(), *outs, C.byref(zer))
    if err:
        raise ValueError(f"Error: area of panel {err} is too small")
    keys = ("xg", "yg", "zg", "xn", "yn", "zn", "aire", "tdis")
    geo = dict(zip(keys, outs))
    geo["zer"] = zer.value
    return geo


class Case:
    """A mesh + environment bound to the oracle's context struct."""

    def __init__(self, vertices, panels, isym, depth=0.0, rho=1025.0, g=9.81, tables=None,
                 geometry=None):
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        p = np.ascontiguousarray(panels, dtype=np.int32)
        self.isym, self.depth, self.rho, self.g = int(isym), float(depth), float(rho), float(g)
        self.nfa, self.np = p.shape[0], v.shape[0]
        self._a = {"x": v[:, 0].copy(), "y": v[:, 1].copy(), "z": v[:, 2].copy(),
                   "m1": p[:, 0].copy(), "m2": p[:, 1].copy(), "m3": p[:, 2].copy(), "m4": p[:, 3].copy()}
        geo = geometry if geometry is not None else prepare_mesh(v, p, isym)
        self.geo = geo
        for k in ("xg", "yg", "zg", "xn", "yn", "zn", "aire", "tdis"):
            self._a[k] = np.ascontiguousarray(geo[k], dtype=np.float64)
        if tables is None:
            xr, xz, (a1x, a1z, a2x, a2z) = crek()
        else:
            xr, xz, (a1x, a1z, a2x, a2z) = tables
        self.tables = (xr, xz, (a1x, a1z, a2x, a2z))
        # oracle layout: (i, j) -> i + 328 j, i.e. Fortran order of (328, 46)
        for k, a in (("xr", xr), ("xz", xz), ("apd1x", a1x), ("apd1z", a1z), ("apd2x", a2x), ("apd2z", a2z)):
            self._a[k] = np.ascontiguousarray(np.asarray(a, dtype=np.float64).T.ravel() if a.ndim == 2 else a)
        c = _Ctx()
        c.np, c.nfa, c.nsym = self.np, self.nfa, self.isym
        for k, a in self._a.items():
            setattr(c, k, a.ctypes.data)
        c.zer, c.depth, c.g, c.rho = float(geo["zer"]), self.depth, self.g, self.rho
        self.ctx = c

    def freq(self, period: float, nq_state: int = 0):
        f = Freq()
        nq = C.c_int(nq_state)
        fd = C.c_int()
        err = lib().ora_freq_setup(C.byref(self.ctx), float(period), C.byref(nq), C.byref(f), C.byref(fd))
        if err:
            raise RuntimeError(f"oracle frequency set-up failed: code {err}")
        return f, bool(fd.value), nq.value

    def pairs(self, period: float, nq_state: int = 0):
        f, fd, _ = self.freq(period, nq_state)
        out = np.zeros(self.nfa * self.nfa * PAIR_N)
        und = lib().ora_pairs_all(C.byref(self.ctx), C.byref(f), int(fd), out)
        return out.reshape(self.nfa, self.nfa, PAIR_N), und

    def pairs_rows(self, period: float, r0: int, nr: int, nq_state: int = 0) -> float:
        f, fd, _ = self.freq(period, nq_state)
        return lib().ora_pairs_rows(C.byref(self.ctx), C.byref(f), int(fd), int(r0), int(nr))

    def session(self):
        return Session(self)


class Session:
    """The reference's cross-call module state (w_previous, NQ, AInv)."""

    def __init__(self, case: Case):
        self.case = case
        self.h = lib().ora_session_new(C.byref(case.ctx))

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_session_free(self.h)
            self.h = None

    def solve_bvp(self, period: float, nvel):
        c = self.case
        ntot = c.nfa * (2 ** c.isym)
        nvel = np.ascontiguousarray(nvel, dtype=np.complex128)
        assert nvel.shape == (ntot,)
        pressure = np.zeros(ntot, np.complex128)
        zigb, zigs, zpb, zps = (np.zeros(c.nfa, np.complex128) for _ in range(4))
        err = lib().ora_solve_bvp(self.h, float(period), nvel, pressure, zigb, zigs, zpb, zps)
        if err:
            raise RuntimeError(f"oracle SOLVE_BVP failed: code {err}")
        return pressure, zigb, zigs, zpb, zps

    @property
    def undefined_pairs(self) -> int:
        return lib().ora_session_undefined_pairs(self.h)


def forces(nds, pressure):
    nds = np.ascontiguousarray(nds, dtype=np.float64)
    pressure = np.ascontiguousarray(pressure, dtype=np.complex128)
    out = np.zeros(nds.shape[0], np.complex128)
    lib().ora_forces(nds.shape[0], nds.shape[1], nds, pressure, out)
    return out


def gaussz_inverse(a):
    """Inverse of a square complex matrix by the GAUSSZ restatement."""
    a = np.asarray(a, dtype=np.complex128)
    n = a.shape[0]
    aug = np.zeros((2 * n, n), np.complex128)  # column-major (n, 2n) stored transposed
    aug[:n, :] = a.T
    aug[n:, :] = np.eye(n)
    flat = np.ascontiguousarray(aug.ravel())
    if lib().ora_gaussz(flat, n, n, 2 * n):
        raise ZeroDivisionError("PIVOT INFERIEUR A 1e-20")
    return flat.reshape(2 * n, n)[n:].T.copy()

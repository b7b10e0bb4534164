// Device-side influence coefficients of the NEMOH boundary-element method.
//
// For a field point (collocation point of panel i) and a source panel j this
// evaluates the Rankine part (exact flat-panel integral near the panel, one
// point far from it) and the wave part of the Green function (tabulated /
// asymptotic / near-free-surface quadrature in infinite depth; four z-images
// plus exponential (Prony) images in finite depth), returning
//   S+- = SP1 + i SP2 / SM1 + i SM2        (potential of unit source density)
//   V+-[k] = VS{X,Y,Z}{P,M}1 + i VS{X,Y,Z}{P,M}2   (its field-point gradient)
// for the symmetric (+) and antisymmetric (-) y-images.
//
// Reference: Solver/Core/COMPUTE_GREEN_INFD.f90 (VAVINFD :33-220, VNSINFD
// :223-503), Solver/Core/COMPUTE_GREEN_FD.f90 (VAVFD :32-219, VNSFD :222-868,
// VSD :1514-1609).  Branch order, thresholds and table-index rules follow the
// reference exactly so that FP64 results agree to rounding; the arithmetic is
// reorganised for a 64-lane wavefront (all images of one pair in registers,
// both symmetry blocks from one evaluation).
#pragma once
#include <hip/hip_runtime.h>

namespace nemoh {

constexpr int kNR = 328;  // XR grid (r-direction) of the wave tables
constexpr int kNZ = 46;   // XZ grid (z-direction)

// Per-frequency constants, built on the host and read uniformly by every lane.
struct FreqParams {
  int fd;          // 0: infinite depth (INFD), 1: finite depth (FD)
  int nsym;        // 0 / 1 (xOz symmetry)
  int nq;          // theta-rule length used by the near-free-surface branch
  int nexp1;       // FD: number of exponential images + 1
  double ak0;      // INFD: w^2/g ; FD: AM0 = AMH/h
  double depth;    // h (FD)
  double zer;      // ZER = -1e-5 * max extent
  double coe1, coe2, coe3, coe4;  // FD scaling constants (VNSFD:264-269)
  double pik, pik3;               // pi / k and pi / k^3 (k = ak0), for the images' 1/r terms
  double ambda[32], ar[32];       // FD: exponents / amplitudes, [nexp] = (0, 2)
  double qq[101], cq[101];        // theta rule
};

// Resident body geometry (structure of arrays, panel-major).
struct Geometry {
  int n;                 // panels
  const double* xg;      // [n] collocation points
  const double* yg;
  const double* zg;
  const double* xn;      // [n] unit normals
  const double* yn;
  const double* zn;
  const double* aire;    // [n] areas
  const double* tdis;    // [n] max vertex distance from the collocation point
  const double* vx;      // [4n] panel vertices, vx[4*j + l]
  const double* vy;
  const double* vz;
};

// Wave tables: node (i, j) of the (XR, XZ) grid holds the 4 integrals
// (APD1X, APD1Z, APD2X, APD2Z) contiguously at tab[4*(i + 328*j) + k].
struct Tables {
  const double* xr;   // [328]
  const double* xz;   // [46]
  const double* tab;  // [4 * 328 * 46]
  // reciprocals of the Lagrange denominators of the 3-node stencil starting
  // at node k (0-based): rr[3k..3k+2] for XR (k < 326), rz likewise for XZ
  const double* rr;
  const double* rz;
};

struct cplx {
  double re, im;
};
__device__ __forceinline__ cplx cmk(double r, double i) { return cplx{r, i}; }
__device__ __forceinline__ cplx operator+(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx operator-(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx operator*(cplx a, cplx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cplx operator*(double s, cplx a) { return {s * a.re, s * a.im}; }
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {  // Smith's algorithm
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const double r = b.re / b.im, d = b.re * r + b.im;
  return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
__device__ __forceinline__ cplx cexp_(cplx z) {
  const double e = exp(z.re);
  double s, c;
  sincos(z.im, &s, &c);
  return {e * c, e * s};
}
__device__ __forceinline__ cplx clog_(cplx z) { return {log(hypot(z.re, z.im)), atan2(z.im, z.re)}; }

// 1/sqrt(x) for x > 0: the hardware estimate (v_rsq_f64, relative error
// about 5e-8) refined by one Newton step y <- y + y (1 - x y^2) / 2, which
// leaves 1.5 (5e-8)^2, about 4e-15 -- four instructions instead of a
// correctly rounded square root followed by divisions.  (A second step
// would reach about an ulp; the influence coefficients need 1e-12.)
__device__ __forceinline__ double rsqrt_pos(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

// 1/x for x > 0: the hardware estimate (v_rcp_f64) refined by two Newton
// steps y <- y + y (1 - x y), which leave the rounding of the last step.
__device__ __forceinline__ double recip_pos(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

// Complex exponential-integral approximation GG(z) ~ e^z E1(z) +- i pi e^z,
// ELEMENTARY_FNS.f90:31-81 (same regimes and coefficients).
__device__ cplx gg(cplx z, cplx cex) {
  if (z.re + 16.0 <= 0.0) {
    const cplx y = cdiv(cmk(1.0, 0.0), z);
    cplx p = cmk(-120.0, 0.0);
    p = cmk(24.0, 0.0) + y * p;
    p = cmk(-6.0, 0.0) + y * p;
    p = cmk(2.0, 0.0) + y * p;
    p = cmk(-1.0, 0.0) + y * p;
    p = cmk(1.0, 0.0) + y * p;
    return y * p;
  }
  const double t = z.im;
  cplx g;
  if (fabs(t) - 10.0 > 0.0) {
    g = cdiv(cmk(0.711093, 0.0), z + cmk(0.415775, 0.0)) + cdiv(cmk(0.278518, 0.0), z + cmk(2.29428, 0.0)) +
        cdiv(cmk(0.010389, 0.0), z + cmk(6.2900, 0.0));
  } else if (z.re + 0.5 > 0.0) {
    cplx a = cmk(0.97087007E-05, 0.0);
    a = cmk(0.763297E-03, 0.0) + z * a;
    a = cmk(0.206543E-01, 0.0) + z * a;
    a = cmk(0.23721365E+00, 0.0) + z * a;
    a = cmk(1.0, 0.0) + z * a;
    cplx b = cmk(-0.51801555E-03, 0.0);
    b = cmk(0.19387339E-02, 0.0) + z * b;
    b = cmk(-0.3000591E-01, 0.0) + z * b;
    b = cmk(0.41806426E-01, 0.0) + z * b;
    b = cmk(-0.149545886E+01, 0.0) + z * b;
    b = cmk(0.99999207E+00, 0.0) + z * b;
    cplx d = cmk(0.2989204E-03, 0.0);
    d = cmk(-0.8700861E-03, 0.0) + z * d;
    d = cmk(0.12982719E-01, 0.0) + z * d;
    d = cmk(-0.66786033E-01, 0.0) + z * d;
    d = cmk(0.28388363E+00, 0.0) + z * d;
    d = cmk(-0.76273617E+00, 0.0) + z * d;
    d = cmk(1.0, 0.0) + z * d;
    const cplx num = clog_(z) * a + 0.5772156649E+00 * b;
    g = cdiv(cmk(-num.re, -num.im), d);
  } else {
    // degree-7/8 complex rational, coefficients conjugated for Im z >= 0
    const double s = (t < 0.0) ? 1.0 : -1.0;
    cplx n = cmk(1.000000, s * 1.3935496E-06);
    n = z * n + cmk(15.82958, -s * 20.14222);
    n = z * n + cmk(-70.52863, -s * 227.9511);
    n = z * n + cmk(-985.4221, -s * 226.6272);
    n = z * n + cmk(-1202.318, s * 1580.907);
    n = z * n + cmk(953.2441, s * 1342.447);
    n = z * n + cmk(417.3716, -s * 196.6665);
    n = z * n + cmk(-9.881266, -s * 24.24952);
    cplx d = cmk(1.0, 0.0);
    d = z * d + cmk(16.83184, -s * 20.14481);
    d = z * d + cmk(-55.66969, -s * 248.1167);
    d = z * d + cmk(-1068.640, -s * 434.4707);
    d = z * d + cmk(-2082.250, s * 1522.471);
    d = z * d + cmk(383.3455, s * 2730.378);
    d = z * d + cmk(1216.791, s * 351.7189);
    d = z * d + cmk(115.3926, -s * 161.2647);
    d = z * d + cmk(-3.777369, -s * 4.510900);
    g = cdiv(n, d);
  }
  const cplx ipc = cmk(-3.14159265 * cex.im, 3.14159265 * cex.re);  // i*pi*cex
  return (t < 0.0) ? g - ipc : g + ipc;
}

// Rankine (1/r) part of one image of source panel j seen from (xo, yo, zo):
// far field one-point rule, near field exact flat-quadrilateral integral.
// my/mz mirror the panel in y / z, dh shifts the z-image, mjj is the image
// orientation sign, self adds the -2 pi n jump.  VAVINFD.f90:85-184.
struct RankineImage {
  double a, vx, vy, vz;
};

// Near field of one image (RO <= 7 TDIS): the exact flat-quadrilateral
// integral over the four edges.  dx, dy, dz: field point minus the image's
// collocation point.
__device__ __forceinline__ RankineImage rankine_near(const Geometry& G, int j, double xo, double yo, double zo,
                                                     double my, double mz, double dh, double mh, double mjj,
                                                     bool self, double dx, double dy, double dz) {
  const double a3j = G.xn[j], a6j = G.yn[j] * my, a9j = G.zn[j] * mz;
  const double tdis = G.tdis[j];
  RankineImage r;
  double tx[5], ty[5], tz[5], rr[5], drx[5], dry[5], drz[5];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    tx[l] = G.vx[4 * j + l];
    ty[l] = G.vy[4 * j + l] * my;
    tz[l] = G.vz[4 * j + l] * mz - dh * mh;
  }
  tx[4] = tx[0];
  ty[4] = ty[0];
  tz[4] = tz[0];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const double ex = xo - tx[l], ey = yo - ty[l], ez = zo - tz[l];
    const double e2 = ex * ex + ey * ey + ez * ez;
    const double ri = rsqrt_pos(e2);
    rr[l] = e2 * ri;
    drx[l] = ex * ri;
    dry[l] = ey * ri;
    drz[l] = ez * ri;
  }
  rr[4] = rr[0];
  drx[4] = drx[0];
  dry[4] = dry[0];
  drz[4] = drz[0];
  const double gz = dx * a3j + dy * a6j + dz * a9j;
  const double agz = fabs(gz);
  const double sgn = copysign(1.0, gz);
  double ai = 0.0, vxs = 0.0, vys = 0.0, vzs = 0.0;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const double ex = tx[l + 1] - tx[l], ey = ty[l + 1] - ty[l], ez = tz[l + 1] - tz[l];
    const double dk2 = ex * ex + ey * ey + ez * ez;
    const double idk = rsqrt_pos(dk2);
    const double dk = dk2 * idk;  // (NaN for a degenerate edge, which the test below skips)
    if (!(dk >= 1.E-3 * tdis)) continue;
    const double pj = ex * idk, qj = ey * idk, rj = ez * idk;
    const double gyx = a6j * rj - a9j * qj;
    const double gyy = a9j * pj - a3j * rj;
    const double gyz = a3j * qj - a6j * pj;
    const double gy = (xo - tx[l]) * gyx + (yo - ty[l]) * gyy + (zo - tz[l]) * gyz;
    const double ddk = 2.0 * dk;
    const double ant = gy * ddk;
    const double rs = rr[l + 1] + rr[l];
    const double dnt = rs * rs - dk * dk + 2.0 * agz * rs;
    const double anl = rs + dk;
    const double dnl = rs - dk;
    const double alden = log(anl / dnl);
    const double at = (agz >= 1.E-4 * tdis) ? atan(ant / dnt) : 0.0;
    ai = ai + gy * alden - 2.0 * agz * at;
    const double dat = 2.0 * at * sgn;
    const double anlx = drx[l + 1] + drx[l], anly = dry[l + 1] + dry[l], anlz = drz[l + 1] + drz[l];
    const double dr = 2.0 * (rs + agz);
    const double ds = 2.0 * rs * sgn;
    const double dntx = dr * anlx + a3j * ds;
    const double dnty = dr * anly + a6j * ds;
    const double dntz = dr * anlz + a9j * ds;
    const double denl = anl * dnl;
    const double dent = ant * ant + dnt * dnt;
    const double rdent = recip_pos(dent);
    const double atx = (gyx * ddk * dnt - dntx * ant) * rdent;
    const double aty = (gyy * ddk * dnt - dnty * ant) * rdent;
    const double atz = (gyz * ddk * dnt - dntz * ant) * rdent;
    const double dlogg = (dnl - anl) * recip_pos(denl);
    vxs = vxs + gyx * alden + gy * anlx * dlogg - 2.0 * agz * atx - dat * a3j;
    vys = vys + gyy * alden + gy * anly * dlogg - 2.0 * agz * aty - dat * a6j;
    vzs = vzs + gyz * alden + gy * anlz * dlogg - 2.0 * agz * atz - dat * a9j;
  }
  if (self) {
    const double dpi = 2.0 * M_PI;
    r.a = ai;
    r.vx = vxs - dpi * a3j;
    r.vy = vys - dpi * a6j;
    r.vz = vzs - dpi * a9j;
  } else {
    r.a = ai * mjj;
    r.vx = vxs * mjj;
    r.vy = vys * mjj;
    r.vz = vzs * mjj;
  }
  return r;
}

__device__ __forceinline__ RankineImage rankine_image(const Geometry& G, int j, double xo, double yo,
                                                      double zo, double my, double mz, double dh,
                                                      double mh, double mjj, bool self) {
  const double xoj = G.xg[j], yoj = G.yg[j] * my, zoj = G.zg[j] * mz - dh * mh;
  const double dx = xo - xoj, dy = yo - yoj, dz = zo - zoj;
  const double ro2 = dx * dx + dy * dy + dz * dz;
  const double tdis = G.tdis[j];
  if (ro2 > 49.0 * tdis * tdis) {  // RO > 7 TDIS
    RankineImage r;
    const double ri = rsqrt_pos(ro2);
    r.a = G.aire[j] * ri;
    const double asro = r.a * ri * ri;
    r.vx = -dx * asro;
    r.vy = -dy * asro;
    r.vz = -dz * asro;
    return r;
  }
  return rankine_near(G, j, xo, yo, zo, my, mz, dh, mh, mjj, self, dx, dy, dz);
}

// Sum of the Rankine images combined into the +/- symmetry parts
// (VAVINFD.f90:187-217): out[0..7] = FSP FSM VSXP VSYP VSZP VSXM VSYM VSZM.
__device__ __forceinline__ void rankine(const Geometry& G, int i, int j, int nsym, double zer, int kkk,
                                        double dh, double out[8]) {
  const double xo = G.xg[i], yo = G.yg[i];
  const double zgi = G.zg[i];
  const double zo = (zgi > zer) ? ((kkk == 1) ? zer : 20.0 * zer) : zgi;
  const double mk = (kkk == 1) ? 1.0 : -1.0;
  // images: 1 direct, 2 z-mirror (mjj -1), 3 y+z mirror (+1), 4 y-mirror (-1)
  const RankineImage r1 = rankine_image(G, j, xo, yo, zo, 1.0, 1.0, dh, 0.0, 1.0, i == j);
  const RankineImage r2 = rankine_image(G, j, xo, yo, zo, 1.0, -1.0, dh, 1.0, -1.0, false);
  const double mq = -1.0 / (4.0 * M_PI);  // the images sum times -1/(4 pi)
  if (nsym == 1) {
    const RankineImage r3 = rankine_image(G, j, xo, yo, zo, -1.0, -1.0, dh, 1.0, 1.0, false);
    const RankineImage r4 = rankine_image(G, j, xo, yo, zo, -1.0, 1.0, dh, 0.0, -1.0, false);
    out[0] = (r1.a - mk * (r2.a + r3.a) + r4.a) * mq;
    out[1] = (r1.a - mk * (r2.a - r3.a) - r4.a) * mq;
    out[2] = (r1.vx - mk * (r2.vx + r3.vx) + r4.vx) * mq;
    out[3] = (r1.vy - mk * (r2.vy + r3.vy) + r4.vy) * mq;
    out[4] = (r1.vz - mk * (r2.vz + r3.vz) + r4.vz) * mq;
    out[5] = (r1.vx - mk * (r2.vx - r3.vx) - r4.vx) * mq;
    out[6] = (r1.vy - mk * (r2.vy - r3.vy) - r4.vy) * mq;
    out[7] = (r1.vz - mk * (r2.vz - r3.vz) - r4.vz) * mq;
  } else {
    out[0] = out[1] = (r1.a - mk * r2.a) * mq;
    out[2] = out[5] = (r1.vx - mk * r2.vx) * mq;
    out[3] = out[6] = (r1.vy - mk * r2.vy) * mq;
    out[4] = out[7] = (r1.vz - mk * r2.vz) * mq;
  }
}

// PL2 (COMPUTE_GREEN_INFD.f90:641-645): ((x-u1)(x-u2)) / ((u3-u1)(u3-u2)),
// with the reciprocal of the node-only denominator taken from the tables.
__device__ __forceinline__ double pl2r(double u1, double u2, double rden, double xu) {
  return ((xu - u1) * (xu - u2)) * rden;
}

// INT(scale * (LOG10(x) + offset) + add) for x > 0, as the reference
// evaluates it in double precision.  The hardware single-precision log2
// gives the value to within 1e-5; unless that lands within 1e-4 of an
// integer the truncation is already certain, otherwise the double-precision
// expression decides, so the index is always the FP64 one.  Negative values
// are only ever clamped up to 2 by the callers, where truncation and floor
// agree.
__device__ __forceinline__ int log10_index(double x, double scale, double offset, double add) {
  const float lf = __builtin_amdgcn_logf((float)x) * 0.30102999566398120f;  // log10 x
  const float yf = (float)scale * (lf + (float)offset) + (float)add;
  const float fr = yf - floorf(yf);
  if (fr > 1e-4f && fr < 1.0f - 1e-4f && fabsf(yf) < 1e6f) return (int)floorf(yf);
  return (int)(scale * (log10(x) + offset) + add);
}

// Table index selection (COMPUTE_GREEN_INFD.f90:292-305), 1-based.
__device__ __forceinline__ int kj_of(double akz) {
  int kj = (akz < -1.E-2) ? log10_index(-akz, 8.0, 4.5, 0.0) : log10_index(-akz, 5.0, 6.0, 0.0);
  return min(max(kj, 2), 45);
}
__device__ __forceinline__ int ki_of(double akr) {
  int ki = (akr < 1.0) ? log10_index(akr + 1.E-20, 5.0, 6.0, 1.0) : (int)(3.0 * akr + 28.0);
  return min(max(ki, 2), 327);
}

// Lagrange weights in r for node KI (1-based) of XR.
struct RWeights {
  int ki;
  double x1, x2, x3;
};
__device__ __forceinline__ RWeights r_weights(const Tables& T, double akr) {
  RWeights w;
  w.ki = ki_of(akr);
  const double a = T.xr[w.ki - 2], b = T.xr[w.ki - 1], c = T.xr[w.ki];
  const double* rd = T.rr + 3 * (w.ki - 2);
  w.x1 = pl2r(b, c, rd[0], akr);
  w.x2 = pl2r(c, a, rd[1], akr);
  w.x3 = pl2r(a, b, rd[2], akr);
  return w;
}

// 3x3 quadratic interpolation of all four tables at (akr, akz):
// t[0..3] = APD1X, APD1Z, APD2X, APD2Z.
__device__ __forceinline__ void interp4(const Tables& T, const RWeights& w, double akz, int kj,
                                        double t[4]) {
  const double a = T.xz[kj - 2], b = T.xz[kj - 1], c = T.xz[kj];
  const double* rd = T.rz + 3 * (kj - 2);
  const double z1 = pl2r(b, c, rd[0], akz), z2 = pl2r(c, a, rd[1], akz), z3 = pl2r(a, b, rd[2], akz);
  const double zl[3] = {z1, z2, z3};
  double f[3][4];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const double* row = T.tab + 4 * ((w.ki - 2) + kNR * (kj - 2 + q));
    const double4 n0 = *reinterpret_cast<const double4*>(row);
    const double4 n1 = *reinterpret_cast<const double4*>(row + 4);
    const double4 n2 = *reinterpret_cast<const double4*>(row + 8);
    f[q][0] = w.x1 * n0.x + w.x2 * n1.x + w.x3 * n2.x;
    f[q][1] = w.x1 * n0.y + w.x2 * n1.y + w.x3 * n2.y;
    f[q][2] = w.x1 * n0.z + w.x2 * n1.z + w.x3 * n2.z;
    f[q][3] = w.x1 * n0.w + w.x2 * n1.w + w.x3 * n2.w;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) t[k] = zl[0] * f[0][k] + zl[1] * f[1][k] + zl[2] * f[2][k];
}

// Wave-part coefficients for one y-image (index jj, sign bx) of the pair,
// before the final scaling: fs1 fs2 vsx1 vsy1 vsz1 vsx2 vsy2 vsz2.
struct WavePart {
  double fs1, fs2, vx1, vy1, vz1, vx2, vy2, vz2;
};

// Infinite depth, VNSINFD.f90:269-453.  kSurf = false leaves out the branch
// for pairs with a point at the free surface (zero result there): the
// assembly runs those pairs in a launch of their own, so that the bulk of the
// matrix runs without that branch's registers.
template <bool kSurf>
__device__ WavePart wave_infd_image(const Geometry& G, const Tables& T, const FreqParams& P, int i, int j,
                                    double bx) {
  WavePart w = {0, 0, 0, 0, 0, 0, 0, 0};
  const double ak0 = P.ak0;
  const double xgi = G.xg[i], ygi = G.yg[i], zgi = G.zg[i];
  const double zmiii = (zgi > P.zer) ? P.zer : zgi;
  const double eps = 0.0001;
  if (zgi < P.zer && G.zg[j] < P.zer) {
    const double ymjjj = G.yg[j] * bx;
    const double ddx = xgi - G.xg[j], ddy = ygi - ymjjj;
    const double rrr = sqrt(ddx * ddx + ddy * ddy);
    const double akr = ak0 * rrr;
    const double zzz = zmiii + G.zg[j];
    const double akz = ak0 * zzz;
    const double dd2 = rrr * rrr + zzz * zzz;
    double psurr = 0.0;
    if (dd2 > eps * eps) {  // DD > EPS; psurr = pi / (k dd)^3
      const double ri = rsqrt_pos(dd2);
      psurr = P.pik3 * (ri * ri * ri);
    }
    double pd1z = 0, pd2z = 0, vz1 = 0, vz2 = 0, vr1 = 0, vr2 = 0;
    if (akz <= -1.5E-6) {  // (the reference leaves values unset above)
      if (akz > -16.0) {
        double pd1x, pd2x;
        if (akr < 99.7) {
          const RWeights rw = r_weights(T, akr);
          double t4[4];
          interp4(T, rw, akz, kj_of(akz), t4);
          pd1x = t4[0];
          pd1z = t4[1];
          pd2x = t4[2];
          pd2z = t4[3];
        } else {
          const double epz = exp(akz);
          const double akp4 = akr - 0.25 * M_PI;
          const double sq = sqrt(2.0 * M_PI / akr);
          double sik, csk;
          sincos(akp4, &sik, &csk);
          pd1z = psurr * akz - M_PI * epz * sq * sik;
          pd2z = epz * sq * csk;
          const double dsk = 0.5 / akr;
          pd1x = -psurr * akr + M_PI * epz * sq * (csk - dsk * sik);
          pd2x = epz * sq * (sik + dsk * csk);
        }
        vz1 = pd1z - psurr * akz;
        vz2 = pd2z;
        vr1 = -pd1x - psurr * akr;
        vr2 = -pd2x;
      } else {
        pd1z = psurr * akz;
        pd2z = 0.0;
      }
    }
    w.fs1 = pd1z;
    w.fs2 = pd2z;
    if (rrr > eps) {
      const double cvx = ddx / rrr, cvy = ddy / rrr;
      w.vx1 = vr1 * cvx;
      w.vx2 = vr2 * cvx;
      w.vy1 = vr1 * cvy;
      w.vy2 = vr2 * cvy;
    }
    w.vz1 = vz1;
    w.vz2 = vz2;
    return w;
  }
  // near the free surface: theta quadrature over the panel edges
  // (VNSINFD.f90:385-453); with NQ = 0 the terms vanish
  if (!kSurf || P.nq == 0) return w;
  double px[5], py[5], pz[5];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    px[l] = G.vx[4 * j + l];
    py[l] = G.vy[4 * j + l];
    pz[l] = G.vz[4 * j + l];
  }
  px[4] = px[0];
  py[4] = py[0];
  pz[4] = pz[0];
  const double xn = G.xn[j], yn = G.yn[j], zn = G.zn[j];
  for (int it = 0; it < P.nq; ++it) {
    double st, ct;
    sincos(P.qq[it], &st, &ct);
    cplx zij[5], cex[5], gz[5];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      zij[l] = cmk(ak0 * (zmiii + pz[l]), ak0 * ((xgi - px[l]) * ct + (ygi - py[l] * bx) * st));
      cex[l] = (zij[l].re > -25.0) ? cexp_(zij[l]) : cmk(0.0, 0.0);
      gz[l] = gg(zij[l], cex[l]);
    }
    zij[4] = zij[0];
    cex[4] = cex[0];
    gz[4] = gz[0];
    cplx za = cmk(0, 0), zb = cmk(0, 0), zvs = cmk(0, 0);
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const cplx c1 = zij[l + 1] - zij[l];
      cplx asd, bsd, csd;
      if (fabs(c1.im) < eps && fabs(c1.re) < eps) {
        asd = 0.5 * (gz[l + 1] + gz[l]);
        bsd = 0.5 * (cex[l + 1] + cex[l]);
        csd = asd - (cdiv(cmk(0.5, 0), zij[l + 1]) + cdiv(cmk(0.5, 0), zij[l]));
      } else {
        csd = cdiv(gz[l + 1] - gz[l], c1);
        asd = cdiv(gz[l + 1] - gz[l] + clog_(cdiv(zij[l + 1], zij[l])), c1);
        bsd = cdiv(cex[l + 1] - cex[l], c1);
      }
      const double dxl = px[l + 1] - px[l], dyl = (py[l + 1] - py[l]) * bx;
      const cplx c2 = cmk(yn * bx * dxl - xn * dyl, -zn * st * dxl + zn * ct * dyl);
      za = za + asd * c2;
      zb = zb + bsd * c2;
      zvs = zvs + csd * c2;
    }
    const double cq = P.cq[it];
    w.fs1 += cq * za.re;
    w.fs2 += cq * zb.re;
    w.vx1 -= cq * ct * zvs.im;
    w.vx2 -= cq * ct * zb.im;
    w.vy1 -= cq * st * zvs.im;
    w.vy2 -= cq * st * zb.im;
    w.vz1 += cq * zvs.re;
    w.vz2 += cq * zb.re;
  }
  const double s = bx / (ak0 * G.aire[j]);
  w.fs1 *= s;
  w.fs2 *= s;
  w.vx1 *= s;
  w.vx2 *= s;
  w.vy1 *= s;
  w.vy2 *= s;
  w.vz1 *= s;
  w.vz2 *= s;
  return w;
}

// One of the four z-images of the finite-depth submerged branch
// (VNSFD.f90:295-518).  akz <= -1.5e-6 is assumed (undefined above).
struct FdImage {
  double pd1z, pd2z, vz2, pd1x, vr2, psr;
};

__device__ __forceinline__ FdImage fd_image(const Tables& T, const RWeights& rw, double am0, double pi_k,
                                            double pi_k3, double rrr, double akr, double zzz, double epz_sq_s,
                                            double sq_c, double scds, double ssds) {
  FdImage m;
  const double akz = am0 * zzz;
  const double dd2 = rrr * rrr + zzz * zzz;
  double psurr = 0.0;
  m.psr = 0.0;
  if (dd2 > 1e-8) {  // DD > 1e-4; psr = pi / (k dd), psurr = pi / (k dd)^3
    const double ri = rsqrt_pos(dd2);
    m.psr = pi_k * ri;
    psurr = pi_k3 * ri * ri * ri;
  }
  if (akz > -16.0) {
    double pd2x;
    if (akr < 99.7) {
      double t4[4];
      interp4(T, rw, akz, kj_of(akz), t4);
      m.pd1x = t4[0];
      m.pd1z = t4[1];
      pd2x = t4[2];
      m.pd2z = t4[3];
    } else {
      const double epz = exp(akz);
      m.pd1z = psurr * akz - epz_sq_s * epz;  // PSK * EPZ
      m.pd2z = epz * sq_c;                    // EPZ * SCK
      m.pd1x = -psurr * akr - epz * scds;     // sign as VNSFD.f90:541
      pd2x = epz * ssds;
    }
    m.vz2 = m.pd2z;
    m.vr2 = -pd2x;
  } else {
    m.pd1z = psurr * akz;
    m.pd2z = 0.0;
    m.vz2 = 0.0;
    m.pd1x = -psurr * akr;
    m.vr2 = 0.0;
  }
  return m;
}

// VSD (COMPUTE_GREEN_FD.f90:1514-1609): Rankine integral of panel j for a
// field point moved to depth zig, y-image sign mjj.
__device__ __forceinline__ void vsd(const Geometry& G, int j, double mjj, double xgi, double ygi, double zig,
                                    double& fs, double& vx, double& vy, double& vz) {
  const double qpi = 4.0 * M_PI;
  const double qmp = mjj / (2.0 * qpi);
  const double xnf = G.xn[j], ynf = G.yn[j], znf = G.zn[j];
  const double td = G.tdis[j];
  const double dx = xgi - G.xg[j], dy = ygi - G.yg[j] * mjj, dz = zig - G.zg[j];
  const double ro2 = dx * dx + dy * dy + dz * dz;
  const double gz = dx * xnf + dy * ynf * mjj + dz * znf;
  if (ro2 > 49.0 * td * td) {  // RO > 7 TDIS
    const double ri = rsqrt_pos(ro2);
    fs = -G.aire[j] * ri / (2.0 * qpi);
    const double asro = fs * ri * ri;
    vx = -dx * asro;
    vy = -dy * asro;
    vz = -dz * asro;
    return;
  }
  double tx[5], ty[5], tz[5], rr[5], drx[5], dry[5], drz[5];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    tx[l] = G.vx[4 * j + l];
    ty[l] = G.vy[4 * j + l];
    tz[l] = G.vz[4 * j + l];
    const double ex = xgi - tx[l], ey = ygi - ty[l] * mjj, ez = zig - tz[l];
    rr[l] = sqrt(ex * ex + ey * ey + ez * ez);
    drx[l] = ex / rr[l];
    dry[l] = ey / rr[l];
    drz[l] = ez / rr[l];
  }
  tx[4] = tx[0];
  ty[4] = ty[0];
  tz[4] = tz[0];
  rr[4] = rr[0];
  drx[4] = drx[0];
  dry[4] = dry[0];
  drz[4] = drz[0];
  const double agz = fabs(gz), sgn = copysign(1.0, gz);
  double ai = 0, vxs = 0, vys = 0, vzs = 0;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const double ex = tx[l + 1] - tx[l], ey = ty[l + 1] - ty[l], ez = tz[l + 1] - tz[l];
    const double dk = sqrt(ex * ex + ey * ey + ez * ez);
    if (!(dk >= 1.E-3 * td)) continue;
    const double pj = ex / dk, qj = ey * mjj / dk, rj = ez / dk;
    const double gyx = ynf * mjj * rj - znf * qj;
    const double gyy = znf * pj - xnf * rj;
    const double gyz = xnf * qj - ynf * mjj * pj;
    const double gy = (xgi - tx[l]) * gyx + (ygi - ty[l] * mjj) * gyy + (zig - tz[l]) * gyz;
    const double ddk = 2.0 * dk;
    const double ant = gy * ddk;
    const double rs = rr[l + 1] + rr[l];
    const double dnt = rs * rs - dk * dk + 2.0 * agz * rs;
    const double anl = rs + dk, dnl = rs - dk;
    const double alden = log(anl / dnl);
    const double at = (agz > 1.E-4 * td) ? atan(ant / dnt) : 0.0;
    ai = ai + gy * alden - 2.0 * agz * at;
    const double dat = 2.0 * at * sgn;
    const double anlx = drx[l + 1] + drx[l], anly = dry[l + 1] + dry[l], anlz = drz[l + 1] + drz[l];
    const double dr = 2.0 * (rs + agz);
    const double ds = 2.0 * rs * sgn;
    const double dntx = dr * anlx + xnf * ds;
    const double dnty = dr * anly + ynf * mjj * ds;
    const double dntz = dr * anlz + znf * ds;
    const double denl = anl * dnl;
    const double dent = ant * ant + dnt * dnt;
    const double atx = (gyx * ddk * dnt - dntx * ant) / dent;
    const double aty = (gyy * ddk * dnt - dnty * ant) / dent;
    const double atz = (gyz * ddk * dnt - dntz * ant) / dent;
    const double dlg = (dnl - anl) / denl;
    vxs = vxs + gyx * alden + gy * anlx * dlg - 2.0 * agz * atx - dat * xnf;
    vys = vys + gyy * alden + gy * anly * dlg - 2.0 * agz * aty - dat * ynf * mjj;
    vzs = vzs + gyz * alden + gy * anlz * dlg - 2.0 * agz * atz - dat * znf;
  }
  fs = -ai * qmp;
  vx = -vxs * qmp;
  vy = -vys * qmp;
  vz = -vzs * qmp;
}

// Finite depth, VNSFD.f90:284-827 for one y-image (bx), already scaled;
// kSurf as for wave_infd_image.
// The near-free-surface theta quadrature is not evaluated: the reference
// runs it with CQ = 0 (CINT_FD fills local arrays, SOLVE_BEM_FD_DIRECT.f90:
// 73-74 zeroes the module ones), so it contributes exactly nothing.
template <bool kSurf>
__device__ WavePart wave_fd_image(const Geometry& G, const Tables& T, const FreqParams& P, int i, int j,
                                  double bx) {
  WavePart w = {0, 0, 0, 0, 0, 0, 0, 0};
  const double am0 = P.ak0, h = P.depth;
  const double xgi = G.xg[i], ygi = G.yg[i], zgi = G.zg[i];
  const double zmiii = (zgi > P.zer) ? 20.0 * P.zer : zgi;
  const double zgj = G.zg[j];
  const double eps = 0.0001;
  if (zgi < P.zer && zgj < P.zer) {
    const double ymjjj = bx * G.yg[j];
    const double aire = G.aire[j];
    const double cof1 = P.coe3 * aire, cof2 = P.coe4 * aire;
    const double cof3 = am0 * cof1, cof4 = am0 * cof2;
    const double xpg = xgi - G.xg[j], ypg = ygi - ymjjj;
    const double rrr = sqrt(xpg * xpg + ypg * ypg);
    const double akr = am0 * rrr;
    // quantities shared by the four images (computed from image 1 in the
    // reference, which is always the shallowest)
    RWeights rw = {2, 0, 0, 0};
    double psk = 0, sck = 0, scds = 0, ssds = 0;
    const double zz1 = zmiii + zgj;
    if (am0 * zz1 > -16.0) {
      if (akr < 99.7) {
        rw = r_weights(T, akr);
      } else {
        const double sq = sqrt(2.0 * M_PI / akr);
        double sik, csk;
        sincos(akr - 0.25 * M_PI, &sik, &csk);
        psk = M_PI * sq * sik;
        sck = sq * csk;
        const double dsk = 0.5 / akr;
        scds = M_PI * sq * (csk - dsk * sik);
        ssds = sq * (sik + dsk * csk);
      }
    }
    const double pi_k = P.pik, pi_k3 = P.pik3;
    // the four images' terms are summed as each image is formed (left to
    // right, as the sums are written in VNSFD.f90:548-600), so that only the
    // running sums stay live, not four images' worth of values
    const double zi[4] = {zz1, zgj - zmiii - 2.0 * h, zmiii - zgj - 2.0 * h, -zgj - zmiii - 4.0 * h};
    double spd1z = 0, spd2z = 0, spd1x = 0, svr2 = 0, dpd1z = 0, dvz2 = 0, psr[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const FdImage m = fd_image(T, rw, am0, pi_k, pi_k3, rrr, akr, zi[k], psk, sck, scds, ssds);
      psr[k] = m.psr;
      if (k == 0) {
        spd1z = m.pd1z;
        spd2z = m.pd2z;
        spd1x = m.pd1x;
        svr2 = m.vr2;
        dpd1z = m.pd1z;
        dvz2 = m.vz2;
      } else {
        spd1z += m.pd1z;
        spd2z += m.pd2z;
        spd1x += m.pd1x;
        svr2 += m.vr2;
        dpd1z = (k & 1) ? dpd1z - m.pd1z : dpd1z + m.pd1z;
        dvz2 = (k & 1) ? dvz2 - m.vz2 : dvz2 + m.vz2;
      }
    }
    w.fs1 = cof1 * (spd1z - psr[0] - psr[1] - psr[2] - psr[3]);
    w.fs2 = cof2 * spd2z;
    if (rrr > eps) {
      const double c1v3 = -cof3 * spd1x;
      const double c2v3 = cof4 * svr2;
      const double cvx = xpg / rrr, cvy = ypg / rrr;
      w.vx1 = c1v3 * cvx;
      w.vx2 = c2v3 * cvx;
      w.vy1 = c1v3 * cvy;
      w.vy2 = c2v3 * cvy;
    }
    w.vz1 = cof3 * dpd1z;
    w.vz2 = cof4 * dvz2;
    // exponential images as one-point sources (VNSFD.f90:635-706).  Every
    // image shares the pair's horizontal offset, so its x / y gradient terms
    // -xpg a, -ypg a are summed as one weight sum scaled once at the end.
    const double act = -0.5 * aire / (4.0 * M_PI);
    const double r2 = rrr * rrr;
    double asum = 0.0, vz = 0.0, fs = 0.0;
    for (int ke = 0; ke < P.nexp1; ++ke) {
      const double aqt = act * P.ar[ke];
      const double hl = h * P.ambda[ke];
      const double zp[4] = {zmiii - 2.0 * h + hl - zgj, -zmiii - hl - zgj, -zmiii - 4.0 * h + hl - zgj,
                            zmiii + 2.0 * h - hl - zgj};
      const double zs[4] = {-1.0, 1.0, 1.0, -1.0};  // signs of the z terms (VZS1-VZS2-VZS3+VZS4)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double rr = r2 + zp[k] * zp[k];
        const double ri = (rr > eps * eps) ? rsqrt_pos(rr) : 0.0;  // RO > 1e-4, else no term
        const double f = aqt * ri, a = f * (ri * ri);
        fs += f;
        asum += a;
        vz = fma(zs[k] * zp[k], a, vz);
      }
    }
    w.fs1 += fs;
    w.vx1 += -xpg * asum;
    w.vy1 += -ypg * asum;
    w.vz1 += vz;
    return w;
  }
  if (!kSurf) return w;
  // near the free surface: exact exponential images (VNSFD.f90:801-825)
  const double zp[4] = {-zmiii, zmiii + 2.0 * h, zmiii - 2.0 * h, -zmiii - 4.0 * h};
  for (int ke = 0; ke < P.nexp1; ++ke) {
    const double hl = h * P.ambda[ke];
    double fs[4], vx[4], vy[4], vz[4];
    vsd(G, j, bx, xgi, ygi, zp[0] - hl, fs[0], vx[0], vy[0], vz[0]);
    vsd(G, j, bx, xgi, ygi, zp[1] - hl, fs[1], vx[1], vy[1], vz[1]);
    vsd(G, j, bx, xgi, ygi, zp[2] + hl, fs[2], vx[2], vy[2], vz[2]);
    vsd(G, j, bx, xgi, ygi, zp[3] + hl, fs[3], vx[3], vy[3], vz[3]);
    const double a = P.ar[ke];
    w.fs1 += (fs[0] + fs[1] + fs[2] + fs[3]) * a;
    w.vx1 += (vx[0] + vx[1] + vx[2] + vx[3]) * a;
    w.vy1 += (vy[0] + vy[1] + vy[2] + vy[3]) * a;
    w.vz1 += (-vz[0] + vz[1] + vz[2] - vz[3]) * a;
  }
  return w;
}

// ---------------------------------------------------------------------------
// The wave part of an unordered pair {i, j} with both points below the free
// surface, once for both orders.  It depends on the pair through the
// horizontal distance and the images' depths only:
//   - infinite depth: one image at z_i + z_j;
//   - finite depth: z-images at z_i + z_j, z_j - z_i - 2h, z_i - z_j - 2h,
//     -z_i - z_j - 4h (the middle two trade places when i and j do) and the
//     exponential images at z_i - z_j - 2h + hl, -z_i - z_j - hl,
//     -z_i - z_j - 4h + hl, z_i - z_j + 2h - hl (the first and last trade
//     places, negated).
// So the potentials and the radial gradient coefficients are the same for
// (i, j) and (j, i), and the vertical gradient splits into a part that is
// the same and a part that changes sign.  Values are per unit source area:
// the order's source panel scales them (VNSINFD.f90:455-503 and VNSFD.f90:
// 548-706 write the same terms per ordered pair).
struct WaveSym {
  double f1, f2;    // potential terms 1 and 2
  double h1, h2;    // horizontal gradient k = hk * (field - source offset)
  double z1s, z1a;  // vertical gradient 1 = z1s + z1a for (i, j), z1s - z1a for (j, i)
  double z2s, z2a;
};

// Infinite depth (VNSINFD.f90:269-384, the submerged branch).  ddx, ddy:
// horizontal offset of field point i from the y-image of source j.
__device__ __forceinline__ WaveSym pair_wave_infd(const Tables& T, const FreqParams& P, double ddx, double ddy,
                                                  double zgi, double zgj) {
  WaveSym w = {0, 0, 0, 0, 0, 0, 0, 0};
  const double ak0 = P.ak0;
  const double eps = 0.0001;
  const double rrr = sqrt(ddx * ddx + ddy * ddy);
  const double akr = ak0 * rrr;
  const double zzz = zgi + zgj;
  const double akz = ak0 * zzz;
  const double dd2 = rrr * rrr + zzz * zzz;
  double psurr = 0.0;
  if (dd2 > eps * eps) {  // DD > EPS; psurr = pi / (k dd)^3
    const double ri = rsqrt_pos(dd2);
    psurr = P.pik3 * (ri * ri * ri);
  }
  double vr1 = 0, vr2 = 0;
  if (akz <= -1.5E-6) {  // (the reference leaves values unset above)
    if (akz > -16.0) {
      double pd1x, pd2x;
      if (akr < 99.7) {
        const RWeights rw = r_weights(T, akr);
        double t4[4];
        interp4(T, rw, akz, kj_of(akz), t4);
        pd1x = t4[0];
        w.f1 = t4[1];
        pd2x = t4[2];
        w.f2 = t4[3];
      } else {
        const double epz = exp(akz);
        const double akp4 = akr - 0.25 * M_PI;
        const double sq = sqrt(2.0 * M_PI / akr);
        double sik, csk;
        sincos(akp4, &sik, &csk);
        w.f1 = psurr * akz - M_PI * epz * sq * sik;
        w.f2 = epz * sq * csk;
        const double dsk = 0.5 / akr;
        pd1x = -psurr * akr + M_PI * epz * sq * (csk - dsk * sik);
        pd2x = epz * sq * (sik + dsk * csk);
      }
      w.z1s = w.f1 - psurr * akz;
      w.z2s = w.f2;
      vr1 = -pd1x - psurr * akr;
      vr2 = -pd2x;
    } else {
      w.f1 = psurr * akz;
    }
  }
  if (rrr > eps) {
    const double rinv = rsqrt_pos(ddx * ddx + ddy * ddy);
    w.h1 = vr1 * rinv;
    w.h2 = vr2 * rinv;
  }
  return w;
}

// Finite depth (VNSFD.f90:284-706, the submerged branch), per unit area.
__device__ __forceinline__ WaveSym pair_wave_fd(const Tables& T, const FreqParams& P, double xpg, double ypg,
                                                double zgi, double zgj) {
  WaveSym w = {0, 0, 0, 0, 0, 0, 0, 0};
  const double am0 = P.ak0, h = P.depth;
  const double eps = 0.0001;
  const double rrr = sqrt(xpg * xpg + ypg * ypg);
  const double akr = am0 * rrr;
  // quantities shared by the four images (from the shallowest, at z_i + z_j)
  RWeights rw = {2, 0, 0, 0};
  double psk = 0, sck = 0, scds = 0, ssds = 0;
  const double zz1 = zgi + zgj;
  if (am0 * zz1 > -16.0) {
    if (akr < 99.7) {
      rw = r_weights(T, akr);
    } else {
      const double sq = sqrt(2.0 * M_PI / akr);
      double sik, csk;
      sincos(akr - 0.25 * M_PI, &sik, &csk);
      psk = M_PI * sq * sik;
      sck = sq * csk;
      const double dsk = 0.5 / akr;
      scds = M_PI * sq * (csk - dsk * sik);
      ssds = sq * (sik + dsk * csk);
    }
  }
  const double pi_k = P.pik, pi_k3 = P.pik3;
  // images 0 and 3 keep their places under i <-> j, 1 and 2 trade them
  const FdImage m0 = fd_image(T, rw, am0, pi_k, pi_k3, rrr, akr, zz1, psk, sck, scds, ssds);
  const FdImage m3 = fd_image(T, rw, am0, pi_k, pi_k3, rrr, akr, -zgj - zgi - 4.0 * h, psk, sck, scds, ssds);
  double s1z = m0.pd1z + m3.pd1z, s2z = m0.pd2z + m3.pd2z, s1x = m0.pd1x + m3.pd1x, svr = m0.vr2 + m3.vr2;
  double spsr = m0.psr + m3.psr;
  const double a1 = m0.pd1z - m3.pd1z, a2 = m0.vz2 - m3.vz2;
  const FdImage m1 = fd_image(T, rw, am0, pi_k, pi_k3, rrr, akr, zgj - zgi - 2.0 * h, psk, sck, scds, ssds);
  const FdImage m2 = fd_image(T, rw, am0, pi_k, pi_k3, rrr, akr, zgi - zgj - 2.0 * h, psk, sck, scds, ssds);
  s1z += m1.pd1z + m2.pd1z;
  s2z += m1.pd2z + m2.pd2z;
  s1x += m1.pd1x + m2.pd1x;
  svr += m1.vr2 + m2.vr2;
  spsr += m1.psr + m2.psr;
  const double b1 = m1.pd1z - m2.pd1z, b2 = m1.vz2 - m2.vz2;
  // exponential images as one-point sources: fs, the radial weight sum, and
  // the vertical gradient of the images that keep their places (v12) and of
  // those that trade them negated (v03)
  const double act = -0.5 / (4.0 * M_PI);
  const double r2 = rrr * rrr;
  const double dz = zgi - zgj, sz = -zgi - zgj;
  double fs = 0.0, asum = 0.0, v12 = 0.0, v03 = 0.0;
  for (int ke = 0; ke < P.nexp1; ++ke) {
    const double aqt = act * P.ar[ke];
    const double hl = h * P.ambda[ke];
    const double c = hl - 2.0 * h;  // the same for every lane
    const double zp[4] = {dz + c, sz - hl, sz + (c - 2.0 * h), dz - c};
    double a[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double rr = r2 + zp[k] * zp[k];
      const double ri = (rr > eps * eps) ? rsqrt_pos(rr) : 0.0;  // RO > 1e-4, else no term
      const double f = aqt * ri;
      a[k] = f * (ri * ri);
      fs += f;
      asum += a[k];
    }
    v12 = fma(zp[1], a[1], fma(zp[2], a[2], v12));
    v03 = fma(-zp[0], a[0], fma(-zp[3], a[3], v03));
  }
  const double c3 = am0 * P.coe3, c4 = am0 * P.coe4;
  w.f1 = P.coe3 * (s1z - spsr) + fs;
  w.f2 = P.coe4 * s2z;
  w.h1 = -asum;
  if (rrr > eps) {
    const double rinv = rsqrt_pos(xpg * xpg + ypg * ypg);
    w.h1 -= c3 * s1x * rinv;
    w.h2 = c4 * svr * rinv;
  }
  w.z1s = c3 * a1 + v12;
  w.z1a = v03 - c3 * b1;
  w.z2s = c4 * a2;
  w.z2a = -c4 * b2;
  return w;
}

// All 16 outputs of one pair (order SP1 SM1 SP2 SM2, VS{X,Y,Z}P1,
// VS{X,Y,Z}M1, VS{X,Y,Z}P2, VS{X,Y,Z}M2).
// kFd: finite depth (P.fd == 1) or infinite depth, a compile-time choice so
// that each assembly kernel carries only its own regime's code.
// kSurf = false: only pairs with both points below the free surface
// (zg < ZER) are handled.
template <bool kFd, bool kSurf = true>
__device__ void pair16(const Geometry& G, const Tables& T, const FreqParams& P, int i, int j, double o[16]) {
  double rk[8];
  if (kFd) rankine(G, i, j, P.nsym, P.zer, 2, 2.0 * P.depth, rk);
  else rankine(G, i, j, P.nsym, P.zer, 1, 0.0, rk);
  const WavePart w1 = kFd ? wave_fd_image<kSurf>(G, T, P, i, j, 1.0) : wave_infd_image<kSurf>(G, T, P, i, j, 1.0);
  WavePart w2 = {0, 0, 0, 0, 0, 0, 0, 0};
  if (P.nsym == 1)
    w2 = kFd ? wave_fd_image<kSurf>(G, T, P, i, j, -1.0) : wave_infd_image<kSurf>(G, T, P, i, j, -1.0);
  double s1, s2, s3;  // scaling of (potential 1, potential 2 / gradient 1, gradient 2)
  double g1, g2;
  if (kFd) {
    s1 = 1.0;
    s2 = 1.0;
    g1 = 1.0;
    g2 = 1.0;
    // FD combination: SP1 = FSP + FS1(1) + FS1(2), ... (VNSFD.f90:830-864)
    o[0] = rk[0] + w1.fs1 + w2.fs1;
    o[1] = rk[1] + w1.fs1 - w2.fs1;
    o[2] = w1.fs2 + w2.fs2;
    o[3] = w1.fs2 - w2.fs2;
  } else {
    const double akair = P.ak0 * G.aire[j];
    s1 = akair / (2.0 * M_PI * M_PI);
    s2 = akair / (2.0 * M_PI);
    g1 = s1 * P.ak0;
    g2 = s2 * P.ak0;
    o[0] = rk[0] - (w1.fs1 + w2.fs1) * s1;
    o[1] = rk[1] - (w1.fs1 - w2.fs1) * s1;
    o[2] = -(w1.fs2 + w2.fs2) * s2;
    o[3] = -(w1.fs2 - w2.fs2) * s2;
  }
  (void)s3;
  const double sg = kFd ? 1.0 : -1.0;  // FD adds, INFD subtracts
  o[4] = rk[2] + sg * (w1.vx1 + w2.vx1) * g1;
  o[5] = rk[3] + sg * (w1.vy1 + w2.vy1) * g1;
  o[6] = rk[4] + sg * (w1.vz1 + w2.vz1) * g1;
  o[7] = rk[5] + sg * (w1.vx1 - w2.vx1) * g1;
  o[8] = rk[6] + sg * (w1.vy1 - w2.vy1) * g1;
  o[9] = rk[7] + sg * (w1.vz1 - w2.vz1) * g1;
  o[10] = sg * (w1.vx2 + w2.vx2) * g2;
  o[11] = sg * (w1.vy2 + w2.vy2) * g2;
  o[12] = sg * (w1.vz2 + w2.vz2) * g2;
  o[13] = sg * (w1.vx2 - w2.vx2) * g2;
  o[14] = sg * (w1.vy2 - w2.vy2) * g2;
  o[15] = sg * (w1.vz2 - w2.vz2) * g2;
}

}  // namespace nemoh

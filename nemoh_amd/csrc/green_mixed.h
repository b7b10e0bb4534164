// Influence coefficients in mixed precision, for the mixed path's optional
// FP32 assembly (NEMOH_OPT_FP32_ASSEMBLY).
//
// The same Green function as green.h, with the work split by what each part
// needs:
//   FP64  every branch selector (RO > 7 TDIS, the table / asymptotic /
//         Rankine-only regimes, the exponential images' RO > 1e-4), the
//         table indices KI/KJ and the distances they come from, and the
//         near-panel Rankine integral (RO <= 7 TDIS, including the self
//         term), which cancels heavily in its log and atan terms;
//   FP32  the far-field one-point Rankine terms, the tabulated wave terms
//         (tables held in FP32, Lagrange weights from FP64 node offsets),
//         the exponential (Prony) images of finite depth, and the sums that
//         combine images and symmetry parts.
// The selectors and indices are those of the FP64 path, so both paths take
// the same branch and the same interpolation stencil for every pair; the
// FP32 parts then differ from FP64 by rounding only (a few units of 6e-8 of
// each term).  The rare asymptotic branch of the wave part (k R >= 99.7) and
// the pairs with a point at the free surface stay in FP64.
//
// Measured against FP64 (forces, normwise per problem): 1e-7..2e-6 in deep
// water, 9e-6 on the 4000-panel barge at h = 50 m and 7e-4 at h = 20 m.  The
// tabulated wave terms are differences of large, nearly equal quantities (the
// table value and the subtracted pi / (k d) of each image, the exponential
// images' sum in shallow water), so their FP32 rounding is amplified; and
// without packed FP32 arithmetic a CDNA wave issues FP32 and FP64 FMAs at
// the same rate, so the split saves only the transcendentals and the bytes
// (about a quarter of the assembly).  The default mixed path therefore
// assembles in FP64.
//
// Mathematics: a far panel acts as a point source of strength AIRE, 1/r and
// its gradient -d/r^3; the wave part is a 3x3 tensor-product quadratic
// Lagrange interpolation of the tabulated integrals in (k R, k Z); each
// exponential image of the finite-depth kernel is a point source of
// amplitude AR at the depth shifted by h AMBDA (Solver/Core/
// COMPUTE_GREEN_INFD.f90:92-97, :290-319; COMPUTE_GREEN_FD.f90:286-706).
#pragma once
#include "green.h"

namespace nemoh {

// The four wave tables in single precision, one float4 per (XR, XZ) node
// (APD1X, APD1Z, APD2X, APD2Z) at node i + 328 j.
struct Tables32 {
  const float4* tab;
};

struct RankineF {
  float a, vx, vy, vz;
};

__device__ __forceinline__ float rsq_f32(float x) { return __builtin_amdgcn_rsqf(x); }

// One Rankine image of panel j (arguments as rankine_image).
__device__ __forceinline__ RankineF rankine_image_m(const Geometry& G, int j, double xo, double yo, double zo,
                                                   double my, double mz, double dh, double mh, double mjj,
                                                   bool self) {
  const double dx = xo - G.xg[j], dy = yo - G.yg[j] * my, dz = zo - (G.zg[j] * mz - dh * mh);
  const double ro2 = dx * dx + dy * dy + dz * dz;
  const double tdis = G.tdis[j];
  if (ro2 > 49.0 * tdis * tdis) {  // far: point source of strength AIRE
    const float ri = rsq_f32((float)ro2);
    const float a = (float)G.aire[j] * ri;
    const float g = a * ri * ri;
    return {a, -(float)dx * g, -(float)dy * g, -(float)dz * g};
  }
  const RankineImage r = rankine_near(G, j, xo, yo, zo, my, mz, dh, mh, mjj, self, dx, dy, dz);
  return {(float)r.a, (float)r.vx, (float)r.vy, (float)r.vz};
}

// Rankine images combined into the +/- symmetry parts, as rankine():
// out = FSP FSM VSXP VSYP VSZP VSXM VSYM VSZM.
__device__ __forceinline__ void rankine_m(const Geometry& G, int i, int j, int nsym, double zer, int kkk, double dh,
                                          float out[8]) {
  const double xo = G.xg[i], yo = G.yg[i];
  const double zgi = G.zg[i];
  const double zo = (zgi > zer) ? ((kkk == 1) ? zer : 20.0 * zer) : zgi;
  const float mk = (kkk == 1) ? 1.0f : -1.0f;
  const float q = (float)(-1.0 / (4.0 * M_PI));
  const RankineF r1 = rankine_image_m(G, j, xo, yo, zo, 1.0, 1.0, dh, 0.0, 1.0, i == j);
  const RankineF r2 = rankine_image_m(G, j, xo, yo, zo, 1.0, -1.0, dh, 1.0, -1.0, false);
  if (nsym == 1) {
    const RankineF r3 = rankine_image_m(G, j, xo, yo, zo, -1.0, -1.0, dh, 1.0, 1.0, false);
    const RankineF r4 = rankine_image_m(G, j, xo, yo, zo, -1.0, 1.0, dh, 0.0, -1.0, false);
    out[0] = q * (r1.a - mk * (r2.a + r3.a) + r4.a);
    out[1] = q * (r1.a - mk * (r2.a - r3.a) - r4.a);
    out[2] = q * (r1.vx - mk * (r2.vx + r3.vx) + r4.vx);
    out[3] = q * (r1.vy - mk * (r2.vy + r3.vy) + r4.vy);
    out[4] = q * (r1.vz - mk * (r2.vz + r3.vz) + r4.vz);
    out[5] = q * (r1.vx - mk * (r2.vx - r3.vx) - r4.vx);
    out[6] = q * (r1.vy - mk * (r2.vy - r3.vy) - r4.vy);
    out[7] = q * (r1.vz - mk * (r2.vz - r3.vz) - r4.vz);
  } else {
    out[0] = out[1] = q * (r1.a - mk * r2.a);
    out[2] = out[5] = q * (r1.vx - mk * r2.vx);
    out[3] = out[6] = q * (r1.vy - mk * r2.vy);
    out[4] = out[7] = q * (r1.vz - mk * r2.vz);
  }
}

// Quadratic Lagrange weights on the 3-node stencil u[0..2] at x, from the
// FP64 offsets x - u (exact to rounding in FP32 even where x sits next to a
// node) and the stencil's reciprocal denominators rd[0..2].
struct Weights3 {
  float w0, w1, w2;
};
__device__ __forceinline__ Weights3 lagrange3(const double* u, const double* rd, double x) {
  const float d0 = (float)(x - u[0]), d1 = (float)(x - u[1]), d2 = (float)(x - u[2]);
  return {d1 * d2 * (float)rd[0], d2 * d0 * (float)rd[1], d0 * d1 * (float)rd[2]};
}

// The four tables at (akr, akz) on the stencils at KI, KJ (1-based, as
// ki_of / kj_of): t = APD1X, APD1Z, APD2X, APD2Z.
__device__ __forceinline__ void interp4_m(const Tables& T, const Tables32& T32, int ki, const Weights3& wr,
                                          double akz, int kj, float t[4]) {
  const Weights3 wz = lagrange3(T.xz + (kj - 2), T.rz + 3 * (kj - 2), akz);
  const float zl[3] = {wz.w0, wz.w1, wz.w2};
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const float4* row = T32.tab + (ki - 2) + kNR * (kj - 2 + q);
    const float4 n0 = row[0], n1 = row[1], n2 = row[2];
    const float f0 = wr.w0 * n0.x + wr.w1 * n1.x + wr.w2 * n2.x;
    const float f1 = wr.w0 * n0.y + wr.w1 * n1.y + wr.w2 * n2.y;
    const float f2 = wr.w0 * n0.z + wr.w1 * n1.z + wr.w2 * n2.z;
    const float f3 = wr.w0 * n0.w + wr.w1 * n1.w + wr.w2 * n2.w;
    acc[0] += zl[q] * f0;
    acc[1] += zl[q] * f1;
    acc[2] += zl[q] * f2;
    acc[3] += zl[q] * f3;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) t[k] = acc[k];
}

struct WaveF {
  float fs1, fs2, vx1, vy1, vz1, vx2, vy2, vz2;
};

// Infinite depth, one y-image (bx) of a pair with both points below the
// free surface (others: zero, redone in FP64 by the surface launch).
// Unscaled, as wave_infd_image.
__device__ WaveF wave_infd_m(const Geometry& G, const Tables& T, const Tables32& T32, const FreqParams& P, int i,
                             int j, double bx) {
  WaveF w = {0, 0, 0, 0, 0, 0, 0, 0};
  const double ak0 = P.ak0;
  const double zgi = G.zg[i];
  if (!(zgi < P.zer && G.zg[j] < P.zer)) return w;
  const double ddx = G.xg[i] - G.xg[j], ddy = G.yg[i] - G.yg[j] * bx;
  const double rrr = sqrt(ddx * ddx + ddy * ddy);
  const double akr = ak0 * rrr;
  const double zzz = zgi + G.zg[j];
  const double akz = ak0 * zzz;
  const double dd2 = rrr * rrr + zzz * zzz;
  const double eps = 0.0001;
  float psurr = 0.f;  // pi / (k dd)^3
  if (dd2 > eps * eps) {
    const float ri = rsq_f32((float)dd2);
    psurr = (float)(M_PI / (ak0 * ak0 * ak0)) * (ri * ri * ri);
  }
  const float akzf = (float)akz, akrf = (float)akr;
  float pd1z = 0.f, pd2z = 0.f, vz1 = 0.f, vz2 = 0.f, vr1 = 0.f, vr2 = 0.f;
  if (akz <= -1.5E-6) {
    if (akz > -16.0) {
      float pd1x, pd2x;
      if (akr < 99.7) {
        const int ki = ki_of(akr);
        const Weights3 wr = lagrange3(T.xr + (ki - 2), T.rr + 3 * (ki - 2), akr);
        float t4[4];
        interp4_m(T, T32, ki, wr, akz, kj_of(akz), t4);
        pd1x = t4[0];
        pd1z = t4[1];
        pd2x = t4[2];
        pd2z = t4[3];
      } else {  // asymptotic expansion (rare: k R >= 99.7), FP64
        const double ps = (double)psurr;
        const double epz = exp(akz);
        const double sq = sqrt(2.0 * M_PI / akr);
        double sik, csk;
        sincos(akr - 0.25 * M_PI, &sik, &csk);
        const double dsk = 0.5 / akr;
        pd1z = (float)(ps * akz - M_PI * epz * sq * sik);
        pd2z = (float)(epz * sq * csk);
        pd1x = (float)(-ps * akr + M_PI * epz * sq * (csk - dsk * sik));
        pd2x = (float)(epz * sq * (sik + dsk * csk));
      }
      vz1 = pd1z - psurr * akzf;
      vz2 = pd2z;
      vr1 = -pd1x - psurr * akrf;
      vr2 = -pd2x;
    } else {
      pd1z = psurr * akzf;
    }
  }
  w.fs1 = pd1z;
  w.fs2 = pd2z;
  if (rrr > eps) {
    const float ir = 1.0f / (float)rrr;
    const float cvx = (float)ddx * ir, cvy = (float)ddy * ir;
    w.vx1 = vr1 * cvx;
    w.vx2 = vr2 * cvx;
    w.vy1 = vr1 * cvy;
    w.vy2 = vr2 * cvy;
  }
  w.vz1 = vz1;
  w.vz2 = vz2;
  return w;
}

// Finite depth, one y-image (bx), scaled, as wave_fd_image for pairs with
// both points below the free surface.
__device__ WaveF wave_fd_m(const Geometry& G, const Tables& T, const Tables32& T32, const FreqParams& P, int i, int j,
                           double bx) {
  WaveF w = {0, 0, 0, 0, 0, 0, 0, 0};
  const double am0 = P.ak0, h = P.depth;
  const double zgi = G.zg[i], zgj = G.zg[j];
  if (!(zgi < P.zer && zgj < P.zer)) return w;
  const double eps = 0.0001;
  const float aire = (float)G.aire[j];
  const float cof1 = (float)P.coe3 * aire, cof2 = (float)P.coe4 * aire;
  const float am0f = (float)am0;
  const float cof3 = am0f * cof1, cof4 = am0f * cof2;
  const double xpg = G.xg[i] - G.xg[j], ypg = G.yg[i] - bx * G.yg[j];
  const double rrr = sqrt(xpg * xpg + ypg * ypg);
  const double akr = am0 * rrr;
  const float akrf = (float)akr;
  // r stencil (or the asymptotic factors) shared by the four z-images,
  // decided by the shallowest image as in the FP64 path
  int ki = 2;
  Weights3 wr = {0.f, 0.f, 0.f};
  float psk = 0.f, sck = 0.f, scds = 0.f, ssds = 0.f;
  if (am0 * (zgi + zgj) > -16.0) {
    if (akr < 99.7) {
      ki = ki_of(akr);
      wr = lagrange3(T.xr + (ki - 2), T.rr + 3 * (ki - 2), akr);
    } else {
      const double sq = sqrt(2.0 * M_PI / akr);
      double sik, csk;
      sincos(akr - 0.25 * M_PI, &sik, &csk);
      const double dsk = 0.5 / akr;
      psk = (float)(M_PI * sq * sik);
      sck = (float)(sq * csk);
      scds = (float)(M_PI * sq * (csk - dsk * sik));
      ssds = (float)(sq * (sik + dsk * csk));
    }
  }
  const float pi_k = (float)(M_PI / am0), pi_k3 = (float)(M_PI / (am0 * am0 * am0));
  const double zi[4] = {zgi + zgj, zgj - zgi - 2.0 * h, zgi - zgj - 2.0 * h, -zgj - zgi - 4.0 * h};
  float spd1z = 0.f, spd2z = 0.f, spd1x = 0.f, svr2 = 0.f, dpd1z = 0.f, dvz2 = 0.f, spsr = 0.f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double akz = am0 * zi[k];
    const double dd2 = rrr * rrr + zi[k] * zi[k];
    float psr = 0.f, psurr = 0.f;
    if (dd2 > 1e-8) {
      const float ri = rsq_f32((float)dd2);
      psr = pi_k * ri;
      psurr = pi_k3 * (ri * ri * ri);
    }
    const float akzf = (float)akz;
    float pd1z, pd2z, pd1x, vr2;
    if (akz > -16.0) {
      if (akr < 99.7) {
        float t4[4];
        interp4_m(T, T32, ki, wr, akz, kj_of(akz), t4);
        pd1x = t4[0];
        pd1z = t4[1];
        vr2 = -t4[2];
        pd2z = t4[3];
      } else {
        const float epz = expf(akzf);
        pd1z = psurr * akzf - psk * epz;
        pd2z = epz * sck;
        pd1x = -psurr * akrf - epz * scds;
        vr2 = -(epz * ssds);
      }
    } else {
      pd1z = psurr * akzf;
      pd2z = 0.f;
      pd1x = -psurr * akrf;
      vr2 = 0.f;
    }
    spd1z += pd1z;
    spd2z += pd2z;
    spd1x += pd1x;
    svr2 += vr2;
    spsr += psr;
    dpd1z += (k & 1) ? -pd1z : pd1z;  // the first and third images add, the second and fourth subtract
    dvz2 += (k & 1) ? -pd2z : pd2z;
  }
  w.fs1 = cof1 * (spd1z - spsr);
  w.fs2 = cof2 * spd2z;
  if (rrr > eps) {
    const float ir = 1.0f / (float)rrr;
    const float c1v3 = -cof3 * spd1x, c2v3 = cof4 * svr2;
    const float cvx = (float)xpg * ir, cvy = (float)ypg * ir;
    w.vx1 = c1v3 * cvx;
    w.vx2 = c2v3 * cvx;
    w.vy1 = c1v3 * cvy;
    w.vy2 = c2v3 * cvy;
  }
  w.vz1 = cof3 * dpd1z;
  w.vz2 = cof4 * dvz2;
  // exponential images: point sources of amplitude -AIRE AR / (8 pi) at the
  // four depths shifted by h AMBDA, all at the pair's horizontal offset
  const float act = (float)(-0.5 / (4.0 * M_PI)) * aire;
  const double r2 = rrr * rrr;
  float asum = 0.f, vz = 0.f, fs = 0.f;
  for (int ke = 0; ke < P.nexp1; ++ke) {
    const float aqt = act * (float)P.ar[ke];
    const double hl = h * P.ambda[ke];
    const double zp[4] = {zgi - 2.0 * h + hl - zgj, -zgi - hl - zgj, -zgi - 4.0 * h + hl - zgj,
                          zgi + 2.0 * h - hl - zgj};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double rr = r2 + zp[k] * zp[k];
      const float ri = (rr > eps * eps) ? rsq_f32((float)rr) : 0.f;
      const float f = aqt * ri, a = f * (ri * ri);
      fs += f;
      asum += a;
      const float zs = (k == 0 || k == 3) ? -(float)zp[k] : (float)zp[k];
      vz += zs * a;
    }
  }
  w.fs1 += fs;
  w.vx1 += -(float)xpg * asum;
  w.vy1 += -(float)ypg * asum;
  w.vz1 += vz;
  return w;
}

// All 16 outputs of one pair with both points below the free surface, in
// the order of pair16.
template <bool kFd>
__device__ void pair16_m(const Geometry& G, const Tables& T, const Tables32& T32, const FreqParams& P, int i, int j,
                         float o[16]) {
  float rk[8];
  if (kFd) rankine_m(G, i, j, P.nsym, P.zer, 2, 2.0 * P.depth, rk);
  else rankine_m(G, i, j, P.nsym, P.zer, 1, 0.0, rk);
  const WaveF w1 = kFd ? wave_fd_m(G, T, T32, P, i, j, 1.0) : wave_infd_m(G, T, T32, P, i, j, 1.0);
  WaveF w2 = {0, 0, 0, 0, 0, 0, 0, 0};
  if (P.nsym == 1) w2 = kFd ? wave_fd_m(G, T, T32, P, i, j, -1.0) : wave_infd_m(G, T, T32, P, i, j, -1.0);
  float g1 = 1.f, g2 = 1.f;
  if (kFd) {
    o[0] = rk[0] + w1.fs1 + w2.fs1;
    o[1] = rk[1] + w1.fs1 - w2.fs1;
    o[2] = w1.fs2 + w2.fs2;
    o[3] = w1.fs2 - w2.fs2;
  } else {
    const double akair = P.ak0 * G.aire[j];
    const float s1 = (float)(akair / (2.0 * M_PI * M_PI)), s2 = (float)(akair / (2.0 * M_PI));
    g1 = s1 * (float)P.ak0;
    g2 = s2 * (float)P.ak0;
    o[0] = rk[0] - (w1.fs1 + w2.fs1) * s1;
    o[1] = rk[1] - (w1.fs1 - w2.fs1) * s1;
    o[2] = -(w1.fs2 + w2.fs2) * s2;
    o[3] = -(w1.fs2 - w2.fs2) * s2;
  }
  const float sg = kFd ? 1.f : -1.f;  // finite depth adds the wave gradient, infinite depth subtracts it
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

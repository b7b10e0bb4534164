// Potential of the last solve's sources at arbitrary field points
// (COMPUTE_POTENTIAL_DOMAIN.f90:24-82), the Solver's free-surface output
// (SOLVE_BEM.f90:104-110 evaluates it on the free-surface mesh, z = 0).
//
// The field-point Green routines differ from the collocation-point ones of
// the assembly and are restated separately (COMPUTE_GREEN_FREESURFACE.f90):
//   * VVV :33-196: the Rankine images with the z-image shifted by 2 Depth
//     for both regimes (infinite-depth solves at finite depth included),
//     the field point not clamped to the free surface;
//   * VNV :199-386 (infinite depth): one table / asymptotic / deep image per
//     y-image with AK0 = w^2/g;
//   * VNVF :388-908 (finite depth): four tabulated z-images, the Prony
//     images, and for the last (constant) term the exact panel integral VVG
//     :911-997 of the free-surface image, whose 4 pi is the literal
//     12.5663706 of the reference.
// A field point whose first image has AK0 (z + ZG) > -1.5e-6 has no value
// (the reference stops): the point is flagged and its result left undefined.
#include <hip/hip_runtime.h>

#include "green.h"
#include "kernels.h"

namespace nemoh {

namespace {

// One z-image of a field point: tabulated / asymptotic / deep value of the
// wave integrals (VNV :233-287, VNVF :466-519 and their copies for images
// 2-4) and the Rankine term PI/(AK DD) that the finite-depth sum subtracts.
struct FieldImage {
  double pd1z, pd2z, psr;
};

__device__ __forceinline__ FieldImage field_image(const Tables& T, double ak, double rrr, double akr, double zzz,
                                                  bool& bad) {
  const double pi = M_PI, eps = 0.0001;
  const double akz = ak * zzz;
  const double dd = sqrt(rrr * rrr + zzz * zzz);
  FieldImage f;
  f.psr = 0.0;
  double psurr = 0.0;
  if (dd > eps) {
    const double rr = ak * dd;
    f.psr = pi / rr;
    psurr = pi / (rr * rr * rr);
  }
  if (akz > -1.5E-6) {
    bad = true;
    f.pd1z = f.pd2z = 0.0;
  } else if (akz > -16.0) {
    if (akr < 99.7) {
      const RWeights rw = r_weights(T, akr);
      double t[4];
      interp4(T, rw, akz, kj_of(akz), t);
      f.pd1z = t[1];
      f.pd2z = t[3];
    } else {
      const double epz = exp(akz), akp4 = akr - 0.25 * pi, sq = sqrt(2.0 * pi / akr);
      double sik, csk;
      sincos(akp4, &sik, &csk);
      f.pd1z = psurr * akz - pi * sq * sik * epz;
      f.pd2z = epz * sq * csk;
    }
  } else {
    f.pd1z = psurr * akz;
    f.pd2z = 0.0;
  }
  return f;
}

// VVV's potentials FSP, FSM of panel j at (xo, yo, zo): images 1 direct,
// 2 z-mirror about -Depth, 3 y- and z-mirror, 4 y-mirror; kkk 1 adds the
// z-images with sign -1, kkk 2 with +1 (MK).
__device__ __forceinline__ void rankine_field(const Geometry& G, int j, double xo, double yo, double zo, int nsym,
                                              bool kkk1, double dhsl, double& fsp, double& fsm) {
  const double mk = kkk1 ? 1.0 : -1.0, qpi = 4.0 * M_PI;
  const double a1 = rankine_image(G, j, xo, yo, zo, 1.0, 1.0, dhsl, 0.0, 1.0, false).a;
  const double a2 = rankine_image(G, j, xo, yo, zo, 1.0, -1.0, dhsl, 1.0, -1.0, false).a;
  if (nsym) {
    const double a3 = rankine_image(G, j, xo, yo, zo, -1.0, -1.0, dhsl, 1.0, 1.0, false).a;
    const double a4 = rankine_image(G, j, xo, yo, zo, -1.0, 1.0, dhsl, 0.0, -1.0, false).a;
    fsp = -(a1 - mk * (a2 + a3) + a4) / qpi;
    fsm = -(a1 - mk * (a2 - a3) - a4) / qpi;
  } else {
    fsp = fsm = -(a1 - mk * a2) / qpi;
  }
}

struct PotArgs {
  int npts, nrhs;
  double dhsl;  // 2 Depth (VVV)
  double a;     // finite depth: (AMH+AKH)^2 / (H (AMH^2 - AKH^2 + AKH))
};

// SP1 SM1 SP2 SM2 of panel j at (xp, yp, zp).
template <bool kFd>
__device__ __forceinline__ void panel_potential(const Geometry& G, const Tables& T, const FreqParams& P,
                                                const PotArgs& A, int j, double xp, double yp, double zp,
                                                double s[4], bool& bad) {
  const double pi = M_PI, eps = 0.0001;
  double fsp, fsm;
  rankine_field(G, j, xp, yp, zp, P.nsym, !kFd, A.dhsl, fsp, fsm);
  const double ak = P.ak0;
  const double xg = G.xg[j], yg = G.yg[j], zg = G.zg[j], aire = G.aire[j];
  double fs1[2] = {0.0, 0.0}, fs2[2] = {0.0, 0.0};
  for (int jj = 0; jj <= P.nsym; ++jj) {
    const double bx = jj ? -1.0 : 1.0;
    const double dx = xp - xg, dy = yp - yg * bx;
    const double rrr = sqrt(dx * dx + dy * dy), akr = ak * rrr;
    if (!kFd) {  // VNV
      const FieldImage f = field_image(T, ak, rrr, akr, zp + zg, bad);
      fs1[jj] = f.pd1z;
      fs2[jj] = f.pd2z;
      continue;
    }
    // VNVF: four z-images
    const double h = P.depth;
    const FieldImage f1 = field_image(T, ak, rrr, akr, zp + zg, bad);
    const FieldImage f2 = field_image(T, ak, rrr, akr, zg - zp - 2.0 * h, bad);
    const FieldImage f3 = field_image(T, ak, rrr, akr, zp - zg - 2.0 * h, bad);
    const FieldImage f4 = field_image(T, ak, rrr, akr, -zg - zp - 4.0 * h, bad);
    const double cof1 = A.a * (-1.0 / (8.0 * pi * pi)) * aire;
    const double cof2 = (-A.a / (8.0 * pi)) * aire;
    double f = cof1 * ((f1.pd1z + f2.pd1z + f3.pd1z + f4.pd1z) - f1.psr - f2.psr - f3.psr - f4.psr);
    fs2[jj] = cof2 * (f1.pd2z + f2.pd2z + f3.pd2z + f4.pd2z);
    // Prony images; the last term (AMBDA 0, AR 2) takes the exact panel
    // integral for its free-surface image (VVG at (xp, yp, -zp))
    const double act = -0.5 * aire / (4.0 * pi);
    const double r2 = rrr * rrr;
    for (int ke = 0; ke < P.nexp1; ++ke) {
      const double aqt = act * P.ar[ke], hl = h * P.ambda[ke];
      const double z1 = zp - 2.0 * h + hl - zg, z3 = -zp - 4.0 * h + hl - zg, z4 = zp + 2.0 * h - hl - zg;
      const double ro1 = sqrt(r2 + z1 * z1), ro3 = sqrt(r2 + z3 * z3), ro4 = sqrt(r2 + z4 * z4);
      const double t1 = ro1 > eps ? aqt / ro1 : 0.0;
      const double t3 = ro3 > eps ? aqt / ro3 : 0.0;
      const double t4 = ro4 > eps ? aqt / ro4 : 0.0;
      double t2;
      if (ke == P.nexp1 - 1) {
        t2 = rankine_image(G, j, xp, yp, -zp, bx, 1.0, 0.0, 0.0, bx, false).a * (-1.0 / 12.5663706);
      } else {
        const double z2 = -zp - hl - zg, ro2 = sqrt(r2 + z2 * z2);
        t2 = ro2 > eps ? aqt / ro2 : 0.0;
      }
      f = f + t1 + t2 + t3 + t4;
    }
    fs1[jj] = f;
  }
  if (!kFd) {  // VNV :338-357
    const double akair = ak * aire, adpi2 = akair / (2.0 * pi * pi), adpi = akair / (2.0 * pi);
    s[0] = fsp - (fs1[0] + fs1[1]) * adpi2;
    s[1] = fsm - (fs1[0] - fs1[1]) * adpi2;
    s[2] = -(fs2[0] + fs2[1]) * adpi;
    s[3] = -(fs2[0] - fs2[1]) * adpi;
  } else {  // VNVF :871-890
    s[0] = fsp + fs1[0] + fs1[1];
    s[1] = fsm + fs1[0] - fs1[1];
    s[2] = fs2[0] + fs2[1];
    s[3] = fs2[0] - fs2[1];
  }
}

// One workgroup per field point; its lanes stride over the panels and the
// per-problem sums are reduced in LDS in a fixed order (results do not
// depend on timing).  phi[p * nrhs + q]; bad[p] = 1 when the point has no
// value.  Sources sig[(i + s n) * nrhs + q]: s = 0 the symmetric part ZIGB,
// s = 1 the antisymmetric part ZIGS (nsym = 1 only).
template <bool kFd>
__global__ __launch_bounds__(256) void potential_kernel(Geometry G, Tables T, const FreqParams* __restrict__ Pg,
                                                        const double* __restrict__ pts, const double2* __restrict__ sig,
                                                        PotArgs A, double2* __restrict__ phi, int* __restrict__ bad) {
  __shared__ double2 red[256];
  __shared__ int anybad;
  const FreqParams& P = *Pg;
  const int t = threadIdx.x, n = G.n, p = blockIdx.x;
  if (t == 0) anybad = 0;
  const double xp = pts[3 * (size_t)p], yp = pts[3 * (size_t)p + 1], zp = pts[3 * (size_t)p + 2];
  cplx acc[kMaxRhs];
#pragma unroll
  for (int q = 0; q < kMaxRhs; ++q) acc[q] = cmk(0.0, 0.0);
  bool b = false;
  for (int j = t; j < n; j += 256) {
    double s[4];
    panel_potential<kFd>(G, T, P, A, j, xp, yp, zp, s, b);
    const double2* zb = sig + (size_t)j * A.nrhs;
    if (P.nsym == 0) {  // :61-67, PHI += ZIGB (SP1 + i SP2)
      const cplx sp = cmk(s[0], s[2]);
#pragma unroll
      for (int q = 0; q < kMaxRhs; ++q)
        if (q < A.nrhs) acc[q] = acc[q] + cmk(zb[q].x, zb[q].y) * sp;
    } else {  // :68-75, PHI += (ZIGB (S+ + S-) + ZIGS (S+ - S-)) / 2
      const double2* zs = sig + ((size_t)j + n) * A.nrhs;
      const cplx sa = cmk(0.5 * (s[0] + s[1]), 0.5 * (s[2] + s[3]));
      const cplx sd = cmk(0.5 * (s[0] - s[1]), 0.5 * (s[2] - s[3]));
#pragma unroll
      for (int q = 0; q < kMaxRhs; ++q)
        if (q < A.nrhs) acc[q] = acc[q] + cmk(zb[q].x, zb[q].y) * sa + cmk(zs[q].x, zs[q].y) * sd;
    }
  }
  __syncthreads();
  if (b) anybad = 1;  // every writer stores the same value
#pragma unroll
  for (int q = 0; q < kMaxRhs; ++q) {
    if (q < A.nrhs) {  // uniform across the workgroup
      red[t] = make_double2(acc[q].re, acc[q].im);
      __syncthreads();
      for (int k = 128; k > 0; k >>= 1) {
        if (t < k) red[t] = make_double2(red[t].x + red[t + k].x, red[t].y + red[t + k].y);
        __syncthreads();
      }
      if (t == 0) phi[(size_t)p * A.nrhs + q] = red[0];
      __syncthreads();
    }
  }
  __syncthreads();
  if (t == 0) bad[p] = anybad;
}

}  // namespace

void launch_potential(const Geometry& G, const Tables& T, const FreqParams* P, int fd, double dhsl, double a,
                      const double* pts, int npts, const double2* sig, int nrhs, double2* phi, int* bad,
                      hipStream_t st) {
  if (npts <= 0) return;
  const PotArgs A{npts, nrhs, dhsl, a};
  if (fd) potential_kernel<true><<<npts, 256, 0, st>>>(G, T, P, pts, sig, A, phi, bad);
  else potential_kernel<false><<<npts, 256, 0, st>>>(G, T, P, pts, sig, A, phi, bad);
}

}  // namespace nemoh

// HIP kernels of the NEMOH solver core for gfx950 (MI355X).
//
//   crek_kernel        wave-Green-function tables (NEMOH.f90:179-274)
//   assemble_kernel    influence matrices D+-, S+- for one frequency
//                      (SOLVE_BEM_INFD_DIRECT.f90:84-104 / FD :105-125)
//   gemv_kernel        Y = alpha*(M X) + beta*Y on FP64 MFMA, streaming M once
//                      (phi = S sigma, and the FP64 refinement residual)
//   rhs/combine/force  right-hand sides, sigma/phi combination of the
//                      symmetry blocks, pressure and force integration
//                      (SOLVE_BEM_*_DIRECT.f90:123-151, SOLVE_BEM.f90:91-96,
//                      NEMOH.f90:137-141)
#include <hip/hip_runtime.h>

#include "green.h"
#include "green_mixed.h"
#include "kernels.h"
#include "linalg.h"

namespace nemoh {

// ---------------------------------------------------------------------------
// Tables.  One thread per (XR, XZ) node; the 251-point Simpson sum runs in
// the reference's order.
__device__ __forceinline__ double simpson_weight(int j) {  // COFINT
  const double pi = M_PI;
  if (j == 1 || j == 251) return pi / (3.0 * 250.0);
  return (j % 2 != 0) ? 2.0 / (3.0 * 250.0) * pi : 4.0 / (3.0 * 250.0) * pi;
}

__global__ void crek_kernel(const double* __restrict__ xr, const double* __restrict__ xz, double* __restrict__ tab) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= kNR * kNZ) return;
  const int i = node % kNR, j = node / kNR;
  const double akz = xz[j], akr = xr[i];
  double f1x = 0, f1z = 0, f2x = 0, f2z = 0;
  for (int it = 1; it <= 251; ++it) {
    const double teta = -M_PI / 2.0 + (it - 1.0) / 250.0 * M_PI;
    const double cq = simpson_weight(it);
    const double ct = cos(teta);
    const cplx zik = cmk(akz, akr * ct);
    const cplx cex = (zik.re + 30.0 <= 0.0) ? cmk(0.0, 0.0) : cexp_(zik);
    const cplx g = gg(zik, cex);
    const cplx c1 = cq * (g - cdiv(cmk(1.0, 0.0), zik));
    const cplx c2 = cq * cex;
    f1x += ct * c1.im;
    f1z += c1.re;
    f2x += ct * c2.im;
    f2z += c2.re;
  }
  double* o = tab + 4 * (size_t)node;
  o[0] = f1x;
  o[1] = f1z;
  o[2] = f2x;
  o[3] = f2z;
}

// ---------------------------------------------------------------------------
// Assembly.  blockIdx.y = field point (row) i, 256 consecutive source panels
// j per workgroup: one wavefront covers 64 panels of one row, so the field
// point is uniform across the wave and panel data loads are coalesced.
//
// The bulk launch handles the pairs with both points below the free surface
// (zg < ZER), which is every pair of a mesh without panels at z = 0.  The
// other branch of the wave part (the reference's exact images or theta
// quadrature for points at the free surface) needs about 90 more registers
// per lane; compiled into the bulk kernel it would leave two waves per SIMD
// (one in infinite depth) instead of three.  Pairs with a point at the free
// surface are redone by assemble_surface_kernel, over the handle's list of
// such panels.
template <bool kSym, bool kF32>
__device__ __forceinline__ void store_pair(const Geometry& G, const MatOut& out, int i, int j, const double o[16]) {
  const double xn = G.xn[i], yn = G.yn[i], zn = G.zn[i];
  const size_t e = (size_t)i * out.ld + j, e32 = (size_t)i * out.ld32 + j;
  const double dpr = o[4] * xn + o[5] * yn + o[6] * zn;
  const double dpi = o[10] * xn + o[11] * yn + o[12] * zn;
  out.dp[e] = make_double2(dpr, dpi);
  out.sp[e] = make_double2(o[0], o[2]);
  if (kF32) out.dp32[e32] = make_float2((float)dpr, (float)dpi);
  if (kSym) {
    const double dmr = o[7] * xn + o[8] * yn + o[9] * zn;
    const double dmi = o[13] * xn + o[14] * yn + o[15] * zn;
    out.dm[e] = make_double2(dmr, dmi);
    out.sm[e] = make_double2(o[1], o[3]);
    if (kF32) out.dm32[e32] = make_float2((float)dmr, (float)dmi);
  }
}

template <bool kFd, bool kSym, bool kF32>
__global__ __launch_bounds__(256, 3) void assemble_kernel(Geometry G, Tables T, const FreqParams* __restrict__ Pp,
                                                          MatOut out) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= G.n) return;
  double o[16];
  // a pair with a point at the free surface gets the Rankine part and a zero
  // wave part here, overwritten by assemble_surface_kernel (an early return
  // for those pairs instead costs this kernel registers)
  pair16<kFd, false>(G, T, *Pp, i, j, o);
  store_pair<kSym, kF32>(G, out, i, j, o);
}

// ---------------------------------------------------------------------------
// Paired assembly: the wave part of each unordered pair {i, j} once
// (green.h, WaveSym), the Rankine part of each order (its near/far switch is
// the source panel's 7 TDIS), both elements written.
//
// Each wavefront takes an 8 x 8 block (ti, tj), ti <= tj, of the 8-panel
// blocks of the matrix: lane l holds i = 8 ti + l / 8, j = 8 tj + l % 8.
// Element (i, j) and element (j, i) are then both written as eight 128-byte
// row segments per store.  Waves take the blocks of the upper triangle in
// row-major order; in a diagonal block (ti == tj) only the lanes with i <= j
// work (i == j: the self term, one element).
//
// Each order's element is accumulated already projected on the field
// point's normal (PairAcc: S+-, D+- = n . grad S+-): the Rankine part of the
// order (rankine(), VAVINFD.f90:187-217), then the wave part of each y-image
// (the combination VNSINFD.f90:455-503 / VNSFD.f90:830-864) scaled by the
// order's source area.  Symmetric meshes (two y-images) hold more values
// live and run at two waves per SIMD, without spills; the others at three.
struct PairAcc {  // one order's element: S+ S- D+ D- (re, im)
  double spr, spi, smr, smi, dpr, dpi, dmr, dmi;
};

// Rankine part of order (f, s) into its element (real parts).
template <bool kFd, bool kSym>
__device__ __forceinline__ void pair_rankine(const Geometry& G, const FreqParams& P, int f, int s, double nx,
                                             double ny, double nz, PairAcc& a) {
  double rk[8];
  if (kFd) rankine(G, f, s, kSym ? 1 : 0, P.zer, 2, 2.0 * P.depth, rk);
  else rankine(G, f, s, kSym ? 1 : 0, P.zer, 1, 0.0, rk);
  a.spr = rk[0];
  a.dpr = rk[2] * nx + rk[3] * ny + rk[4] * nz;
  a.spi = a.dpi = 0.0;
  if (kSym) {
    a.smr = rk[1];
    a.dmr = rk[5] * nx + rk[6] * ny + rk[7] * nz;
    a.smi = a.dmi = 0.0;
  }
}

// Wave part of one y-image into an order's element: sd = +1 for (i, j), -1
// for (j, i); ey: the y offset of this image for this order; mb = +1 for the
// direct y-image, -1 for the mirrored one (its sign in the - block).
template <bool kSym>
__device__ __forceinline__ void pair_wave_add(const WaveSym& w, double sd, double ex, double ey, double nx,
                                              double ny, double nz, double pot1, double pot2, double grad1,
                                              double grad2, double mb, PairAcc& a) {
  const double s1 = pot1 * w.f1, s2 = pot2 * w.f2;
  const double d1 = grad1 * (w.h1 * (ex * nx + ey * ny) + (w.z1s + sd * w.z1a) * nz);
  const double d2 = grad2 * (w.h2 * (ex * nx + ey * ny) + (w.z2s + sd * w.z2a) * nz);
  a.spr += s1;
  a.spi += s2;
  a.dpr += d1;
  a.dpi += d2;
  if (kSym) {
    a.smr += mb * s1;
    a.smi += mb * s2;
    a.dmr += mb * d1;
    a.dmi += mb * d2;
  }
}

template <bool kSym, bool kF32>
__device__ __forceinline__ void store_acc(const MatOut& out, int f, int s, const PairAcc& a) {
  const size_t e = (size_t)f * out.ld + s, e32 = (size_t)f * out.ld32 + s;
  out.dp[e] = make_double2(a.dpr, a.dpi);
  out.sp[e] = make_double2(a.spr, a.spi);
  if (kF32) out.dp32[e32] = make_float2((float)a.dpr, (float)a.dpi);
  if (kSym) {
    out.dm[e] = make_double2(a.dmr, a.dmi);
    out.sm[e] = make_double2(a.smr, a.smi);
    if (kF32) out.dm32[e32] = make_float2((float)a.dmr, (float)a.dmi);
  }
}

// Block (ti, tj) of wave number g in the row-major upper triangle of an
// nt x nt block grid (row t starts at t nt - t (t - 1) / 2).
__device__ __forceinline__ void tri_block(int g, int nt, int& ti, int& tj) {
  const double b = 2.0 * nt + 1.0;
  int t = (int)((b - sqrt(b * b - 8.0 * g)) * 0.5);
  t = max(0, min(t, nt - 1));
  while (t > 0 && (long long)t * nt - (long long)t * (t - 1) / 2 > g) --t;
  while (t + 1 < nt && (long long)(t + 1) * nt - (long long)(t + 1) * t / 2 <= g) ++t;
  ti = t;
  tj = t + (g - (int)((long long)t * nt - (long long)t * (t - 1) / 2));
}

template <bool kFd, bool kSym, bool kF32>
__device__ __forceinline__ void assemble_pair_body(const Geometry& G, const Tables& T, const FreqParams* __restrict__ Pp,
                                                   const MatOut& out, int nt, int g) {
  int ti, tj;
  tri_block(g, nt, ti, tj);
  const int lane = threadIdx.x & 63;
  const int i = 8 * ti + (lane >> 3), j = 8 * tj + (lane & 7);
  if (i >= G.n || j >= G.n || (ti == tj && i > j)) return;
  const FreqParams& P = *Pp;
  const bool two = i != j;  // the self term is one element
  PairAcc a0, a1;
  pair_rankine<kFd, kSym>(G, P, i, j, G.xn[i], G.yn[i], G.zn[i], a0);
  if (two) pair_rankine<kFd, kSym>(G, P, j, i, G.xn[j], G.yn[j], G.zn[j], a1);
  const double xi = G.xg[i], yi = G.yg[i], zi = G.zg[i];
  const double xj = G.xg[j], yj = G.yg[j], zj = G.zg[j];
  // a pair with a point at the free surface keeps a zero wave part here; it
  // is overwritten by assemble_surface_kernel
  if (zi < P.zer && zj < P.zer) {
    const double dx = xi - xj, dy1 = yi - yj, dy2 = yi + yj;
    const double ai = G.aire[i], aj = G.aire[j];
    double p1[2], p2[2], g1[2], g2[2];  // scaling by the order's source area: [0] (i, j), [1] (j, i)
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      const double aire = d ? ai : aj;
      if (kFd) {
        p1[d] = p2[d] = g1[d] = g2[d] = aire;
      } else {
        const double akair = P.ak0 * aire;
        p1[d] = akair * (-1.0 / (2.0 * M_PI * M_PI));
        p2[d] = akair * (-1.0 / (2.0 * M_PI));
        g1[d] = p1[d] * P.ak0;
        g2[d] = p2[d] * P.ak0;
      }
    }
    const double nxi = G.xn[i], nyi = G.yn[i], nzi = G.zn[i];
    const double nxj = G.xn[j], nyj = G.yn[j], nzj = G.zn[j];
    {
      const WaveSym w = kFd ? pair_wave_fd(T, P, dx, dy1, zi, zj) : pair_wave_infd(T, P, dx, dy1, zi, zj);
      pair_wave_add<kSym>(w, 1.0, dx, dy1, nxi, nyi, nzi, p1[0], p2[0], g1[0], g2[0], 1.0, a0);
      pair_wave_add<kSym>(w, -1.0, -dx, -dy1, nxj, nyj, nzj, p1[1], p2[1], g1[1], g2[1], 1.0, a1);
    }
    if (kSym) {
      const WaveSym w = kFd ? pair_wave_fd(T, P, dx, dy2, zi, zj) : pair_wave_infd(T, P, dx, dy2, zi, zj);
      pair_wave_add<kSym>(w, 1.0, dx, dy2, nxi, nyi, nzi, p1[0], p2[0], g1[0], g2[0], -1.0, a0);
      pair_wave_add<kSym>(w, -1.0, -dx, dy2, nxj, nyj, nzj, p1[1], p2[1], g1[1], g2[1], -1.0, a1);
    }
  }
  store_acc<kSym, kF32>(out, i, j, a0);
  if (two) store_acc<kSym, kF32>(out, j, i, a1);
}

template <bool kFd, bool kSym, bool kF32>
__global__ __launch_bounds__(256, kSym ? 2 : 3) void assemble_pair_kernel(Geometry G, Tables T,
                                                                          const FreqParams* __restrict__ Pp,
                                                                          MatOut out, int nt, int nwaves) {
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= nwaves) return;
  assemble_pair_body<kFd, kSym, kF32>(G, T, Pp, out, nt, g);
}

// Grouped paired assembly: the matrices of many bodies in one launch
// (AsmJob j: workgroups [wg0, wg1), four waves of 8 x 8 pair blocks each).
template <bool kFd, bool kSym, bool kF32>
__global__ __launch_bounds__(256, kSym ? 2 : 3) void assemble_group_kernel(const AsmJob* __restrict__ jobs, int count) {
  const int b = blockIdx.x;
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].wg0 <= b) lo = mid;
    else hi = mid - 1;
  }
  const AsmJob& J = jobs[lo];
  const int g = (b - J.wg0) * 4 + (threadIdx.x >> 6);
  if (b >= J.wg1 || g >= J.nwaves) return;
  assemble_pair_body<kFd, kSym, kF32>(J.G, J.T, J.P, J.out, J.nt, g);
}


// The pairs with a point at the free surface: blockIdx.z = 0 takes the rows
// of the listed panels (every j), blockIdx.z = 1 their columns (every i that
// is not itself listed, so that no element is written twice).
template <bool kFd, bool kSym, bool kF32>
__global__ __launch_bounds__(256) void assemble_surface_kernel(Geometry G, Tables T,
                                                               const FreqParams* __restrict__ Pp, MatOut out,
                                                               const int* __restrict__ surf) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= G.n) return;
  const FreqParams& P = *Pp;
  const int s = surf[blockIdx.y];
  int i = s, j = k;
  if (blockIdx.z == 1) {
    i = k;
    j = s;
    if (!(G.zg[i] < P.zer)) return;  // a listed row: done by blockIdx.z = 0
  }
  double o[16];
  pair16<kFd, true>(G, T, P, i, j, o);
  store_pair<kSym, kF32>(G, out, i, j, o);
}

// Mixed path: D = n_i . grad G and S = G of the pair in single precision,
// D twice (the residual's copy and the one the LU factors in place).
template <bool kSym, typename R>
__device__ __forceinline__ void store_pair_m(const Geometry& G, const MatOut32& out, int i, int j, const R o[16]) {
  const float xn = (float)G.xn[i], yn = (float)G.yn[i], zn = (float)G.zn[i];
  const size_t e = (size_t)i * out.ld + j;
  const float2 dp = make_float2((float)o[4] * xn + (float)o[5] * yn + (float)o[6] * zn,
                                (float)o[10] * xn + (float)o[11] * yn + (float)o[12] * zn);
  out.dp[e] = dp;
  out.dp_fac[e] = dp;
  out.sp[e] = make_float2((float)o[0], (float)o[2]);
  if (kSym) {
    const float2 dm = make_float2((float)o[7] * xn + (float)o[8] * yn + (float)o[9] * zn,
                                  (float)o[13] * xn + (float)o[14] * yn + (float)o[15] * zn);
    out.dm[e] = dm;
    out.dm_fac[e] = dm;
    out.sm[e] = make_float2((float)o[1], (float)o[3]);
  }
}

// As assemble_kernel, in mixed precision (green_mixed.h); pairs with a point
// at the free surface are left to assemble_surface_mixed_kernel.
// Three waves per SIMD: 150 registers, no spills (four spill 20 registers of
// the inlined near-panel integral and ran within 5 % of three).
constexpr int kMixedWaves = 3;
template <bool kFd, bool kSym>
__global__ __launch_bounds__(256, kMixedWaves) void assemble_mixed_kernel(Geometry G, Tables T, Tables32 T32,
                                                                const FreqParams* __restrict__ Pp, MatOut32 out) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= G.n) return;
  float o[16];
  pair16_m<kFd>(G, T, T32, *Pp, i, j, o);
  store_pair_m<kSym>(G, out, i, j, o);
}

// The free-surface pairs of the mixed path, in FP64 (as
// assemble_surface_kernel), stored in single precision.
template <bool kFd, bool kSym>
__global__ __launch_bounds__(256) void assemble_surface_mixed_kernel(Geometry G, Tables T,
                                                                     const FreqParams* __restrict__ Pp, MatOut32 out,
                                                                     const int* __restrict__ surf) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= G.n) return;
  const FreqParams& P = *Pp;
  const int s = surf[blockIdx.y];
  int i = s, j = k;
  if (blockIdx.z == 1) {
    i = k;
    j = s;
    if (!(G.zg[i] < P.zer)) return;  // a listed row: done by blockIdx.z = 0
  }
  double o[16];
  pair16<kFd, true>(G, T, P, i, j, o);
  store_pair_m<kSym>(G, out, i, j, o);
}

// Raw 16-value dump of every pair (tests only; small meshes).
template <bool kFd>
__global__ void pairs_kernel(Geometry G, Tables T, const FreqParams* __restrict__ Pp, double* __restrict__ out) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= G.n) return;
  double o[16];
  pair16<kFd>(G, T, *Pp, i, j, o);
  double* d = out + ((size_t)i * G.n + j) * 16;
#pragma unroll
  for (int k = 0; k < 16; ++k) d[k] = o[k];
}

// ---------------------------------------------------------------------------
// Y(n x nrhs) = beta*Y + alpha*M(n x n) X(n x nrhs) in FP64 (phi = S sigma and
// the refinement residual b - D x).  Bound by streaming M once from HBM.
// One workgroup per 16 rows; its kGemvWaves waves split the columns, each
// running FP64 16x16x4 MFMAs with A = 16 rows x 4 columns of M and B = the
// matching 4 rows x 16 right-hand sides of X (nrhs <= 16), 8 k-steps per
// batch of loads; the partial blocks are summed through LDS in wave order.
// Sixteen waves (four per SIMD, one workgroup per CU) keep 128 KB of M in
// flight per CU, which HBM's latency needs at full bandwidth; four waves
// streamed at 2.9 TB/s.
constexpr int kGemvWaves = 16;
typedef double GemvRed[kGemvWaves][2][16][17];
__device__ __forceinline__ void gemv_body(GemvRed& red, const double2* __restrict__ M, int ld, int n,
                                          const double2* __restrict__ X, int nrhs, double2* __restrict__ Y,
                                          double alpha, double beta, int blk) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
  const int r0 = blk * 16;
  const int row = r0 + li;
  const int nk = (n + 3) / 4;  // k-steps of 4 columns
  const int per = (nk + kGemvWaves - 1) / kGemvWaves;
  const int s0 = wave * per, s1 = min(nk, s0 + per);
  // M and X through global (address space 1) views of their base pointers:
  // a grouped launch takes them from its job table, and through those
  // generic pointers the loads would be flat loads, whose waits also count
  // LDS accesses and every load in flight
  typedef const __attribute__((address_space(1))) double2 GD2;
  GD2* gm = (GD2*)M;
  GD2* gx = (GD2*)X;
  GD2* mrow = gm + (size_t)min(row, n - 1) * ld;
  v4d cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
  for (int s = s0; s < s1; s += 8) {
    double2 a[8], x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = 4 * (s + u) + lk;
      const bool kin = (s + u < s1) && k < n;
      a[u] = (kin && row < n) ? make_double2(mrow[k].x, mrow[k].y) : make_double2(0.0, 0.0);
      x[u] = (kin && li < nrhs) ? make_double2(gx[(size_t)k * nrhs + li].x, gx[(size_t)k * nrhs + li].y)
                                : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      cr = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].x, x[u].x, cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].x, x[u].y, ci, 0, 0, 0);
      cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[u].y, x[u].y, cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].y, x[u].x, ci, 0, 0, 0);
    }
  }
  // FP64 16x16x4 output: lane (li, lk) holds rows lk + 4 q, column li
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    red[wave][0][lk + 4 * q][li] = cr[q];
    red[wave][1][lk + 4 * q][li] = ci[q];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = lk + 4 * q;
      if (r0 + r < n && li < nrhs) {
        double vr = red[0][0][r][li], vi = red[0][1][r][li];
#pragma unroll
        for (int w = 1; w < kGemvWaves; ++w) {
          vr += red[w][0][r][li];
          vi += red[w][1][r][li];
        }
        double2* y = Y + (size_t)(r0 + r) * nrhs + li;
        const double2 old = (beta == 0.0) ? make_double2(0.0, 0.0) : *y;
        *y = make_double2(beta * old.x + alpha * vr, beta * old.y + alpha * vi);
      }
    }
  }
}
__global__ __launch_bounds__(64 * kGemvWaves) void gemv_kernel(const double2* __restrict__ M, int ld, int n,
                                                               const double2* __restrict__ X, int nrhs,
                                                               double2* __restrict__ Y, double alpha, double beta,
                                                               const RefineState* __restrict__ rs) {
  if (rs && rs->done) return;
  __shared__ GemvRed red;
  gemv_body(red, M, ld, n, X, nrhs, Y, alpha, beta, (int)blockIdx.x);
}

// Y = beta Y + alpha M X with M complex64 (the mixed path's D and S) and X,
// Y and the accumulation in FP64.  As gemv_kernel, one workgroup per 16 rows
// and 16 waves splitting the columns, but each lane streams 8 consecutive
// complex64 of its row (64 bytes, 256 bytes per row and wave) per 32-column
// batch: the MFMA k-step t of a batch takes columns 8 lk + t of the four
// lane groups lk, and the x operand is loaded in the same permuted order, so
// the product is unchanged.  ld: a multiple of 4 (16-byte aligned rows).
constexpr int kGemv32Waves = 16;
__global__ __launch_bounds__(64 * kGemv32Waves) void gemv32_kernel(const float2* __restrict__ M, int ld, int n,
                                                                   const double2* __restrict__ X, int nrhs,
                                                                   double2* __restrict__ Y, double alpha,
                                                                   double beta, const RefineState* __restrict__ rs) {
  if (rs && rs->done) return;
  __shared__ double red[kGemv32Waves][2][16][17];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
  const int r0 = blockIdx.x * 16;
  const int row = r0 + li;
  const int nb = (n + 31) / 32;  // 32-column batches
  const int per = (nb + kGemv32Waves - 1) / kGemv32Waves;
  const int b0 = wave * per, b1 = min(nb, b0 + per);
  const float2* mrow = M + (size_t)min(row, n - 1) * ld;
  v4d cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
  for (int bb = b0; bb < b1; ++bb) {
    const int c0 = bb * 32 + 8 * lk;
    float2 a[8];
    if (row < n && c0 + 8 <= n) {
      const float4* q = reinterpret_cast<const float4*>(mrow + c0);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float4 v = q[u];
        a[2 * u] = make_float2(v.x, v.y);
        a[2 * u + 1] = make_float2(v.z, v.w);
      }
    } else {
#pragma unroll
      for (int t = 0; t < 8; ++t) a[t] = (row < n && c0 + t < n) ? mrow[c0 + t] : make_float2(0.f, 0.f);
    }
    double2 x[8];
#pragma unroll
    for (int t = 0; t < 8; ++t)
      x[t] = (c0 + t < n && li < nrhs) ? X[(size_t)(c0 + t) * nrhs + li] : make_double2(0.0, 0.0);
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const double ar = a[t].x, ai = a[t].y;
      cr = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, x[t].x, cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, x[t].y, ci, 0, 0, 0);
      cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, x[t].y, cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, x[t].x, ci, 0, 0, 0);
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    red[wave][0][lk + 4 * q][li] = cr[q];
    red[wave][1][lk + 4 * q][li] = ci[q];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = lk + 4 * q;
      if (r0 + r < n && li < nrhs) {
        double vr = red[0][0][r][li], vi = red[0][1][r][li];
#pragma unroll
        for (int w = 1; w < kGemv32Waves; ++w) {
          vr += red[w][0][r][li];
          vi += red[w][1][r][li];
        }
        double2* y = Y + (size_t)(r0 + r) * nrhs + li;
        const double2 old = (beta == 0.0) ? make_double2(0.0, 0.0) : *y;
        *y = make_double2(beta * old.x + alpha * vr, beta * old.y + alpha * vi);
      }
    }
  }
}

// B_b = (V[:n] + bx V[n:]) / 2 (symmetric) or V (SOLVE_BEM_INFD_DIRECT.f90:
// 125-131); V is (ntot x nrhs) row-major.
__global__ void rhs_kernel(const double2* __restrict__ V, int n, int nsym, int nrhs, double bx,
                           double2* __restrict__ B) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * nrhs) return;
  const int i = e / nrhs, q = e % nrhs;
  const double2 a = V[(size_t)i * nrhs + q];
  if (nsym) {
    const double2 b = V[(size_t)(i + n) * nrhs + q];
    B[e] = make_double2((a.x + bx * b.x) * 0.5, (a.y + bx * b.y) * 0.5);
  } else {
    B[e] = a;
  }
}

// sigma and phi of the two symmetry blocks -> ZIGB, ZIGS, ZPB, ZPS and the
// pressure p = rho i w ZP on the full (2n) surface (SOLVE_BEM.f90:91-96).
__global__ void combine_kernel(const double2* __restrict__ Z1, const double2* __restrict__ Z2,
                               const double2* __restrict__ F1, const double2* __restrict__ F2, int n, int nsym,
                               int nrhs, double rho_w, double2* __restrict__ sig, double2* __restrict__ pot,
                               double2* __restrict__ pres) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * nrhs) return;
  const int i = e / nrhs, q = e % nrhs;
  double2 zb, zs, pb, ps;
  if (nsym) {
    const double2 a = Z1[e], b = Z2[e], c = F1[e], d = F2[e];
    zb = make_double2(a.x + b.x, a.y + b.y);
    zs = make_double2(a.x - b.x, a.y - b.y);
    pb = make_double2(c.x + d.x, c.y + d.y);
    ps = make_double2(c.x - d.x, c.y - d.y);
  } else {
    zb = Z1[e];
    zs = make_double2(0.0, 0.0);
    pb = F1[e];
    ps = make_double2(0.0, 0.0);
  }
  if (sig) {
    sig[(size_t)i * nrhs + q] = zb;
    sig[(size_t)(i + n) * nrhs + q] = zs;
  }
  if (pot) {
    pot[(size_t)i * nrhs + q] = pb;
    pot[(size_t)(i + n) * nrhs + q] = ps;
  }
  // p = rho * i * w * phi
  pres[(size_t)i * nrhs + q] = make_double2(-rho_w * pb.y, rho_w * pb.x);
  if (nsym) pres[(size_t)(i + n) * nrhs + q] = make_double2(-rho_w * ps.y, rho_w * ps.x);
}

// F(k, q) = -sum_c p(c, q) NDS(k, c); one workgroup per (k, q), fixed-order
// tree reduction (deterministic, batch-invariant).
__device__ __forceinline__ void force_body(double* sr, double* si, const double2* __restrict__ pres,
                                           const double* __restrict__ nds, int ntot, int nrhs, double2* __restrict__ F,
                                           int blk) {
  const int k = blk / nrhs, q = blk % nrhs;
  double ar = 0.0, ai = 0.0;
  for (int c = threadIdx.x; c < ntot; c += 256) {
    const double2 p = pres[(size_t)c * nrhs + q];
    const double w = nds[(size_t)k * ntot + c];
    ar -= p.x * w;
    ai -= p.y * w;
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) F[(size_t)k * nrhs + q] = make_double2(sr[0], si[0]);
}
__global__ __launch_bounds__(256) void force_kernel(const double2* __restrict__ pres, const double* __restrict__ nds,
                                                    int ntot, int nrhs, double2* __restrict__ F) {
  __shared__ double sr[256], si[256];
  force_body(sr, si, pres, nds, ntot, nrhs, F, (int)blockIdx.x);
}

// Mixed-path refinement on the device.  The sweep's kernels read the
// block's RefineState: the residual scale rmax, and `done`, after which they
// return at once, so the host can queue sweeps without waiting for each
// convergence test.
//
// max that keeps NaN (fmax would drop it, and an all-NaN residual would
// then look converged)
__device__ __forceinline__ double nan_max(double m, double v) { return (v > m || v != v) ? v : m; }

// R32 = R / rmax (R32 = 0 when rmax = 0: an all-zero right-hand side)
__global__ void f64_to_f32_kernel(const double2* __restrict__ a, float2* __restrict__ b, size_t count,
                                  const RefineState* __restrict__ rs) {
  if (rs->done) return;
  const double scale = rs->rmax > 0.0 ? 1.0 / rs->rmax : 0.0;
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e < count) b[e] = make_float2((float)(a[e].x * scale), (float)(a[e].y * scale));
}
// Z (+)= rmax * X32
__global__ void axpy_f32_to_f64_kernel(const float2* __restrict__ a, double2* __restrict__ b, size_t count,
                                       const RefineState* __restrict__ rs, int accumulate) {
  if (rs->done) return;
  const double scale = rs->rmax;
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e < count) {
    const double2 v = make_double2((double)a[e].x * scale, (double)a[e].y * scale);
    b[e] = accumulate ? make_double2(b[e].x + v.x, b[e].y + v.y) : v;
  }
}
__global__ void copy_c128_kernel(const double2* __restrict__ a, double2* __restrict__ b, size_t count,
                                 const RefineState* __restrict__ rs) {
  if (rs && rs->done) return;
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e < count) b[e] = a[e];
}
// max |a| over count complex entries into out[blockIdx.x] (per-block maxima,
// NaN kept)
__global__ __launch_bounds__(256) void absmax_kernel(const double2* __restrict__ a, size_t count,
                                                     double* __restrict__ out, const RefineState* __restrict__ rs) {
  if (rs && rs->done) return;
  __shared__ double s[256];
  double m = 0.0;
  for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < count; e += (size_t)gridDim.x * 256)
    m = nan_max(nan_max(m, fabs(a[e].x)), fabs(a[e].y));
  s[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) s[threadIdx.x] = nan_max(s[threadIdx.x], s[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}
// Convergence test after a sweep (first: the right-hand side's max before
// any sweep).  One wave reduces the partial maxima of absmax_kernel.
__global__ void refine_check_kernel(const double* __restrict__ part, int nparts, RefineState* __restrict__ rs,
                                    double tol, int first) {
  if (rs->done) return;
  double m = 0.0;
  for (int k = threadIdx.x; k < nparts; k += 64) m = nan_max(m, part[k]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_xor(m, off, 64));
  if (threadIdx.x != 0) return;
  if (!(m <= 1.79e308)) {  // NaN or infinite: stop, the host reports it
    rs->bad = 1;
    rs->rmax = m;
    rs->done = 1;
    return;
  }
  rs->rmax = m;
  if (first) {
    rs->bmax = m;
    rs->iters = 0;
    return;
  }
  rs->iters += 1;
  if (m <= tol * rs->bmax) rs->done = 1;
}
__global__ void refine_reset_kernel(RefineState* __restrict__ rs) {
  rs->bmax = rs->rmax = 0.0;
  rs->done = rs->iters = rs->bad = 0;
}

// Largest squared horizontal vertex distance (GM^2 of
// SOLVE_BEM_INFD_DIRECT.f90:66-75, SOLVE_BEM_FD_DIRECT.f90:86-95): over the
// pairs i < j, and with nsym = 1 also from vertex i to the y-image of j.
// Workgroup b takes one pair of 256-vertex tiles (ti <= tj) in triangular
// order; thread t holds vertex ti*256+t against tile tj staged in LDS.  Max
// is exact and order-free, and the squares are formed without contraction,
// so the result equals the sequential loop bit for bit.
__global__ __launch_bounds__(256) void extent_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                     int np, int nsym, double* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double sx[256], sy[256], red[256];
  const int ntile = (np + 255) / 256;
  int b = blockIdx.x, ti = 0;
  while (b >= ntile - ti) {
    b -= ntile - ti;
    ++ti;
  }
  const int tj = ti + b;
  const int t = threadIdx.x, i = ti * 256 + t, j0 = tj * 256;
  const int jn = min(256, np - j0);
  sx[t] = (t < jn) ? x[j0 + t] : 0.0;
  sy[t] = (t < jn) ? y[j0 + t] : 0.0;
  __syncthreads();
  double m = 0.0;
  if (i < np) {
    const double xi = x[i], yi = y[i];
    for (int q = (tj == ti) ? t + 1 : 0; q < jn; ++q) {
      const double dx = sx[q] - xi, dy = yi - sy[q];
      m = fmax(dx * dx + dy * dy, m);
      if (nsym) {
        const double dz = yi + sy[q];
        m = fmax(dx * dx + dz * dz, m);
      }
    }
  }
  red[t] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (t < k) red[t] = fmax(red[t], red[t + k]);
    __syncthreads();
  }
  if (t == 0) out[blockIdx.x] = red[0];
}

// Kochin function (COMPUTE_KOCHIN.f90:25-153) of every problem of the last
// solve: one workgroup per angle; thread t walks panels t, t+256, ..., forms
// each panel's exact integral of exp(k (z - i wbar)) as a sum over its four
// edges from the vertex exponentials CEP = exp(ZJ), CEM = exp(-ZJ - 2kh),
// ZJ = k (z - i wbar), for the body and (NSYMY = 1) its y-image, and
// accumulates sigma * ZS for every problem; the panel sums are reduced in a
// fixed order.  H[theta * nrhs + q].
struct KochinArgs {
  double k, depth, xeff, yeff;
  int nsym, nrhs;
};
__global__ __launch_bounds__(256) void kochin_kernel(Geometry G, const double2* __restrict__ sig,
                                                     const double* __restrict__ theta, KochinArgs A,
                                                     double2* __restrict__ H) {
  __shared__ double2 red[256];
  const int t = threadIdx.x, n = G.n;
  const double th = theta[blockIdx.x];
  const double ct = cos(th), st = sin(th);
  const double kh = A.k * A.depth;
  const double scale = (kh > 18.0 || kh <= 0.0) ? 1.0 / A.k : 0.5 * exp(kh) / sinh(kh) * (1.0 / A.k);
  cplx acc[kMaxRhs];
#pragma unroll
  for (int q = 0; q < kMaxRhs; ++q) acc[q] = cmk(0.0, 0.0);
  for (int i = t; i < n; i += 256) {
    for (int img = 0; img <= A.nsym; ++img) {
      const double sy = img ? -1.0 : 1.0;
      cplx zj[5], cep[5], cem[5];
      double vx[5], vy[5];
      for (int l = 0; l < 5; ++l) {
        const int v = 4 * i + (l & 3);
        vx[l] = G.vx[v];
        vy[l] = G.vy[v];
        const double z = G.vz[v];
        const double wbar = (vx[l] - A.xeff) * ct + (sy * vy[l] - A.yeff) * st;
        zj[l] = cmk(A.k * z, -A.k * wbar);
        cep[l] = (zj[l].re < -18.0) ? cmk(0.0, 0.0) : cexp_(zj[l]);
        const cplx em = cmk(-zj[l].re - 2.0 * kh, -zj[l].im);
        cem[l] = (em.re < -18.0 || -kh >= 0.0) ? cmk(0.0, 0.0) : cexp_(em);
      }
      const cplx zms = cmk(sy * G.yn[i], -G.zn[i] * st);
      const cplx zns = cmk(G.xn[i], -G.zn[i] * ct);
      cplx zs = cmk(0.0, 0.0);
      for (int l = 0; l < 4; ++l) {
        const double dx = vx[l + 1] - vx[l], dy = vy[l + 1] - vy[l];
        const cplx zcs = img ? (dx * zms + dy * zns) : (dx * zms - dy * zns);
        const cplx zc = zj[l + 1] - zj[l];
        cplx zhs, zhd;
        if (fabs(zc.im) < 1.E-04 && fabs(zc.re) < 1.E-04) {
          zhs = 0.5 * (cep[l + 1] + cep[l]);
          zhd = -0.5 * (cem[l + 1] + cem[l]);
        } else {
          zhs = cdiv(cep[l + 1] - cep[l], zc);
          zhd = cdiv(cem[l + 1] - cem[l], zc);
        }
        const cplx a = zcs * zhs, b = zcs * zhd;
        zs = zs + a + cmk(b.re, -b.im);
      }
      zs = (img ? -scale : scale) * zs;
      const double2* s = sig + (size_t)(i + img * n) * A.nrhs;
#pragma unroll
      for (int q = 0; q < kMaxRhs; ++q)
        if (q < A.nrhs) {
          const double2 v = s[q];
          acc[q] = acc[q] + cmk(v.x, v.y) * zs;
        }
    }
  }
  const double c4 = 0.25 / M_PI;
#pragma unroll
  for (int q = 0; q < kMaxRhs; ++q) {
    if (q < A.nrhs) {  // uniform across the workgroup
      red[t] = make_double2(acc[q].re, acc[q].im);
      __syncthreads();
      for (int k = 128; k > 0; k >>= 1) {
        if (t < k) red[t] = make_double2(red[t].x + red[t + k].x, red[t].y + red[t + k].y);
        __syncthreads();
      }
      if (t == 0) H[(size_t)blockIdx.x * A.nrhs + q] = make_double2(c4 * red[0].x, c4 * red[0].y);
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Grouped per-body kernels of a batch: the same per-element operations over
// every body's buffers in one launch (BodyJob j: its workgroup range of each
// kind; bodies found by binary search).  Each body gets the arithmetic of
// its own launches, so results do not depend on the grouping.
template <int kKind>
__device__ __forceinline__ int body_of(const BodyJob* __restrict__ jobs, int count, int& w) {
  const int g = blockIdx.x;
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].wg[kKind] <= g) lo = mid;
    else hi = mid - 1;
  }
  w = g - jobs[lo].wg[kKind];
  return (w < jobs[lo].nwg[kKind]) ? lo : -1;
}

// elementwise steps over n x nrhs entries (kind 0)
__global__ void group_elementwise_kernel(const BodyJob* __restrict__ jobs, int count, int op, double bx) {
  int w;
  const int j = body_of<0>(jobs, count, w);
  if (j < 0) return;
  const BodyJob& J = jobs[j];
  const size_t e = (size_t)w * blockDim.x + threadIdx.x, cnt = (size_t)J.n * J.nrhs;
  if (e >= cnt) return;
  switch (op) {
    case kGroupRhs: {  // B = the block's right-hand sides from V
      const int i = (int)(e / J.nrhs), q = (int)(e % J.nrhs);
      const double2 a = J.V[(size_t)i * J.nrhs + q];
      if (J.nsym) {
        const double2 b = J.V[(size_t)(i + J.n) * J.nrhs + q];
        J.B[e] = make_double2((a.x + bx * b.x) * 0.5, (a.y + bx * b.y) * 0.5);
      } else {
        J.B[e] = a;
      }
      break;
    }
    case kGroupCopyBZ: J.Z[e] = J.B[e]; break;
    case kGroupCopyBR: J.R[e] = J.B[e]; break;
    case kGroupCopyBRSkip:
      if (!J.rs->done) J.R[e] = J.B[e];
      break;
    case kGroupToF32:
      if (!J.rs->done) {
        const double scale = J.rs->rmax > 0.0 ? 1.0 / J.rs->rmax : 0.0;
        J.R32[e] = make_float2((float)(J.R[e].x * scale), (float)(J.R[e].y * scale));
      }
      break;
    case kGroupToF64Set:
    case kGroupToF64Add:
      if (!J.rs->done) {
        const double scale = J.rs->rmax;
        const double2 v = make_double2((double)J.R32[e].x * scale, (double)J.R32[e].y * scale);
        J.Z[e] = op == kGroupToF64Add ? make_double2(J.Z[e].x + v.x, J.Z[e].y + v.y) : v;
      }
      break;
    default: break;
  }
}

// per-body maxima of |R| (or |B| for the first test) in 64 parts each, then
// the convergence test (one wave per body), as absmax_kernel and
// refine_check_kernel do for one body
__global__ __launch_bounds__(256) void group_absmax_kernel(const BodyJob* __restrict__ jobs, int count, int first) {
  const int j = blockIdx.x / 64, part = blockIdx.x % 64;
  if (j >= count) return;
  const BodyJob& J = jobs[j];
  if (!first && J.rs->done) return;
  const double2* a = first ? J.B : J.R;
  const size_t cnt = (size_t)J.n * J.nrhs;
  __shared__ double s[256];
  double m = 0.0;
  for (size_t e = part * (size_t)256 + threadIdx.x; e < cnt; e += (size_t)64 * 256)
    m = nan_max(nan_max(m, fabs(a[e].x)), fabs(a[e].y));
  s[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) s[threadIdx.x] = nan_max(s[threadIdx.x], s[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) J.red[part] = s[0];
}
__global__ void group_refine_check_kernel(const BodyJob* __restrict__ jobs, int count, int first) {
  const int j = blockIdx.x;
  if (j >= count) return;
  const BodyJob& J = jobs[j];
  RefineState* rs = J.rs;
  if (rs->done) return;
  double m = nan_max(0.0, J.red[threadIdx.x]);  // 64 partial maxima, one per lane
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_xor(m, off, 64));
  if (threadIdx.x != 0) return;
  if (!(m <= 1.79e308)) {  // NaN or infinite: stop, the host reports it
    rs->bad = 1;
    rs->rmax = m;
    rs->done = 1;
    return;
  }
  rs->rmax = m;
  if (first) {
    rs->bmax = m;
    rs->iters = 0;
    return;
  }
  rs->iters += 1;
  if (m <= J.tol * rs->bmax) rs->done = 1;
}
__global__ void group_refine_reset_kernel(const BodyJob* __restrict__ jobs, int count) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  RefineState* rs = jobs[j].rs;
  rs->bmax = rs->rmax = 0.0;
  rs->done = rs->iters = rs->bad = 0;
}

// the residual R -= D Z (skipped once converged) or phi F = S Z (kind 1:
// ceil(n / 16) workgroups per body)
__global__ __launch_bounds__(64 * kGemvWaves) void group_gemv_kernel(const BodyJob* __restrict__ jobs, int count,
                                                                     int phi) {
  int w;
  const int j = body_of<1>(jobs, count, w);
  if (j < 0) return;
  const BodyJob& J = jobs[j];
  if (!phi && J.rs->done) return;
  __shared__ GemvRed red;
  if (phi) gemv_body(red, J.S, J.ld, J.n, J.Z, J.nrhs, J.F, 1.0, 0.0, w);
  else gemv_body(red, J.D, J.ld, J.n, J.Z, J.nrhs, J.R, -1.0, 1.0, w);
}

// sigma, phi and pressure of both symmetry blocks (block 0's jobs carry
// block 1's Z and F in Z2, F2), as combine_kernel; omega[j]: body j's
// frequency
__global__ void group_combine_kernel(const BodyJob* __restrict__ jobs, int count, const double* __restrict__ omega) {
  int w;
  const int j = body_of<0>(jobs, count, w);
  if (j < 0) return;
  const BodyJob& J = jobs[j];
  const int e = w * blockDim.x + threadIdx.x;
  if (e >= J.n * J.nrhs) return;
  const int i = e / J.nrhs, q = e % J.nrhs, nrhs = J.nrhs, n = J.n;
  double2 zb, zs, pb, ps;
  if (J.nsym) {
    const double2 a = J.Z[e], b = J.Z2[e], c = J.F[e], d = J.F2[e];
    zb = make_double2(a.x + b.x, a.y + b.y);
    zs = make_double2(a.x - b.x, a.y - b.y);
    pb = make_double2(c.x + d.x, c.y + d.y);
    ps = make_double2(c.x - d.x, c.y - d.y);
  } else {
    zb = J.Z[e];
    zs = make_double2(0.0, 0.0);
    pb = J.F[e];
    ps = make_double2(0.0, 0.0);
  }
  J.sig[(size_t)i * nrhs + q] = zb;
  J.sig[(size_t)(i + n) * nrhs + q] = zs;
  J.pot[(size_t)i * nrhs + q] = pb;
  J.pot[(size_t)(i + n) * nrhs + q] = ps;
  const double rho_w = J.rho * omega[j];
  J.pres[(size_t)i * nrhs + q] = make_double2(-rho_w * pb.y, rho_w * pb.x);
  if (J.nsym) J.pres[(size_t)(i + n) * nrhs + q] = make_double2(-rho_w * ps.y, rho_w * ps.x);
}

// forces -NDS . p of every body (kind 2: nforce x nrhs workgroups) and the
// Froude-Krylov forces from the incident pressure (kind 3: nforce x nbeta),
// into force[0, nforce nrhs) and after it
__global__ __launch_bounds__(256) void group_force_kernel(const BodyJob* __restrict__ jobs, int count, int fk) {
  int w;
  const int j = fk ? body_of<3>(jobs, count, w) : body_of<2>(jobs, count, w);
  if (j < 0) return;
  const BodyJob& J = jobs[j];
  __shared__ double sr[256], si[256];
  const int ntot = J.n * (J.nsym ? 2 : 1);
  if (fk) force_body(sr, si, J.pinc, J.nds, ntot, J.nbeta, J.force + (size_t)J.nforce * J.nrhs, w);
  else force_body(sr, si, J.pres, J.nds, ntot, J.nrhs, J.force, w);
}

// every body's pivot report reset (before the factorisation), and every
// body's report and refinement states gathered into one array (after the
// solves): out[j] = {info[0], info[1]}, then the block states
__global__ void group_info_reset_kernel(const BodyJob* __restrict__ jobs, int count) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  jobs[j].info[0] = 0;
  jobs[j].info[1] = 0;
}
__global__ void group_status_kernel(const BodyJob* __restrict__ jobs, int count, int nb, BatchStatus* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  BatchStatus s;
  s.info[0] = jobs[j].info[0];
  s.info[1] = jobs[j].info[1];
  s.rs[0] = jobs[j].rs[0];
  if (nb > 1) s.rs[1] = jobs[j].rs[1];
  out[j] = s;
}

// ---------------------------------------------------------------------------
// host-side launchers
void launch_crek(const double* xr, const double* xz, double* tab, hipStream_t st) {
  crek_kernel<<<(kNR * kNZ + 127) / 128, 128, 0, st>>>(xr, xz, tab);
}

template <bool kFd, bool kSym, bool kF32>
static void launch_assemble_variant(const Geometry& G, const Tables& T, const FreqParams* P, const MatOut& out,
                                    const int* surf, int nsurf, bool paired, hipStream_t st) {
  const int nx = (G.n + 255) / 256;
  if (paired) {
    const int nt = (G.n + 7) / 8, nwaves = nt * (nt + 1) / 2;
    assemble_pair_kernel<kFd, kSym, kF32><<<(nwaves + 3) / 4, 256, 0, st>>>(G, T, P, out, nt, nwaves);
  } else {
    assemble_kernel<kFd, kSym, kF32><<<dim3(nx, G.n), 256, 0, st>>>(G, T, P, out);
  }
  if (nsurf > 0) assemble_surface_kernel<kFd, kSym, kF32><<<dim3(nx, nsurf, 2), 256, 0, st>>>(G, T, P, out, surf);
}

template <bool kFd>
static void launch_assemble_regime(const Geometry& G, const Tables& T, const FreqParams* P, int nsym,
                                   const MatOut& out, const int* surf, int nsurf, bool paired, hipStream_t st) {
  const bool f32 = out.dp32 != nullptr;
  if (nsym) {
    if (f32) launch_assemble_variant<kFd, true, true>(G, T, P, out, surf, nsurf, paired, st);
    else launch_assemble_variant<kFd, true, false>(G, T, P, out, surf, nsurf, paired, st);
  } else {
    if (f32) launch_assemble_variant<kFd, false, true>(G, T, P, out, surf, nsurf, paired, st);
    else launch_assemble_variant<kFd, false, false>(G, T, P, out, surf, nsurf, paired, st);
  }
}

void launch_assemble(const Geometry& G, const Tables& T, const FreqParams* P, int fd, int nsym, const MatOut& out,
                     const int* surf, int nsurf, bool paired, hipStream_t st) {
  if (fd) launch_assemble_regime<true>(G, T, P, nsym, out, surf, nsurf, paired, st);
  else launch_assemble_regime<false>(G, T, P, nsym, out, surf, nsurf, paired, st);
}

template <bool kFd, bool kSym, bool kF32>
static void launch_surface_variant(const Geometry& G, const Tables& T, const FreqParams* P, const MatOut& out,
                                   const int* surf, int nsurf, hipStream_t st) {
  const int nx = (G.n + 255) / 256;
  assemble_surface_kernel<kFd, kSym, kF32><<<dim3(nx, nsurf, 2), 256, 0, st>>>(G, T, P, out, surf);
}

void launch_assemble_surface(const Geometry& G, const Tables& T, const FreqParams* P, int fd, int nsym,
                             const MatOut& out, const int* surf, int nsurf, hipStream_t st) {
  if (nsurf <= 0) return;
  const bool f32 = out.dp32 != nullptr;
  if (fd) {
    if (nsym) f32 ? launch_surface_variant<true, true, true>(G, T, P, out, surf, nsurf, st)
              : launch_surface_variant<true, true, false>(G, T, P, out, surf, nsurf, st);
    else f32 ? launch_surface_variant<true, false, true>(G, T, P, out, surf, nsurf, st)
             : launch_surface_variant<true, false, false>(G, T, P, out, surf, nsurf, st);
  } else {
    if (nsym) f32 ? launch_surface_variant<false, true, true>(G, T, P, out, surf, nsurf, st)
              : launch_surface_variant<false, true, false>(G, T, P, out, surf, nsurf, st);
    else f32 ? launch_surface_variant<false, false, true>(G, T, P, out, surf, nsurf, st)
             : launch_surface_variant<false, false, false>(G, T, P, out, surf, nsurf, st);
  }
}

template <bool kFd, bool kSym, bool kF32>
static void launch_group_variant(const AsmJob* jobs, int count, int grid, hipStream_t st) {
  assemble_group_kernel<kFd, kSym, kF32><<<grid, 256, 0, st>>>(jobs, count);
}

void launch_assemble_group(const AsmJob* jobs, int count, int grid, int fd, int nsym, bool f32, hipStream_t st) {
  if (count <= 0) return;
  if (fd) {
    if (nsym) f32 ? launch_group_variant<true, true, true>(jobs, count, grid, st)
              : launch_group_variant<true, true, false>(jobs, count, grid, st);
    else f32 ? launch_group_variant<true, false, true>(jobs, count, grid, st)
             : launch_group_variant<true, false, false>(jobs, count, grid, st);
  } else {
    if (nsym) f32 ? launch_group_variant<false, true, true>(jobs, count, grid, st)
              : launch_group_variant<false, true, false>(jobs, count, grid, st);
    else f32 ? launch_group_variant<false, false, true>(jobs, count, grid, st)
             : launch_group_variant<false, false, false>(jobs, count, grid, st);
  }
}

template <bool kFd, bool kSym>
static void launch_assemble_mixed_variant(const Geometry& G, const Tables& T, const float4* tab32,
                                          const FreqParams* P, const MatOut32& out, const int* surf, int nsurf,
                                          hipStream_t st) {
  const int nx = (G.n + 255) / 256;
  assemble_mixed_kernel<kFd, kSym><<<dim3(nx, G.n), 256, 0, st>>>(G, T, Tables32{tab32}, P, out);
  if (nsurf > 0) assemble_surface_mixed_kernel<kFd, kSym><<<dim3(nx, nsurf, 2), 256, 0, st>>>(G, T, P, out, surf);
}

void launch_assemble_mixed(const Geometry& G, const Tables& T, const float4* tab32, const FreqParams* P, int fd,
                           int nsym, const MatOut32& out, const int* surf, int nsurf, hipStream_t st) {
  if (fd) {
    if (nsym) launch_assemble_mixed_variant<true, true>(G, T, tab32, P, out, surf, nsurf, st);
    else launch_assemble_mixed_variant<true, false>(G, T, tab32, P, out, surf, nsurf, st);
  } else {
    if (nsym) launch_assemble_mixed_variant<false, true>(G, T, tab32, P, out, surf, nsurf, st);
    else launch_assemble_mixed_variant<false, false>(G, T, tab32, P, out, surf, nsurf, st);
  }
}

void launch_pairs(const Geometry& G, const Tables& T, const FreqParams* P, int fd, double* out, hipStream_t st) {
  dim3 grid((G.n + 63) / 64, G.n);
  if (fd) pairs_kernel<true><<<grid, 64, 0, st>>>(G, T, P, out);
  else pairs_kernel<false><<<grid, 64, 0, st>>>(G, T, P, out);
}

void launch_matvec(const double2* M, int ld, int n, const double2* X, int nrhs, double2* Y, double alpha,
                   double beta, hipStream_t st, const RefineState* skip) {
  gemv_kernel<<<(n + 15) / 16, 64 * kGemvWaves, 0, st>>>(M, ld, n, X, nrhs, Y, alpha, beta, skip);
}

void launch_matvec32(const float2* M, int ld, int n, const double2* X, int nrhs, double2* Y, double alpha,
                     double beta, hipStream_t st, const RefineState* skip) {
  gemv32_kernel<<<(n + 15) / 16, 64 * kGemv32Waves, 0, st>>>(M, ld, n, X, nrhs, Y, alpha, beta, skip);
}

void launch_group_elementwise(const BodyJob* jobs, int count, int grid, int op, double bx, hipStream_t st) {
  if (count > 0 && grid > 0) group_elementwise_kernel<<<grid, 256, 0, st>>>(jobs, count, op, bx);
}
void launch_group_residual_max(const BodyJob* jobs, int count, int first, hipStream_t st) {
  if (count <= 0) return;
  group_absmax_kernel<<<64 * count, 256, 0, st>>>(jobs, count, first);
  group_refine_check_kernel<<<count, 64, 0, st>>>(jobs, count, first);
}
void launch_group_refine_reset(const BodyJob* jobs, int count, hipStream_t st) {
  if (count > 0) group_refine_reset_kernel<<<(count + 255) / 256, 256, 0, st>>>(jobs, count);
}
void launch_group_gemv(const BodyJob* jobs, int count, int grid, int phi, hipStream_t st) {
  if (count > 0 && grid > 0) group_gemv_kernel<<<grid, 64 * kGemvWaves, 0, st>>>(jobs, count, phi);
}
void launch_group_combine(const BodyJob* jobs, int count, int grid, const double* omega, hipStream_t st) {
  if (count > 0 && grid > 0) group_combine_kernel<<<grid, 256, 0, st>>>(jobs, count, omega);
}
void launch_group_force(const BodyJob* jobs, int count, int grid, int fk, hipStream_t st) {
  if (count > 0 && grid > 0) group_force_kernel<<<grid, 256, 0, st>>>(jobs, count, fk);
}
void launch_group_info_reset(const BodyJob* jobs, int count, hipStream_t st) {
  if (count > 0) group_info_reset_kernel<<<(count + 255) / 256, 256, 0, st>>>(jobs, count);
}
void launch_group_status(const BodyJob* jobs, int count, int nb, BatchStatus* out, hipStream_t st) {
  if (count > 0) group_status_kernel<<<(count + 255) / 256, 256, 0, st>>>(jobs, count, nb, out);
}

void launch_rhs(const double2* V, int n, int nsym, int nrhs, double bx, double2* B, hipStream_t st) {
  rhs_kernel<<<(n * nrhs + 255) / 256, 256, 0, st>>>(V, n, nsym, nrhs, bx, B);
}

void launch_combine(const double2* Z1, const double2* Z2, const double2* F1, const double2* F2, int n, int nsym,
                    int nrhs, double rho_w, double2* sig, double2* pot, double2* pres, hipStream_t st) {
  combine_kernel<<<(n * nrhs + 255) / 256, 256, 0, st>>>(Z1, Z2, F1, F2, n, nsym, nrhs, rho_w, sig, pot, pres);
}

void launch_force(const double2* pres, const double* nds, int nint, int ntot, int nrhs, double2* F, hipStream_t st) {
  force_kernel<<<nint * nrhs, 256, 0, st>>>(pres, nds, ntot, nrhs, F);
}

void launch_f64_to_f32(const double2* a, float2* b, size_t count, const RefineState* rs, hipStream_t st) {
  f64_to_f32_kernel<<<(unsigned)((count + 255) / 256), 256, 0, st>>>(a, b, count, rs);
}
void launch_f32_to_f64(const float2* a, double2* b, size_t count, const RefineState* rs, int accumulate,
                       hipStream_t st) {
  axpy_f32_to_f64_kernel<<<(unsigned)((count + 255) / 256), 256, 0, st>>>(a, b, count, rs, accumulate);
}
void launch_copy(const double2* a, double2* b, size_t count, const RefineState* skip, hipStream_t st) {
  copy_c128_kernel<<<(unsigned)((count + 255) / 256), 256, 0, st>>>(a, b, count, skip);
}
void launch_refine_residual_max(const double2* r, size_t count, double* part, RefineState* rs, double tol, int first,
                                hipStream_t st) {
  constexpr int kParts = 64;
  absmax_kernel<<<kParts, 256, 0, st>>>(r, count, part, first ? nullptr : rs);
  refine_check_kernel<<<1, 64, 0, st>>>(part, kParts, rs, tol, first);
}
void launch_refine_reset(RefineState* rs, hipStream_t st) { refine_reset_kernel<<<1, 1, 0, st>>>(rs); }
__global__ void info_reset_kernel(int* __restrict__ info) {
  info[0] = 0;
  info[1] = 0;
}
void launch_info_reset(int* info, hipStream_t st) { info_reset_kernel<<<1, 1, 0, st>>>(info); }

void launch_kochin(const Geometry& G, const double2* sig, int nsym, int nrhs, double k, double depth, double xeff,
                   double yeff, const double* theta, int ntheta, double2* H, hipStream_t st) {
  kochin_kernel<<<ntheta, 256, 0, st>>>(G, sig, theta, KochinArgs{k, depth, xeff, yeff, nsym, nrhs}, H);
}

// out[q * rows + r] = in[r * cols + q]
__global__ void transpose_c128_kernel(const double2* __restrict__ in, int rows, int cols, double2* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * cols) return;
  const int r = e / cols, q = e - r * cols;
  out[(size_t)q * rows + r] = in[e];
}
void launch_transpose_c128(const double2* in, int rows, int cols, double2* out, hipStream_t st) {
  transpose_c128_kernel<<<(rows * cols + 255) / 256, 256, 0, st>>>(in, rows, cols, out);
}

void launch_extent(const double* x, const double* y, int np, int nsym, double* out, hipStream_t st) {
  const int ntile = (np + 255) / 256;
  extent_kernel<<<(unsigned)((size_t)ntile * (ntile + 1) / 2), 256, 0, st>>>(x, y, np, nsym, out);
}

}  // namespace nemoh

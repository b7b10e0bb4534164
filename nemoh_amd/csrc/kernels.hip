// HIP kernels of the NEMOH solver core for gfx950 (MI355X).
//
//   crek_kernel        wave-Green-function tables (NEMOH.f90:179-274)
//   assemble_kernel    influence matrices D+-, S+- for one frequency
//                      (SOLVE_BEM_INFD_DIRECT.f90:84-104 / FD :105-125)
//   lu_panel_kernel    diagonal-block LU + both panel triangular solves
//   lu_update_kernel   trailing update on MFMA
//   trsv_*_kernel      multi-RHS forward / backward substitution
//   matvec_kernel      Y = alpha*(M X) + beta*Y, streaming M once (phi = S
//                      sigma, and the FP64 refinement residual)
//   rhs/combine/force  right-hand sides, sigma/phi combination of the
//                      symmetry blocks, pressure and force integration
//                      (SOLVE_BEM_*_DIRECT.f90:123-151, SOLVE_BEM.f90:91-96,
//                      NEMOH.f90:137-141)
#include <hip/hip_runtime.h>

#include "green.h"
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
template <bool kSym, bool kF32>
__global__ __launch_bounds__(256) void assemble_kernel(Geometry G, Tables T, const FreqParams* __restrict__ Pp,
                                                       MatOut out) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= G.n) return;
  const FreqParams& P = *Pp;
  double o[16];
  pair16(G, T, P, i, j, o);
  const double xn = G.xn[i], yn = G.yn[i], zn = G.zn[i];
  const size_t e = (size_t)i * out.ld + j;
  const double dpr = o[4] * xn + o[5] * yn + o[6] * zn;
  const double dpi = o[10] * xn + o[11] * yn + o[12] * zn;
  out.dp[e] = make_double2(dpr, dpi);
  out.sp[e] = make_double2(o[0], o[2]);
  if (kF32) out.dp32[e] = make_float2((float)dpr, (float)dpi);
  if (kSym) {
    const double dmr = o[7] * xn + o[8] * yn + o[9] * zn;
    const double dmi = o[13] * xn + o[14] * yn + o[15] * zn;
    out.dm[e] = make_double2(dmr, dmi);
    out.sm[e] = make_double2(o[1], o[3]);
    if (kF32) out.dm32[e] = make_float2((float)dmr, (float)dmi);
  }
}

// Raw 16-value dump of every pair (tests only; small meshes).
__global__ void pairs_kernel(Geometry G, Tables T, const FreqParams* __restrict__ Pp, double* __restrict__ out) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= G.n) return;
  double o[16];
  pair16(G, T, *Pp, i, j, o);
  double* d = out + ((size_t)i * G.n + j) * 16;
#pragma unroll
  for (int k = 0; k < 16; ++k) d[k] = o[k];
}

// ---------------------------------------------------------------------------
// LU panel step.  A 256-thread workgroup holds a 64x64 complex tile in
// registers, thread (ty, tx) = (tid / 16, tid % 16) owning the 4x4 strided
// elements (ty + 16 i, tx + 16 j).  A sequential sweep over k publishes the
// pivot row / column of step k through double-buffered LDS vectors, so each
// step costs one barrier and 16 complex FMAs per thread.
__device__ __forceinline__ bool tid_is_zero(int ty, int tx) { return (ty | tx) == 0; }

template <typename T>
struct PanelLds {
  C2<T> d[kNB][kNB + 1];  // factored diagonal block
  C2<T> col[2][kNB];      // pivot column of the current step (double buffered)
  C2<T> row[2][kNB];      // pivot row of the current step
  C2<T> lp[kNB][17];      // 16 finished columns of a sub-panel (batched update)
  C2<T> up[16][kNB + 1];  // 16 finished rows of a sub-panel
  C2<T> ipiv[kNB];        // reciprocal pivots
};

template <typename T>
__device__ __forceinline__ C2<T> crecip(C2<T> p) {
  const T s = T(1) / (p.re * p.re + p.im * p.im);
  return {p.re * s, -p.im * s};
}

template <typename T>
__device__ __forceinline__ void load_tile(C2<T> a[4][4], const C2<T>* __restrict__ A, int ld, int r0, int c0, int nr,
                                          int nc, int ty, int tx) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = ty + 16 * i, c = tx + 16 * j;
      a[i][j] = (r < nr && c < nc) ? A[(size_t)(r0 + r) * ld + c0 + c] : C2<T>{0, 0};
    }
}

template <typename T>
__device__ __forceinline__ void store_tile(const C2<T> a[4][4], C2<T>* __restrict__ A, int ld, int r0, int c0, int nr,
                                           int nc, int ty, int tx) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = ty + 16 * i, c = tx + 16 * j;
      if (r < nr && c < nc) A[(size_t)(r0 + r) * ld + c0 + c] = a[i][j];
    }
}

// In-register LU without pivoting of the nb x nb leading part of the tile,
// blocked by 16-column sub-panels: a step k touches only its sub-panel
// (column block S = k / 16) and the rows of that sub-panel to the right;
// the rest of the trailing tile gets one rank-16 update per sub-panel.
// Register indices are compile-time constants (S unrolled).
template <typename T>
__device__ __forceinline__ void reg_factor(C2<T> a[4][4], int nb, PanelLds<T>& L, int ty, int tx) {
#pragma unroll
  for (int S = 0; S < 4; ++S) {
    if (16 * S >= nb) break;
    for (int kl = 0; kl < 16; ++kl) {
      const int k = 16 * S + kl;
      if (k >= nb) break;
      const int buf = k & 1;
      if (tx == kl) {  // column k: rows below the pivot
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = ty + 16 * i;
          if (r > k && r < nb) L.col[buf][r] = a[i][S];
        }
      }
      if (ty == kl) {  // row k from the pivot on
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int c = tx + 16 * j;
          if (c >= k && c < nb) L.row[buf][c] = a[S][j];
        }
      }
      __syncthreads();
      const C2<T> inv = crecip(L.row[buf][k]);
      if (tid_is_zero(ty, tx)) L.ipiv[k] = inv;
      C2<T> lr[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = ty + 16 * i;
        lr[i] = (r > k && r < nb) ? cmul(L.col[buf][r], inv) : C2<T>{0, 0};
      }
      if (tx == kl) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (ty + 16 * i > k && ty + 16 * i < nb) a[i][S] = lr[i];
      }
      // inside the sub-panel: column block S, columns after k
      if (tx > kl && 16 * S + tx < nb) {
        const C2<T> u = L.row[buf][16 * S + tx];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i][S] = csub_mul(a[i][S], lr[i], u);
      }
      // rows of the sub-panel below k, column blocks to the right
      if (ty > kl) {
#pragma unroll
        for (int j = S + 1; j < 4; ++j) {
          const int c = tx + 16 * j;
          if (c < nb) a[S][j] = csub_mul(a[S][j], lr[S], L.row[buf][c]);
        }
      }
    }
    if (S < 3 && 16 * (S + 1) < nb) {
      // rank-16 update of rows and columns beyond the sub-panel
      if (tx < 16) {
#pragma unroll
        for (int i = S + 1; i < 4; ++i) L.lp[ty + 16 * i][tx] = a[i][S];
      }
#pragma unroll
      for (int j = S + 1; j < 4; ++j) L.up[ty][tx + 16 * j] = a[S][j];
      __syncthreads();
#pragma unroll
      for (int i = S + 1; i < 4; ++i)
#pragma unroll
        for (int j = S + 1; j < 4; ++j) {
          C2<T> acc = a[i][j];
          for (int q = 0; q < 16; ++q) acc = csub_mul(acc, L.lp[ty + 16 * i][q], L.up[q][tx + 16 * j]);
          a[i][j] = acc;
        }
      __syncthreads();
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void lu_panel_kernel(C2<T>* __restrict__ A, int ld, int n, int k0,
                                                       int* __restrict__ info) {
  __shared__ PanelLds<T> L;
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  const int nb = min(kNB, n - k0);
  C2<T> a[4][4];
  load_tile(a, A, ld, k0, k0, nb, nb, ty, tx);
  reg_factor(a, nb, L, ty, tx);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) L.d[ty + 16 * i][tx + 16 * j] = a[i][j];
  const int rest = n - k0 - nb;
  const int nblk = (rest + kNB - 1) / kNB;
  const int bid = blockIdx.x;
  if (bid == 0) {
    store_tile(a, A, ld, k0, k0, nb, nb, ty, tx);
    if (tx == ty) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = ty + 16 * i;
        const C2<T> p = a[i][i];
        if (r < nb && !(sqrt((double)p.re * p.re + (double)p.im * p.im) > 1e-20)) atomicMax(info, k0 + r + 1);
      }
    }
  }
  __syncthreads();
  if (bid >= 2 * nblk) return;
  C2<T> b[4][4];
  if (bid < nblk) {
    // 64 rows of L21 = A21 U11^-1: column sweep x_k = b_k / u_kk,
    // b_c -= x_k u_kc, blocked by 16 columns like reg_factor
    const int r0 = k0 + nb + bid * kNB;
    const int m = min(kNB, n - r0);
    load_tile(b, A, ld, r0, k0, m, nb, ty, tx);
#pragma unroll
    for (int S = 0; S < 4; ++S) {
      if (16 * S >= nb) break;
      for (int kl = 0; kl < 16; ++kl) {
        const int k = 16 * S + kl;
        if (k >= nb) break;
        const int buf = k & 1;
        if (tx == kl) {
          const C2<T> inv = L.ipiv[k];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            b[i][S] = cmul(b[i][S], inv);
            L.col[buf][ty + 16 * i] = b[i][S];
          }
        }
        __syncthreads();
        if (tx > kl && 16 * S + tx < nb) {
          const C2<T> u = L.d[k][16 * S + tx];
#pragma unroll
          for (int i = 0; i < 4; ++i) b[i][S] = csub_mul(b[i][S], L.col[buf][ty + 16 * i], u);
        }
      }
      if (S < 3 && 16 * (S + 1) < nb) {
#pragma unroll
        for (int i = 0; i < 4; ++i) L.lp[ty + 16 * i][tx] = b[i][S];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = S + 1; j < 4; ++j) {
            const int c = tx + 16 * j;
            if (c < nb) {
              C2<T> acc = b[i][j];
              for (int q = 0; q < 16; ++q) acc = csub_mul(acc, L.lp[ty + 16 * i][q], L.d[16 * S + q][c]);
              b[i][j] = acc;
            }
          }
        __syncthreads();
      }
    }
    store_tile(b, A, ld, r0, k0, m, nb, ty, tx);
  } else {
    // 64 columns of U12 = L11^-1 A12 (unit lower): b_r -= l_rk b_k, blocked
    const int c0 = k0 + nb + (bid - nblk) * kNB;
    const int m = min(kNB, n - c0);
    load_tile(b, A, ld, k0, c0, nb, m, ty, tx);
#pragma unroll
    for (int S = 0; S < 4; ++S) {
      if (16 * S >= nb) break;
      for (int kl = 0; kl < 16; ++kl) {
        const int k = 16 * S + kl;
        if (k >= nb) break;
        const int buf = k & 1;
        if (ty == kl) {
#pragma unroll
          for (int j = 0; j < 4; ++j) L.row[buf][tx + 16 * j] = b[S][j];
        }
        __syncthreads();
        if (ty > kl && 16 * S + ty < nb) {
          const C2<T> l = L.d[16 * S + ty][k];
#pragma unroll
          for (int j = 0; j < 4; ++j) b[S][j] = csub_mul(b[S][j], l, L.row[buf][tx + 16 * j]);
        }
      }
      if (S < 3 && 16 * (S + 1) < nb) {
#pragma unroll
        for (int j = 0; j < 4; ++j) L.up[ty][tx + 16 * j] = b[S][j];
        __syncthreads();
#pragma unroll
        for (int i = S + 1; i < 4; ++i) {
          const int r = ty + 16 * i;
          if (r < nb) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              C2<T> acc = b[i][j];
              for (int q = 0; q < 16; ++q) acc = csub_mul(acc, L.d[r][16 * S + q], L.up[q][tx + 16 * j]);
              b[i][j] = acc;
            }
          }
        }
        __syncthreads();
      }
    }
    store_tile(b, A, ld, k0, c0, nb, m, ty, tx);
  }
}

// Inverses of the diagonal blocks of the factors, one workgroup per block:
// inv[kb][0] = L_kk^-1 (unit lower), inv[kb][1] = U_kk^-1, each 64x64
// row-major.  They turn each block step of the triangular solves into a
// small matrix product.
template <typename T>
__global__ __launch_bounds__(256) void diag_inverse_kernel(const C2<T>* __restrict__ A, int ld, int n,
                                                           C2<T>* __restrict__ inv) {
  __shared__ PanelLds<T> L;
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  const int kb = blockIdx.x, k0 = kb * kNB;
  const int nb = min(kNB, n - k0);
  {
    C2<T> a[4][4];
    load_tile(a, A, ld, k0, k0, nb, nb, ty, tx);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) L.d[ty + 16 * i][tx + 16 * j] = a[i][j];
  }
  __syncthreads();
  C2<T> x[4][4];
  // L X = I, forward sweep over rows
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = ty + 16 * i, c = tx + 16 * j;
      x[i][j] = C2<T>{(r == c && r < nb) ? T(1) : T(0), T(0)};
    }
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    for (int kl = 0; kl < 16; ++kl) {
      const int k = 16 * kk + kl;
      if (k >= nb) break;
      const int buf = k & 1;
      if (ty == kl) {
#pragma unroll
        for (int j = 0; j < 4; ++j) L.row[buf][tx + 16 * j] = x[kk][j];
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = ty + 16 * i;
        if (r > k && r < nb) {
          const C2<T> l = L.d[r][k];
#pragma unroll
          for (int j = 0; j < 4; ++j) x[i][j] = csub_mul(x[i][j], l, L.row[buf][tx + 16 * j]);
        }
      }
    }
  }
  C2<T>* o = inv + (size_t)kb * 2 * kNB * kNB;
  store_tile(x, o, kNB, 0, 0, kNB, kNB, ty, tx);
  __syncthreads();  // the backward sweep reuses the row buffers
  // U X = I, backward sweep
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = ty + 16 * i, c = tx + 16 * j;
      x[i][j] = C2<T>{(r == c && r < nb) ? T(1) : T(0), T(0)};
    }
#pragma unroll
  for (int kk = 3; kk >= 0; --kk) {
    for (int kl = 15; kl >= 0; --kl) {
      const int k = 16 * kk + kl;
      if (k >= nb) continue;
      const int buf = k & 1;
      if (ty == kl) {
        const C2<T> d = crecip(L.d[k][k]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          x[kk][j] = cmul(x[kk][j], d);
          L.row[buf][tx + 16 * j] = x[kk][j];
        }
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = ty + 16 * i;
        if (r < k) {
          const C2<T> u = L.d[r][k];
#pragma unroll
          for (int j = 0; j < 4; ++j) x[i][j] = csub_mul(x[i][j], u, L.row[buf][tx + 16 * j]);
        }
      }
    }
  }
  store_tile(x, o + kNB * kNB, kNB, 0, 0, kNB, kNB, ty, tx);
}

// Trailing update A22 -= L21 U12 (K = 64), 64x64 output tile per workgroup,
// 4 waves each owning a 32x32 quadrant = 2x2 MFMA 16x16 tiles; the complex
// product is four real MFMAs, Cr += (-Ar) Br + Ai Bi, Ci += (-Ar) Bi +
// (-Ai) Br, with the accumulators initialised from C (so C is read once, at
// the start, and written once).  K streams through LDS in chunks of 16 with
// double buffering: the global loads of chunk c+1 are in flight while chunk
// c feeds the MFMAs.  LDS keeps the operands k-major, split into real and
// imaginary planes, A already negated.
template <typename T>
struct UpdLds {
  T ar[2][16][64], ai[2][16][64], br[2][16][64], bi[2][16][64];
};

template <typename T>
__global__ __launch_bounds__(256) void lu_update_kernel(C2<T>* __restrict__ A, int ld, int n, int k0) {
  typedef typename MfmaTraits<T>::acc_t acc_t;
  __shared__ UpdLds<T> S;
  const int s0 = k0 + kNB;  // first row / column of the trailing matrix
  const int row0 = s0 + blockIdx.y * 64, col0 = s0 + blockIdx.x * 64;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  // accumulators <- C
  acc_t cr[2][2], ci[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = row0 + wr + a * 16 + MfmaTraits<T>::row(lk, q);
        const int c = col0 + wc + b * 16 + li;
        C2<T> v = {0, 0};
        if (r < n && c < n) v = A[(size_t)r * ld + c];
        cr[a][b][q] = v.re;
        ci[a][b][q] = v.im;
      }
  // staging map: A chunk = 64 rows x 16 k, B chunk = 16 k x 64 cols; each
  // thread moves 4 elements of each
  C2<T> pa[4], pb[4];
  auto fetch = [&](int chunk) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u;
      const int ar_ = e >> 4, ak = e & 15;  // A: row, k (16 consecutive k per row)
      const int bk = e >> 6, bc = e & 63;   // B: k, col
      pa[u] = (row0 + ar_ < n) ? A[(size_t)(row0 + ar_) * ld + k0 + chunk * 16 + ak] : C2<T>{0, 0};
      pb[u] = (col0 + bc < n) ? A[(size_t)(k0 + chunk * 16 + bk) * ld + col0 + bc] : C2<T>{0, 0};
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u;
      const int ar_ = e >> 4, ak = e & 15;
      const int bk = e >> 6, bc = e & 63;
      S.ar[buf][ak][ar_] = -pa[u].re;
      S.ai[buf][ak][ar_] = -pa[u].im;
      S.br[buf][bk][bc] = pb[u].re;
      S.bi[buf][bk][bc] = pb[u].im;
    }
  };
  fetch(0);
  stash(0);
  __syncthreads();
#pragma unroll
  for (int chunk = 0; chunk < 4; ++chunk) {
    const int buf = chunk & 1;
    if (chunk < 3) fetch(chunk + 1);
#pragma unroll
    for (int k = 0; k < 16; k += 4) {
      T xr[2], xi[2], yr[2], yi[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        xr[a] = S.ar[buf][k + lk][wr + a * 16 + li];
        xi[a] = S.ai[buf][k + lk][wr + a * 16 + li];
        yr[a] = S.br[buf][k + lk][wc + a * 16 + li];
        yi[a] = S.bi[buf][k + lk][wc + a * 16 + li];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          cr[a][b] = MfmaTraits<T>::mma(xr[a], yr[b], cr[a][b]);
          ci[a][b] = MfmaTraits<T>::mma(xr[a], yi[b], ci[a][b]);
        }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          cr[a][b] = MfmaTraits<T>::mma(-xi[a], yi[b], cr[a][b]);
          ci[a][b] = MfmaTraits<T>::mma(xi[a], yr[b], ci[a][b]);
        }
    }
    if (chunk < 3) {
      stash(buf ^ 1);
      __syncthreads();
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = row0 + wr + a * 16 + MfmaTraits<T>::row(lk, q);
        const int c = col0 + wc + b * 16 + li;
        if (r < n && c < n) A[(size_t)r * ld + c] = C2<T>{cr[a][b][q], ci[a][b][q]};
      }
}

// MFMA layout check (tests): D = A(16x4) B(4x16) with the lane mapping the
// update kernel assumes; in-bounds for any layout.
template <typename T>
__global__ void mfma_probe_kernel(const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ d) {
  typedef typename MfmaTraits<T>::acc_t acc_t;
  const int lane = threadIdx.x;
  const int li = lane & 15, lk = lane >> 4;
  acc_t c = acc_t{0, 0, 0, 0};
  c = MfmaTraits<T>::mma(a[li * 4 + lk], b[lk * 16 + li], c);
  for (int q = 0; q < 4; ++q) d[MfmaTraits<T>::row(lk, q) * 16 + li] = c[q];
}

// ---------------------------------------------------------------------------
// Multi-RHS triangular solves, one 64-row block step per launch.  B (n x
// nrhs, row-major) holds the right-hand sides being updated, X receives the
// solved block.  Every workgroup forms x_k = inv_kk b_k (inverse diagonal
// block from diag_inverse_kernel) from B, which this launch never writes;
// workgroup 0 stores x_k, the others update their 64-row block of B.
template <typename T, bool kLower>
__global__ __launch_bounds__(256) void trsv_kernel(const C2<T>* __restrict__ A, int ld, int n, int kb,
                                                   const C2<T>* __restrict__ inv, C2<T>* __restrict__ B,
                                                   C2<T>* __restrict__ X, int nrhs) {
  __shared__ C2<T> t[kNB][kNB + 1];
  __shared__ C2<T> sb[kNB][kMaxRhs];
  __shared__ C2<T> sx[kNB][kMaxRhs];
  const int tid = threadIdx.x;
  const int k0 = kb * kNB;
  const int nb = min(kNB, n - k0);
  const C2<T>* iv = inv + ((size_t)kb * 2 + (kLower ? 0 : 1)) * kNB * kNB;
  for (int e = tid; e < kNB * kNB; e += 256) t[e >> 6][e & 63] = iv[e];
  for (int e = tid; e < kNB * kMaxRhs; e += 256) {
    const int r = e / kMaxRhs, q = e % kMaxRhs;
    sb[r][q] = (r < nb && q < nrhs) ? B[(size_t)(k0 + r) * nrhs + q] : C2<T>{0, 0};
  }
  __syncthreads();
  for (int e = tid; e < kNB * kMaxRhs; e += 256) {
    const int r = e / kMaxRhs, q = e % kMaxRhs;
    C2<T> acc = {0, 0};
    for (int m = 0; m < nb; ++m) {
      const C2<T> p = cmul(t[r][m], sb[m][q]);
      acc.re += p.re;
      acc.im += p.im;
    }
    sx[r][q] = acc;
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    for (int e = tid; e < nb * nrhs; e += 256) {
      const int r = e / nrhs, q = e % nrhs;
      X[(size_t)(k0 + r) * nrhs + q] = sx[r][q];
    }
    return;
  }
  const int r0 = kLower ? k0 + nb + (blockIdx.x - 1) * kNB : (blockIdx.x - 1) * kNB;
  const int m = min(kNB, (kLower ? n : k0) - r0);
  if (m <= 0) return;
  for (int e = tid; e < kNB * kNB; e += 256) {
    const int r = e >> 6, c = e & 63;
    t[r][c] = (r < m && c < nb) ? A[(size_t)(r0 + r) * ld + k0 + c] : C2<T>{0, 0};
  }
  __syncthreads();
  for (int e = tid; e < m * nrhs; e += 256) {
    const int r = e / nrhs, q = e % nrhs;
    C2<T> acc = B[(size_t)(r0 + r) * nrhs + q];
    for (int k = 0; k < nb; ++k) acc = csub_mul(acc, t[r][k], sx[k][q]);
    B[(size_t)(r0 + r) * nrhs + q] = acc;
  }
}

// Single-launch triangular solve.  Workgroup w owns row block b (lower:
// b = w, upper: b = nblk-1-w, so blocks without dependencies are dispatched
// first) and accumulates r_b -= A_bj x_j for every solved block j in
// dependency order, waiting for block j's flag (set with release semantics
// after x_j is stored), then forms x_b = inv_bb r_b and publishes it.
// The grid never exceeds the device's CU count, so all workgroups are
// resident.  Every wait is bounded: after spin_limit polls the workgroup
// records a timeout in *status and goes on, so the kernel always ends (the
// host then reports an error instead of the result).
template <typename T, bool kLower>
__global__ __launch_bounds__(256) void trsv_persistent_kernel(const C2<T>* __restrict__ A, int ld, int n,
                                                              const C2<T>* __restrict__ inv,
                                                              const C2<T>* __restrict__ R, C2<T>* __restrict__ X,
                                                              int nrhs, int* __restrict__ flags, int epoch,
                                                              int* __restrict__ status, int spin_limit) {
  __shared__ C2<T> t[kNB][kNB + 1];
  __shared__ C2<T> sx[kNB][kMaxRhs];
  __shared__ C2<T> acc[kNB][kMaxRhs];
  __shared__ int timed_out;
  const int tid = threadIdx.x;
  const int nblk = (n + kNB - 1) / kNB;
  const int b = kLower ? (int)blockIdx.x : nblk - 1 - (int)blockIdx.x;
  const int k0 = b * kNB;
  const int nb = min(kNB, n - k0);
  for (int e = tid; e < kNB * kMaxRhs; e += 256) {
    const int r = e / kMaxRhs, q = e % kMaxRhs;
    acc[r][q] = (r < nb && q < nrhs) ? R[(size_t)(k0 + r) * nrhs + q] : C2<T>{0, 0};
  }
  if (tid == 0) timed_out = 0;
  const int ndep = kLower ? b : nblk - 1 - b;
  for (int d = 0; d < ndep; ++d) {
    const int j = kLower ? d : nblk - 1 - d;
    const int j0 = j * kNB, nj = min(kNB, n - j0);
    __syncthreads();
    // the coupling tile does not depend on x_j: load it before waiting
    for (int e = tid; e < kNB * kNB; e += 256) {
      const int r = e >> 6, c = e & 63;
      t[r][c] = (r < nb && c < nj) ? A[(size_t)(k0 + r) * ld + j0 + c] : C2<T>{0, 0};
    }
    if (tid == 0) {
      int spins = 0;
      while (__hip_atomic_load(&flags[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        if (++spins > spin_limit) {
          timed_out = 1;
          atomicMax(status, 1);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (int e = tid; e < kNB * kMaxRhs; e += 256) {
      const int r = e / kMaxRhs, q = e % kMaxRhs;
      sx[r][q] = (r < nj && q < nrhs) ? X[(size_t)(j0 + r) * nrhs + q] : C2<T>{0, 0};
    }
    __syncthreads();
    for (int e = tid; e < kNB * kMaxRhs; e += 256) {
      const int r = e / kMaxRhs, q = e % kMaxRhs;
      if (q < nrhs) {
        C2<T> a = acc[r][q];
#pragma unroll 8
        for (int k = 0; k < kNB; ++k) a = csub_mul(a, t[r][k], sx[k][q]);
        acc[r][q] = a;
      }
    }
  }
  __syncthreads();
  // x_b = inv_bb r_b
  const C2<T>* iv = inv + ((size_t)b * 2 + (kLower ? 0 : 1)) * kNB * kNB;
  for (int e = tid; e < kNB * kNB; e += 256) t[e >> 6][e & 63] = iv[e];
  __syncthreads();
  for (int e = tid; e < nb * nrhs; e += 256) {
    const int r = e / nrhs, q = e % nrhs;
    C2<T> x = {0, 0};
#pragma unroll 8
    for (int m = 0; m < kNB; ++m) {
      const C2<T> p = cmul(t[r][m], acc[m][q]);
      x.re += p.re;
      x.im += p.im;
    }
    X[(size_t)(k0 + r) * nrhs + q] = x;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (tid == 0) __hip_atomic_store(&flags[b], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Y(n x nrhs) = beta*Y + alpha*M(n x n) X(n x nrhs), FP64 accumulation.  One
// wavefront per matrix row: lanes stride the columns (coalesced row reads of
// M), each lane accumulates all right-hand sides, then a wave reduction.
template <typename TM>
__global__ __launch_bounds__(256) void matvec_kernel(const C2<TM>* __restrict__ M, int ld, int n,
                                                     const double2* __restrict__ X, int nrhs,
                                                     double2* __restrict__ Y, double alpha, double beta) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + wave;
  if (r >= n) return;
  double accr[kMaxRhs], acci[kMaxRhs];
#pragma unroll
  for (int q = 0; q < kMaxRhs; ++q) accr[q] = acci[q] = 0.0;
  const C2<TM>* row = M + (size_t)r * ld;
  for (int c = lane; c < n; c += 64) {
    const C2<TM> m = row[c];
    const double mr = m.re, mi = m.im;
    const double2* xc = X + (size_t)c * nrhs;
#pragma unroll
    for (int q = 0; q < kMaxRhs; ++q) {
      if (q < nrhs) {
        const double2 v = xc[q];
        accr[q] += mr * v.x - mi * v.y;
        acci[q] += mr * v.y + mi * v.x;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kMaxRhs; ++q) {
    if (q < nrhs) {
      double vr = accr[q], vi = acci[q];
      for (int off = 32; off > 0; off >>= 1) {
        vr += __shfl_xor(vr, off, 64);
        vi += __shfl_xor(vi, off, 64);
      }
      if (lane == 0) {
        double2* y = Y + (size_t)r * nrhs + q;
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
__global__ __launch_bounds__(256) void force_kernel(const double2* __restrict__ pres, const double* __restrict__ nds,
                                                    int ntot, int nrhs, double2* __restrict__ F) {
  const int k = blockIdx.x / nrhs, q = blockIdx.x % nrhs;
  __shared__ double sr[256], si[256];
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

// Precision conversions for the mixed path.
__global__ void f64_to_f32_kernel(const double2* __restrict__ a, float2* __restrict__ b, size_t count, double scale) {
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e < count) b[e] = make_float2((float)(a[e].x * scale), (float)(a[e].y * scale));
}
__global__ void axpy_f32_to_f64_kernel(const float2* __restrict__ a, double2* __restrict__ b, size_t count,
                                       double scale, int accumulate) {
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e < count) {
    const double2 v = make_double2((double)a[e].x * scale, (double)a[e].y * scale);
    b[e] = accumulate ? make_double2(b[e].x + v.x, b[e].y + v.y) : v;
  }
}
// max |a| over count complex entries into out[blockIdx.x] (per-block maxima)
__global__ __launch_bounds__(256) void absmax_kernel(const double2* __restrict__ a, size_t count, double* __restrict__ out) {
  __shared__ double s[256];
  double m = 0.0;
  for (size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < count; e += (size_t)gridDim.x * 256)
    m = fmax(m, fmax(fabs(a[e].x), fabs(a[e].y)));
  s[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// ---------------------------------------------------------------------------
// host-side launchers
void launch_crek(const double* xr, const double* xz, double* tab, hipStream_t st) {
  crek_kernel<<<(kNR * kNZ + 127) / 128, 128, 0, st>>>(xr, xz, tab);
}

void launch_assemble(const Geometry& G, const Tables& T, const FreqParams* P, int nsym, const MatOut& out,
                     hipStream_t st) {
  dim3 grid((G.n + 255) / 256, G.n);
  const bool f32 = out.dp32 != nullptr;
  if (nsym) {
    if (f32) assemble_kernel<true, true><<<grid, 256, 0, st>>>(G, T, P, out);
    else assemble_kernel<true, false><<<grid, 256, 0, st>>>(G, T, P, out);
  } else {
    if (f32) assemble_kernel<false, true><<<grid, 256, 0, st>>>(G, T, P, out);
    else assemble_kernel<false, false><<<grid, 256, 0, st>>>(G, T, P, out);
  }
}

void launch_pairs(const Geometry& G, const Tables& T, const FreqParams* P, double* out, hipStream_t st) {
  dim3 grid((G.n + 63) / 64, G.n);
  pairs_kernel<<<grid, 64, 0, st>>>(G, T, P, out);
}

template <typename T>
void launch_lu(C2<T>* A, int ld, int n, int* info, hipStream_t st) {
  for (int k0 = 0; k0 < n; k0 += kNB) {
    const int nb = std::min(kNB, n - k0);
    const int rest = n - k0 - nb;
    const int nblk = (rest + kNB - 1) / kNB;
    lu_panel_kernel<T><<<std::max(1, 2 * nblk), 256, 0, st>>>(A, ld, n, k0, info);
    if (rest > 0) {
      dim3 grid(nblk, nblk);
      lu_update_kernel<T><<<grid, 256, 0, st>>>(A, ld, n, k0);
    }
  }
}
template void launch_lu<double>(C2<double>*, int, int, int*, hipStream_t);
template void launch_lu<float>(C2<float>*, int, int, int*, hipStream_t);

template <typename T>
void launch_diag_inverse(const C2<T>* A, int ld, int n, C2<T>* inv, hipStream_t st) {
  diag_inverse_kernel<T><<<(n + kNB - 1) / kNB, 256, 0, st>>>(A, ld, n, inv);
}
template void launch_diag_inverse<double>(const C2<double>*, int, int, C2<double>*, hipStream_t);
template void launch_diag_inverse<float>(const C2<float>*, int, int, C2<float>*, hipStream_t);

template <typename T>
void launch_lu_solve(const C2<T>* A, int ld, int n, const C2<T>* inv, C2<T>* B, C2<T>* X, int nrhs, hipStream_t st) {
  const int nblk = (n + kNB - 1) / kNB;
  for (int kb = 0; kb < nblk; ++kb) {
    const int below = nblk - kb - 1;
    trsv_kernel<T, true><<<1 + below, 256, 0, st>>>(A, ld, n, kb, inv, B, X, nrhs);
  }
  // X now holds y = L^-1 b; backward sweep reads it as the right-hand side
  for (int kb = nblk - 1; kb >= 0; --kb) {
    trsv_kernel<T, false><<<1 + kb, 256, 0, st>>>(A, ld, n, kb, inv, X, B, nrhs);
  }
}
template void launch_lu_solve<double>(const C2<double>*, int, int, const C2<double>*, C2<double>*, C2<double>*, int,
                                      hipStream_t);

template <typename T>
void launch_lu_solve_persistent(const C2<T>* A, int ld, int n, const C2<T>* inv, C2<T>* B, C2<T>* X, int nrhs,
                                int* flags, int epoch, int* status, int spin_limit, hipStream_t st) {
  const int nblk = (n + kNB - 1) / kNB;
  // forward: B -> X (= L^-1 b); backward: X -> B (= U^-1 y); epochs 2e, 2e+1
  trsv_persistent_kernel<T, true><<<nblk, 256, 0, st>>>(A, ld, n, inv, B, X, nrhs, flags, 2 * epoch, status,
                                                         spin_limit);
  trsv_persistent_kernel<T, false><<<nblk, 256, 0, st>>>(A, ld, n, inv, X, B, nrhs, flags + nblk, 2 * epoch + 1,
                                                          status, spin_limit);
}
template void launch_lu_solve_persistent<double>(const C2<double>*, int, int, const C2<double>*, C2<double>*,
                                                 C2<double>*, int, int*, int, int*, int, hipStream_t);
template void launch_lu_solve_persistent<float>(const C2<float>*, int, int, const C2<float>*, C2<float>*, C2<float>*,
                                                int, int*, int, int*, int, hipStream_t);
template void launch_lu_solve<float>(const C2<float>*, int, int, const C2<float>*, C2<float>*, C2<float>*, int,
                                     hipStream_t);

template <typename TM>
void launch_matvec(const C2<TM>* M, int ld, int n, const double2* X, int nrhs, double2* Y, double alpha,
                   double beta, hipStream_t st) {
  matvec_kernel<TM><<<(n + 3) / 4, 256, 0, st>>>(M, ld, n, X, nrhs, Y, alpha, beta);
}
template void launch_matvec<double>(const C2<double>*, int, int, const double2*, int, double2*, double, double,
                                    hipStream_t);
template void launch_matvec<float>(const C2<float>*, int, int, const double2*, int, double2*, double, double,
                                   hipStream_t);

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

void launch_f64_to_f32(const double2* a, float2* b, size_t count, double scale, hipStream_t st) {
  f64_to_f32_kernel<<<(unsigned)((count + 255) / 256), 256, 0, st>>>(a, b, count, scale);
}
void launch_f32_to_f64(const float2* a, double2* b, size_t count, double scale, int accumulate, hipStream_t st) {
  axpy_f32_to_f64_kernel<<<(unsigned)((count + 255) / 256), 256, 0, st>>>(a, b, count, scale, accumulate);
}
void launch_absmax(const double2* a, size_t count, double* out, int nblocks, hipStream_t st) {
  absmax_kernel<<<nblocks, 256, 0, st>>>(a, count, out);
}

void launch_mfma_probe(const double* a, const double* b, double* d, hipStream_t st) {
  mfma_probe_kernel<double><<<1, 64, 0, st>>>(a, b, d);
}
void launch_mfma_probe_f32(const float* a, const float* b, float* d, hipStream_t st) {
  mfma_probe_kernel<float><<<1, 64, 0, st>>>(a, b, d);
}

}  // namespace nemoh

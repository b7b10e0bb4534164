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
// LU panel step.  LDS holds the 64x64 diagonal block (factored in place by
// every workgroup) and one 64x64 panel block.
template <typename T>
__device__ __forceinline__ void lds_factor(C2<T> (*a)[kNB + 1], int nb) {
  const int tid = threadIdx.x;
  for (int k = 0; k < nb; ++k) {
    __syncthreads();
    const C2<T> piv = a[k][k];
    for (int r = k + 1 + tid; r < nb; r += blockDim.x) a[r][k] = cdiv2(a[r][k], piv);
    __syncthreads();
    const int w = nb - k - 1;
    for (int e = tid; e < w * w; e += blockDim.x) {
      const int r = k + 1 + e / w, c = k + 1 + e % w;
      a[r][c] = csub_mul(a[r][c], a[r][k], a[k][c]);
    }
  }
  __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(256) void lu_panel_kernel(C2<T>* __restrict__ A, int ld, int n, int k0,
                                                       int* __restrict__ info) {
  __shared__ C2<T> d[kNB][kNB + 1];
  __shared__ C2<T> b[kNB][kNB + 1];
  const int nb = min(kNB, n - k0);
  const int tid = threadIdx.x;
  for (int e = tid; e < nb * nb; e += blockDim.x) {
    const int r = e / nb, c = e % nb;
    d[r][c] = A[(size_t)(k0 + r) * ld + k0 + c];
  }
  lds_factor(d, nb);
  const int rest = n - k0 - nb;
  const int nblk = (rest + kNB - 1) / kNB;
  const int bid = blockIdx.x;
  if (bid == 0) {
    for (int e = tid; e < nb * nb; e += blockDim.x) {
      const int r = e / nb, c = e % nb;
      A[(size_t)(k0 + r) * ld + k0 + c] = d[r][c];
    }
    if (tid == 0) {
      for (int k = 0; k < nb; ++k) {
        const C2<T> p = d[k][k];
        if (!(sqrt((double)p.re * p.re + (double)p.im * p.im) > 1e-20)) atomicMax(info, k0 + k + 1);
      }
    }
  }
  if (bid >= 2 * nblk) return;
  if (bid < nblk) {
    // rows r0..r0+m of L21 = A21 U11^-1 : column sweep x_k = a_k / u_kk,
    // a_c -= x_k u_kc
    const int r0 = k0 + nb + bid * kNB;
    const int m = min(kNB, n - r0);
    for (int e = tid; e < m * nb; e += blockDim.x) {
      const int r = e / nb, c = e % nb;
      b[r][c] = A[(size_t)(r0 + r) * ld + k0 + c];
    }
    for (int k = 0; k < nb; ++k) {
      __syncthreads();
      for (int r = tid; r < m; r += blockDim.x) b[r][k] = cdiv2(b[r][k], d[k][k]);
      __syncthreads();
      const int w = nb - k - 1;
      for (int e = tid; e < m * w; e += blockDim.x) {
        const int r = e / w, c = k + 1 + e % w;
        b[r][c] = csub_mul(b[r][c], b[r][k], d[k][c]);
      }
    }
    __syncthreads();
    for (int e = tid; e < m * nb; e += blockDim.x) {
      const int r = e / nb, c = e % nb;
      A[(size_t)(r0 + r) * ld + k0 + c] = b[r][c];
    }
  } else {
    // columns c0..c0+m of U12 = L11^-1 A12 (unit lower)
    const int c0 = k0 + nb + (bid - nblk) * kNB;
    const int m = min(kNB, n - c0);
    for (int e = tid; e < nb * m; e += blockDim.x) {
      const int r = e / m, c = e % m;
      b[r][c] = A[(size_t)(k0 + r) * ld + c0 + c];
    }
    for (int k = 0; k < nb; ++k) {
      __syncthreads();
      const int w = nb - k - 1;
      for (int e = tid; e < w * m; e += blockDim.x) {
        const int r = k + 1 + e / m, c = e % m;
        b[r][c] = csub_mul(b[r][c], d[r][k], b[k][c]);
      }
    }
    __syncthreads();
    for (int e = tid; e < nb * m; e += blockDim.x) {
      const int r = e / m, c = e % m;
      A[(size_t)(k0 + r) * ld + c0 + c] = b[r][c];
    }
  }
}

// Trailing update A22 -= L21 U12 (K = 64), 64x64 output tile per workgroup,
// 4 waves each owning a 32x32 quadrant = 2x2 MFMA 16x16 tiles, complex
// product as Cr += Ar Br + (-Ai) Bi, Ci += Ar Bi + Ai Br.
// LDS keeps the operands k-major and split into real / imaginary planes.
template <typename T>
__global__ __launch_bounds__(256) void lu_update_kernel(C2<T>* __restrict__ A, int ld, int n, int k0) {
  typedef typename MfmaTraits<T>::acc_t acc_t;
  __shared__ T ar[kNB][kNB], ai[kNB][kNB], br[kNB][kNB], bi[kNB][kNB];
  const int s = k0 + kNB;  // first row / column of the trailing matrix
  const int row0 = s + blockIdx.y * 64, col0 = s + blockIdx.x * 64;
  const int tid = threadIdx.x;
  // stage A rows row0.. (columns k0..k0+63) and B rows k0.. (columns col0..)
  for (int e = tid; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63;
    C2<T> va = {0, 0}, vb = {0, 0};
    if (row0 + r < n) va = A[(size_t)(row0 + r) * ld + k0 + c];
    if (col0 + c < n) vb = A[(size_t)(k0 + r) * ld + col0 + c];
    ar[c][r] = va.re;
    ai[c][r] = va.im;
    br[r][c] = vb.re;
    bi[r][c] = vb.im;
  }
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  acc_t cr[2][2], ci[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      cr[a][b] = acc_t{0, 0, 0, 0};
      ci[a][b] = acc_t{0, 0, 0, 0};
    }
#pragma unroll 4
  for (int k = 0; k < kNB; k += 4) {
    T xr[2], xi[2], yr[2], yi[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      xr[a] = ar[k + lk][wr + a * 16 + li];
      xi[a] = ai[k + lk][wr + a * 16 + li];
      yr[a] = br[k + lk][wc + a * 16 + li];
      yi[a] = bi[k + lk][wc + a * 16 + li];
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        cr[a][b] = MfmaTraits<T>::mma(xr[a], yr[b], cr[a][b]);
        cr[a][b] = MfmaTraits<T>::mma(-xi[a], yi[b], cr[a][b]);
        ci[a][b] = MfmaTraits<T>::mma(xr[a], yi[b], ci[a][b]);
        ci[a][b] = MfmaTraits<T>::mma(xi[a], yr[b], ci[a][b]);
      }
  }
  // D layout of the 16x16 tile: lane l, register q -> row 4*(l/16)+q, col l%16
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = row0 + wr + a * 16 + 4 * lk + q;
        const int c = col0 + wc + b * 16 + li;
        if (r < n && c < n) {
          C2<T>* p = A + (size_t)r * ld + c;
          C2<T> v = *p;
          v.re -= cr[a][b][q];
          v.im -= ci[a][b][q];
          *p = v;
        }
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
  for (int q = 0; q < 4; ++q) d[(4 * lk + q) * 16 + li] = c[q];
}

// ---------------------------------------------------------------------------
// Multi-RHS triangular solves, one 64-row block per launch.  B (n x nrhs,
// row-major) holds the right-hand sides being updated; X receives the solved
// block.  Every workgroup solves the diagonal block redundantly from B (never
// written in the same launch), workgroup 0 stores it, the others update their
// 64-row block of B.
template <typename T, bool kLower>
__global__ __launch_bounds__(256) void trsv_kernel(const C2<T>* __restrict__ A, int ld, int n, int kb,
                                                   C2<T>* __restrict__ B, C2<T>* __restrict__ X, int nrhs) {
  __shared__ C2<T> d[kNB][kNB + 1];
  __shared__ C2<T> x[kNB][kMaxRhs];
  const int tid = threadIdx.x;
  const int k0 = kb * kNB;
  const int nb = min(kNB, n - k0);
  for (int e = tid; e < nb * nb; e += blockDim.x) {
    const int r = e / nb, c = e % nb;
    d[r][c] = A[(size_t)(k0 + r) * ld + k0 + c];
  }
  for (int e = tid; e < nb * nrhs; e += blockDim.x) {
    const int r = e / nrhs, q = e % nrhs;
    x[r][q] = B[(size_t)(k0 + r) * nrhs + q];
  }
  __syncthreads();
  if (kLower) {  // unit lower: forward
    for (int k = 0; k < nb; ++k) {
      for (int e = tid; e < (nb - k - 1) * nrhs; e += blockDim.x) {
        const int r = k + 1 + e / nrhs, q = e % nrhs;
        x[r][q] = csub_mul(x[r][q], d[r][k], x[k][q]);
      }
      __syncthreads();
    }
  } else {  // upper: backward
    for (int k = nb - 1; k >= 0; --k) {
      for (int q = tid; q < nrhs; q += blockDim.x) x[k][q] = cdiv2(x[k][q], d[k][k]);
      __syncthreads();
      for (int e = tid; e < k * nrhs; e += blockDim.x) {
        const int r = e / nrhs, q = e % nrhs;
        x[r][q] = csub_mul(x[r][q], d[r][k], x[k][q]);
      }
      __syncthreads();
    }
  }
  if (blockIdx.x == 0) {
    for (int e = tid; e < nb * nrhs; e += blockDim.x) {
      const int r = e / nrhs, q = e % nrhs;
      X[(size_t)(k0 + r) * nrhs + q] = x[r][q];
    }
    return;
  }
  // rows of the other blocks: below (lower) or above (upper) the diagonal
  int r0;
  if (kLower) r0 = k0 + nb + (blockIdx.x - 1) * kNB;
  else r0 = (blockIdx.x - 1) * kNB;
  const int rend = kLower ? n : k0;
  const int m = min(kNB, rend - r0);
  if (m <= 0) return;
  __syncthreads();
  // reuse d for the off-diagonal tile A[r0.., k0..]
  for (int e = tid; e < m * nb; e += blockDim.x) {
    const int r = e / nb, c = e % nb;
    d[r][c] = A[(size_t)(r0 + r) * ld + k0 + c];
  }
  __syncthreads();
  for (int e = tid; e < m * nrhs; e += blockDim.x) {
    const int r = e / nrhs, q = e % nrhs;
    C2<T> acc = B[(size_t)(r0 + r) * nrhs + q];
    for (int k = 0; k < nb; ++k) acc = csub_mul(acc, d[r][k], x[k][q]);
    B[(size_t)(r0 + r) * nrhs + q] = acc;
  }
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
void launch_lu_solve(const C2<T>* A, int ld, int n, C2<T>* B, C2<T>* X, int nrhs, hipStream_t st) {
  const int nblk = (n + kNB - 1) / kNB;
  for (int kb = 0; kb < nblk; ++kb) {
    const int below = nblk - kb - 1;
    trsv_kernel<T, true><<<1 + below, 256, 0, st>>>(A, ld, n, kb, B, X, nrhs);
  }
  // X now holds y = L^-1 b; backward sweep reads it as the right-hand side
  for (int kb = nblk - 1; kb >= 0; --kb) {
    trsv_kernel<T, false><<<1 + kb, 256, 0, st>>>(A, ld, n, kb, X, B, nrhs);
  }
}
template void launch_lu_solve<double>(const C2<double>*, int, int, C2<double>*, C2<double>*, int, hipStream_t);
template void launch_lu_solve<float>(const C2<float>*, int, int, C2<float>*, C2<float>*, int, hipStream_t);

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

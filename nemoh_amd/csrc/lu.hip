// Dense complex block LU without pivoting and the multi-RHS triangular solves
// of the influence-matrix system, on CDNA4 matrix cores.
//
// The reference inverts the matrix by Gauss-Jordan elimination without row
// interchanges (GAUSSZ, Solver/Core/M_SOLVER.f90:31-73; its row search never
// swaps).  Here the same unpivoted elimination is organised in 64x64 blocks
// (nblk = ceil(n/64), block step k):
//
//   P_k   = (A_kk)^-1            Gauss-Jordan on the updated diagonal block,
//                                one workgroup
//   L_ik  = A_ik P_k    (i > k)  stored over A_ik          (panel GEMM)
//   A_ij -= L_ik A_kj   (i,j > k)                          (update GEMM)
//   U_kj  = P_k A_kj    (j > k)  stored over A_kj after update k has read it
//
// so A = L D U with L, U unit block triangular and D = diag(A_kk), and
//   forward   y_b = b_b - sum_{j<b} L_bj y_j
//   backward  x_b = P_b y_b - sum_{j>b} U_bj x_j
// are both plain block recurrences without a diagonal solve on the critical
// path.  The Gauss-Jordan pivots are the pivots of the pointwise elimination,
// so the reference's "pivot < 1e-20" check keeps its meaning.
//
// Launch sequence of one factorisation: block_inverse(0), then for each k:
// panel(k) [L_ik for i > k and U_{k-1,j} for j >= k], update(k) [A_ij for
// i,j > k; workgroup 0 updates the next diagonal block first and inverts it
// while the other workgroups finish the trailing matrix (look-ahead)].
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include <utility>

#include "kernels.h"
#include "linalg.h"

namespace nemoh {

constexpr int kTile = kNB * kNB;  // elements of one 64x64 block

// ---------------------------------------------------------------------------
// 64x64x64 complex tile product on MFMA.  Four waves each own a 32x32
// quadrant = 2x2 MFMA 16x16 tiles; the complex product is four real MFMAs.
// K streams through LDS in chunks of 16, double buffered (the global loads of
// chunk c+2 are in flight while chunk c feeds the MFMAs), in separate real /
// imaginary planes: A row-major with a row stride of 20 elements, B k-major
// with 72, so that the MFMA operand reads (lane (li, lk): row li, k lk of
// A; k lk, column li of B) touch each LDS bank at most twice and the
// staging writes at most three times.

template <typename T>
struct GemmLds {
  T ar[2][64][20], ai[2][64][20];  // [buf][row][k]
  T br[2][16][72], bi[2][16][72];  // [buf][k][col]
};

template <typename T>
struct Acc64 {
  typename MfmaTraits<T>::acc_t re[2][2], im[2][2];
};

// 64x64 operand window: element (r, c) = p[r * ld + c] for r < nr, c < nc,
// zero outside.
template <typename T>
struct View {
  const C2<T>* p;
  int ld, nr, nc;
};

// Stores and loads of data that other workgroups of the same launch read or
// wrote: system-scope relaxed atomics, one per 64-bit word, which are
// coherent across the XCDs' L2 caches without a cache-wide write-back or
// invalidation.
__device__ __forceinline__ void st_coherent(C2<double>* p, C2<double> v) {
  double* q = reinterpret_cast<double*>(p);
  __hip_atomic_store(q, v.re, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(q + 1, v.im, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st_coherent(C2<float>* p, C2<float> v) {
  unsigned long long u;
  __builtin_memcpy(&u, &v, sizeof u);
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ C2<float> ld_coherent(const C2<float>* p) {
  const unsigned long long u =
      __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  C2<float> v;
  __builtin_memcpy(&v, &u, sizeof v);
  return v;
}
__device__ __forceinline__ C2<double> ld_coherent(const C2<double>* p) {
  const double* q = reinterpret_cast<const double*>(p);
  return {__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM),
          __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)};
}

template <typename T>
__device__ __forceinline__ void acc_zero(Acc64<T>& acc) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc.re[a][b][q] = acc.im[a][b][q] = T(0);
}

// element (row, col) of the tile held by this lane in accumulator [a][b][q]
__device__ __forceinline__ void acc_coord(int a, int b, int row_in16, int& r, int& c) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  r = (wave >> 1) * 32 + a * 16 + row_in16;
  c = (wave & 1) * 32 + b * 16 + (lane & 15);
}

// The block LU's operands all live in device memory.  Read and written
// through a global (address space 1) pointer they are vector-memory
// accesses, whose completion the compiler's wait counts track one by one;
// through a generic pointer taken from a job table they would be flat
// accesses, whose waits also cover LDS traffic and every access in flight,
// so that a stash of one operand chunk would wait for the loads of the
// chunks after it.  The cast is taken of a block's base pointer, before any
// per-lane offset.
template <typename T>
using GlobalC2 = __attribute__((address_space(1))) C2<T>;
template <typename T>
__device__ __forceinline__ const GlobalC2<T>* global_view(const C2<T>* p) {
  return (const GlobalC2<T>*)p;
}
template <typename T>
__device__ __forceinline__ GlobalC2<T>* global_view(C2<T>* p) {
  return (GlobalC2<T>*)p;
}

template <typename T>
__device__ __forceinline__ void acc_load(Acc64<T>& acc, const C2<T>* __restrict__ Cp, int ld, int nr, int nc) {
  const GlobalC2<T>* C = global_view(Cp);
  const int lk = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int r, c;
        acc_coord(a, b, MfmaTraits<T>::row(lk, q), r, c);
        C2<T> v = {0, 0};
        if (r < nr && c < nc) {
          const GlobalC2<T>& g = C[(size_t)r * ld + c];
          v = C2<T>{g.re, g.im};
        }
        acc.re[a][b][q] = v.re;
        acc.im[a][b][q] = v.im;
      }
}

template <typename T>
__device__ __forceinline__ void acc_store(const Acc64<T>& acc, C2<T>* __restrict__ Cp, int ld, int nr, int nc) {
  GlobalC2<T>* C = global_view(Cp);
  const int lk = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int r, c;
        acc_coord(a, b, MfmaTraits<T>::row(lk, q), r, c);
        if (r < nr && c < nc) {
          GlobalC2<T>& g = C[(size_t)r * ld + c];
          g.re = acc.re[a][b][q];
          g.im = acc.im[a][b][q];
        }
      }
}

// acc = (kNegA ? -A : A) B, acc zero on entry.  All global reads of both
// operands complete before the last barrier, so the caller may store over an
// operand.  With
// C given, the tile C (nr x nc at C, leading dimension ldc) is loaded while
// the last chunk runs and added at the end: acc = C + (+-A) B, so the
// tile's first MFMA waits only for the first operand chunk.
template <typename T, bool kNegA, bool kFull, int kChunks = 4, bool kCoherentB = false>
__device__ __forceinline__ void gemm64_impl(Acc64<T>& acc, View<T> A, View<T> B, GemmLds<T>& S,
                                            const C2<T>* __restrict__ C, int ldc, int cnr, int cnc) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  // staging map: A chunk = 64 rows x 16 k, B chunk = 16 k x 64 cols; each
  // thread moves 4 elements of each.  Chunk c is loaded into register set
  // c & 1 two chunks before it is needed (the loads of chunk c + 2 go out
  // while chunk c feeds the MFMAs and are stashed in LDS at the end of chunk
  // c + 1), so each load has two chunks' MFMA time to land.
  C2<T> pa[2][4], pb[2][4];
  // complex64 full blocks: each thread moves two adjacent elements per load
  // (pairs along k for A, along columns for B), which the loads fetch as one
  // 16-byte word: half the load instructions.  The LDS contents, and so the
  // products, are the same as with the element map below.
  constexpr bool kPairs = sizeof(T) == 4 && kFull && !kCoherentB;
  const GlobalC2<T>* ga = global_view(A.p);
  const GlobalC2<T>* gb = global_view(B.p);
  auto gload = [](const GlobalC2<T>& g) { return C2<T>{g.re, g.im}; };
  auto fetch = [&](int chunk, int set) {
    if constexpr (kPairs) {
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const int e = tid + 256 * v;
        const int ar_ = e >> 3, ak = chunk * 16 + 2 * (e & 7);
        const int bk = chunk * 16 + (e >> 5), bc = 2 * (e & 31);
        const size_t ia = (size_t)ar_ * A.ld + ak, ib = (size_t)bk * B.ld + bc;
        pa[set][2 * v] = gload(ga[ia]);
        pa[set][2 * v + 1] = gload(ga[ia + 1]);
        pb[set][2 * v] = gload(gb[ib]);
        pb[set][2 * v + 1] = gload(gb[ib + 1]);
      }
      return;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u;
      const int ar_ = e >> 4, ak = chunk * 16 + (e & 15);
      const int bk = chunk * 16 + (e >> 6), bc = e & 63;
      // kCoherentB: B was written during this launch by another workgroup
      if (kFull) {  // unguarded loads: no select waits on their data
        pa[set][u] = gload(ga[(size_t)ar_ * A.ld + ak]);
        pb[set][u] = kCoherentB ? ld_coherent(B.p + (size_t)bk * B.ld + bc) : gload(gb[(size_t)bk * B.ld + bc]);
      } else {
        pa[set][u] = (ar_ < A.nr && ak < A.nc) ? gload(ga[(size_t)ar_ * A.ld + ak]) : C2<T>{0, 0};
        pb[set][u] = (bk < B.nr && bc < B.nc)
                         ? (kCoherentB ? ld_coherent(B.p + (size_t)bk * B.ld + bc) : gload(gb[(size_t)bk * B.ld + bc]))
                         : C2<T>{0, 0};
      }
    }
  };
  auto stash = [&](int buf, int set) {
    if constexpr (kPairs) {
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const int e = tid + 256 * v;
        const int ar_ = e >> 3, ak = 2 * (e & 7);
        const int bk = e >> 5, bc = 2 * (e & 31);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          S.ar[buf][ar_][ak + h] = pa[set][2 * v + h].re;
          S.ai[buf][ar_][ak + h] = pa[set][2 * v + h].im;
          S.br[buf][bk][bc + h] = pb[set][2 * v + h].re;
          S.bi[buf][bk][bc + h] = pb[set][2 * v + h].im;
        }
      }
      return;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u;
      const int ar_ = e >> 4, ak = e & 15;
      const int bk = e >> 6, bc = e & 63;
      S.ar[buf][ar_][ak] = pa[set][u].re;
      S.ai[buf][ar_][ak] = pa[set][u].im;
      S.br[buf][bk][bc] = pb[set][u].re;
      S.bi[buf][bk][bc] = pb[set][u].im;
    }
  };
  fetch(0, 0);
  fetch(1, 1);
  stash(0, 0);  // waits for chunk 0's loads only (chunk 1's were issued after)
  __syncthreads();
  Acc64<T> cv;
  // Single precision forms the complex product with three real products
  // (Gauss): K1 = (Ar + Ai) Br, K2 = Ar (Bi - Br), K3 = Ai (Br + Bi),
  // Re = K1 - K3, Im = K1 + K2 -- a quarter fewer MFMAs.  Its rounding is a
  // small multiple of the four-product form's, which the FP64 refinement
  // that follows every single-precision solve absorbs.  Double precision
  // keeps the four products: its update is bound by moving the 16-byte
  // elements of C, not by the MFMAs.
  constexpr bool kGauss = sizeof(T) == 4;
  typename MfmaTraits<T>::acc_t k1[2][2];
  if (kGauss) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) k1[a][b][q] = T(0);
  }
#pragma unroll
  for (int chunk = 0; chunk < kChunks; ++chunk) {
    const int buf = chunk & 1;
    if (chunk + 2 < kChunks) fetch(chunk + 2, chunk & 1);  // set chunk & 1 was stashed already
    // in flight under the last chunk's MFMAs (issued with the chunk before
    // instead, the 12k-panel factorisation ran 4 % slower)
    if (chunk == kChunks - 1 && C) acc_load(cv, C, ldc, cnr, cnc);
#pragma unroll
    for (int k = 0; k < 16; k += 4) {
      T xr[2], xi[2], yr[2], yi[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        xr[a] = S.ar[buf][wr + a * 16 + li][k + lk];
        xi[a] = S.ai[buf][wr + a * 16 + li][k + lk];
        yr[a] = S.br[buf][k + lk][wc + a * 16 + li];
        yi[a] = S.bi[buf][k + lk][wc + a * 16 + li];
      }
      if (kGauss) {  // acc.re accumulates K3, acc.im K2
        T sx[2], dy[2], ey[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          sx[a] = xr[a] + xi[a];
          dy[a] = yi[a] - yr[a];
          ey[a] = yr[a] + yi[a];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            k1[a][b] = MfmaTraits<T>::mma(sx[a], yr[b], k1[a][b]);
            acc.im[a][b] = MfmaTraits<T>::mma(xr[a], dy[b], acc.im[a][b]);
            acc.re[a][b] = MfmaTraits<T>::mma(xi[a], ey[b], acc.re[a][b]);
          }
      } else {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            acc.re[a][b] = MfmaTraits<T>::mma(xr[a], yr[b], acc.re[a][b]);
            acc.im[a][b] = MfmaTraits<T>::mma(xr[a], yi[b], acc.im[a][b]);
          }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            acc.re[a][b] = MfmaTraits<T>::mma(-xi[a], yi[b], acc.re[a][b]);
            acc.im[a][b] = MfmaTraits<T>::mma(xi[a], yr[b], acc.im[a][b]);
          }
      }
    }
    if (chunk < kChunks - 1) {
      stash(buf ^ 1, (chunk + 1) & 1);
      __syncthreads();
    }
  }
  // acc = AB (Gauss: K1 - K3, K1 + K2; acc entered as zero), then -AB for
  // kNegA: the sign is applied here, once per element, not to the operands
  // at every k-step.  Round-to-nearest is symmetric, so every product and
  // partial sum of (-A)B is the negation of AB's and the result is bit for
  // bit that of negating the operands.
  if (kGauss) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (kNegA) {
            acc.re[a][b][q] = acc.re[a][b][q] - k1[a][b][q];
            acc.im[a][b][q] = -k1[a][b][q] - acc.im[a][b][q];
          } else {
            acc.re[a][b][q] = k1[a][b][q] - acc.re[a][b][q];
            acc.im[a][b][q] = k1[a][b][q] + acc.im[a][b][q];
          }
        }
  } else if (kNegA) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          acc.re[a][b][q] = -acc.re[a][b][q];
          acc.im[a][b][q] = -acc.im[a][b][q];
        }
  }
  if (C) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          acc.re[a][b][q] = cv.re[a][b][q] + acc.re[a][b][q];
          acc.im[a][b][q] = cv.im[a][b][q] + acc.im[a][b][q];
        }
  }
}

template <typename T, bool kNegA, bool kCoherentB = false>
__device__ __forceinline__ void gemm64(Acc64<T>& acc, View<T> A, View<T> B, GemmLds<T>& S,
                                       const C2<T>* __restrict__ C = nullptr, int ldc = 0, int cnr = 0,
                                       int cnc = 0) {
  if (A.nr == kNB && A.nc == kNB && B.nr == kNB && B.nc == kNB)  // uniform over the workgroup
    gemm64_impl<T, kNegA, true, 4, kCoherentB>(acc, A, B, S, C, ldc, cnr, cnc);
  else
    gemm64_impl<T, kNegA, false, 4, kCoherentB>(acc, A, B, S, C, ldc, cnr, cnc);
}

// The same with K = 128 (A: 64 x 128, B: 128 x 64): the rank-128 trailing
// update of a pair of block steps.
template <typename T, bool kNegA>
__device__ __forceinline__ void gemm128(Acc64<T>& acc, View<T> A, View<T> B, GemmLds<T>& S, const C2<T>* __restrict__ C,
                                        int ldc, int cnr, int cnc) {
  if (A.nr == kNB && A.nc == 2 * kNB && B.nr == 2 * kNB && B.nc == kNB)
    gemm64_impl<T, kNegA, true, 8>(acc, A, B, S, C, ldc, cnr, cnc);
  else
    gemm64_impl<T, kNegA, false, 8>(acc, A, B, S, C, ldc, cnr, cnc);
}

// ---------------------------------------------------------------------------
// In-place inversion without pivoting of a 64x64 block held in LDS, by
// Gauss-Jordan elimination grouped in 16x16 blocks (level S = 0..3, K the
// rows / columns of block S):
//   D = (A_KK)^-1                          one wave, in registers
//   A_KJ <- D A_KJ             (J != K)    MFMA
//   A_IJ <- A_IJ - A_IK A_KJ   (I, J != K) MFMA, with the new A_KJ
//   A_IK <- -A_IK D            (I != K)    MFMA, with the old A_IK
//   A_KK <- D
// This is the scalar elimination of GAUSSZ (M_SOLVER.f90:37-60) taken 16
// pivots at a time: the pivots met inside the 16x16 inversions are the
// pointwise pivots, in order.

template <typename T>
struct InvLds {
  C2<T> m[kNB][kNB + 1];
};

template <typename T>
__device__ __forceinline__ C2<T> crecip(C2<T> p) {
  const T s = T(1) / (p.re * p.re + p.im * p.im);
  return {p.re * s, -p.im * s};
}

template <typename T>
__device__ __forceinline__ C2<T> shfl_c(C2<T> v, int src) {
  return {__shfl(v.re, src, 64), __shfl(v.im, src, 64)};
}

// Reciprocal of a positive real to working precision: the hardware estimate
// (v_rcp) refined by Newton steps y <- y + y (1 - d y) (two in FP64; the
// FP32 estimate is already within 1 ulp).
__device__ __forceinline__ float rcp_pos(float d) { return __builtin_amdgcn_rcpf(d); }
__device__ __forceinline__ double rcp_pos(double d) {
  double y = __builtin_amdgcn_rcp(d);
  double e = fma(-d, y, 1.0);
  y = fma(y, e, y);
  e = fma(-d, y, 1.0);
  return fma(y, e, y);
}

// Lane movement inside a wave.  bcast16<K>: every lane receives the value of
// lane K of its own row of 16 lanes (DPP row_newbcast, no LDS traffic);
// lane_value: the value of one lane, uniform (readlane).
template <int K>
__device__ __forceinline__ float bcast16(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + K, 0xF, 0xF, false));
}
template <int K>
__device__ __forceinline__ double bcast16(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x150 + K, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x150 + K, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double lane_value(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// One wave inverts a 16x16 block held 4 elements per lane: lane l owns
// column l & 15, rows (l >> 4) + 4 e.  Pivot step K: the pivot column value
// of each of the lane's rows sits in lane K of the lane's own row of 16 lanes
// (DPP broadcast), the pivot-row value of the lane's column in lane
// 16 (K & 3) + column (one shuffle), the pivot itself in lane
// 16 (K & 3) + K (readlane).  Every element then takes a <- base - l u with
// u = row K / p (u_K = 1/p), l = column K, base = a except 0 in column K,
// and for the pivot row l = -1 and base = 0 (so a <- u).
template <typename T, int K>
__device__ __forceinline__ void gj16_step(C2<T> a[4], int nvalid, int& first_bad) {
  const int lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
  const C2<T> sel = a[K >> 2];
  const C2<T> rowv = shfl_c(sel, 16 * (K & 3) + c);
  C2<T> colv[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) colv[e] = C2<T>{bcast16<K>(a[e].re), bcast16<K>(a[e].im)};
  const C2<T> p = {lane_value(sel.re, 16 * (K & 3) + K), lane_value(sel.im, 16 * (K & 3) + K)};
  const double pm = (double)p.re * (double)p.re + (double)p.im * (double)p.im;
  if (K < nvalid && !(pm > 1e-40) && first_bad == 16) first_bad = K;
  const T s = rcp_pos(p.re * p.re + p.im * p.im);
  const C2<T> pinv = {p.re * s, -p.im * s};
  const bool pivot_col = (c == K);
  const C2<T> u = pivot_col ? pinv : cmul(rowv, pinv);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const bool pivot_row = (e == (K >> 2)) && (g == (K & 3));
    const C2<T> l = pivot_row ? C2<T>{T(-1), T(0)} : colv[e];
    const C2<T> base = (pivot_col || pivot_row) ? C2<T>{0, 0} : a[e];
    a[e] = csub_mul(base, l, u);
  }
}

template <typename T, int... Ks>
__device__ __forceinline__ void gj16_steps(C2<T> a[4], int nvalid, int& first_bad, std::integer_sequence<int, Ks...>) {
  (gj16_step<T, Ks>(a, nvalid, first_bad), ...);
}

// Pivots K < nvalid with |p| <= 1e-20 are reported as row row0 + K + 1 in
// *info (K >= nvalid is identity padding).
template <typename T>
__device__ __forceinline__ void gj16_wave(C2<T> a[4], int nvalid, int row0, int* __restrict__ info) {
  int first_bad = 16;
  gj16_steps<T>(a, nvalid, first_bad, std::make_integer_sequence<int, 16>{});
  if ((threadIdx.x & 63) == 0 && first_bad < 16) atomicMax(info, row0 + first_bad + 1);
}

// acc += (kNeg ? -1 : 1) A B for the 16x16 blocks A = m[ra.., ka..] and
// B = m[ka.., cb..] of the LDS matrix (one wave).
template <typename T, bool kNeg>
__device__ __forceinline__ void blk16_mma(typename MfmaTraits<T>::acc_t& cr, typename MfmaTraits<T>::acc_t& ci,
                                          const InvLds<T>& L, int ra, int ka, int cb) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int k = ka + 4 * s + lk;
    const C2<T> x = L.m[ra + li][k];
    const C2<T> y = L.m[k][cb + li];
    const T xr = kNeg ? -x.re : x.re, xi = kNeg ? -x.im : x.im;
    cr = MfmaTraits<T>::mma(xr, y.re, cr);
    ci = MfmaTraits<T>::mma(xr, y.im, ci);
    cr = MfmaTraits<T>::mma(-xi, y.im, cr);
    ci = MfmaTraits<T>::mma(xi, y.re, ci);
  }
}

template <typename T>
__device__ __forceinline__ void blk16_load(typename MfmaTraits<T>::acc_t& cr, typename MfmaTraits<T>::acc_t& ci,
                                           const InvLds<T>& L, int r0, int c0) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const C2<T> v = L.m[r0 + MfmaTraits<T>::row(lk, q)][c0 + li];
    cr[q] = v.re;
    ci[q] = v.im;
  }
}

template <typename T>
__device__ __forceinline__ void blk16_store(const typename MfmaTraits<T>::acc_t& cr,
                                            const typename MfmaTraits<T>::acc_t& ci, InvLds<T>& L, int r0, int c0) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) L.m[r0 + MfmaTraits<T>::row(lk, q)][c0 + li] = C2<T>{cr[q], ci[q]};
}

// t-th block index other than S (t = 0, 1, 2)
__device__ __forceinline__ int other_block(int t, int S) { return t < S ? t : t + 1; }

// Wave 0: inverts the 16x16 diagonal block at (K, K) of L.m in place.
template <typename T>
__device__ __forceinline__ void gj16_block(InvLds<T>& L, int K, int nvalid, int row0, int* __restrict__ info) {
  const int lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
  C2<T> a[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) a[e] = L.m[K + g + 4 * e][K + c];
  gj16_wave(a, nvalid, row0, info);
#pragma unroll
  for (int e = 0; e < 4; ++e) L.m[K + g + 4 * e][K + c] = a[e];
}

// Inverts L.m in place (256 threads; the caller synchronises before and
// after).  nb: order of the real block (rows / columns >= nb hold the
// identity); row0: global row of m[0] for the pivot report.  Look-ahead
// inside: in the trailing phase of level S, wave 0 first updates the next
// diagonal block (S+1, S+1) -- which no other job of that phase reads -- and
// inverts it while waves 1-3 do the other 11 block products.
template <typename T>
__device__ __forceinline__ void lds_invert(InvLds<T>& L, int nb, int row0, int* __restrict__ info) {
  typedef typename MfmaTraits<T>::acc_t acc_t;
  const int wave = threadIdx.x >> 6;
  if (wave == 0) gj16_block(L, 0, nb, row0, info);
  __syncthreads();
#pragma unroll
  for (int S = 0; S < 4; ++S) {
    const int K = 16 * S;
    // row blocks A_KJ <- D A_KJ, waves 0..2
    acc_t cr = {}, ci = {};
    if (wave < 3) blk16_mma<T, false>(cr, ci, L, K, K, 16 * other_block(wave, S));
    __syncthreads();
    if (wave < 3) blk16_store<T>(cr, ci, L, K, 16 * other_block(wave, S));
    __syncthreads();
    // trailing blocks A_IJ -= A_IK A_KJ and column blocks A_IK <- -A_IK D
    acc_t jr[4], ji[4];
    int jib[4], jjb[4];  // target block of each job (-1: none)
    if (S < 3 && wave == 0) {
      const int K1 = K + 16;
      acc_t nr_, ni_;
      blk16_load<T>(nr_, ni_, L, K1, K1);
      blk16_mma<T, true>(nr_, ni_, L, K1, K, K1);
      blk16_store<T>(nr_, ni_, L, K1, K1);
      // the same wave reads the block back in the 16x16-inversion layout
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      gj16_block(L, K1, nb - K1, row0 + K1, info);
#pragma unroll
      for (int t = 0; t < 4; ++t) jib[t] = -1;
    } else {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        // S < 3: waves 1..3 share the 11 jobs other than (S+1, S+1);
        // S == 3: the 12 jobs over the four waves
        const int jj = (S < 3) ? (wave - 1) + 3 * t : 3 * wave + t;
        const int njobs = (S < 3) ? 11 : 12;
        const int ntrail = (S < 3) ? 8 : 9;
        jib[t] = -1;
        if (t >= ((S < 3) ? 4 : 3) || jj >= njobs) continue;
        if (jj < ntrail) {
          const int t9 = (S < 3 && jj >= 4 * S) ? jj + 1 : jj;
          jib[t] = other_block(t9 / 3, S);
          jjb[t] = other_block(t9 % 3, S);
          blk16_load<T>(jr[t], ji[t], L, 16 * jib[t], 16 * jjb[t]);
          blk16_mma<T, true>(jr[t], ji[t], L, 16 * jib[t], K, 16 * jjb[t]);
        } else {
          jib[t] = other_block(jj - ntrail, S);
          jjb[t] = S;
          jr[t] = acc_t{};
          ji[t] = acc_t{};
          blk16_mma<T, true>(jr[t], ji[t], L, 16 * jib[t], K, K);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 4; ++t)
      if (jib[t] >= 0) blk16_store<T>(jr[t], ji[t], L, 16 * jib[t], 16 * jjb[t]);
    __syncthreads();
  }
}

// P (64x64 row-major) <- L.m, zero outside the nb x nb block (kCoherent:
// with system-scope stores, for readers in the same launch).
template <typename T, bool kCoherent = false>
__device__ __forceinline__ void store_inverse(const InvLds<T>& L, C2<T>* __restrict__ P, int nb) {
  for (int e = threadIdx.x; e < kTile; e += 256) {
    const int r = e >> 6, c = e & 63;
    const C2<T> v = (r < nb && c < nb) ? L.m[r][c] : C2<T>{0, 0};
    if (kCoherent)
      st_coherent(P + e, v);
    else
      P[e] = v;
  }
}

// P_kb for one diagonal block read from A (the first block of a factorisation).
template <typename T>
__device__ __forceinline__ void block_inverse_body(InvLds<T>& L, const C2<T>* __restrict__ A, int ld, int n, int kb,
                                                   C2<T>* __restrict__ P, int* __restrict__ info) {
  const int k0 = kb * kNB, nb = min(kNB, n - k0);
  for (int e = threadIdx.x; e < kTile; e += 256) {
    const int r = e >> 6, c = e & 63;
    L.m[r][c] = (r < nb && c < nb) ? A[(size_t)(k0 + r) * ld + k0 + c] : C2<T>{T(r == c ? 1 : 0), 0};
  }
  __syncthreads();
  lds_invert(L, nb, k0, info);
  store_inverse(L, P + (size_t)kb * kTile, nb);
}
template <typename T>
__global__ __launch_bounds__(256) void block_inverse_kernel(const C2<T>* __restrict__ A, int ld, int n, int kb,
                                                            C2<T>* __restrict__ P, int* __restrict__ info) {
  __shared__ InvLds<T> L;
  block_inverse_body(L, A, ld, n, kb, P, info);
}

// Panel step kb: workgroups [0, ncol) form L_ib = A_ib P_kb for the row
// blocks i = kb+1+w below the diagonal, the rest form U_{kb-1,j} = P_{kb-1}
// A_{kb-1,j} for j = kb + (w - ncol), in place.
template <typename T>
__global__ __launch_bounds__(256) void lu_panel_kernel(C2<T>* __restrict__ A, int ld, int n, int kb,
                                                       const C2<T>* __restrict__ P, int ncol) {
  __shared__ GemmLds<T> S;
  Acc64<T> acc;
  acc_zero(acc);
  const int w = blockIdx.x;
  if (w < ncol) {
    const int r0 = (kb + 1 + w) * kNB, c0 = kb * kNB;
    const int nr = min(kNB, n - r0), nk = min(kNB, n - c0);
    C2<T>* blk = A + (size_t)r0 * ld + c0;
    gemm64<T, false>(acc, View<T>{blk, ld, nr, nk}, View<T>{P + (size_t)kb * kTile, kNB, kNB, kNB}, S);
    acc_store(acc, blk, ld, nr, nk);
  } else {
    const int r0 = (kb - 1) * kNB, c0 = (kb + w - ncol) * kNB;
    const int nk = min(kNB, n - r0), nc = min(kNB, n - c0);
    C2<T>* blk = A + (size_t)r0 * ld + c0;
    gemm64<T, false>(acc, View<T>{P + (size_t)(kb - 1) * kTile, kNB, kNB, kNB}, View<T>{blk, ld, nk, nc}, S);
    acc_store(acc, blk, ld, nk, nc);
  }
}

template <typename T>
union UpdLds {
  GemmLds<T> g;
  InvLds<T> inv;
};

// Workgroup -> trailing tile (ti, tj) of an m x m tile grid.  Workgroups are
// dealt to the eight XCDs in turn (workgroup w to XCD w mod 8), each XCD
// behind its own L2.  Each XCD gets one contiguous run of a tile order that
// walks bands of kBand tile rows column by column: going down a column reuses
// one U column panel kBand times, and the band's kBand L row panels (16 of
// them: 0.5 MB in single precision, 1 MB in double) stay in the L2 from one
// column to the next.  Bands of 16, 32 and 64 rows run the single-precision
// update within 0.5 % of each other at 4000 panels; 64 is 2 % slower at
// 12,048.  An XCD's run (m^2 / 8 tiles) then reads about one
// band's worth of L panels and a fraction m / (8 kBand) of the U panels (when
// below one) instead of all of both, as the row-major order does on every
// XCD.  Workgroup 0 gets tile (0, 0).  The
// map is a bijection for every m.
__device__ __forceinline__ void update_tile(int bid, int m, int band, int& ti, int& tj) {
  constexpr int kXcds = 8;
  const int total = m * m, per = total / kXcds, rem = total % kXcds;
  const int x = bid % kXcds, local = bid / kXcds;
  const int idx = (x < rem) ? x * (per + 1) + local : rem * (per + 1) + (x - rem) * per + local;
  const int g = idx / (band * m), w = idx - g * band * m;
  const int rows = min(band, m - g * band);
  tj = w / rows;
  ti = g * band + (w - tj * rows);
}

// Look-ahead of the updates: workgroup 0 inverts the diagonal block it has
// just updated (acc, nr x nr at row r0) into P.
template <typename T, bool kCoherent = false>
__device__ __forceinline__ void invert_tile(UpdLds<T>& S, const Acc64<T>& acc, int nr, int r0, C2<T>* __restrict__ P,
                                            int* __restrict__ info) {
  __syncthreads();  // every wave is done with the GEMM staging area
  {
    const int lk = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          int r, c;
          acc_coord(a, b, MfmaTraits<T>::row(lk, q), r, c);
          S.inv.m[r][c] = (r < nr && c < nr) ? C2<T>{acc.re[a][b][q], acc.im[a][b][q]} : C2<T>{T(r == c ? 1 : 0), 0};
        }
  }
  __syncthreads();
  lds_invert(S.inv, nr, r0, info);
  store_inverse<T, kCoherent>(S.inv, P, nr);
}

// Trailing update of step kb: A_ij -= L_i,kb A_kb,j for i, j > kb, one 64x64
// tile per workgroup.  Workgroup 0 owns tile (kb+1, kb+1), the next diagonal
// block, and inverts it after its update (P_{kb+1}).
template <typename T>
__global__ __launch_bounds__(256) void lu_update_kernel(C2<T>* __restrict__ A, int ld, int n, int kb,
                                                        C2<T>* __restrict__ P, int* __restrict__ info, int band) {
  __shared__ UpdLds<T> S;
  const int s0 = (kb + 1) * kNB, k0 = kb * kNB;
  const int m = (n - s0 + kNB - 1) / kNB;
  int ti, tj;
  update_tile(blockIdx.x, m, band, ti, tj);
  const int r0 = s0 + ti * kNB, c0 = s0 + tj * kNB;
  const int nr = min(kNB, n - r0), nc = min(kNB, n - c0);
  C2<T>* blk = A + (size_t)r0 * ld + c0;
  Acc64<T> acc;
  acc_zero(acc);
  gemm64<T, true>(acc, View<T>{A + (size_t)r0 * ld + k0, ld, nr, kNB}, View<T>{A + (size_t)k0 * ld + c0, ld, kNB, nc},
                  S.g, blk, ld, nr, nc);
  acc_store(acc, blk, ld, nr, nc);
  if (blockIdx.x != 0) return;
  invert_tile(S, acc, nr, r0, P + (size_t)(kb + 1) * kTile, info);
}

// Paired block steps (k, k+1), for the rank-128 form of the factorisation:
//
//   panel     L_ik = A_ik P_k (i > k), and the U rows of the previous pair
//   cross     A_{k+1,j} -= L_{k+1,k} A_kj (j > k), A_{i,k+1} -= L_ik A_{k,k+1}
//             (i > k+1): block row and column k+1 take step k; workgroup 0
//             updates (k+1, k+1) and inverts it (P_{k+1})
//   panel     L_{i,k+1} = A_{i,k+1} P_{k+1} (i > k+1)
//   update128 A_ij -= [L_ik L_{i,k+1}] [A_kj; A_{k+1,j}] (i, j > k+1), K = 128:
//             both steps' updates of the trailing matrix in one pass over it;
//             workgroup 0 inverts (k+2, k+2) (P_{k+2})
//
// Rows k and k+1 keep A_kj and the cross-updated A_{k+1,j} until update128
// has read them; the next pair's first panel turns them into U.  The
// elimination is the same as two rank-64 steps (the two products of each
// trailing element are summed in one accumulation instead of two), with half
// the passes over the trailing matrix and twice the work per tile load.

// Panel launch of the paired form (PanelJob, kernels.h): workgroups [0, nL)
// form L_{lc+1+w, lc} = A P_lc; the next nU0 form U_{ur0, ur0+1+w} = P_ur0
// A, the next nU1 U_{ur1, ur1+1+w} = P_ur1 A (in place).
template <typename T>
__device__ __forceinline__ void lu_panel2_body(GemmLds<T>& S, C2<T>* __restrict__ A, int ld, int n, const PanelJob& J,
                                               const C2<T>* __restrict__ P, int w) {
  Acc64<T> acc;
  acc_zero(acc);
  if (w < J.nL) {
    const int r0 = (J.lc + 1 + w) * kNB, c0 = J.lc * kNB;
    const int nr = min(kNB, n - r0), nk = min(kNB, n - c0);
    C2<T>* blk = A + (size_t)r0 * ld + c0;
    gemm64<T, false>(acc, View<T>{blk, ld, nr, nk}, View<T>{P + (size_t)J.lc * kTile, kNB, kNB, kNB}, S);
    acc_store(acc, blk, ld, nr, nk);
    return;
  }
  w -= J.nL;
  int ur = J.ur0;
  if (w >= J.nU0) {
    w -= J.nU0;
    ur = J.ur1;
  }
  const int r0 = ur * kNB, c0 = (ur + 1 + w) * kNB;
  const int nk = min(kNB, n - r0), nc = min(kNB, n - c0);
  C2<T>* blk = A + (size_t)r0 * ld + c0;
  gemm64<T, false>(acc, View<T>{P + (size_t)ur * kTile, kNB, kNB, kNB}, View<T>{blk, ld, nk, nc}, S);
  acc_store(acc, blk, ld, nk, nc);
}
template <typename T>
__global__ __launch_bounds__(256) void lu_panel2_kernel(C2<T>* __restrict__ A, int ld, int n, PanelJob J,
                                                        const C2<T>* __restrict__ P) {
  __shared__ GemmLds<T> S;
  lu_panel2_body(S, A, ld, n, J, P, (int)blockIdx.x);
}

// Cross update of the pair (k, k+1): workgroup w < nr takes tile (k+1,
// k+1+w) of block row k+1 (nr = blocks from k+1 on), the others tile (k+2+w',
// k+1) of block column k+1; workgroup 0 (the diagonal tile) then inverts it.
template <typename T>
__device__ __forceinline__ void lu_cross_body(UpdLds<T>& S, C2<T>* __restrict__ A, int ld, int n, int k,
                                              C2<T>* __restrict__ P, int* __restrict__ info, int w) {
  const int nblk = (n + kNB - 1) / kNB, s = k + 1, nrow = nblk - s;
  const int bi = w < nrow ? s : s + 1 + (w - nrow), bj = w < nrow ? s + w : s;
  const int r0 = bi * kNB, c0 = bj * kNB, k0 = k * kNB;
  const int nr = min(kNB, n - r0), nc = min(kNB, n - c0);
  C2<T>* blk = A + (size_t)r0 * ld + c0;
  Acc64<T> acc;
  acc_zero(acc);
  gemm64<T, true>(acc, View<T>{A + (size_t)r0 * ld + k0, ld, nr, kNB}, View<T>{A + (size_t)k0 * ld + c0, ld, kNB, nc},
                  S.g, blk, ld, nr, nc);
  acc_store(acc, blk, ld, nr, nc);
  if (w != 0) return;
  invert_tile(S, acc, nr, r0, P + (size_t)s * kTile, info);
}
template <typename T>
__global__ __launch_bounds__(256) void lu_cross_kernel(C2<T>* __restrict__ A, int ld, int n, int k,
                                                       C2<T>* __restrict__ P, int* __restrict__ info) {
  __shared__ UpdLds<T> S;
  lu_cross_body(S, A, ld, n, k, P, info, (int)blockIdx.x);
}

// The same cross update with the panel of block column k+1 folded in (one
// launch instead of two).  Workgroups take tickets from ctl[0] (0 when the
// launch starts): ticket 0 takes the diagonal tile (k+1, k+1) and inverts
// it, publishing P_{k+1} through ctl[1] (set to 1); tickets [1, nrow) the
// rest of block row k+1; the others the tiles (i, k+1) of block column k+1,
// which, once their cross update is stored, wait for P_{k+1} and turn the
// tile into L_{i,k+1} = A_{i,k+1} P_{k+1} in place.  A workgroup only waits
// for ticket 0, which was taken by a running workgroup, so the wait ends
// however the launch's workgroups are placed; it is bounded all the same
// (spin_limit polls, then info[1] |= 2 and the tile is left unfinished; the
// caller reports the factorisation as failed).
template <typename T>
__global__ __launch_bounds__(256) void lu_cross_panel_kernel(C2<T>* __restrict__ A, int ld, int n, int k,
                                                             C2<T>* __restrict__ P, int* __restrict__ info,
                                                             int* __restrict__ ctl, int spin_limit) {
  __shared__ UpdLds<T> S;
  __shared__ int ticket_s, ready_s;
  const int tid = threadIdx.x;
  const int nblk = (n + kNB - 1) / kNB, s = k + 1, nrow = nblk - s, m = nblk - s - 1;
  if (tid == 0) ticket_s = (int)atomicAdd(reinterpret_cast<unsigned*>(ctl), 1u);
  __syncthreads();
  const int w = ticket_s;
  if (w < 0 || w >= nrow + m) {  // a counter out of step with the launch: never wait on it
    if (tid == 0) atomicOr(info + 1, 2);
    return;
  }
  const int bi = w < nrow ? s : s + 1 + (w - nrow), bj = w < nrow ? s + w : s;
  const int r0 = bi * kNB, c0 = bj * kNB, k0 = k * kNB;
  const int nr = min(kNB, n - r0), nc = min(kNB, n - c0);
  C2<T>* blk = A + (size_t)r0 * ld + c0;
  Acc64<T> acc;
  acc_zero(acc);
  gemm64<T, true>(acc, View<T>{A + (size_t)r0 * ld + k0, ld, nr, kNB}, View<T>{A + (size_t)k0 * ld + c0, ld, kNB, nc},
                  S.g, blk, ld, nr, nc);
  acc_store(acc, blk, ld, nr, nc);
  if (w == 0) {
    // P_{k+1} with system-scope stores, every one performed before the flag
    invert_tile<T, true>(S, acc, nr, r0, P + (size_t)s * kTile, info);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (tid == 0) __hip_atomic_store(ctl + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    return;
  }
  if (w < nrow) return;
  if (tid == 0) {
    int spins = 0, ok = 1;
    while (__hip_atomic_load(ctl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) == 0) {
      if (++spins > spin_limit) {
        ok = 0;
        atomicOr(info + 1, 2);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    ready_s = ok;
  }
  __syncthreads();
  if (!ready_s) return;
  // P_{k+1} read with system-scope loads (no cache-wide invalidation); the
  // tile itself was stored by this workgroup
  acc_zero(acc);
  gemm64<T, false, true>(acc, View<T>{blk, ld, nr, nc}, View<T>{P + (size_t)s * kTile, kNB, kNB, kNB}, S.g);
  acc_store(acc, blk, ld, nr, nc);
}

// Control words of the fused cross launches (2 per block pair: a ticket
// counter and the P_{k+1} flag) set to 0 at the start of a factorisation.
__global__ void lu_ctl_reset_kernel(int* __restrict__ ctl, int count) {
  for (int e = threadIdx.x; e < count; e += blockDim.x) ctl[e] = 0;
}

// Rank-128 trailing update of the pair (k, k+1): A_ij -= L_{i,k:k+1}
// A_{k:k+1,j} for i, j > k+1 (m trailing blocks from s = k+2), one 64x64 tile
// per workgroup.  region 0: every trailing tile, in the XCD-banded order of
// update_tile; region 1 (look-ahead): only the next pair's block rows s, s+1
// (workgroups [0, 2m)) and block columns s, s+1 below them ([2m, 4m-4));
// region 2: the rest (i, j >= s+2), banded.  Workgroup 0 of regions 0 and 1
// owns (s, s) and inverts it (P_s).
template <typename T>
__device__ __forceinline__ void lu_update128_body(UpdLds<T>& S, C2<T>* __restrict__ A, int ld, int n, int k,
                                                  C2<T>* __restrict__ P, int* __restrict__ info, int region, int band,
                                                  int w) {
  const int s0 = (k + 2) * kNB, k0 = k * kNB;
  const int m = (n - s0 + kNB - 1) / kNB;
  int ti, tj;
  if (region == 1) {
    if (w < 2 * m) {
      ti = w / m;
      tj = w - ti * m;
    } else {
      const int t = w - 2 * m;
      ti = 2 + (t >> 1);
      tj = t & 1;
    }
  } else if (region == 2) {
    update_tile(w, m - 2, band, ti, tj);
    ti += 2;
    tj += 2;
  } else {
    update_tile(w, m, band, ti, tj);
  }
  const int r0 = s0 + ti * kNB, c0 = s0 + tj * kNB;
  const int nr = min(kNB, n - r0), nc = min(kNB, n - c0);
  C2<T>* blk = A + (size_t)r0 * ld + c0;
  Acc64<T> acc;
  acc_zero(acc);
  gemm128<T, true>(acc, View<T>{A + (size_t)r0 * ld + k0, ld, nr, 2 * kNB},
                   View<T>{A + (size_t)k0 * ld + c0, ld, 2 * kNB, nc}, S.g, blk, ld, nr, nc);
  acc_store(acc, blk, ld, nr, nc);
  if (w != 0 || region == 2) return;
  invert_tile(S, acc, nr, r0, P + (size_t)(k + 2) * kTile, info);
}
template <typename T>
__global__ __launch_bounds__(256) void lu_update128_kernel(C2<T>* __restrict__ A, int ld, int n, int k,
                                                           C2<T>* __restrict__ P, int* __restrict__ info,
                                                           int region, int band) {
  __shared__ UpdLds<T> S;
  lu_update128_body(S, A, ld, n, k, P, info, region, band, (int)blockIdx.x);
}

// ---------------------------------------------------------------------------
// Grouped launches: one launch does the same step for the matrices of many
// independent bodies (LuJob j: workgroups [wg0, wg1) of the launch, wg0 a
// multiple of 8 so that a job's workgroups meet the XCDs in the order the
// single-matrix launch deals them, and its tile map stays XCD-banded).  Each
// workgroup finds its job by binary search over wg0 and runs the body of the
// single-matrix kernel with its index within the job, so every matrix gets
// the arithmetic of its own launch sequence, bit for bit.
// the job of this workgroup (-1: padding up to the next multiple of 8) and
// the workgroup's index within it
template <typename J>
__device__ __forceinline__ int group_job(const J* __restrict__ jobs, int count, int& w) {
  const int g = blockIdx.x;
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].wg0 <= g) lo = mid;
    else hi = mid - 1;
  }
  w = g - jobs[lo].wg0;
  return (g < jobs[lo].wg1) ? lo : -1;
}
template <typename T>
__global__ __launch_bounds__(256) void group_inverse_kernel(const LuJob<T>* __restrict__ jobs, int count) {
  __shared__ InvLds<T> L;
  int w;
  const int j = group_job(jobs, count, w);
  if (j < 0) return;
  const LuJob<T> J = jobs[j];
  block_inverse_body(L, J.A, J.ld, J.n, J.k, J.P, J.info);
}
template <typename T>
__global__ __launch_bounds__(256) void group_panel2_kernel(const LuJob<T>* __restrict__ jobs, int count) {
  __shared__ GemmLds<T> S;
  int w;
  const int j = group_job(jobs, count, w);
  if (j < 0) return;
  const LuJob<T> J = jobs[j];
  lu_panel2_body(S, J.A, J.ld, J.n, J.J, J.P, w);
}
template <typename T>
__global__ __launch_bounds__(256) void group_cross_kernel(const LuJob<T>* __restrict__ jobs, int count) {
  __shared__ UpdLds<T> S;
  int w;
  const int j = group_job(jobs, count, w);
  if (j < 0) return;
  const LuJob<T> J = jobs[j];
  lu_cross_body(S, J.A, J.ld, J.n, J.k, J.P, J.info, w);
}
template <typename T>
__global__ __launch_bounds__(256) void group_update128_kernel(const LuJob<T>* __restrict__ jobs, int count) {
  __shared__ UpdLds<T> S;
  int w;
  const int j = group_job(jobs, count, w);
  if (j < 0) return;
  const LuJob<T> J = jobs[j];
  lu_update128_body(S, J.A, J.ld, J.n, J.k, J.P, J.info, 0, J.band, w);
}

// ---------------------------------------------------------------------------
// Triangular solves.  One block step is r_b -= M_bj x_j with M_bj a 64x64
// block of L or U and x_j (64 x nrhs, nrhs <= 16): each wave computes 16 rows
// with 16 MFMA k-steps (N = 16 right-hand-side columns).  The wave's operand
// registers for M_bj are loaded before x_j is needed (HopTile), with the sign
// folded in: pr = s Re M, pi = s Im M, npi = -s Im M.

template <typename T>
struct HopTile {
  T pr[16], pi[16], npi[16];
};

// Column of the block (row of x) that MFMA k-step s takes from lane group
// lk; the product sums over the 64 columns in this order, and M and x use
// the same map.  complex64: steps 2p and 2p + 1 take two adjacent columns,
// so a lane loads each pair as one 16-byte word, and the four lane groups of
// a row read 64 contiguous bytes per load (8 loads per lane for a block).
// complex128: one element is already 16 bytes; four consecutive columns per
// step, 64 contiguous bytes per row per load (16 loads).
template <typename T>
__device__ __forceinline__ int hop_k(int s, int lk) {
  return sizeof(T) == 4 ? 8 * (s >> 1) + 2 * lk + (s & 1) : 4 * s + lk;
}

// The loads below take a clamped address (element 0, always inside the
// block, for an element outside nr x nc) and select zero afterwards, so that
// they carry no branch: a load under a divergent branch is waited for inside
// it, and the 16 loads of a lane would go out one at a time.
template <typename T>
__device__ __forceinline__ void hop_load_tile(HopTile<T>& h, const C2<T>* __restrict__ M, int ld, int nr, int nc,
                                              T sign) {
  const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) & 3;  // wave within its 64-row block
  const int r = 16 * wave + (lane & 15), lk = lane >> 4;
  C2<T> v[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int c = hop_k<T>(s, lk);
    v[s] = M[(r < nr && c < nc) ? (size_t)r * ld + c : 0];
  }
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const bool in = r < nr && hop_k<T>(s, lk) < nc;
    h.pr[s] = in ? sign * v[s].re : T(0);
    h.pi[s] = in ? sign * v[s].im : T(0);
    h.npi[s] = in ? -sign * v[s].im : T(0);
  }
}

// An element of a chained solve's x that is not solved yet holds a marker: a
// signalling-NaN bit pattern in every 64-bit word (both halves of a complex64
// word).  Arithmetic never yields a signalling NaN, so a solved element never
// equals it, and since each word is stored and loaded whole (64-bit atomics)
// a reader sees either the marker or the final value.  Readers therefore poll
// the data itself: no flag, no wait for the stores to complete before one.
constexpr unsigned long long kUnsolved32 = 0x7FA5A5A57FA5A5A5ull;
constexpr unsigned long long kUnsolved64 = 0x7FF5A5A5A5A5A5A5ull;

__device__ __forceinline__ void st_unsolved(C2<float>* p) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), kUnsolved32, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st_unsolved(C2<double>* p) {
  unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
  __hip_atomic_store(q, kUnsolved64, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(q + 1, kUnsolved64, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// false while v is (partly) the marker
__device__ __forceinline__ bool is_solved(C2<float> v) {
  unsigned long long u;
  __builtin_memcpy(&u, &v, sizeof u);
  return u != kUnsolved32;
}
__device__ __forceinline__ bool is_solved(C2<double> v) {
  unsigned long long a, b;
  __builtin_memcpy(&a, &v.re, sizeof a);
  __builtin_memcpy(&b, &v.im, sizeof b);
  return a != kUnsolved64 && b != kUnsolved64;
}

// x operand of one block step: lane (li, lk) holds x[hop_k(s, lk)][li], s < 16.
template <typename T>
__device__ __forceinline__ void hop_load_x(T xr[16], T xi[16], const C2<T>* __restrict__ X, int nj, int nrhs) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  C2<T> v[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int r = hop_k<T>(s, lk);
    v[s] = X[(r < nj && li < nrhs) ? (size_t)r * nrhs + li : 0];
  }
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const bool in = hop_k<T>(s, lk) < nj && li < nrhs;
    xr[s] = in ? v[s].re : T(0);
    xi[s] = in ? v[s].im : T(0);
  }
}

// The chained solve's operands are loaded ahead of their use in two steps:
// hop_issue_* sends the loads (clamped addresses, no branch), and the
// operand is formed after the loads land.  Tile: lane (li, lk) of wave w
// holds M[16 w + li][hop_k(s, lk)]; elements outside nr x nc hold M[0][0],
// which only ever meets a zero x element (columns) or fills accumulator rows
// that are not stored (rows).  A full complex64 block (the same for the
// whole workgroup) takes unclamped addresses, which lets the loads of
// adjacent columns merge (complex128 measured slower that way).
template <typename T>
__device__ __forceinline__ void hop_issue_tile(C2<T> t[16], const C2<T>* __restrict__ M, int ld, int nr, int nc) {
  const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) & 3;
  const int r = 16 * wave + (lane & 15), lk = lane >> 4;
  if (sizeof(T) == 4 && nr == kNB && nc == kNB) {
    const C2<T>* row = M + (size_t)r * ld;
#pragma unroll
    for (int s = 0; s < 16; ++s) t[s] = row[hop_k<T>(s, lk)];
  } else {
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int c = hop_k<T>(s, lk);
      t[s] = M[(r < nr && c < nc) ? (size_t)r * ld + c : 0];
    }
  }
}
// x_j as other workgroups store it (lane (li, lk): x[hop_k(s, lk)][li])
template <typename T>
__device__ __forceinline__ void hop_issue_x(C2<T> v[16], const C2<T>* __restrict__ X, int nj, int nrhs) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int r = hop_k<T>(s, lk);
    v[s] = ld_coherent(X + ((r < nj && li < nrhs) ? (size_t)r * nrhs + li : 0));
  }
}
// x operand from hop_issue_x's values (zero outside nj x nrhs), and nxi =
// -xi; true when every element this lane holds is solved
template <typename T>
__device__ __forceinline__ bool hop_x_operand(T xr[16], T xi[16], T nxi[16], const C2<T> v[16], int nj, int nrhs) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  bool ok = true;
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const bool in = hop_k<T>(s, lk) < nj && li < nrhs;
    ok = ok && (!in || is_solved(v[s]));
    xr[s] = in ? v[s].re : T(0);
    xi[s] = in ? v[s].im : T(0);
    nxi[s] = -xi[s];
  }
  return ok;
}
// (cr, ci) += M x with M as hop_issue_tile loaded it
template <typename T>
__device__ __forceinline__ void hop_mma_raw(typename MfmaTraits<T>::acc_t& cr, typename MfmaTraits<T>::acc_t& ci,
                                            const C2<T> m[16], const T xr[16], const T xi[16], const T nxi[16]) {
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    cr = MfmaTraits<T>::mma(m[s].re, xr[s], cr);
    ci = MfmaTraits<T>::mma(m[s].re, xi[s], ci);
    cr = MfmaTraits<T>::mma(m[s].im, nxi[s], cr);
    ci = MfmaTraits<T>::mma(m[s].im, xr[s], ci);
  }
}

template <typename T>
__device__ __forceinline__ void hop_mma(typename MfmaTraits<T>::acc_t& cr, typename MfmaTraits<T>::acc_t& ci,
                                        const HopTile<T>& h, const T xr[16], const T xi[16]) {
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    cr = MfmaTraits<T>::mma(h.pr[s], xr[s], cr);
    ci = MfmaTraits<T>::mma(h.pr[s], xi[s], ci);
    cr = MfmaTraits<T>::mma(h.npi[s], xi[s], cr);
    ci = MfmaTraits<T>::mma(h.pi[s], xr[s], ci);
  }
}

// rows of the wave's 16-row slab held in accumulator slot q
template <typename T>
__device__ __forceinline__ int hop_row(int q) {
  const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) & 3;
  return 16 * wave + MfmaTraits<T>::row(lane >> 4, q);
}

template <typename T>
__device__ __forceinline__ void hop_acc_load(typename MfmaTraits<T>::acc_t& cr, typename MfmaTraits<T>::acc_t& ci,
                                             const C2<T>* __restrict__ R, int nb, int nrhs) {
  const int li = threadIdx.x & 15;
  C2<T> v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = hop_row<T>(q);
    v[q] = R[(r < nb && li < nrhs) ? (size_t)r * nrhs + li : 0];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const bool in = hop_row<T>(q) < nb && li < nrhs;
    cr[q] = in ? v[q].re : T(0);
    ci[q] = in ? v[q].im : T(0);
  }
}

template <typename T, bool kCoherent>
__device__ __forceinline__ void hop_acc_store(const typename MfmaTraits<T>::acc_t& cr,
                                              const typename MfmaTraits<T>::acc_t& ci, C2<T>* __restrict__ X, int nb,
                                              int nrhs) {
  const int li = threadIdx.x & 15;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = hop_row<T>(q);
    if (r < nb && li < nrhs) {
      const C2<T> v = {cr[q], ci[q]};
      if (kCoherent)
        st_coherent(X + (size_t)r * nrhs + li, v);
      else
        X[(size_t)r * nrhs + li] = v;
    }
  }
}

// the elements of a 64-row block that hop_acc_store writes, set to the marker
template <typename T>
__device__ __forceinline__ void hop_mark_unsolved(C2<T>* __restrict__ X, int nb, int nrhs) {
  const int li = threadIdx.x & 15;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = hop_row<T>(q);
    if (r < nb && li < nrhs) st_unsolved(X + (size_t)r * nrhs + li);
  }
}

// The workgroup at chain position w (0 <= w < row blocks; the forward chain
// solves row block w, the backward chain row block nblk - 1 - w) of one
// triangle: it waits only for the positions before w.
//
// x of a dependency is loaded where it is polled, not with the coupling
// block one dependency early: workgroups keep pace with the chain's front,
// so an early load mostly returns the unsolved marker, and issuing it (16
// uncached loads per lane in each wave) delays the product.  Clock stamps of
// a 4000-panel chain: 3.7 us per dependency with the early load, 3.0
// without; a batch of 40 chains is slightly faster without it too.
template <typename T, bool kLower>
__device__ __forceinline__ void trsv_chain_run(int w, const C2<T>* __restrict__ A, int ld, int n,
                                               const C2<T>* __restrict__ P, C2<T>* __restrict__ R,
                                               C2<T>* __restrict__ X, int nrhs, int* __restrict__ status,
                                               int spin_limit) {
  typedef typename MfmaTraits<T>::acc_t acc_t;
  const int lane = threadIdx.x & 63;
  const int nblk = (n + kNB - 1) / kNB;
  const int b = kLower ? w : nblk - 1 - w;
  const int k0 = b * kNB, nb = min(kNB, n - k0);
  // (cr, ci) = -(the block's running value): -r_b (forward) or -P_b y_b
  // (backward), then + M_bj x_j for every dependency j
  acc_t cr, ci;
  T xr[16], xi[16], nxi[16];
  if (kLower) {
    hop_acc_load<T>(cr, ci, R + (size_t)k0 * nrhs, nb, nrhs);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      cr[q] = -cr[q];
      ci[q] = -ci[q];
    }
  } else {  // P_b (-y_b)
#pragma unroll
    for (int q = 0; q < 4; ++q) cr[q] = ci[q] = T(0);
    C2<T> m[16];
    hop_issue_tile(m, P + (size_t)b * kTile, kNB, kNB, kNB);
    hop_load_x<T>(xr, xi, R + (size_t)k0 * nrhs, nb, nrhs);
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      nxi[s] = xi[s];
      xr[s] = -xr[s];
      xi[s] = -xi[s];
    }
    hop_mma_raw<T>(cr, ci, m, xr, xi, nxi);
  }
  // wave-uniform: a wave that gave up waiting (at once when spin_limit is 0,
  // which forces the stepwise fallback) does no further waits
  bool gave_up = spin_limit == 0;
  if (gave_up && lane == 0) atomicMax(status, 1);
  const int ndep = kLower ? b : nblk - 1 - b;
  auto dep_rows = [&](int d, int& j0, int& nj) {
    const int j = kLower ? d : nblk - 1 - d;
    j0 = j * kNB;
    nj = min(kNB, n - j0);
  };
  // the coupling block of dependency d + 1 is loaded while the product of
  // dependency d runs
  C2<T> t[16], xv[16];
  if (ndep > 0) {
    int j0, nj;
    dep_rows(0, j0, nj);
    hop_issue_tile(t, A + (size_t)k0 * ld + j0, ld, nb, nj);
  }
  for (int d = 0; d < ndep; ++d) {
    int j0, nj;
    dep_rows(d, j0, nj);
    hop_issue_x(xv, X + (size_t)j0 * nrhs, nj, nrhs);
    bool ok = hop_x_operand(xr, xi, nxi, xv, nj, nrhs);
    int spins = 0;
    while (!__all(ok) && !gave_up) {
      if (++spins > spin_limit) {
        gave_up = true;
        if (lane == 0) atomicMax(status, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      hop_issue_x(xv, X + (size_t)j0 * nrhs, nj, nrhs);
      ok = hop_x_operand(xr, xi, nxi, xv, nj, nrhs);
    }
    C2<T> m[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) m[s] = t[s];
    if (d + 1 < ndep) {
      int j1, n1;
      dep_rows(d + 1, j1, n1);
      hop_issue_tile(t, A + (size_t)k0 * ld + j1, ld, nb, n1);
    }
    hop_mma_raw<T>(cr, ci, m, xr, xi, nxi);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    cr[q] = -cr[q];
    ci[q] = -ci[q];
  }
  hop_acc_store<T, true>(cr, ci, X + (size_t)k0 * nrhs, nb, nrhs);
  // R_b was read into the accumulators above (every MFMA and the store of x_b
  // waited for those loads): mark it for the backward launch, which solves
  // into it
  if (kLower) hop_mark_unsolved<T>(R + (size_t)k0 * nrhs, nb, nrhs);
}

// Single-launch triangular solve.  Each workgroup takes the next ticket w of
// the launch from a counter (atomic add) and owns row block b (forward:
// b = w, backward: b = nblk-1-w, so blocks without dependencies come first);
// for every solved block j, in dependency order, it loads the coupling block
// (which does not depend on x_j), polls x_j until no element holds the
// unsolved marker and subtracts M_bj x_j, then stores x_b.  Each wave polls
// and proceeds on its own.  A workgroup only ever waits for blocks whose
// tickets were taken before its own, i.e. by workgroups that are already
// running, so the chain makes progress however few of its workgroups are
// resident (other streams' kernels and other chains may hold the rest of the
// GPU): the lowest unsolved block never waits.  Every wait is bounded: after
// spin_limit polls (at once when it is 0) the wave records a timeout in
// *status and stops waiting, so the kernel always ends (its result is then
// discarded by the caller).
// R: right-hand sides (forward) or the forward solution y (backward); X gets
// the solution and holds the marker in every element when the launch starts.
// The forward launch also marks its R (the backward launch's X) block by
// block once it has read it.  ctl->ticket[tri] is 0 when the launch starts
// (set, with X's marker, by chain_begin_kernel for the forward triangle), so
// the launches take no per-call argument and replay from a HIP graph.
//
// The critical path is one hand-over per block (x_b's stores reaching memory
// and the next workgroup's poll seeing them) plus the product with the block
// just solved.  Superblocks of two blocks per workgroup (128 rows, 8 waves,
// the inner coupling through LDS) halve the hand-overs but make the product
// on the critical path four times larger, 128 x 128 against 64 x 64, i.e. 3.4
// against 0.85 us of one CU's FP32 MFMA peak: measured 480 us per triangle
// at 4000 panels against 307 us for the 64-row form (both with flags).
template <typename T, bool kLower>
__device__ __forceinline__ void trsv_chain_body(int& my_ticket, const C2<T>* __restrict__ A, int ld, int n,
                                                const C2<T>* __restrict__ P, C2<T>* __restrict__ R,
                                                C2<T>* __restrict__ X, int nrhs, ChainCtl* __restrict__ ctl, int tri,
                                                int* __restrict__ status, int spin_limit,
                                                const RefineState* __restrict__ skip) {
  const int tid = threadIdx.x;
  const int nblk = (n + kNB - 1) / kNB;
  // a converged refinement: nothing is done (every workgroup reads the same
  // state, and no ticket is taken)
  if (skip && skip->done) return;
  if (tid == 0) my_ticket = (int)atomicAdd(&ctl->ticket[tri], 1u);
  __syncthreads();
  const int w = my_ticket;
  if (w < 0 || w >= nblk) {  // a counter out of step with the launches: never wait on it
    if (tid == 0) atomicMax(status, 1);
    return;
  }
  trsv_chain_run<T, kLower>(w, A, ld, n, P, R, X, nrhs, status, spin_limit);
}

template <typename T, bool kLower>
__global__ __launch_bounds__(256) void trsv_chain_kernel(const C2<T>* __restrict__ A, int ld, int n,
                                                         const C2<T>* __restrict__ P, C2<T>* __restrict__ R,
                                                         C2<T>* __restrict__ X, int nrhs, ChainCtl* __restrict__ ctl,
                                                         int tri, int* __restrict__ status, int spin_limit,
                                                         const RefineState* __restrict__ skip) {
  __shared__ int my_ticket;
  trsv_chain_body<T, kLower>(my_ticket, A, ld, n, P, R, X, nrhs, ctl, tri, status, spin_limit, skip);
}

// Grouped chained solves: one launch per triangle for many systems, one
// workgroup per row block of each.  The workgroups take tickets from one
// counter for the whole launch (ticket[tri], 0 when the launch starts), dealt
// in chain order across the systems: level b (chain position b) of every
// system that has one, then level b + 1, with the systems sorted by size so
// that those having level b are jobs [0, cum[b+1] - cum[b]) and ticket t at
// level b (cum[b] <= t < cum[b+1]) is system t - cum[b].  A workgroup waits
// only for earlier positions of its own system, whose tickets are smaller,
// i.e. taken by running workgroups: the argument of the single-system launch
// holds.  Dealt system by system instead, the launch ran its chains a few
// residents' worth at a time (40 chains of 63 blocks at 4000 panels: 2 ms a
// triangle, ten times one chain's time).  Each position runs the single
// system's arithmetic, so results equal the single launches' bit for bit.
template <typename T, bool kLower>
__global__ __launch_bounds__(256) void group_trsv_kernel(const TrsvJob<T>* __restrict__ jobs, int count,
                                                         const int* __restrict__ cum, int nlevel,
                                                         unsigned* __restrict__ ticket, int spin_limit) {
  __shared__ int my_ticket;
  if (threadIdx.x == 0) my_ticket = (int)atomicAdd(ticket + (kLower ? 0 : 1), 1u);
  __syncthreads();
  const int t = my_ticket;
  if (t < 0 || t >= cum[nlevel]) {  // a counter out of step with the launches: never wait on it
    if (threadIdx.x == 0)
      for (int j = 0; j < count; ++j) atomicMax(jobs[j].status, 1);
    return;
  }
  int lo = 0, hi = nlevel - 1;  // the level: the last b with cum[b] <= t
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cum[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  const TrsvJob<T> J = jobs[t - cum[lo]];
  if (J.skip && J.skip->done) return;  // a converged refinement: nothing to do
  if (kLower)
    trsv_chain_run<T, true>(lo, J.A, J.ld, J.n, J.P, J.B, J.W, J.nrhs, J.status, spin_limit);
  else
    trsv_chain_run<T, false>(lo, J.A, J.ld, J.n, J.P, J.W, J.B, J.nrhs, J.status, spin_limit);
}

// Stepwise fallback (one launch per block step s, for matrices with more row
// blocks than can be resident): every remaining row block i does
// R_i -= M_is R_s in place (R_s is final).
template <typename T, bool kLower>
__global__ __launch_bounds__(256) void trsv_step_kernel(const C2<T>* __restrict__ A, int ld, int n,
                                                        C2<T>* __restrict__ R, int nrhs, int s,
                                                        const RefineState* __restrict__ skip) {
  typedef typename MfmaTraits<T>::acc_t acc_t;
  if (skip && skip->done) return;
  const int i = kLower ? s + 1 + (int)blockIdx.x : (int)blockIdx.x;
  const int i0 = i * kNB, ni = min(kNB, n - i0);
  const int s0 = s * kNB, ns = min(kNB, n - s0);
  acc_t cr, ci;
  HopTile<T> h;
  T xr[16], xi[16];
  hop_acc_load<T>(cr, ci, R + (size_t)i0 * nrhs, ni, nrhs);
  hop_load_tile(h, A + (size_t)i0 * ld + s0, ld, ni, ns, T(-1));
  hop_load_x<T>(xr, xi, R + (size_t)s0 * nrhs, ns, nrhs);
  hop_mma<T>(cr, ci, h, xr, xi);
  hop_acc_store<T, false>(cr, ci, R + (size_t)i0 * nrhs, ni, nrhs);
}

// X_b = P_b Y_b for every block (start of the stepwise backward solve).
template <typename T>
__global__ __launch_bounds__(256) void block_apply_kernel(const C2<T>* __restrict__ P, int n,
                                                          const C2<T>* __restrict__ Y, C2<T>* __restrict__ X,
                                                          int nrhs, const RefineState* __restrict__ skip) {
  typedef typename MfmaTraits<T>::acc_t acc_t;
  if (skip && skip->done) return;
  const int b = blockIdx.x, k0 = b * kNB, nb = min(kNB, n - k0);
  acc_t cr, ci;
#pragma unroll
  for (int q = 0; q < 4; ++q) cr[q] = ci[q] = T(0);
  HopTile<T> h;
  T xr[16], xi[16];
  hop_load_tile(h, P + (size_t)b * kTile, kNB, kNB, kNB, T(1));
  hop_load_x<T>(xr, xi, Y + (size_t)k0 * nrhs, nb, nrhs);
  hop_mma<T>(cr, ci, h, xr, xi);
  hop_acc_store<T, false>(cr, ci, X + (size_t)k0 * nrhs, nb, nrhs);
}

// Start of a chained solve: both triangles' tickets back to 0, and every
// element of the forward launch's X (count elements) set to the unsolved
// marker.
template <typename T>
__global__ void chain_begin_kernel(ChainCtl* __restrict__ ctl, C2<T>* __restrict__ X, size_t count,
                                   const RefineState* __restrict__ skip) {
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e == 0) {
    ctl->ticket[0] = 0u;
    ctl->ticket[1] = 0u;
  }
  if (skip && skip->done) return;
  if (e < count) st_unsolved(X + e);
}

// The grouped form: every system's forward X marked, and the launch-wide
// ticket counters of both grouped triangles back to 0.
template <typename T>
__global__ void group_chain_begin_kernel(const TrsvJob<T>* __restrict__ jobs, int count,
                                         unsigned* __restrict__ ticket) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ticket[0] = 0u;
    ticket[1] = 0u;
  }
  int w;
  const int j = group_job(jobs, count, w);
  if (j < 0) return;
  const TrsvJob<T> J = jobs[j];
  const size_t e = (size_t)w * blockDim.x + threadIdx.x;
  if (J.skip && J.skip->done) return;
  if (e < (size_t)J.n * J.nrhs) st_unsolved(J.W + e);
}

// A well-conditioned order-n test matrix for timing factorisations: 4 + 0i
// on the diagonal, a hashed value in [-1/n, 1/n] (each part) elsewhere, so
// that every pivot stays near 4.
template <typename T>
__global__ void fill_test_matrix_kernel(C2<T>* __restrict__ A, int ld, int n) {
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e >= (size_t)n * n) return;
  const int i = (int)(e / n), j = (int)(e % n);
  unsigned h = (unsigned)e * 2654435761u;
  h ^= h >> 15;
  h *= 2246822519u;
  h ^= h >> 13;
  const T s = T(1) / T(n);
  const T re = (i == j) ? T(4) : s * (T(h & 0xFFFF) / T(32767.5) - T(1));
  const T im = (i == j) ? T(0) : s * (T(h >> 16) / T(32767.5) - T(1));
  A[(size_t)i * ld + j] = C2<T>{re, im};
}

template <typename T>
__global__ void copy_c2_kernel(const C2<T>* __restrict__ a, C2<T>* __restrict__ b, size_t count,
                               const RefineState* __restrict__ skip) {
  if (skip && skip->done) return;
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e < count) b[e] = a[e];
}

// MFMA layout check (tests): D = A(16x4) B(4x16) with the lane mapping the
// kernels above assume; in-bounds for any layout.
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
// launchers


template <typename T>
int launch_lu(C2<T>* A, int ld, int n, C2<T>* P, int* info, hipStream_t st, int rank, hipStream_t st2,
              hipEvent_t* dep, hipEvent_t* upd_ev, double* upd_flops, int* luctl, int spin_limit, int band) {
  const int nblk = (n + kNB - 1) / kNB;
  int timed = 0;
  // trailing-update launch: events around it when timing (stamped by the
  // runtime at its start and end), and its algorithmic flops
  auto update = [&](auto kernel, hipStream_t s, int grid, int kk, double flops, auto... extra) {
    if (upd_ev) {
      hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(256), 0, s, upd_ev[2 * timed], upd_ev[2 * timed + 1], 0, A, ld, n,
                            kk, P, info, extra...);
      upd_flops[timed++] = flops;
    } else {
      kernel<<<grid, 256, 0, s>>>(A, ld, n, kk, P, info, extra...);
    }
  };
  block_inverse_kernel<T><<<1, 256, 0, st>>>(A, ld, n, 0, P, info);
  if (rank == 64) {
    for (int kb = 0; kb < nblk; ++kb) {
      const int ncol = nblk - kb - 1;
      const int nrow = kb >= 1 ? nblk - kb : 0;
      if (ncol + nrow > 0) lu_panel_kernel<T><<<ncol + nrow, 256, 0, st>>>(A, ld, n, kb, P, ncol);
      if (ncol > 0) {
        const double m = n - (kb + 1.0) * kNB;
        update(lu_update_kernel<T>, st, ncol * ncol, kb, 8.0 * kNB * m * m, band);
      }
    }
    return timed;
  }
  // Paired steps, rank-128 trailing updates.  With st2 (look-ahead), the
  // update of each pair is split: the next pair's block rows and columns
  // (and its pivot inversion) on st, the bulk on st2, which then runs beside
  // the next pair's panels and cross update; the bulk of pair k must be done
  // before the look-ahead update of pair k+2 (it updates the same tiles
  // first) and before rows k, k+1 are turned into U (it reads them), which
  // the panel launch of pair k+4 does.  dep: one event per pair for st -> st2
  // (its panels are done) and one for st2 -> st (its bulk is done).
  const bool ahead = st2 != nullptr;
  // the panel of each pair's second block column inside its cross launch
  const bool fused = luctl != nullptr && !ahead;
  if (fused) lu_ctl_reset_kernel<<<1, 256, 0, st>>>(luctl, 2 * lu_pairs(n));
  PanelJob pend[2] = {{0, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0}};  // U rows waiting for a bulk update to end
  int pair = 0;
  int bulk_done = -1;  // event of the latest bulk update (-1: none yet)
  for (int k = 0; k < nblk; k += 2, ++pair) {
    PanelJob J = pend[0];
    J.lc = k;
    J.nL = nblk - k - 1;
    if (J.nL + J.nU0 + J.nU1 > 0) lu_panel2_kernel<T><<<J.nL + J.nU0 + J.nU1, 256, 0, st>>>(A, ld, n, J, P);
    pend[0] = pend[1];
    pend[1] = PanelJob{0, 0, 0, 0, 0, 0};
    if (k + 1 >= nblk) break;  // odd count: the last block has no pair
    const int nrow = nblk - (k + 1), m = nblk - (k + 2);
    if (fused) {
      lu_cross_panel_kernel<T><<<nrow + m, 256, 0, st>>>(A, ld, n, k, P, info, luctl + 2 * pair, spin_limit);
    } else {
      lu_cross_kernel<T><<<nrow + m, 256, 0, st>>>(A, ld, n, k, P, info);
      if (m > 0) {
        const PanelJob L1 = {k + 1, m, 0, 0, 0, 0};
        lu_panel2_kernel<T><<<m, 256, 0, st>>>(A, ld, n, L1, P);
      }
    }
    if (m > 0) {
      const double mr = n - (k + 2.0) * kNB;
      if (ahead && m > 2) {
        const double mb = mr - 2.0 * kNB;
        (void)hipEventRecord(dep[2 * pair], st);  // this pair's panels are done
        (void)hipStreamWaitEvent(st2, dep[2 * pair], 0);
        update(lu_update128_kernel<T>, st2, (m - 2) * (m - 2), k, 8.0 * 2 * kNB * mb * mb, 2, band);
        if (bulk_done >= 0) (void)hipStreamWaitEvent(st, dep[bulk_done], 0);  // the previous bulk is done
        update(lu_update128_kernel<T>, st, 4 * m - 4, k, 8.0 * 2 * kNB * (mr * mr - mb * mb), 1, band);
        bulk_done = 2 * pair + 1;
        (void)hipEventRecord(dep[bulk_done], st2);  // this pair's bulk is done
      } else {
        if (bulk_done >= 0) (void)hipStreamWaitEvent(st, dep[bulk_done], 0);
        update(lu_update128_kernel<T>, st, m * m, k, 8.0 * 2 * kNB * mr * mr, 0, band);
        bulk_done = -1;  // on st itself: nothing to wait for
      }
    }
    // rows k and k+1 become U once the bulk update that reads them is done:
    // two pairs on without look-ahead ordering them, one on with it
    pend[ahead ? 1 : 0] = PanelJob{0, 0, k, nblk - k - 1, k + 1, nblk - k - 2};
  }
  if (bulk_done >= 0) (void)hipStreamWaitEvent(st, dep[bulk_done], 0);  // the last bulk update
  for (const PanelJob& q : pend)
    if (q.nU0 + q.nU1 > 0) lu_panel2_kernel<T><<<q.nU0 + q.nU1, 256, 0, st>>>(A, ld, n, q, P);
  return timed;
}
template int launch_lu<double>(C2<double>*, int, int, C2<double>*, int*, hipStream_t, int, hipStream_t, hipEvent_t*,
                               hipEvent_t*, double*, int*, int, int);
template int launch_lu<float>(C2<float>*, int, int, C2<float>*, int*, hipStream_t, int, hipStream_t, hipEvent_t*,
                              hipEvent_t*, double*, int*, int, int);

template <typename T>
void launch_lu_solve(const C2<T>* A, int ld, int n, const C2<T>* P, C2<T>* B, C2<T>* W, int nrhs, bool chained,
                     ChainCtl* ctl, int* status, int spin_limit, hipStream_t st, const RefineState* skip) {
  const int nblk = (n + kNB - 1) / kNB;
  const size_t cnt = (size_t)n * nrhs;
  if (chained) {
    chain_begin_kernel<T><<<(unsigned)((cnt + 255) / 256), 256, 0, st>>>(ctl, W, cnt, skip);
    trsv_chain_kernel<T, true><<<nblk, 256, 0, st>>>(A, ld, n, P, B, W, nrhs, ctl, 0, status, spin_limit, skip);
    trsv_chain_kernel<T, false><<<nblk, 256, 0, st>>>(A, ld, n, P, W, B, nrhs, ctl, 1, status, spin_limit, skip);
    return;
  }
  for (int s = 0; s + 1 < nblk; ++s)
    trsv_step_kernel<T, true><<<nblk - 1 - s, 256, 0, st>>>(A, ld, n, B, nrhs, s, skip);
  block_apply_kernel<T><<<nblk, 256, 0, st>>>(P, n, B, W, nrhs, skip);
  for (int s = nblk - 1; s >= 1; --s) trsv_step_kernel<T, false><<<s, 256, 0, st>>>(A, ld, n, W, nrhs, s, skip);
  copy_c2_kernel<T><<<(unsigned)((cnt + 255) / 256), 256, 0, st>>>(W, B, cnt, skip);
}
template void launch_lu_solve<double>(const C2<double>*, int, int, const C2<double>*, C2<double>*, C2<double>*, int,
                                      bool, ChainCtl*, int*, int, hipStream_t, const RefineState*);
template void launch_lu_solve<float>(const C2<float>*, int, int, const C2<float>*, C2<float>*, C2<float>*, int, bool,
                                     ChainCtl*, int*, int, hipStream_t, const RefineState*);

// ---------------------------------------------------------------------------
// Grouped launches (host side)

namespace {
// appends a launch of the given jobs, each on a workgroup range starting at a
// multiple of 8
template <typename T>
void add_launch(LuGroup<T>* g, std::vector<LuJob<T>>& all, int kind, std::vector<LuJob<T>>& jobs) {
  if (jobs.empty()) return;
  int wg = 0;
  for (auto& j : jobs) {
    j.wg0 = wg;
    j.wg1 = wg + j.wg1;  // (wg1 holds the job's workgroup count until here)
    wg = (j.wg1 + 7) & ~7;
  }
  double flops = 0.0, bytes = 0.0;
  if (kind == 3)
    for (const auto& j : jobs) {
      const double m = (double)(j.n - (j.k + 2) * kNB);
      flops += 8.0 * 128.0 * m * m;
      bytes += (double)sizeof(C2<T>) * (2.0 * m * m + 256.0 * m);
    }
  g->launches.push_back({kind, (int)all.size(), (int)jobs.size(), jobs.back().wg1, flops, bytes});
  all.insert(all.end(), jobs.begin(), jobs.end());
  jobs.clear();
}
}  // namespace

template <typename T>
hipError_t build_lu_group(const std::vector<LuJob<T>>& mats, LuGroup<T>* g) {
  g->launches.clear();
  std::vector<LuJob<T>> all, jobs;
  const int nm = (int)mats.size();
  std::vector<int> nblk(nm);
  std::vector<PanelJob> pend(nm, PanelJob{0, 0, 0, 0, 0, 0});
  int maxblk = 0;
  for (int i = 0; i < nm; ++i) {
    nblk[i] = (mats[i].n + kNB - 1) / kNB;
    maxblk = std::max(maxblk, nblk[i]);
  }
  auto job = [&](int i, int k, PanelJob J, int grid) {
    LuJob<T> j = mats[i];
    if (j.band <= 0) j.band = 16;
    j.k = k;
    j.J = J;
    j.wg1 = grid;
    jobs.push_back(j);
  };
  for (int i = 0; i < nm; ++i) job(i, 0, PanelJob{0, 0, 0, 0, 0, 0}, 1);
  add_launch(g, all, 0, jobs);
  // launch_lu's paired loop, step k for every matrix that has it
  for (int k = 0; k < maxblk; k += 2) {
    for (int i = 0; i < nm; ++i) {
      if (k >= nblk[i]) continue;
      PanelJob J = pend[i];
      J.lc = k;
      J.nL = nblk[i] - k - 1;
      pend[i] = PanelJob{0, 0, 0, 0, 0, 0};
      if (J.nL + J.nU0 + J.nU1 > 0) job(i, k, J, J.nL + J.nU0 + J.nU1);
    }
    add_launch(g, all, 1, jobs);
    for (int i = 0; i < nm; ++i)
      if (k + 1 < nblk[i]) job(i, k, PanelJob{0, 0, 0, 0, 0, 0}, (nblk[i] - (k + 1)) + (nblk[i] - (k + 2)));
    add_launch(g, all, 2, jobs);
    for (int i = 0; i < nm; ++i)
      if (k + 2 < nblk[i]) job(i, k, PanelJob{k + 1, nblk[i] - (k + 2), 0, 0, 0, 0}, nblk[i] - (k + 2));
    add_launch(g, all, 1, jobs);
    for (int i = 0; i < nm; ++i)
      if (k + 2 < nblk[i]) job(i, k, PanelJob{0, 0, 0, 0, 0, 0}, (nblk[i] - (k + 2)) * (nblk[i] - (k + 2)));
    add_launch(g, all, 3, jobs);
    for (int i = 0; i < nm; ++i)
      if (k + 1 < nblk[i]) pend[i] = PanelJob{0, 0, k, nblk[i] - k - 1, k + 1, nblk[i] - k - 2};
  }
  for (int i = 0; i < nm; ++i)  // the U rows of each matrix's last pair
    if (pend[i].nU0 + pend[i].nU1 > 0) job(i, 0, pend[i], pend[i].nU0 + pend[i].nU1);
  add_launch(g, all, 1, jobs);
  if (g->d_jobs) (void)hipFree(g->d_jobs);
  g->d_jobs = nullptr;
  hipError_t e = hipMalloc((void**)&g->d_jobs, std::max<size_t>(1, all.size()) * sizeof(LuJob<T>));
  if (e != hipSuccess) return e;
  return hipMemcpy(g->d_jobs, all.data(), all.size() * sizeof(LuJob<T>), hipMemcpyHostToDevice);
}

template <typename T>
int launch_lu_group(const LuGroup<T>& g, hipStream_t st, const hipEvent_t* ev) {
  int updates = 0;
  for (const auto& L : g.launches) {
    const LuJob<T>* j = g.d_jobs + L.first;
    switch (L.kind) {
      case 0: group_inverse_kernel<T><<<L.grid, 256, 0, st>>>(j, L.count); break;
      case 1: group_panel2_kernel<T><<<L.grid, 256, 0, st>>>(j, L.count); break;
      case 2: group_cross_kernel<T><<<L.grid, 256, 0, st>>>(j, L.count); break;
      default:
        if (ev) (void)hipEventRecord(ev[2 * updates], st);
        group_update128_kernel<T><<<L.grid, 256, 0, st>>>(j, L.count);
        if (ev) (void)hipEventRecord(ev[2 * updates + 1], st);
        ++updates;
        break;
    }
  }
  return updates;
}

template <typename T>
void free_trsv_group(TrsvGroup<T>* g) {
  for (void* p : {(void*)g->d_begin, (void*)g->d_tri, (void*)g->d_cum, (void*)g->d_ticket})
    if (p) (void)hipFree(p);
  g->d_begin = g->d_tri = nullptr;
  g->d_cum = nullptr;
  g->d_ticket = nullptr;
  g->count = g->grid_begin = g->grid_tri = g->nlevel = 0;
}

template <typename T>
hipError_t build_trsv_group(const std::vector<TrsvJob<T>>& in, TrsvGroup<T>* g) {
  // largest systems first: the systems that have chain position b are then
  // a prefix of the list, for every b
  std::vector<TrsvJob<T>> sys(in);
  std::stable_sort(sys.begin(), sys.end(), [](const TrsvJob<T>& a, const TrsvJob<T>& b) { return a.n > b.n; });
  std::vector<TrsvJob<T>> begin(sys), tri(sys);
  int wb = 0, wt = 0;
  for (size_t i = 0; i < sys.size(); ++i) {
    begin[i].wg0 = wb;
    begin[i].wg1 = wb + (int)(((size_t)sys[i].n * sys[i].nrhs + 255) / 256);
    wb = begin[i].wg1;
    tri[i].wg0 = wt;
    tri[i].wg1 = wt + (sys[i].n + kNB - 1) / kNB;
    wt = tri[i].wg1;
  }
  // cum[b]: tickets dealt before level b (the systems with more than b row
  // blocks, summed over the levels below b)
  const int nlevel = sys.empty() ? 0 : (sys[0].n + kNB - 1) / kNB;
  std::vector<int> cum(nlevel + 1, 0);
  for (int b = 0; b < nlevel; ++b) {
    int have = 0;
    for (const TrsvJob<T>& j : sys) have += (j.n + kNB - 1) / kNB > b ? 1 : 0;
    cum[b + 1] = cum[b] + have;
  }
  free_trsv_group(g);
  g->count = (int)sys.size();
  g->grid_begin = wb;
  g->grid_tri = wt;
  g->nlevel = nlevel;
  if (cum[nlevel] != wt) return hipErrorInvalidValue;  // every row block gets one ticket
  const size_t bytes = std::max<size_t>(1, sys.size()) * sizeof(TrsvJob<T>);
  hipError_t e = hipMalloc((void**)&g->d_begin, bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&g->d_tri, bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&g->d_cum, cum.size() * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void**)&g->d_ticket, 2 * sizeof(unsigned));
  if (e == hipSuccess) e = hipMemcpy(g->d_begin, begin.data(), sys.size() * sizeof(TrsvJob<T>), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->d_tri, tri.data(), sys.size() * sizeof(TrsvJob<T>), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(g->d_cum, cum.data(), cum.size() * sizeof(int), hipMemcpyHostToDevice);
  return e;
}

template <typename T>
void launch_trsv_group(const TrsvGroup<T>& g, int spin_limit, hipStream_t st) {
  if (g.count == 0) return;
  group_chain_begin_kernel<T><<<g.grid_begin, 256, 0, st>>>(g.d_begin, g.count, g.d_ticket);
  group_trsv_kernel<T, true><<<g.grid_tri, 256, 0, st>>>(g.d_tri, g.count, g.d_cum, g.nlevel, g.d_ticket,
                                                          spin_limit);
  group_trsv_kernel<T, false><<<g.grid_tri, 256, 0, st>>>(g.d_tri, g.count, g.d_cum, g.nlevel, g.d_ticket,
                                                           spin_limit);
}

template hipError_t build_lu_group<float>(const std::vector<LuJob<float>>&, LuGroup<float>*);
template hipError_t build_lu_group<double>(const std::vector<LuJob<double>>&, LuGroup<double>*);
template int launch_lu_group<float>(const LuGroup<float>&, hipStream_t, const hipEvent_t*);
template int launch_lu_group<double>(const LuGroup<double>&, hipStream_t, const hipEvent_t*);
template hipError_t build_trsv_group<float>(const std::vector<TrsvJob<float>>&, TrsvGroup<float>*);
template hipError_t build_trsv_group<double>(const std::vector<TrsvJob<double>>&, TrsvGroup<double>*);
template void launch_trsv_group<float>(const TrsvGroup<float>&, int, hipStream_t);
template void launch_trsv_group<double>(const TrsvGroup<double>&, int, hipStream_t);
template void free_trsv_group<float>(TrsvGroup<float>*);
template void free_trsv_group<double>(TrsvGroup<double>*);

template <typename T>
void launch_fill_test_matrix(C2<T>* A, int ld, int n, hipStream_t st) {
  const size_t cnt = (size_t)n * n;
  fill_test_matrix_kernel<T><<<(unsigned)((cnt + 255) / 256), 256, 0, st>>>(A, ld, n);
}
template void launch_fill_test_matrix<float>(C2<float>*, int, int, hipStream_t);
template void launch_fill_test_matrix<double>(C2<double>*, int, int, hipStream_t);

void launch_mfma_probe(const double* a, const double* b, double* d, hipStream_t st) {
  mfma_probe_kernel<double><<<1, 64, 0, st>>>(a, b, d);
}
void launch_mfma_probe_f32(const float* a, const float* b, float* d, hipStream_t st) {
  mfma_probe_kernel<float><<<1, 64, 0, st>>>(a, b, d);
}

}  // namespace nemoh

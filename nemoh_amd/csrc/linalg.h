// Dense complex linear algebra for the influence-matrix solve on CDNA4.
//
// Matrices are row-major, element = {re, im} (double2 / float2), leading
// dimension ld.  Shared definitions of the block size, the complex element
// type and the CDNA4 MFMA operand / result layouts; the factorisation and the
// solves built on them are in lu.hip.
#pragma once
#include <hip/hip_runtime.h>

namespace nemoh {

constexpr int kNB = 64;  // LU / solve block size

template <typename T>
struct C2 {
  T re, im;
};

template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ C2<T> csub_mul(C2<T> acc, C2<T> a, C2<T> b) {  // acc - a*b
  return {acc.re - (a.re * b.re - a.im * b.im), acc.im - (a.re * b.im + a.im * b.re)};
}
template <typename T>
__device__ __forceinline__ C2<T> cdiv2(C2<T> a, C2<T> b) {
  const T d = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

template <typename T>
struct MfmaTraits;

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA operands: lane l holds A[l%16][l/16] and B[l/16][l%16];
// the 16x16 result is spread as 4 registers per lane, column l%16, with the
// row given by row(l/16, q) -- the FP64 and FP32 instructions differ here.
template <>
struct MfmaTraits<double> {
  typedef v4d acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int q) { return lk + 4 * q; }
};
template <>
struct MfmaTraits<float> {
  typedef v4f acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int q) { return 4 * lk + q; }
};

}  // namespace nemoh

// Complex arithmetic on double2 / float2 for the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace lsw {

template <typename T2>
struct CTraits;
template <>
struct CTraits<double2> {
  using R = double;
  __host__ __device__ static double2 make(double a, double b) { return make_double2(a, b); }
};
template <>
struct CTraits<float2> {
  using R = float;
  __host__ __device__ static float2 make(float a, float b) { return make_float2(a, b); }
};

template <typename T2>
__device__ __forceinline__ T2 czero() {
  return CTraits<T2>::make(0, 0);
}
template <typename T2>
__device__ __forceinline__ T2 cadd(T2 a, T2 b) {
  return CTraits<T2>::make(a.x + b.x, a.y + b.y);
}
template <typename T2>
__device__ __forceinline__ T2 csub(T2 a, T2 b) {
  return CTraits<T2>::make(a.x - b.x, a.y - b.y);
}
template <typename T2>
__device__ __forceinline__ T2 cneg(T2 a) {
  return CTraits<T2>::make(-a.x, -a.y);
}
template <typename T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  return CTraits<T2>::make(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// acc + a*b
template <typename T2>
__device__ __forceinline__ T2 cfma(T2 a, T2 b, T2 acc) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
  return acc;
}
// acc - a*b
template <typename T2>
__device__ __forceinline__ T2 cfms(T2 a, T2 b, T2 acc) {
  acc.x = fma(-a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(-a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
  return acc;
}
// conj(a)*b accumulated
template <typename T2>
__device__ __forceinline__ T2 cfma_conj(T2 a, T2 b, T2 acc) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
  return acc;
}
__device__ __forceinline__ double2 cinv(double2 a) {
  const double d = a.x * a.x + a.y * a.y;
  return make_double2(a.x / d, -a.y / d);
}
template <typename T2>
__device__ __forceinline__ T2 shfl_xor2(T2 v, int mask) {
  v.x = __shfl_xor(v.x, mask, 64);
  v.y = __shfl_xor(v.y, mask, 64);
  return v;
}

}  // namespace lsw

// orad_dual.hpp -- forward-mode dual numbers for the element chains.
//
// Dual<K> carries a double value and K tangents.  Instantiating an element chain
// on Dual<K> evaluates the exact derivative of the arithmetic that chain performs
// (the same branches, taken on the values) along K seeded input directions; the
// adjoint kernels contract those directional derivatives with output adjoints.
// Comparisons look at values only, so min/max/select pick a whole operand, as the
// branch they stand for does.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "orad_math.hpp"

namespace orad {

template <int K>
struct Dual {
  double v;
  double d[K];
  __device__ __forceinline__ Dual() : v(0.0) {
#pragma unroll
    for (int i = 0; i < K; ++i) d[i] = 0.0;
  }
  __device__ __forceinline__ Dual(double x) : v(x) {  // NOLINT: constants convert implicitly
#pragma unroll
    for (int i = 0; i < K; ++i) d[i] = 0.0;
  }
};

// A dual number as the kinematics type: itself, or its value when the kinematics run
// in plain FP64 (used only for inputs without a seeded tangent).
template <typename TG, int K> __device__ __forceinline__ TG as_kin(const Dual<K> &x) {
  if constexpr (std::is_same<TG, Dual<K>>::value) return x;
  else return TG(x.v);
}

template <int K>
__device__ __forceinline__ Dual<K> operator+(const Dual<K> &a, const Dual<K> &b) {
  Dual<K> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K> &a, const Dual<K> &b) {
  Dual<K> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K> &a) {
  Dual<K> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K> &a, const Dual<K> &b) {
  Dual<K> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K> &a, const Dual<K> &b) {
  Dual<K> r;
  const double q = a.v / b.v, ib = 1.0 / b.v;
  r.v = q;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = (a.d[i] - q * b.d[i]) * ib;
  return r;
}
template <int K> __device__ __forceinline__ Dual<K> &operator+=(Dual<K> &a, const Dual<K> &b) {
  a = a + b;
  return a;
}
template <int K> __device__ __forceinline__ bool operator<(const Dual<K> &a, const Dual<K> &b) { return a.v < b.v; }
template <int K> __device__ __forceinline__ bool operator>(const Dual<K> &a, const Dual<K> &b) { return a.v > b.v; }
template <int K> __device__ __forceinline__ bool operator<=(const Dual<K> &a, const Dual<K> &b) { return a.v <= b.v; }
template <int K> __device__ __forceinline__ bool operator>=(const Dual<K> &a, const Dual<K> &b) { return a.v >= b.v; }
template <int K> __device__ __forceinline__ bool operator==(const Dual<K> &a, const Dual<K> &b) { return a.v == b.v; }
template <int K> __device__ __forceinline__ bool operator!=(const Dual<K> &a, const Dual<K> &b) { return a.v != b.v; }

// f(a) with f' = df at a.v
template <int K>
__device__ __forceinline__ Dual<K> chain1(const Dual<K> &a, double f, double df) {
  Dual<K> r;
  r.v = f;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = df * a.d[i];
  return r;
}

template <int K>
struct M<Dual<K>> {
  typedef Dual<K> D;
  // sqrt: the derivative is taken as 0 at 0 (only reached where the result is
  // discarded by a following select, e.g. the von Mises norm of a zero stress)
  static __device__ __forceinline__ D sqrt_(const D &x) {
    const double s = __builtin_sqrt(x.v);
    return chain1(x, s, s > 0.0 ? 0.5 / s : 0.0);
  }
  static __device__ __forceinline__ D log_(const D &x) { return chain1(x, log(x.v), 1.0 / x.v); }
  static __device__ __forceinline__ D exp_(const D &x) {
    const double e = exp(x.v);
    return chain1(x, e, e);
  }
  static __device__ __forceinline__ D powe(const D &x, const D &y) { return exp_(y * log_(x)); }
  static __device__ __forceinline__ D pow_(const D &x, const D &y) { return exp_(y * log_(x)); }
  static __device__ __forceinline__ D cbrt_pow(const D &x) { return pow_(x, D(1.0 / 3.0)); }
  static __device__ __forceinline__ D two_third_pow(const D &x) { return pow_(x, D(2.0 / 3.0)); }
  static __device__ __forceinline__ D rcp(const D &x) { return D(1.0) / x; }
  static __device__ __forceinline__ D div(const D &a, const D &b) { return a / b; }
  static __device__ __forceinline__ D rnorm(const D &x) {
    const D a = sqrt_(x);
    return (a.v != 0.0) ? D(1.0) / a : a;
  }
};

}  // namespace orad

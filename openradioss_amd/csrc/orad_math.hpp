// orad_math.hpp -- precision-dispatched device math for the element chains.
//
// M<double> keeps IEEE operations (the files instantiating it are compiled
// without FMA contraction, so the reference's double-precision operation order
// carries over); M<float> is the mixed form's fast path: hardware reciprocal,
// square root, reciprocal square root, exp2 and log2 (about 1 ulp).
#pragma once
#include <hip/hip_runtime.h>

namespace orad {

template <typename T> struct M;
template <> struct M<double> {
  static __device__ __forceinline__ double sqrt_(double x) { return __builtin_sqrt(x); }
  static __device__ __forceinline__ double pow_(double x, double y) { return pow(x, y); }
  static __device__ __forceinline__ double cbrt_pow(double x) { return pow(x, 1.0 / 3.0); }
  static __device__ __forceinline__ double two_third_pow(double x) { return pow(x, 2.0 / 3.0); }
  static __device__ __forceinline__ double log_(double x) { return log(x); }
  static __device__ __forceinline__ double exp_(double x) { return exp(x); }
  // EXP(y*LOG(x)), the reference's spelling of x**y
  static __device__ __forceinline__ double powe(double x, double y) { return exp(y * log(x)); }
  static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
  static __device__ __forceinline__ double div(double a, double b) { return a / b; }
  // aa = SQRT(x); IF (aa /= 0) aa = 1/aa  (sortho3.F)
  static __device__ __forceinline__ double rnorm(double x) {
    double a = __builtin_sqrt(x);
    return (a != 0.0) ? 1.0 / a : a;
  }
};
template <> struct M<float> {
  static __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float pow_(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
  }
  static __device__ __forceinline__ float cbrt_pow(float x) {
    return __builtin_amdgcn_exp2f((1.0f / 3.0f) * __builtin_amdgcn_logf(x));
  }
  static __device__ __forceinline__ float two_third_pow(float x) {
    return __builtin_amdgcn_exp2f((2.0f / 3.0f) * __builtin_amdgcn_logf(x));
  }
  static __device__ __forceinline__ float log_(float x) { return __logf(x); }
  static __device__ __forceinline__ float exp_(float x) { return __expf(x); }
  static __device__ __forceinline__ float powe(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
  }
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  // FP32 form: hardware reciprocal / square root (1 ulp) instead of the IEEE sequences.
  static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  static __device__ __forceinline__ float rnorm(float x) {
    return (x > 0.0f) ? __builtin_amdgcn_rsqf(x) : 0.0f;
  }
};

// x as the kinematics type TG (a dual number's overload, orad_dual.hpp, gives its value
// when TG is a plain type).
template <typename TG, typename T> __device__ __forceinline__ TG as_kin(const T &x) {
  return TG(x);
}

template <typename T> __device__ __forceinline__ T tmin(T a, T b) { return (a < b) ? a : b; }
template <typename T> __device__ __forceinline__ T tmax(T a, T b) { return (a > b) ? a : b; }

}  // namespace orad

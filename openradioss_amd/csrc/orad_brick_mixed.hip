// orad_brick_mixed.hip -- mixed form: FP32 element arithmetic on coordinates and
// velocities taken relative to corner 1 in FP64; nodes stay FP64.
#include "orad_brick_kernel.hpp"

// Minimum waves per SIMD requested from the register allocator.
#ifndef ORAD_MINW
#define ORAD_MINW 2
#endif

namespace orad {

hipError_t launch_brick_mixed(const BrickArgs<float> &a, int e_begin, int e_end,
                              const double *dt1p, unsigned long long *dt2t, double *ehou_part,
                              int xcd_remap, hipStream_t s) {
  const int n = e_end - e_begin;
  if (n <= 0) return hipSuccess;
  const int nb = (n + BRICK_BLOCK - 1) / BRICK_BLOCK;
  hipLaunchKernelGGL((brick_forces_kernel<float, true, ORAD_MINW>), dim3(nb), dim3(BRICK_BLOCK), 0, s,
                     a, e_begin, e_end, dt1p, dt2t, ehou_part, xcd_remap);
  return hipGetLastError();
}

}  // namespace orad

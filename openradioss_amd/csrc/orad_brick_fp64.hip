// orad_brick_fp64.hip -- FP64 element form.  Built with -ffp-contract=off so its
// arithmetic follows the reference's double-precision build operation by operation.
#include "orad_brick_kernel.hpp"

// Minimum waves per SIMD requested from the register allocator.
#ifndef ORAD_MINW
#define ORAD_MINW 1
#endif

namespace orad {

hipError_t launch_brick_fp64(const BrickArgs<double> &a, int e_begin, int e_end,
                             const double *dt1p, unsigned long long *dt2t, double *ehou_part,
                             int xcd_remap, hipStream_t s) {
  const int n = e_end - e_begin;
  if (n <= 0) return hipSuccess;
  const int nb = (n + BRICK_BLOCK - 1) / BRICK_BLOCK;
  hipLaunchKernelGGL((brick_forces_kernel<double, false, ORAD_MINW>), dim3(nb), dim3(BRICK_BLOCK), 0, s,
                     a, e_begin, e_end, dt1p, dt2t, ehou_part, xcd_remap);
  return hipGetLastError();
}

}  // namespace orad

// orad_brick_fp64.hip -- FP64 element form.  Built with -ffp-contract=off so its
// arithmetic follows the reference's double-precision build operation by operation.
#include "orad_brick_kernel.hpp"

// Instances (threads per workgroup, minimum waves per SIMD) the launch configuration
// and the autotuner choose from.  The default, 64 threads at one wave per SIMD and no
// LDS reservation, was the fastest measured on the 200^3 deck.
namespace orad {

int brick_block_fp64() { return 64; }
int brick_waves_fp64() { return 1; }

bool brick_shape_ok_fp64(int block, int waves) {
  return (waves == 1 && (block == 64 || block == 128 || block == 256)) ||
         (block == 64 && waves == 2);
}

hipError_t launch_brick_fp64(const BrickArgs<double> &a, int e_begin, int e_end,
                             const double *dt1p, unsigned long long *dt2t, double *ehou_part,
                             int xcd_remap, int dyn_lds, int block, int waves, hipStream_t s) {
  if (block <= 0) block = brick_block_fp64();
  if (waves <= 0) waves = brick_waves_fp64();
  const unsigned lds = dyn_lds < 0 ? 0u : (unsigned)dyn_lds;
#define ORAD_CASE(B, W)                                                                       \
  if (block == B && waves == W)                                                               \
    return launch_brick_instance<double, false, W, B>(a, e_begin, e_end, dt1p, dt2t, ehou_part, \
                                                      xcd_remap, lds, s);
  ORAD_CASE(64, 1)
  ORAD_CASE(128, 1)
  ORAD_CASE(256, 1)
  ORAD_CASE(64, 2)
#undef ORAD_CASE
  return hipErrorInvalidConfiguration;
}

}  // namespace orad

// orad_brick_mixed.hip -- mixed form: FP32 element arithmetic on coordinates and
// velocities taken relative to corner 1 in FP64; nodes stay FP64.
#include "orad_brick_kernel.hpp"

// Instances (threads per workgroup, minimum waves per SIMD) the launch configuration
// and the autotuner choose from.  The default, 256 threads at two waves per SIMD with
// the LDS reservation that keeps a CU at two workgroups, was the fastest measured on
// the 200^3 deck: more resident waves evict the corner-node records that neighbouring
// elements reuse before they are reused.
namespace orad {

int brick_block_mixed() { return 256; }
int brick_waves_mixed() { return 2; }

bool brick_shape_ok_mixed(int block, int waves) {
  return (block == 256 && (waves == 2 || waves == 3)) || (block == 128 && waves == 2) ||
         (block == 64 && waves == 2);
}

hipError_t launch_brick_mixed(const BrickArgs<float> &a, int e_begin, int e_end,
                              const double *dt1p, unsigned long long *dt2t, double *ehou_part,
                              int xcd_remap, int dyn_lds, int block, int waves, hipStream_t s) {
  if (block <= 0) block = brick_block_mixed();
  if (waves <= 0) waves = brick_waves_mixed();
  const unsigned lds = dyn_lds < 0 ? lds_for_waves(block, waves) : (unsigned)dyn_lds;
#define ORAD_CASE(B, W)                                                                      \
  if (block == B && waves == W)                                                              \
    return launch_brick_instance<float, true, W, B>(a, e_begin, e_end, dt1p, dt2t, ehou_part, \
                                                    xcd_remap, lds, s);
  ORAD_CASE(256, 2)
  ORAD_CASE(256, 3)
  ORAD_CASE(128, 2)
  ORAD_CASE(64, 2)
#undef ORAD_CASE
  return hipErrorInvalidConfiguration;
}

}  // namespace orad

// orad_brick_mixed.hip -- mixed form: FP32 element arithmetic on coordinates and
// velocities taken relative to corner 1 in FP64; nodes stay FP64.
#include "orad_brick_kernel.hpp"

// Launch shape (measured on the 200^3 deck): MINW = minimum waves per SIMD asked of
// the register allocator, BLOCK = threads per workgroup, LDS_KB = dynamic LDS
// reserved per workgroup, which caps the groups resident on a CU (160 KB of LDS):
// fewer groups in flight keep more of the shared corner-node records in cache.
#ifndef ORAD_MINW
#define ORAD_MINW 2
#endif
#ifndef ORAD_BLOCK
#define ORAD_BLOCK 256
#endif
#ifndef ORAD_LDS_KB
#define ORAD_LDS_KB 54
#endif

namespace orad {

int brick_block_mixed() { return ORAD_BLOCK; }

hipError_t launch_brick_mixed(const BrickArgs<float> &a, int e_begin, int e_end,
                             const double *dt1p, unsigned long long *dt2t, double *ehou_part,
                             int xcd_remap, int dyn_lds, hipStream_t s) {
  const int n = e_end - e_begin;
  if (n <= 0) return hipSuccess;
  const int nb = (n + ORAD_BLOCK - 1) / ORAD_BLOCK;
  const unsigned lds = dyn_lds < 0 ? 1024u * ORAD_LDS_KB : (unsigned)dyn_lds;
  hipLaunchKernelGGL((brick_forces_kernel<float, true, ORAD_MINW, ORAD_BLOCK>), dim3(nb),
                     dim3(ORAD_BLOCK), lds, s, a, e_begin, e_end, dt1p, dt2t, ehou_part, xcd_remap);
  return hipGetLastError();
}

}  // namespace orad

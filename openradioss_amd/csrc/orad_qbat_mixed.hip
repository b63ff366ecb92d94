// orad_qbat_mixed.hip -- QBAT shell element kernel, mixed form: FP32 state, law,
// resultants and forces; nodes stay FP64.  The frame, warped-element geometry,
// rigid-body projection and strain-rate stage runs in FP64 on coordinates and
// translational velocities taken relative to corner 1.
#include "orad_qbat_kernel.hpp"

#ifndef ORAD_QBAT_BLOCK_MIXED
#define ORAD_QBAT_BLOCK_MIXED 64
#endif

namespace orad {

int qbat_block_mixed() { return ORAD_QBAT_BLOCK_MIXED; }

hipError_t launch_qbat_mixed(const ShellArgs<float> &a, int e_begin, int e_end, const double *dt1p,
                             unsigned long long *dt2t, double *ehou_part, int xcd_remap,
                             hipStream_t s) {
  const int n = e_end - e_begin;
  if (n <= 0) return hipSuccess;
  const int nb = (n + ORAD_QBAT_BLOCK_MIXED - 1) / ORAD_QBAT_BLOCK_MIXED;
  hipLaunchKernelGGL((qbat_forces_kernel<float, double, true, ORAD_QBAT_BLOCK_MIXED>), dim3(nb),
                     dim3(ORAD_QBAT_BLOCK_MIXED), 0, s, a, e_begin, e_end, dt1p, dt2t, ehou_part,
                     xcd_remap);
  return hipGetLastError();
}

}  // namespace orad

// orad_qbat_fp64.hip -- QBAT shell element kernel, FP64 form.  Built with
// -ffp-contract=off so its arithmetic follows the reference's double-precision
// build operation by operation.
#include "orad_qbat_kernel.hpp"

#ifndef ORAD_QBAT_BLOCK
#define ORAD_QBAT_BLOCK 64
#endif

namespace orad {

int qbat_block_fp64() { return ORAD_QBAT_BLOCK; }

hipError_t launch_qbat_fp64(const ShellArgs<double> &a, int e_begin, int e_end, const double *dt1p,
                            unsigned long long *dt2t, double *ehou_part, int xcd_remap,
                            hipStream_t s) {
  const int n = e_end - e_begin;
  if (n <= 0) return hipSuccess;
  const int nb = (n + ORAD_QBAT_BLOCK - 1) / ORAD_QBAT_BLOCK;
  hipLaunchKernelGGL((qbat_forces_kernel<double, double, false, ORAD_QBAT_BLOCK>), dim3(nb),
                     dim3(ORAD_QBAT_BLOCK), 0, s, a, e_begin, e_end, dt1p, dt2t, ehou_part,
                     xcd_remap);
  return hipGetLastError();
}

}  // namespace orad

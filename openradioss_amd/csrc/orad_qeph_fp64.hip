// orad_qeph_fp64.hip -- QEPH shell element kernel, FP64 form.  Built with
// -ffp-contract=off so its arithmetic follows the reference's double-precision
// build operation by operation.
#include "orad_qeph_kernel.hpp"

#ifndef ORAD_QEPH_BLOCK
#define ORAD_QEPH_BLOCK 64
#endif

namespace orad {

int qeph_block_fp64() { return ORAD_QEPH_BLOCK; }

hipError_t launch_qeph_fp64(const ShellArgs<double> &a, int e_begin, int e_end, const double *dt1p,
                            unsigned long long *dt2t, double *ehou_part, int xcd_remap,
                            hipStream_t s) {
  const int n = e_end - e_begin;
  if (n <= 0) return hipSuccess;
  const int nb = (n + ORAD_QEPH_BLOCK - 1) / ORAD_QEPH_BLOCK;
  hipLaunchKernelGGL((qeph_forces_kernel<double, double, false, ORAD_QEPH_BLOCK>), dim3(nb),
                     dim3(ORAD_QEPH_BLOCK), 0, s, a, e_begin, e_end, dt1p, dt2t, ehou_part,
                     xcd_remap);
  return hipGetLastError();
}

}  // namespace orad

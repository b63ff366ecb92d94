// orad_qeph_fp64.hip -- QEPH shell element kernel, FP64 form.  Built with
// -ffp-contract=off so its arithmetic follows the reference's double-precision
// build operation by operation.
#include "orad_batch_kernel.hpp"
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

// Batched QEPH models: one workgroup per tile, or (persist) a resident grid taking
// tiles from the counter; dyn_lds reserves LDS per workgroup to cap residency per CU.
hipError_t launch_qeph_batch_fp64(const ShellArgs<double> &a, const BatchArgs &b, int parity,
                                   int persist, int grid, int dyn_lds, int xcd_remap,
                                   hipStream_t s) {
  if (b.ntiles <= 0) return hipSuccess;
  const size_t lds = dyn_lds > 0 ? (size_t)dyn_lds : 0;
  if (persist)
    hipLaunchKernelGGL((shell_batch_kernel<double, double, false, ORAD_QEPH_BLOCK, true, true>),
                       dim3(grid), dim3(ORAD_QEPH_BLOCK), lds, s, a, b, parity, xcd_remap);
  else
    hipLaunchKernelGGL((shell_batch_kernel<double, double, false, ORAD_QEPH_BLOCK, false, true>),
                       dim3(b.ntiles), dim3(ORAD_QEPH_BLOCK), lds, s, a, b, parity, xcd_remap);
  return hipGetLastError();
}

// Resident workgroups per CU of the persistent batched QEPH kernel at dyn_lds bytes.
int qeph_batch_occupancy_fp64(int dyn_lds) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &nb, shell_batch_kernel<double, double, false, ORAD_QEPH_BLOCK, true, true>,
          ORAD_QEPH_BLOCK, dyn_lds > 0 ? (size_t)dyn_lds : 0) != hipSuccess)
    return 0;
  return nb;
}

}  // namespace orad

// orad_shell_mixed.hip -- shell element kernel, mixed form: FP32 element arithmetic and state;
// nodes stay FP64.  The frame and kinematics stage (local frame, shape derivatives,
// strain and curvature rates) runs on coordinates and translational velocities
// taken relative to corner 1 in FP64, in the precision ORAD_SHELL_KIN_FP64 selects.
#include "orad_batch_kernel.hpp"

#ifndef ORAD_SHELL_KIN_FP64
#define ORAD_SHELL_KIN_FP64 1
#endif

#ifndef ORAD_SHELL_BLOCK
#define ORAD_SHELL_BLOCK 256
#endif

namespace orad {

#if ORAD_SHELL_KIN_FP64
typedef double ShellKinT;
#else
typedef float ShellKinT;
#endif

int shell_block_mixed() { return ORAD_SHELL_BLOCK; }

hipError_t launch_shell_mixed(const ShellArgs<float> &a, int e_begin, int e_end, const double *dt1p,
                             unsigned long long *dt2t, double *ehou_part, int xcd_remap,
                             hipStream_t s) {
  const int n = e_end - e_begin;
  if (n <= 0) return hipSuccess;
  const int nb = (n + ORAD_SHELL_BLOCK - 1) / ORAD_SHELL_BLOCK;
  hipLaunchKernelGGL((shell_forces_kernel<float, ShellKinT, true, ORAD_SHELL_BLOCK>), dim3(nb),
                     dim3(ORAD_SHELL_BLOCK), 0, s, a, e_begin, e_end, dt1p, dt2t, ehou_part,
                     xcd_remap);
  return hipGetLastError();
}

// Batched models: one workgroup per tile, or (persist) a resident grid taking tiles
// from the counter; dyn_lds reserves LDS per workgroup to cap residency per CU.
hipError_t launch_shell_batch_mixed(const ShellArgs<float> &a, const BatchArgs &b, int parity,
                                    int persist, int grid, int dyn_lds, int xcd_remap,
                                    hipStream_t s) {
  if (b.ntiles <= 0) return hipSuccess;
  const size_t lds = dyn_lds > 0 ? (size_t)dyn_lds : 0;
  if (persist)
    hipLaunchKernelGGL((shell_batch_kernel<float, ShellKinT, true, ORAD_SHELL_BLOCK, true, false>), dim3(grid),
                       dim3(ORAD_SHELL_BLOCK), lds, s, a, b, parity, xcd_remap);
  else
    hipLaunchKernelGGL((shell_batch_kernel<float, ShellKinT, true, ORAD_SHELL_BLOCK, false, false>),
                       dim3(b.ntiles), dim3(ORAD_SHELL_BLOCK), lds, s, a, b, parity, xcd_remap);
  return hipGetLastError();
}

// Resident workgroups per CU of the persistent batch kernel at dyn_lds bytes.
int shell_batch_occupancy_mixed(int dyn_lds) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &nb, shell_batch_kernel<float, ShellKinT, true, ORAD_SHELL_BLOCK, true, false>, ORAD_SHELL_BLOCK,
          dyn_lds > 0 ? (size_t)dyn_lds : 0) != hipSuccess)
    return 0;
  return nb;
}

}  // namespace orad

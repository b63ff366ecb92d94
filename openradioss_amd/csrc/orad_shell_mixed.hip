// orad_shell_mixed.hip -- shell element kernel, mixed form: FP32 element arithmetic and state;
// nodes stay FP64.  The frame and kinematics stage (local frame, shape derivatives,
// strain and curvature rates) runs on coordinates and translational velocities
// taken relative to corner 1 in FP64, in the precision ORAD_SHELL_KIN_FP64 selects.
#include "orad_shell_kernel.hpp"

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

}  // namespace orad

// orad_adjoint_args.hpp -- arguments of the element adjoint kernel (shared by the
// kernels and the host driver).
#pragma once
#include "orad_common.hpp"

namespace orad {

constexpr int ADJ_NIN = 73;      // element inputs with an adjoint
constexpr int ADJ_NST = 33;      // element state columns with an adjoint

// state column -> adjoint index (SIG 0..24, PLA 25..29, HOUR(1:3) 30..32), or -1
__host__ __device__ __forceinline__ int adj_state_index(int col) {
  if (col >= SC_SIG && col < SC_SIG + 25) return col - SC_SIG;
  if (col >= SC_PLA && col < SC_PLA + 5) return 25 + col - SC_PLA;
  if (col >= SC_HOUR && col < SC_HOUR + 3) return 30 + col - SC_HOUR;
  return -1;
}
__host__ __device__ __forceinline__ int adj_state_col(int j) {
  return j < 25 ? SC_SIG + j : (j < 30 ? SC_PLA + j - 25 : SC_HOUR + j - 30);
}

template <typename S>
struct AdjArgs {
  int numelc;
  const int32_t *conn;     // [4][ne]
  const double4 *X4, *V4, *VR4;
  const S *st;             // [SC_NCOL][ne] pre-cycle element state
  const double *thke;
  const double *lamS, *lamR;   // [3 * numnod] adjoints of the slot sums
  const double *lst_in;    // [ADJ_NST][ne] adjoint of the post-cycle state
  double *lst_out;         // [ADJ_NST][ne] adjoint of the pre-cycle state
  double *lcorner;         // [36][ne] corner c, 9c + (x, y, z, vx, vy, vz, vrx, vry, vrz)
  double *g_thk;           // [ne]
  double *g_jc;            // [3][ne]
  ShMat mat;
  ShProp prop;
  ShConst k;
  double dt1;
};

}  // namespace orad

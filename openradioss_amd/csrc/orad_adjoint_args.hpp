// orad_adjoint_args.hpp -- arguments of the element adjoint kernel (shared by the
// kernels and the host driver).
#pragma once
#include "orad_common.hpp"

namespace orad {

// Element inputs with an adjoint, in this order: 4 corners x (x, y, z, vx, vy, vz,
// vrx, vry, vrz) = 36, the state columns carried from cycle to cycle (SIG of points
// 1..5, PLA(1..5), then the incremental hourglass forces: HOUR(1:3) for
// Belytschko-Tsay, HOURG(1:12) for QEPH), the thickness, then A, B, n.
template <bool QEPH> struct AdjShape {
  static constexpr int NST = QEPH ? 42 : 33;   // state columns with an adjoint
  static constexpr int I_THK = 36 + NST;
  static constexpr int I_JC = I_THK + 1;
  static constexpr int NIN = I_JC + 3;          // 73 (BT) or 82 (QEPH)
};
constexpr int ADJ_NST_MAX = AdjShape<true>::NST;

// state column -> adjoint index (SIG 0..24, PLA 25..29, hourglass 30..), or -1
template <bool QEPH>
__host__ __device__ __forceinline__ int adj_state_index(int col) {
  if (col >= SC_SIG && col < SC_SIG + 25) return col - SC_SIG;
  if (col >= SC_PLA && col < SC_PLA + 5) return 25 + col - SC_PLA;
  if (QEPH) {
    if (col >= SC_HOUR && col < SC_HOUR + 5) return 30 + col - SC_HOUR;
    if (col >= SC_HOURX && col < SC_HOURX + 7) return 35 + col - SC_HOURX;
  } else {
    if (col >= SC_HOUR && col < SC_HOUR + 3) return 30 + col - SC_HOUR;
  }
  return -1;
}

template <typename S>
struct AdjArgs {
  int numelc;
  const int32_t *conn;     // [4][ne]
  const double4 *X4, *V4, *VR4;
  const S *st;             // [SC_NCOL or SC_NCOL_QEPH][ne] pre-cycle element state
  const double *thke;
  const double *lamS, *lamR;   // [3 * numnod] adjoints of the slot sums
  const double *lst_in;    // [NST][ne] adjoint of the post-cycle state
  double *lst_out;         // [NST][ne] adjoint of the pre-cycle state
  double *lcorner;         // [36][ne] corner c, 9c + (x, y, z, vx, vy, vz, vrx, vry, vrz)
  double *g_thk;           // [ne]
  double *g_jc;            // [3][ne]
  ShMat mat;
  ShProp prop;
  ShConst k;
  double dt1;
};

// Brick element inputs with an adjoint, in this order: 8 corners x (x, y, z, vx, vy,
// vz) = 48, the carried state that reaches the forces (SIG(6), PLA) = 7, then A, B, n.
constexpr int ADJB_NST = 7, ADJB_I_JC = 55, ADJB_NIN = 58;

template <typename S>
struct AdjBrickArgs {
  int numels;
  const int32_t *conn;     // [8][ne], library element order
  const double4 *X4, *V4;
  const S *st;             // [12][ne] pre-cycle element state: SIG(6) EINT RHO PLA QVIS OFF EPSD
  const double *vol0;
  const double *lamS;      // [3 * numnod] adjoints of the slot sums
  const double *lst_in;    // [7][ne] adjoint of the post-cycle SIG, PLA
  double *lst_out;         // [7][ne] adjoint of the pre-cycle SIG, PLA
  double *lcorner;         // [48][ne] corner k, 6k + (x, y, z, vx, vy, vz)
  double *g_jc;            // [3][ne]
  Law2 mat;
  SProp prop;
  double dt1;
};

}  // namespace orad

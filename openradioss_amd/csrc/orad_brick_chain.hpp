// orad_brick_chain.hpp -- one under-integrated 8-node brick through the explicit
// force chain of SFORC3 (engine/source/elements/solid/solide/sforc3.F:125) for
// the default solid: Isolid=1 (JHBE=1), co-rotational frame (JCVT=1, the starter's
// automatic choice for 8-node Lagrangian bricks, sgrtails.F:857-859), Ismstr=4,
// law 2 Johnson-Cook, viscous hourglass, PARITH/ON slot output.
//
// T = double reproduces the reference's double-precision operation order (the
// file instantiating it is compiled without FMA contraction).  T = float is the
// mixed form: the caller passes coordinates and velocities relative to corner 1,
// formed in FP64, so every later difference is exact to FP32 rounding of O(h).
// Any other T (a dual number, orad_dual.hpp) takes the FP64 form's steps.
#pragma once
#include <type_traits>

#include "orad_common.hpp"
#include "orad_math.hpp"

namespace orad {

// Element outputs besides the corner forces, which go out through `emit(k, fx, fy, fz)`
// (global frame) as soon as each corner is complete.
template <typename T>
struct BrickOut {
  T sti;           // STI/4 (SCUMU3P)
  T dtx;           // DTFAC*DTX, or a huge value when the element gives no candidate
  T ehou;          // hourglass energy increment
  int mstop;
};

// Reference constants (common_source/modules/constant_mod.F) in precision T.
template <typename T> struct K {
  static constexpr T THIRD = T(0x1.5555555555555p-2);
  static constexpr T TWO_THIRD = T(0x1.5555555555555p-1);
  static constexpr T ONEP333 = T(0x1.553f7ced91687p+0);
  static constexpr T EM15 = T(0x1.203af9ee75616p-50);
  static constexpr T EM20 = T(0x1.79ca10c924223p-67);
  static constexpr T ONE_OVER_64 = T(0x1.0p-6);
};

// Hourglass base vectors Gamma_m (shvis3.F:317-353 without the PxHy terms).
__device__ __forceinline__ int gamma_sign(int m, int k) {
  const int g[3][8] = {{1, -1, 1, -1, 1, -1, 1, -1},
                       {1, 1, -1, -1, -1, -1, 1, 1},
                       {1, -1, -1, 1, -1, 1, 1, -1}};
  return g[m][k];
}
__device__ __forceinline__ float gamma_of(int m, int k) { return float(gamma_sign(m, k)); }

// Reference gamma vector G(k, m) = Gamma_m[k] -/+ PX(a)H(m) (shvis3.F:317-353):
// corners 1-4 subtract PX1..PX4 H, corners 5..8 add PX3, PX4, PX1, PX2 H.
template <typename T>
__device__ __forceinline__ T G_of(int m, int k, const T PH[3][4]) {
  const int ak[8] = {0, 1, 2, 3, 2, 3, 0, 1};
  const T one = T(1);
  const T base = (gamma_sign(m, k) > 0) ? one : -one;
  return (k < 4) ? base - PH[m][ak[k]] : base + PH[m][ak[k]];
}

// st_* are the element's state values (in/out).  xd/yd/zd, vx/vy/vz: the eight
// corners' coordinates and velocities (absolute for T=double, relative to corner
// 1 for the mixed form).  vol0d: initial volume (FP64).  voldp: the mixed form's
// FP64 Jacobian determinant (VOLDP), used, as the reference's single-precision
// build does (srho3.F, IRESP=1 branch), for density, volume change and the
// volumetric strain mu, where FP32 would cancel; ignored when T = double.
// Corner values of one direction: plain registers, or a column of an LDS tile
// (thread-major, stride = workgroup size) to keep them out of registers.
template <typename T, int STRIDE>
struct Col {
  T *p;
  __device__ __forceinline__ T &operator[](int k) const { return p[k * STRIDE]; }
};

// jc = (A, B, n) of the hardening curve (mat.ca, mat.cb, mat.cn; separate so that an
// adjoint can seed them).
template <typename T, typename VA, typename Emit>
__device__ __forceinline__ void brick_chain(T xd[8], T yd[8], T zd[8], VA vx, VA vy, VA vz,
                                            T sig[6], T &st_eint, T &st_rho, double vol0d,
                                            double voldp, T &st_pla,
                                            T &st_qvis, T &st_offg, T &st_epsd, T dt1,
                                            const Law2 &mat, const T jc[3], const SProp &prop,
                                            BrickOut<T> &out, Emit emit) {
  constexpr bool FP32 = std::is_same<T, float>::value;   // the mixed form's factored steps
  const T ZERO = T(0), ONE = T(1), TWO = T(2), THREE = T(3), FOUR = T(4), HALF = T(0.5);
  const T THIRD = T(K<double>::THIRD), ONEP333 = T(K<double>::ONEP333);
  const T EM15 = T(K<double>::EM15), EM20 = T(K<double>::EM20);
  const T ONE_OVER_64 = T(K<double>::ONE_OVER_64);
  T offg = st_offg;
  int mstop = 0;

  // ---- SREPISO3 (srepiso3.F:60-92) ----
  T r[3], s[3], t[3];
  {
    T X17 = xd[6] - xd[0], X28 = xd[7] - xd[1], X35 = xd[4] - xd[2], X46 = xd[5] - xd[3];
    T Y17 = yd[6] - yd[0], Y28 = yd[7] - yd[1], Y35 = yd[4] - yd[2], Y46 = yd[5] - yd[3];
    T Z17 = zd[6] - zd[0], Z28 = zd[7] - zd[1], Z35 = zd[4] - zd[2], Z46 = zd[5] - zd[3];
    T A17 = X17 + X46, A28 = X28 + X35, B17 = Y17 + Y46, B28 = Y28 + Y35;
    T C17 = Z17 + Z46, C28 = Z28 + Z35;
    r[0] = X17 + X28 - X35 - X46; r[1] = Y17 + Y28 - Y35 - Y46; r[2] = Z17 + Z28 - Z35 - Z46;
    s[0] = A17 + A28; s[1] = B17 + B28; s[2] = C17 + C28;
    t[0] = A17 - A28; t[1] = B17 - B28; t[2] = C17 - C28;
  }
  // ---- SORTHO3 (sortho3.F:70-160), 3 iterations ----
  T R[3][3];  // R[i][j] = R(i+1,j+1); column j is e_(j+1)
  {
    T aa, bb, U[3], V[3], W[3], e1[3], e2[3], e3[3];
    aa = M<T>::rnorm(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    U[0] = r[0] * aa; U[1] = r[1] * aa; U[2] = r[2] * aa;
    aa = M<T>::rnorm(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    V[0] = s[0] * aa; V[1] = s[1] * aa; V[2] = s[2] * aa;
    aa = M<T>::rnorm(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    W[0] = t[0] * aa; W[1] = t[1] * aa; W[2] = t[2] * aa;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      e1[0] = V[1] * W[2] - V[2] * W[1] + U[0];
      e1[1] = V[2] * W[0] - V[0] * W[2] + U[1];
      e1[2] = V[0] * W[1] - V[1] * W[0] + U[2];
      e2[0] = W[1] * U[2] - W[2] * U[1] + V[0];
      e2[1] = W[2] * U[0] - W[0] * U[2] + V[1];
      e2[2] = W[0] * U[1] - W[1] * U[0] + V[2];
      e3[0] = U[1] * V[2] - U[2] * V[1] + W[0];
      e3[1] = U[2] * V[0] - U[0] * V[2] + W[1];
      e3[2] = U[0] * V[1] - U[1] * V[0] + W[2];
      bb = M<T>::rnorm(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
      U[0] = e1[0] * bb; U[1] = e1[1] * bb; U[2] = e1[2] * bb;
      bb = M<T>::rnorm(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
      V[0] = e2[0] * bb; V[1] = e2[1] * bb; V[2] = e2[2] * bb;
      bb = M<T>::rnorm(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2]);
      W[0] = e3[0] * bb; W[1] = e3[1] * bb; W[2] = e3[2] * bb;
    }
    e1[0] = U[0]; e1[1] = U[1]; e1[2] = U[2];
    e3[0] = e1[1] * V[2] - e1[2] * V[1];
    e3[1] = e1[2] * V[0] - e1[0] * V[2];
    e3[2] = e1[0] * V[1] - e1[1] * V[0];
    aa = M<T>::rnorm(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2]);
    e3[0] = e3[0] * aa; e3[1] = e3[1] * aa; e3[2] = e3[2] * aa;
    e2[0] = e3[1] * e1[2] - e3[2] * e1[1];
    e2[1] = e3[2] * e1[0] - e3[0] * e1[2];
    e2[2] = e3[0] * e1[1] - e3[1] * e1[0];
#pragma unroll
    for (int i = 0; i < 3; ++i) { R[i][0] = e1[i]; R[i][1] = e2[i]; R[i][2] = e3[i]; }
  }
  // ---- SRCOOR3: coordinates and velocities to the convected frame (srcoor3.F:415-447,668-678) ----
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    T XDL = R[0][0] * xd[k] + R[1][0] * yd[k] + R[2][0] * zd[k];
    T YDL = R[0][1] * xd[k] + R[1][1] * yd[k] + R[2][1] * zd[k];
    T ZDL = R[0][2] * xd[k] + R[1][2] * yd[k] + R[2][2] * zd[k];
    xd[k] = XDL; yd[k] = YDL; zd[k] = ZDL;
  }
  T off = (offg < ZERO) ? -offg : offg;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    T X = R[0][0] * vx[k] + R[1][0] * vy[k] + R[2][0] * vz[k];
    T Y = R[0][1] * vx[k] + R[1][1] * vy[k] + R[2][1] * vz[k];
    T Z = R[0][2] * vx[k] + R[1][2] * vy[k] + R[2][2] * vz[k];
    vx[k] = X; vy[k] = Y; vz[k] = Z;
  }
  if (offg < ZERO) {
#pragma unroll
    for (int k = 0; k < 8; ++k) { vx[k] = ZERO; vy[k] = ZERO; vz[k] = ZERO; }
  }

  // ---- SDERI3 (sderi3.F:104-155, 265-335) + SCHKJABT3 ----
  T PX[4], PY[4], PZ[4], PH[3][4], voln;
  {
    T X17 = xd[6] - xd[0], X28 = xd[7] - xd[1], X35 = xd[4] - xd[2], X46 = xd[5] - xd[3];
    T Y17 = yd[6] - yd[0], Y28 = yd[7] - yd[1], Y35 = yd[4] - yd[2], Y46 = yd[5] - yd[3];
    T Z17 = zd[6] - zd[0], Z28 = zd[7] - zd[1], Z35 = zd[4] - zd[2], Z46 = zd[5] - zd[3];
    T JAC1 = X17 + X28 - X35 - X46, JAC2 = Y17 + Y28 - Y35 - Y46, JAC3 = Z17 + Z28 - Z35 - Z46;
    T X_17_46 = X17 + X46, X_28_35 = X28 + X35;
    T Y_17_46 = Y17 + Y46, Y_28_35 = Y28 + Y35;
    T Z_17_46 = Z17 + Z46, Z_28_35 = Z28 + Z35;
    T JAC4 = X_17_46 + X_28_35, JAC5 = Y_17_46 + Y_28_35, JAC6 = Z_17_46 + Z_28_35;
    T JAC7 = X_17_46 - X_28_35, JAC8 = Y_17_46 - Y_28_35, JAC9 = Z_17_46 - Z_28_35;
    T JAC_59_68 = JAC5 * JAC9 - JAC6 * JAC8;
    T JAC_67_49 = JAC6 * JAC7 - JAC4 * JAC9;
    T JAC_48_57 = JAC4 * JAC8 - JAC5 * JAC7;
    T det;
    if constexpr (!FP32) det = ONE_OVER_64 * (JAC1 * JAC_59_68 + JAC2 * JAC_67_49 + JAC3 * JAC_48_57);
    else det = T(voldp);
    if (off == ZERO) {
      det = ONE;
    } else if (offg > ONE) {
      // small-strain switch state, not produced on this path
    } else if (det <= ZERO) {
      det = ONE; off = ZERO; offg = ZERO; mstop = 1;
    }
    voln = det;
    T DETT = M<T>::div(ONE_OVER_64, det);
    T JACI1 = DETT * JAC_59_68, JACI4 = DETT * JAC_67_49, JACI7 = DETT * JAC_48_57;
    T JACI2 = DETT * (JAC3 * JAC8 - JAC2 * JAC9);
    T JACI5 = DETT * (JAC1 * JAC9 - JAC3 * JAC7);
    T JACI8 = DETT * (JAC2 * JAC7 - JAC1 * JAC8);
    T JACI3 = DETT * (JAC2 * JAC6 - JAC3 * JAC5);
    T JACI6 = DETT * (JAC3 * JAC4 - JAC1 * JAC6);
    T JACI9 = DETT * (JAC1 * JAC5 - JAC2 * JAC4);
    T JAC12 = JACI1 + JACI2, JAC45 = JACI4 + JACI5, JAC78 = JACI7 + JACI8;
    PX[0] = -JAC12 - JACI3; PY[0] = -JAC45 - JACI6; PZ[0] = -JAC78 - JACI9;
    PX[1] = -JAC12 + JACI3; PY[1] = -JAC45 + JACI6; PZ[1] = -JAC78 + JACI9;
    JAC12 = JACI1 - JACI2; JAC45 = JACI4 - JACI5; JAC78 = JACI7 - JACI8;
    PX[2] = JAC12 + JACI3; PY[2] = JAC45 + JACI6; PZ[2] = JAC78 + JACI9;
    PX[3] = JAC12 - JACI3; PY[3] = JAC45 - JACI6; PZ[3] = JAC78 - JACI9;
    // hourglass modes {1-11-11-11-1}, {11-1-1-1-111}, {1-1-11-111-1}
    T HX, HY, HZ;
    HX = xd[0] - xd[1] + xd[2] - xd[3] + xd[4] - xd[5] + xd[6] - xd[7];
    HY = yd[0] - yd[1] + yd[2] - yd[3] + yd[4] - yd[5] + yd[6] - yd[7];
    HZ = zd[0] - zd[1] + zd[2] - zd[3] + zd[4] - zd[5] + zd[6] - zd[7];
#pragma unroll
    for (int a = 0; a < 4; ++a) PH[0][a] = PX[a] * HX + PY[a] * HY + PZ[a] * HZ;
    HX = xd[0] + xd[1] - xd[2] - xd[3] - xd[4] - xd[5] + xd[6] + xd[7];
    HY = yd[0] + yd[1] - yd[2] - yd[3] - yd[4] - yd[5] + yd[6] + yd[7];
    HZ = zd[0] + zd[1] - zd[2] - zd[3] - zd[4] - zd[5] + zd[6] + zd[7];
#pragma unroll
    for (int a = 0; a < 4; ++a) PH[1][a] = PX[a] * HX + PY[a] * HY + PZ[a] * HZ;
    HX = xd[0] - xd[1] - xd[2] + xd[3] - xd[4] + xd[5] + xd[6] - xd[7];
    HY = yd[0] - yd[1] - yd[2] + yd[3] - yd[4] + yd[5] + yd[6] - yd[7];
    HZ = zd[0] - zd[1] - zd[2] + zd[3] - zd[4] + zd[5] + zd[6] - zd[7];
#pragma unroll
    for (int a = 0; a < 4; ++a) PH[2][a] = PX[a] * HX + PY[a] * HY + PZ[a] * HZ;
  }

  // ---- SDLEN3 / SLEN (sdlen3.F:109-167, slen.F:67-87) ----
  T deltax;
  {
    T aream = EM20;
    const int fc[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                          {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const int a = fc[f][0], b = fc[f][1], c = fc[f][2], d = fc[f][3];
      T X13 = xd[c] - xd[a], X24 = xd[d] - xd[b], Y13 = yd[c] - yd[a], Y24 = yd[d] - yd[b];
      T Z13 = zd[c] - zd[a], Z24 = zd[d] - zd[b];
      T FS1 = X13 - X24, FT1 = X13 + X24, FS2 = Y13 - Y24, FT2 = Y13 + Y24;
      T FS3 = Z13 - Z24, FT3 = Z13 + Z24;
      T E = FS1 * FS1 + FS2 * FS2 + FS3 * FS3;
      T F = FS1 * FT1 + FS2 * FT2 + FS3 * FT3;
      T G = FT1 * FT1 + FT2 * FT2 + FT3 * FT3;
      T ar = E * G - F * F;
      aream = tmax(ar, aream);
    }
    deltax = M<T>::div(FOUR * voln * ONE, M<T>::sqrt_(aream));
  }

  // ---- SDEFO3, JCVT/=0 branch (sdefo3.F:114-138,164-194) ----
  T DXX, DYY, DZZ, D4, D5, D6;
  {
    T VX17 = vx[0] - vx[6], VX28 = vx[1] - vx[7], VX35 = vx[2] - vx[4], VX46 = vx[3] - vx[5];
    T VY17 = vy[0] - vy[6], VY28 = vy[1] - vy[7], VY35 = vy[2] - vy[4], VY46 = vy[3] - vy[5];
    T VZ17 = vz[0] - vz[6], VZ28 = vz[1] - vz[7], VZ35 = vz[2] - vz[4], VZ46 = vz[3] - vz[5];
    DXX = PX[0] * VX17 + PX[1] * VX28 + PX[2] * VX35 + PX[3] * VX46;
    DYY = PY[0] * VY17 + PY[1] * VY28 + PY[2] * VY35 + PY[3] * VY46;
    DZZ = PZ[0] * VZ17 + PZ[1] * VZ28 + PZ[2] * VZ35 + PZ[3] * VZ46;
    T DXY = PY[0] * VX17 + PY[1] * VX28 + PY[2] * VX35 + PY[3] * VX46;
    T DXZ = PZ[0] * VX17 + PZ[1] * VX28 + PZ[2] * VX35 + PZ[3] * VX46;
    T DYX = PX[0] * VY17 + PX[1] * VY28 + PX[2] * VY35 + PX[3] * VY46;
    T DYZ = PZ[0] * VY17 + PZ[1] * VY28 + PZ[2] * VY35 + PZ[3] * VY46;
    T DZX = PX[0] * VZ17 + PX[1] * VZ28 + PX[2] * VZ35 + PX[3] * VZ46;
    T DZY = PY[0] * VZ17 + PY[1] * VZ28 + PY[2] * VZ35 + PY[3] * VZ46;
    T DT1D2 = HALF * dt1;
    T EXX = DXX, EYY = DYY, EZZ = DZZ, EXY = DXY, EYX = DYX;
    T EXZ = DXZ, EZX = DZX, EYZ = DYZ, EZY = DZY, AAA;
    DXX = DXX - DT1D2 * (EXX * EXX + EYX * EYX + EZX * EZX);
    DYY = DYY - DT1D2 * (EYY * EYY + EZY * EZY + EXY * EXY);
    DZZ = DZZ - DT1D2 * (EZZ * EZZ + EXZ * EXZ + EYZ * EYZ);
    AAA = DT1D2 * (EXX * EXY + EYX * EYY + EZX * EZY);
    DXY = DXY - AAA; DYX = DYX - AAA; D4 = DXY + DYX;
    AAA = DT1D2 * (EYY * EYZ + EZY * EZZ + EXY * EXZ);
    DYZ = DYZ - AAA; DZY = DZY - AAA; D5 = DYZ + DZY;
    AAA = DT1D2 * (EZZ * EZX + EXZ * EXX + EYZ * EYX);
    DXZ = DXZ - AAA; DZX = DZX - AAA; D6 = DXZ + DZX;
  }

  // ---- SRHO3 (Lagrangian), MMAIN prologue (mmain.F90:678-737) ----
  const T rho0 = T(mat.rho0);
  const T vol0 = T(vol0d);
  T dvol, rho, amu;
  if constexpr (!FP32) {
    if (offg == ZERO && voln == ONE) voln = vol0;
    dvol = voln - (rho0 / st_rho) * vol0;
    rho = rho0 * (vol0 / voln);
    amu = rho / rho0 - ONE;
  } else {
    double vdp = (off == ZERO || mstop) ? 1.0 : voldp;   // SCHKJABT3 set DET = 1
    if (offg == ZERO) { vdp = vol0d; voln = vol0; }
    dvol = T(vdp - double(M<T>::div(rho0, st_rho)) * vol0d);
    rho = T(mat.rho0 * (vol0d / vdp));
    amu = T(vol0d / vdp - 1.0);
  }
  T eint = st_eint * vol0;
  T vol_avg = voln - HALF * dvol;

  // ---- M2LAW (m2law.F:162-353) ----
  T sold[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) sold[i] = sig[i];
  const T G = T(mat.shear) * off;
  const T bulk = T(mat.bulk);
  T SIGMX = T(mat.sigmx);
  T P = -THIRD * (sig[0] + sig[1] + sig[2]);
  T DAV = -THIRD * (DXX + DYY + DZZ);
  T G1 = dt1 * G, G2 = TWO * G1;
  T SSP = M<T>::sqrt_(M<T>::div(ONEP333 * G + bulk, rho0));
  sig[0] = sig[0] + P + G2 * (DXX + DAV);
  sig[1] = sig[1] + P + G2 * (DYY + DAV);
  sig[2] = sig[2] + P + G2 * (DZZ + DAV);
  sig[3] = sig[3] + G1 * D4;
  sig[4] = sig[4] + G1 * D5;
  sig[5] = sig[5] + G1 * D6;
  T AJ2 = HALF * (sig[0] * sig[0] + sig[1] * sig[1] + sig[2] * sig[2]) + sig[3] * sig[3] +
          sig[4] * sig[4] + sig[5] * sig[5];
  AJ2 = M<T>::sqrt_(THREE * AJ2);
  // MSTRAIN_RATE (mstrain_rate.F), IDEV = VP-2
  T EPSDOT;
  if (mat.vp == 2) {
    T E4 = HALF * D4, E5 = HALF * D5, E6 = HALF * D6;
    EPSDOT = M<T>::sqrt_(DXX * DXX + DYY * DYY + DZZ * DZZ + TWO * (E4 * E4 + E5 * E5 + E6 * E6));
  } else {
    T dav = (DXX + DYY + DZZ) * THIRD;
    T E1 = DXX - dav, E2 = DYY - dav, E3 = DZZ - dav;
    T E4 = HALF * D4, E5 = HALF * D5, E6 = HALF * D6;
    T EPSP = HALF * (E1 * E1 + E2 * E2 + E3 * E3) + E4 * E4 + E5 * E5 + E6 * E6;
    EPSDOT = M<T>::div(M<T>::sqrt_(THREE * EPSP), T(1.5));
  }
  T epsd = EPSDOT;
  if (mat.israte > 0) {
    T asrate = tmin(ONE, T(1.0e20) * dt1);
    epsd = asrate * EPSDOT + (ONE - asrate) * st_epsd;
  }
  // Temperature stays at TREF on this path (TSTAR = 0), so 1 - TSTAR**m = 1.
  T EPD = ONE;
  if (mat.cc != 0.0) {
    EPD = (mat.vp == 1) ? tmax(epsd, T(mat.epdr)) : tmax(epsd, EM15);
    EPD = M<T>::log_(M<T>::div(EPD, T(mat.epdr)));
    EPD = tmax(ZERO, EPD);
    EPD = (ONE + T(mat.cc) * EPD) * ONE;
    if (mat.icc == 1) SIGMX = T(mat.sigmx) * EPD;
  } else {
    if (mat.icc == 1) SIGMX = T(mat.sigmx) * EPD;
  }
  T EPXE = st_pla, AK, QH;
  const T CA = jc[0], CB = jc[1], CN = jc[2];
  if (CN == ONE) {
    AK = CA + CB * EPXE;
    QH = CB * EPD;
  } else if (EPXE > ZERO) {
    AK = CA + CB * M<T>::pow_(EPXE, CN);
    if (CN > ONE) QH = (CB * CN * M<T>::pow_(EPXE, CN - ONE)) * EPD;
    else          QH = M<T>::div(CB * CN, M<T>::pow_(EPXE, ONE - CN)) * EPD;
  } else {
    AK = CA; QH = ZERO;
  }
  AK = AK * EPD;
  if (SIGMX < AK) { AK = SIGMX; QH = ZERO; }
  if (EPXE > T(mat.epmx)) { AK = ZERO; QH = ZERO; }
  T SCALE = tmin(ONE, M<T>::div(AK, tmax(AJ2, EM15)));
  T DPLA = M<T>::div((ONE - SCALE) * AJ2, tmax(THREE * G + QH, EM15));
  AK = AK + ONE * DPLA * QH;
  SCALE = tmin(ONE, M<T>::div(AK, tmax(AJ2, EM15)));
#pragma unroll
  for (int i = 0; i < 6; ++i) sig[i] = SCALE * sig[i];
  EPXE = EPXE + DPLA;

  // ---- MQVISCB (mqviscb.F:139-263,382-416); its VOL is the current volume ----
  T DD = -DXX - DYY - DZZ;
  T AD = ZERO, AL = ZERO;
  T CX = SSP;
  if (off == ONE) {
    if (voln > ZERO) AL = M<T>::cbrt_pow(voln);
    AD = tmax(ZERO, DD);
  }
  const T QA = T(prop.qa), QB = T(prop.qb);
  T QAA = QA * QA * AD;
  T QX = QB * SSP + AL * QAA;
  T qnew = rho * AD * AL * (QAA * AL + QB * SSP);
  T SSP_EQ = tmax(EM20, QX + M<T>::sqrt_(QX * QX + CX * CX));
  T DTX = M<T>::div(deltax, SSP_EQ);
  T STI = ZERO;
  if (!(off == ZERO || offg < ZERO)) {
    T TIDT = M<T>::div(ONE, DTX);
    T TRHO = rho * TIDT;
    T TVOL = voln * TIDT;
    STI = TRHO * TVOL;
  }
  DTX = T(prop.dtfac) * DTX;
  out.dtx = (voln > ZERO && off != ZERO && offg >= ZERO) ? DTX : T(1.0e30);

  // ---- M2LAW epilogue (m2law.F:390-426) ----
  T DTA = HALF * dt1;
  T PNEW = bulk * amu;
  sig[0] = (sig[0] - PNEW) * off;
  sig[1] = (sig[1] - PNEW) * off;
  sig[2] = (sig[2] - PNEW) * off;
  sig[3] = sig[3] * off; sig[4] = sig[4] * off; sig[5] = sig[5] * off;
  {
    T E1 = DXX * (sold[0] + sig[0]);
    T E2 = DYY * (sold[1] + sig[1]);
    T E3 = DZZ * (sold[2] + sig[2]);
    T E4 = D4 * (sold[3] + sig[3]);
    T E5 = D5 * (sold[4] + sig[4]);
    T E6 = D6 * (sold[5] + sig[5]);
    T EINC = vol_avg * (E1 + E2 + E3 + E4 + E5 + E6) * DTA - HALF * dvol * (st_qvis + qnew);
    eint = M<T>::div(eint + EINC * off, tmax(EM15, vol0));
  }

  // ---- SHVIS3 (shvis3.F:163-238,317-428) ----
  T CAQ = T(0.25) * off * T(prop.hvis);
  T FCL = CAQ * rho * M<T>::two_third_pow(voln);
  T FCQ = FCL * CAQ * T(100);
  FCL = FCL * SSP;
  // Hourglass rates HG and viscous hourglass "forces" HH per direction and mode;
  // the corner forces are formed corner by corner below.
  T HG[3][4], HH[3][4], Q[3][4];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const VA w = (d == 0) ? vx : (d == 1) ? vy : vz;
    if constexpr (!FP32) {
      // FP64 form: the reference's G = Gamma -/+ PxHy vectors and operation order.
#pragma unroll
      for (int m = 0; m < 3; ++m)
        HG[d][m] = G_of<T>(m, 0, PH) * w[0] + G_of<T>(m, 1, PH) * w[1] + G_of<T>(m, 2, PH) * w[2] +
                   G_of<T>(m, 3, PH) * w[3] + G_of<T>(m, 4, PH) * w[4] + G_of<T>(m, 5, PH) * w[5] +
                   G_of<T>(m, 6, PH) * w[6] + G_of<T>(m, 7, PH) * w[7];
    } else {
      // FP32 form, factored: G_m = Gamma_m + c_k PH[m][a(k)], a = (0,1,2,3,2,3,0,1),
      // c = (-,-,-,-,+,+,+,+), so Gamma_m.v is adds and PH acts on v1-v7, v2-v8, v3-v5, v4-v6.
      const T p0 = w[0] - w[6], p1 = w[1] - w[7], p2 = w[2] - w[4], p3 = w[3] - w[5];
      const T g[3] = {w[0] - w[1] + w[2] - w[3] + w[4] - w[5] + w[6] - w[7],
                      w[0] + w[1] - w[2] - w[3] - w[4] - w[5] + w[6] + w[7],
                      w[0] - w[1] - w[2] + w[3] - w[4] + w[5] + w[6] - w[7]};
#pragma unroll
      for (int m = 0; m < 3; ++m)
        HG[d][m] = g[m] - PH[m][0] * p0 - PH[m][1] * p1 - PH[m][2] * p2 - PH[m][3] * p3;
    }
    HG[d][3] = w[0] - w[1] + w[2] - w[3] - w[4] + w[5] - w[6] + w[7];
#pragma unroll
    for (int m = 0; m < 4; ++m) HH[d][m] = HG[d][m] * (FCL + ((HG[d][m] < ZERO) ? -HG[d][m] : HG[d][m]) * FCQ);
    if constexpr (FP32) {
#pragma unroll
      for (int a = 0; a < 4; ++a) Q[d][a] = PH[0][a] * HH[d][0] + PH[1][a] * HH[d][1] + PH[2][a] * HH[d][2];
    }
  }
  out.ehou = dt1 * (HH[2][0] * HG[2][0] + HH[2][1] * HG[2][1] + HH[2][2] * HG[2][2] + HH[2][3] * HG[2][3] +
                    HH[0][0] * HG[0][0] + HH[0][1] * HG[0][1] + HH[0][2] * HG[0][2] + HH[0][3] * HG[0][3] +
                    HH[1][0] * HG[1][0] + HH[1][1] * HG[1][1] + HH[1][2] * HG[1][2] + HH[1][3] * HG[1][3]);

  // SFINT3 (sfint3.F:257-322), no viscous stress for law 2: the pair term FINT(d,a)
  // is subtracted at corner a and added at its opposite corner (1-7, 2-8, 3-5, 4-6).
  T FI[3][4];
  {
    const T S1 = (sig[0] - qnew) * voln, S2 = (sig[1] - qnew) * voln, S3 = (sig[2] - qnew) * voln;
    const T S4 = sig[3] * voln, S5 = sig[4] * voln, S6 = sig[5] * voln;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      FI[0][a] = S1 * PX[a] + S4 * PY[a] + S6 * PZ[a];
      FI[1][a] = S2 * PY[a] + S4 * PX[a] + S5 * PZ[a];
      FI[2][a] = S3 * PZ[a] + S6 * PX[a] + S5 * PY[a];
    }
  }
  // Corner by corner: SHVIS3 force, +/- SFINT3 term, SRROTA3 back to the global
  // frame (sforc3.F:1604-1614), SCUMU3P's zeroing of deleted elements, emit.
  const bool zero_f = (offg < ZERO);
  const int s4[8] = {-1, 1, -1, 1, 1, -1, 1, -1};
  const int ak[8] = {0, 1, 2, 3, 2, 3, 0, 1};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    T Fl[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T h0 = HH[d][0], h1 = HH[d][1], h2 = HH[d][2];
      const T h4 = (s4[k] > 0) ? HH[d][3] : -HH[d][3];
      T hg;
      if constexpr (!FP32) {
        hg = -G_of<T>(0, k, PH) * h0 - G_of<T>(1, k, PH) * h1 - G_of<T>(2, k, PH) * h2 + h4;
      } else {
        const T gh = gamma_of(0, k) * h0 + gamma_of(1, k) * h1 + gamma_of(2, k) * h2;
        hg = -gh + ((k < 4) ? Q[d][ak[k]] : -Q[d][ak[k]]) + h4;
      }
      Fl[d] = (k < 4) ? hg - FI[d][ak[k]] : hg + FI[d][ak[k]];
    }
    const T X = R[0][0] * Fl[0] + R[0][1] * Fl[1] + R[0][2] * Fl[2];
    const T Y = R[1][0] * Fl[0] + R[1][1] * Fl[1] + R[1][2] * Fl[2];
    const T Z = R[2][0] * Fl[0] + R[2][1] * Fl[1] + R[2][2] * Fl[2];
    emit(k, zero_f ? ZERO : X, zero_f ? ZERO : Y, zero_f ? ZERO : Z);
  }
  out.sti = T(0.25) * STI;
  out.mstop = mstop;

  st_eint = eint; st_rho = rho; st_pla = EPXE; st_qvis = qnew;
  st_offg = offg; st_epsd = epsd;
}

}  // namespace orad

// orad_qbat_chain.hpp -- one QBAT shell (Ishell=12, IHBE=11) through the explicit
// force chain of CBAFORC3 (engine/source/elements/shell/coqueba/cbaforc3.F:440-1140):
// 2 x 2 Gauss points, IGTYP=1, Ismstr=2, ISROT=0, no pinching, law 2 at NPT >= 2
// layers, PARITH/ON slots.
//
// Stages and the routines they follow:
//   CBACOOR (cbacoor.F:37-1112) with CLSKEW3 IREP=0 (cdkcoor3.F:303-380): local
//   frame, corner coordinates, relative velocities with the mid-step rotation
//   correction; flat elements (ZL1^2 < 1e-12 LL) take the constant-Jacobian branch
//   (Gauss-point Jacobians and hourglass vectors HX, HY), warped ones build the
//   nodal frames VQN, the Gauss-point frames VQG with their tangent inverses VJFI,
//   the mid-edge normals VNRM and tying terms VASTN of the assumed transverse shear,
//   and the rigid-body projection DI of the 6-DOF velocities; the characteristic
//   length factors.  CNCOEF3 (sh3n/coquedk/cncoef3.F:295-440) moduli and DN;
//   CBADEFSH (cbadef.F:722-767) constant membrane shear; per Gauss point CBADEF
//   (cbadef.F:28-716) strain rates, CBASTRA3 increments, CBAENER (cbaener.F:24-50),
//   the Gauss point's strain rate (cbaforc3.F:759-767), MULAWC's law-2 layer loop
//   for NPG > 1 (mulawc.F90:530-560,2968-3020), CBAVISC (cbavisc.F:24-95) and CBAFORI
//   (cbafori.F:24-300); then FOR/MOM as the Gauss-point mean (cbaforc3.F:960-980),
//   CBAFORCT (cbafori.F:305-350), CBAENERS, CBAPROJ (cbaproj.F:24-330), CNDT3
//   (sh3n/coquedk/cndt3.F:84-315) and CUPDTN3P (coque/cupdtn3.F:488-700).
//
// The B-matrices of a Gauss point (BM, BMF, BF, BC: 136 values) are not kept: the
// strain pass and the force pass each recompute a corner's entries from the
// element's frames when they need them, which gives the same values operation for
// operation and keeps them out of the register file across the law.
//
// T = double follows the reference's double-precision operation order (its file
// is compiled without FMA contraction).  T = float is the mixed form: the frame,
// projection and strain-rate stage runs in TG on coordinates and translational
// velocities taken relative to corner 1 in FP64 (every quantity of that stage is
// invariant to the shift), the law, resultants and forces in FP32.
#pragma once
#include "orad_shell_chain.hpp"

namespace orad {

template <typename T>
struct QbatOut {
  T sti13, sti24;  // FSKY(7) of corners 1,3 (STI*FACN(1)) and 2,4 (STI*FACN(2))
  T dt;            // element time step candidate
  T ehou;          // CBAVISC's damping energy increment (PARTSAV(8))
  int alive;       // the candidate enters DT2T
};

// Gauss-abscissa literal of CBACOOR's DATA and CBADEF's PARAMETER: a default-REAL
// constant, held widened from single precision in the double build.
__device__ __forceinline__ double qbat_pg() { return double(0.577350269189626f); }

// CBACOOR's element data that the Gauss points and the projection read.
template <typename TG>
struct QbatGeo {
  TG e1[3], e2[3], e3[3];
  TG AREA, X13, X24, Y13, Y24;   // X13 .. Y24 scaled by 1/AREA (CBACOOR X13_T ..)
  TG LL, FACN[2];
  TG VXYZ[3][4], RXYZ[2][4];     // flat: V13, V24, VHI / R13, R24, RHI, RTI
  TG JAC[4];
  TG VSH;                        // CBADEFSH: constant membrane shear rate VDEF(3)
  bool flat;
  bool degenerate;               // CLSKEW3 found a zero-area frame and set OFFG = 0
  // flat branch
  TG VCF[12], HX[4], HY[4];
  // warped branch
  TG VQN[9][4], VQG[6][4], VJFI[6][4], VNRM[3][4], VASTN[4][4], DI[6], VC[3][4];
};

template <typename TG>
__device__ __forceinline__ void qbat_shape_tables(TG VKSI[4][4], TG VETA[4][4]) {
  const TG ONE = TG(1), FOURTH = TG(0.25), PG = TG(qbat_pg());
  VKSI[0][0] = -FOURTH * (ONE + PG);
  VKSI[1][0] = -VKSI[0][0];
  VKSI[2][0] = FOURTH * (ONE - PG);
  VKSI[3][0] = -VKSI[2][0];
  VETA[0][0] = -FOURTH * (ONE + PG);
  VETA[1][0] = -FOURTH * (ONE - PG);
  VETA[2][0] = -VETA[1][0];
  VETA[3][0] = -VETA[0][0];
  VKSI[0][1] = VKSI[0][0];
  VKSI[1][1] = -VKSI[0][1];
  VKSI[2][1] = VKSI[2][0];
  VKSI[3][1] = -VKSI[2][1];
  VETA[0][1] = VETA[1][0];
  VETA[1][1] = VETA[0][0];
  VETA[2][1] = -VETA[1][1];
  VETA[3][1] = -VETA[0][1];
  VKSI[0][2] = -VKSI[2][0];
  VKSI[1][2] = -VKSI[0][2];
  VKSI[2][2] = -VKSI[0][0];
  VKSI[3][2] = -VKSI[2][2];
  VETA[0][2] = VETA[0][1];
  VETA[1][2] = VETA[1][1];
  VETA[2][2] = -VETA[1][2];
  VETA[3][2] = -VETA[0][2];
  VKSI[0][3] = VKSI[0][2];
  VKSI[1][3] = -VKSI[0][3];
  VKSI[2][3] = VKSI[2][2];
  VKSI[3][3] = -VKSI[2][3];
  VETA[0][3] = VETA[0][0];
  VETA[1][3] = VETA[1][0];
  VETA[2][3] = -VETA[1][3];
  VETA[3][3] = -VETA[0][3];
}

// n-th frame's third column as the cross product of its first two (VQN/VQG rows 4..6).
template <typename TG, int R>
__device__ __forceinline__ void qbat_cross(TG Q[R][4], const TG n7, const TG n8, const TG n9,
                                           int n) {
  Q[3][n] = n8 * Q[2][n] - n9 * Q[1][n];
  Q[4][n] = n9 * Q[0][n] - n7 * Q[2][n];
  Q[5][n] = n7 * Q[1][n] - n8 * Q[0][n];
}

// ---- CBACOOR + CBADEFSH ----
template <typename TG>
__device__ __forceinline__ void qbat_geometry(const TG xg[3][4], const TG vg[3][4],
                                              const TG vrg[3][4], TG dt1, TG offg,
                                              const ShConst &k, QbatGeo<TG> &o) {
  const TG ZERO = TG(0), ONE = TG(1), TWO = TG(2), FOUR = TG(4), HALF = TG(0.5),
           FOURTH = TG(0.25);
  const TG EM20 = TG(k.em20), EM10 = TG(k.em10);
  const TG PG = TG(qbat_pg());
  const TG RX = xg[0][1] + xg[0][2] - xg[0][0] - xg[0][3];
  const TG SX = xg[0][2] + xg[0][3] - xg[0][0] - xg[0][1];
  const TG RY = xg[1][1] + xg[1][2] - xg[1][0] - xg[1][3];
  const TG SY = xg[1][2] + xg[1][3] - xg[1][0] - xg[1][1];
  const TG RZ = xg[2][1] + xg[2][2] - xg[2][0] - xg[2][3];
  const TG SZ = xg[2][2] + xg[2][3] - xg[2][0] - xg[2][1];
  TG E3X = RY * SZ - RZ * SY, E3Y = RZ * SX - RX * SZ, E3Z = RX * SY - RY * SX;
  TG DET = M<TG>::sqrt_(E3X * E3X + E3Y * E3Y + E3Z * E3Z);
  o.degenerate = (DET < EM20 && offg != ZERO);
  if (o.degenerate) offg = ZERO;
  {
    const TG off_loc = (tabs(offg) != ZERO) ? ONE : ZERO;
    DET = tmax(EM20, DET);
    TG CC = M<TG>::div(off_loc, DET);
    CC = tmax(CC, EM20);
    E3X = E3X * CC; E3Y = E3Y * CC; E3Z = E3Z * CC;
  }
  TG E1X, E1Y, E1Z;
  {
    const TG C1C1 = RX * RX + RY * RY + RZ * RZ;
    const TG C2C2 = SX * SX + SY * SY + SZ * SZ;
    TG c2_1 = ZERO, c1_1 = ZERO;
    if (C1C1 != ZERO) {
      c2_1 = M<TG>::sqrt_(M<TG>::div(C2C2, tmax(EM20, C1C1)));
      c1_1 = ONE;
    } else if (C2C2 != ZERO) {
      c2_1 = ONE;
      c1_1 = M<TG>::sqrt_(M<TG>::div(C1C1, tmax(EM20, C2C2)));
    }
    E1X = RX * c2_1 + (SY * E3Z - SZ * E3Y) * c1_1;
    E1Y = RY * c2_1 + (SZ * E3X - SX * E3Z) * c1_1;
    E1Z = RZ * c2_1 + (SX * E3Y - SY * E3X) * c1_1;
    TG C1 = M<TG>::sqrt_(E1X * E1X + E1Y * E1Y + E1Z * E1Z);
    if (C1 != ZERO) C1 = M<TG>::div(ONE, tmax(EM20, C1));
    E1X = E1X * C1; E1Y = E1Y * C1; E1Z = E1Z * C1;
  }
  const TG E2X = E3Y * E1Z - E3Z * E1Y, E2Y = E3Z * E1X - E3X * E1Z, E2Z = E3X * E1Y - E3Y * E1X;
  o.e1[0] = E1X; o.e1[1] = E1Y; o.e1[2] = E1Z;
  o.e2[0] = E2X; o.e2[1] = E2Y; o.e2[2] = E2Z;
  o.e3[0] = E3X; o.e3[1] = E3Y; o.e3[2] = E3Z;
  TG AREA = FOURTH * DET;
  const TG AREA_I = M<TG>::div(ONE, AREA);
  TG XL2, YL2, XL3, YL3, XL4, YL4, ZL1;
  {
    TG L0[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) L0[d] = FOURTH * (xg[d][2] + xg[d][3] + xg[d][0] + xg[d][1]);
    TG XX = xg[0][1] - xg[0][0], YY = xg[1][1] - xg[1][0], ZZ = xg[2][1] - xg[2][0];
    XL2 = E1X * XX + E1Y * YY + E1Z * ZZ;
    YL2 = E2X * XX + E2Y * YY + E2Z * ZZ;
    XX = xg[0][0] - L0[0]; YY = xg[1][0] - L0[1]; ZZ = xg[2][0] - L0[2];
    ZL1 = E3X * XX + E3Y * YY + E3Z * ZZ;
    XX = xg[0][2] - xg[0][0]; YY = xg[1][2] - xg[1][0]; ZZ = xg[2][2] - xg[2][0];
    XL3 = E1X * XX + E1Y * YY + E1Z * ZZ;
    YL3 = E2X * XX + E2Y * YY + E2Z * ZZ;
    XX = xg[0][3] - xg[0][0]; YY = xg[1][3] - xg[1][0]; ZZ = xg[2][3] - xg[2][0];
    XL4 = E1X * XX + E1Y * YY + E1Z * ZZ;
    YL4 = E2X * XX + E2Y * YY + E2Z * ZZ;
  }
  // corner coordinates about the centre (VCORE), diagonals, characteristic lengths
  TG VC[2][4];
  TG X13, X24, Y13, Y24, L13, L24, LL, LM;
  {
    const TG LX = FOURTH * (XL2 + XL3 + XL4), LY = FOURTH * (YL2 + YL3 + YL4);
    VC[0][0] = -LX; VC[0][1] = XL2 - LX; VC[0][2] = XL3 - LX; VC[0][3] = XL4 - LX;
    VC[1][0] = -LY; VC[1][1] = YL2 - LY; VC[1][2] = YL3 - LY; VC[1][3] = YL4 - LY;
    X13 = (VC[0][0] - VC[0][2]) * HALF;
    X24 = (VC[0][1] - VC[0][3]) * HALF;
    Y13 = (VC[1][0] - VC[1][2]) * HALF;
    Y24 = (VC[1][1] - VC[1][3]) * HALF;
    L13 = X13 * X13 + Y13 * Y13;
    L24 = X24 * X24 + Y24 * Y24;
    LL = tmax(L13, L24);
    const TG C1 = VC[0][1] * VC[1][3] - VC[1][1] * VC[0][3];
    const TG C2 = VC[0][0] * VC[1][2] - VC[1][0] * VC[0][2];
    LM = tmax(tabs(C1), tabs(C2));
  }
  // relative velocities in the element frame with the mid-step rotation correction
  TG V13[3], V24[3], VHI[3];
  {
    TG G13[3], G24[3], GHI[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      G13[d] = vg[d][0] - vg[d][2];
      G24[d] = vg[d][1] - vg[d][3];
      GHI[d] = vg[d][0] - vg[d][1] + vg[d][2] - vg[d][3];
    }
    const TG *E[3] = {o.e1, o.e2, o.e3};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      V13[q] = E[q][0] * G13[0] + E[q][1] * G13[1] + E[q][2] * G13[2];
      V24[q] = E[q][0] * G24[0] + E[q][1] * G24[1] + E[q][2] * G24[2];
      VHI[q] = E[q][0] * GHI[0] + E[q][1] * GHI[1] + E[q][2] * GHI[2];
    }
    const TG DT05 = HALF * dt1, DT025 = FOURTH * dt1;
    const TG EXZ = Y24 * V13[2] - Y13 * V24[2];
    const TG EYZ = -X24 * V13[2] + X13 * V24[2];
    const TG DDRY = DT05 * EXZ * AREA_I;
    const TG DDRX = DT05 * EYZ * AREA_I;
    const TG V13X = V13[0], V24X = V24[0], VHIX = VHI[0];
    TG DDRZ1 = M<TG>::div(DT025 * (V13[1] - V24[1]), X13 - X24);
    if (tabs(X13 - X24) < EM10) DDRZ1 = ZERO;
    V13[0] = V13[0] - DDRY * V13[2] - DDRZ1 * V13[1];
    V24[0] = V24[0] - DDRY * V24[2] - DDRZ1 * V24[1];
    VHI[0] = VHI[0] - DDRY * VHI[2] - DDRZ1 * VHI[1];
    TG DDRZ2 = M<TG>::div(DT025 * (V13X + V24X), Y13 + Y24);
    if (tabs(Y13 + Y24) < EM10) DDRZ2 = ZERO;
    V13[1] = V13[1] - DDRX * V13[2] - DDRZ2 * V13X;
    V24[1] = V24[1] - DDRX * V24[2] - DDRZ2 * V24X;
    VHI[1] = VHI[1] - DDRX * VHI[2] - DDRZ2 * VHIX;
  }
  TG RR[3][4];
  {
    const TG *E[3] = {o.e1, o.e2, o.e3};
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int c = 0; c < 4; ++c)
        RR[q][c] = E[q][0] * vrg[0][c] + E[q][1] * vrg[1][c] + E[q][2] * vrg[2][c];
  }
  o.flat = (ZL1 * ZL1 < TG(k.em12) * LL);
  const TG MX13 = (VC[0][0] + VC[0][2]) * HALF, MY13 = (VC[1][0] + VC[1][2]) * HALF;
  const TG MX23 = (VC[0][1] + VC[0][2]) * HALF, MY23 = (VC[1][1] + VC[1][2]) * HALF;
  const TG MX34 = (VC[0][2] + VC[0][3]) * HALF, MY34 = (VC[1][2] + VC[1][3]) * HALF;
  o.X13 = X13 * AREA_I; o.X24 = X24 * AREA_I; o.Y13 = Y13 * AREA_I; o.Y24 = Y24 * AREA_I;
  const TG GAMA1 = -MX13 * Y24 + MY13 * X24;
  const TG GAMA2 = MX13 * Y13 - MY13 * X13;
  if (o.flat) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      o.VXYZ[d][0] = V13[d];
      o.VXYZ[d][1] = V24[d];
      o.VXYZ[d][2] = VHI[d];
      o.VXYZ[d][3] = ZERO;
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      o.RXYZ[q][0] = RR[q][0] - RR[q][2];
      o.RXYZ[q][1] = RR[q][1] - RR[q][3];
      o.RXYZ[q][2] = RR[q][0] - RR[q][1] + RR[q][2] - RR[q][3];
      o.RXYZ[q][3] = RR[q][0] + RR[q][1] + RR[q][2] + RR[q][3];
    }
    o.VCF[0] = o.Y24; o.VCF[1] = -o.Y13; o.VCF[2] = -o.X24; o.VCF[3] = o.X13;
    o.VCF[4] = GAMA1 * AREA_I; o.VCF[5] = GAMA2 * AREA_I;
    o.VCF[6] = MX23; o.VCF[7] = MY23; o.VCF[8] = MX34; o.VCF[9] = MY34;
    o.VCF[10] = MX13; o.VCF[11] = MY13;
    TG J1 = (MX23 * MY13 - MX13 * MY23) * PG;
    TG J2 = -(MX13 * MY34 - MX34 * MY13) * PG;
    const TG J0 = AREA * FOURTH;
    o.JAC[0] = tabs(J0 + J2 - J1);
    o.JAC[1] = tabs(J0 + J2 + J1);
    o.JAC[2] = tabs(J0 - J2 + J1);
    o.JAC[3] = tabs(J0 - J2 - J1);
    J1 = (MY23 - MY34) * PG;
    J2 = -(MY23 + MY34) * PG;
    o.HX[0] = M<TG>::div(J1, o.JAC[0]);
    o.HX[1] = M<TG>::div(J2, o.JAC[1]);
    o.HX[2] = M<TG>::div(-J1, o.JAC[2]);
    o.HX[3] = M<TG>::div(-J2, o.JAC[3]);
    J1 = (MX34 - MX23) * PG;
    J2 = (MX34 + MX23) * PG;
    o.HY[0] = M<TG>::div(J1, o.JAC[0]);
    o.HY[1] = M<TG>::div(J2, o.JAC[1]);
    o.HY[2] = M<TG>::div(-J1, o.JAC[2]);
    o.HY[3] = M<TG>::div(-J2, o.JAC[3]);
  } else {
    const TG Z1 = ZL1, Z2 = Z1 * Z1;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      o.VC[0][c] = VC[0][c];
      o.VC[1][c] = VC[1][c];
    }
    o.VC[2][0] = Z1; o.VC[2][1] = -Z1; o.VC[2][2] = Z1; o.VC[2][3] = -Z1;
    // edges of [F]
    const TG X21 = MX23 - MX13, X34 = (VC[0][2] - VC[0][3]) * HALF;
    const TG Y21 = MY23 - MY13, Y34 = (VC[1][2] - VC[1][3]) * HALF;
    const TG Z21 = -Z1, Z34 = Z1;
    const TG L12 = M<TG>::sqrt_(X21 * X21 + Y21 * Y21 + Z2);
    const TG L34 = M<TG>::sqrt_(X34 * X34 + Y34 * Y34 + Z2);
    const TG X41 = MX34 - MX13, X32 = (VC[0][2] - VC[0][1]) * HALF;
    const TG Y41 = MY34 - MY13, Y32 = (VC[1][2] - VC[1][1]) * HALF;
    const TG Z41 = -Z1, Z32 = Z1;
    // nodal frames [QN]
    TG A_4 = AREA * FOURTH;
    TG (&Q)[9][4] = o.VQN;
    TG SL = M<TG>::div(ONE, tmax(L12, EM20));
    Q[0][0] = X21 * SL; Q[1][0] = Y21 * SL; Q[2][0] = Z21 * SL;
    TG SZ2 = A_4 - GAMA1;
    TG SZv = Z2 * L24;
    SL = M<TG>::div(ONE, M<TG>::sqrt_(SZv + SZ2 * SZ2));
    Q[6][0] = -Z1 * Y24; Q[7][0] = Z1 * X24; Q[8][0] = SZ2 * SL;
    Q[6][2] = -Q[6][0]; Q[7][2] = -Q[7][0];
    Q[6][0] = Q[6][0] * SL; Q[7][0] = Q[7][0] * SL;
    qbat_cross<TG, 9>(Q, Q[6][0], Q[7][0], Q[8][0], 0);
    SL = M<TG>::div(ONE, tmax(L34, EM20));
    Q[0][2] = X34 * SL; Q[1][2] = Y34 * SL; Q[2][2] = Z34 * SL;
    SZ2 = A_4 + GAMA1;
    SL = M<TG>::div(ONE, M<TG>::sqrt_(SZv + SZ2 * SZ2));
    Q[6][2] = Q[6][2] * SL; Q[7][2] = Q[7][2] * SL; Q[8][2] = SZ2 * SL;
    qbat_cross<TG, 9>(Q, Q[6][2], Q[7][2], Q[8][2], 2);
    Q[0][1] = Q[0][0]; Q[1][1] = Q[1][0]; Q[2][1] = Q[2][0];
    SZ2 = A_4 + GAMA2;
    SZv = Z2 * L13;
    SL = M<TG>::div(ONE, M<TG>::sqrt_(SZv + SZ2 * SZ2));
    Q[6][1] = -Z1 * Y13; Q[7][1] = Z1 * X13; Q[8][1] = SZ2 * SL;
    Q[6][3] = -Q[6][1]; Q[7][3] = -Q[7][1];
    Q[6][1] = Q[6][1] * SL; Q[7][1] = Q[7][1] * SL;
    qbat_cross<TG, 9>(Q, Q[6][1], Q[7][1], Q[8][1], 1);
    Q[0][3] = Q[0][2]; Q[1][3] = Q[1][2]; Q[2][3] = Q[2][2];
    SZ2 = A_4 - GAMA2;
    SL = M<TG>::div(ONE, M<TG>::sqrt_(SZv + SZ2 * SZ2));
    Q[6][3] = Q[6][3] * SL; Q[7][3] = Q[7][3] * SL; Q[8][3] = SZ2 * SL;
    qbat_cross<TG, 9>(Q, Q[6][3], Q[7][3], Q[8][3], 3);
    // mid-edge normals and tying terms of the assumed transverse shear
    {
      const int na[4] = {0, 3, 0, 1}, nb[4] = {1, 2, 3, 2};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int d = 0; d < 3; ++d) o.VNRM[d][j] = Q[6 + d][na[j]] + Q[6 + d][nb[j]];
        TG C1 = M<TG>::sqrt_(o.VNRM[0][j] * o.VNRM[0][j] + o.VNRM[1][j] * o.VNRM[1][j] +
                             o.VNRM[2][j] * o.VNRM[2][j]);
        C1 = tmax(EM20, C1);
#pragma unroll
        for (int d = 0; d < 3; ++d) o.VNRM[d][j] = M<TG>::div(o.VNRM[d][j], C1);
      }
    }
    o.VASTN[0][0] = ZERO; o.VASTN[1][0] = L12; o.VASTN[2][0] = ZERO; o.VASTN[3][0] = L12;
    o.VASTN[0][1] = ZERO; o.VASTN[1][1] = L34; o.VASTN[2][1] = ZERO; o.VASTN[3][1] = L34;
    o.VASTN[0][2] = -(X41 * Q[3][0] + Y41 * Q[4][0] + Z41 * Q[5][0]);
    o.VASTN[1][2] = X41 * Q[0][0] + Y41 * Q[1][0] + Z41 * Q[2][0];
    o.VASTN[2][2] = -(X41 * Q[3][3] + Y41 * Q[4][3] + Z41 * Q[5][3]);
    o.VASTN[3][2] = X41 * Q[0][3] + Y41 * Q[1][3] + Z41 * Q[2][3];
    o.VASTN[0][3] = -(X32 * Q[3][1] + Y32 * Q[4][1] + Z32 * Q[5][1]);
    o.VASTN[1][3] = X32 * Q[0][1] + Y32 * Q[1][1] + Z32 * Q[2][1];
    o.VASTN[2][3] = -(X32 * Q[3][2] + Y32 * Q[4][2] + Z32 * Q[5][2]);
    o.VASTN[3][3] = X32 * Q[0][2] + Y32 * Q[1][2] + Z32 * Q[2][2];
    // Gauss-point frames [QG] (rows 1..6) and the tangent inverses VJFI
    A_4 = M<TG>::div(A_4, PG);
    const TG JMX13 = MX13 * PG, JMY13 = MY13 * PG, JMZ13 = Z1 * PG;
    const TG J2MYZ = JMZ13 * JMZ13;
    const TG G1X1 = MX23 - JMX13, G1Y1 = MY23 - JMY13;
    const TG C1g = M<TG>::sqrt_(G1X1 * G1X1 + G1Y1 * G1Y1 + J2MYZ);
    const TG G2X1 = MX34 - JMX13, G2Y1 = MY34 - JMY13;
    const TG C2g = M<TG>::sqrt_(G2X1 * G2X1 + G2Y1 * G2Y1 + J2MYZ);
    const TG G1X3 = MX23 + JMX13, G1Y3 = MY23 + JMY13;
    const TG J1g = M<TG>::sqrt_(G1X3 * G1X3 + G1Y3 * G1Y3 + J2MYZ);
    const TG G2X2 = MX34 + JMX13, G2Y2 = MY34 + JMY13;
    const TG J2g = M<TG>::sqrt_(G2X2 * G2X2 + G2Y2 * G2Y2 + J2MYZ);
    // per Gauss point: (SZ2, L, sign of the z-terms, first tangent G1 with its
    // length, second tangent G2 with its length); NG = 1, 2, 3, 4 in storage order
    TG (&V)[6][4] = o.VJFI;
    TG (&G)[6][4] = o.VQG;
#pragma unroll
    for (int ng = 0; ng < 4; ++ng) {
      const TG sz2 = (ng == 0) ? A_4 - GAMA1 : (ng == 1) ? A_4 + GAMA2
                   : (ng == 2) ? A_4 + GAMA1 : A_4 - GAMA2;
      const TG szv = (ng == 0 || ng == 2) ? Z2 * L24 : Z2 * L13;
      TG sl = M<TG>::sqrt_(szv + sz2 * sz2);
      o.JAC[ng] = sl * PG;
      sl = M<TG>::div(ONE, tmax(sl, EM20));
      // VQG rows 7..9 (the frame normal): NG 1/3 from (-Z1*Y24, Z1*X24), NG 2/4 from
      // (-Z1*Y13, Z1*X13), NG 3 and 4 with the opposite sign
      TG n7, n8;
      if (ng == 0) { n7 = -Z1 * Y24; n8 = Z1 * X24; }
      else if (ng == 2) { n7 = -(-Z1 * Y24); n8 = -(Z1 * X24); }
      else if (ng == 1) { n7 = -Z1 * Y13; n8 = Z1 * X13; }
      else { n7 = -(-Z1 * Y13); n8 = -(Z1 * X13); }
      n7 = n7 * sl;
      n8 = n8 * sl;
      const TG n9 = sz2 * sl;
      const bool minus = (ng == 0 || ng == 3);   // G2 = (G2X1, G2Y1) with +JMZ13 terms
      const TG g2x = (ng == 0 || ng == 3) ? G2X1 : G2X2;
      const TG g2y = (ng == 0 || ng == 3) ? G2Y1 : G2Y2;
      const TG g1x = (ng == 0 || ng == 1) ? G1X1 : G1X3;
      const TG g1y = (ng == 0 || ng == 1) ? G1Y1 : G1Y3;
      const TG l1 = (ng == 0 || ng == 1) ? C1g : J1g;   // length of G1
      const TG l2 = (ng == 0 || ng == 3) ? C2g : J2g;   // length of G2
      TG v1, v2, v3;
      if (minus) {
        v1 = (g2y * n9 + JMZ13 * n8);
        v2 = (-JMZ13 * n7 - g2x * n9);
      } else {
        v1 = (g2y * n9 - JMZ13 * n8);
        v2 = (JMZ13 * n7 - g2x * n9);
      }
      v3 = (g2x * n8 - g2y * n7);
      TG q1 = g1x * l2 + v1 * l1;
      TG q2 = g1y * l2 + v2 * l1;
      const TG gz = (ng == 0 || ng == 1) ? -JMZ13 : JMZ13;
      TG q3 = gz * l2 + v3 * l1;
      sl = M<TG>::div(sl, PG);
      V[0][ng] = v1 * sl;
      V[1][ng] = v2 * sl;
      V[2][ng] = v3 * sl;
      if (ng == 0 || ng == 1) {
        V[3][ng] = (-JMZ13 * n8 - g1y * n9) * sl;
        V[4][ng] = (g1x * n9 + JMZ13 * n7) * sl;
      } else {
        V[3][ng] = (JMZ13 * n8 - g1y * n9) * sl;
        V[4][ng] = (g1x * n9 - JMZ13 * n7) * sl;
      }
      V[5][ng] = (g1y * n7 - g1x * n8) * sl;
      TG s_ = M<TG>::sqrt_(q1 * q1 + q2 * q2 + q3 * q3);
      if (s_ != ZERO) s_ = M<TG>::div(ONE, s_);
      q1 = q1 * s_; q2 = q2 * s_; q3 = q3 * s_;
      G[0][ng] = q1; G[1][ng] = q2; G[2][ng] = q3;
      G[3][ng] = n8 * q3 - n9 * q2;
      G[4][ng] = n9 * q1 - n7 * q3;
      G[5][ng] = n7 * q2 - n8 * q1;
    }
    AREA = o.JAC[0] + o.JAC[1] + o.JAC[2] + o.JAC[3];
    // rigid-body projection of the nodal velocities (6 DOF -> 5 DOF)
    {
      TG AR[3], ALR[3], D[6];
      AR[0] = -Z1 * VHI[1] + Y13 * V13[2] + Y24 * V24[2] + MY13 * VHI[2] + RR[0][0] + RR[0][1] +
              RR[0][2] + RR[0][3];
      AR[1] = Z1 * VHI[0] - X13 * V13[2] - X24 * V24[2] - MX13 * VHI[2] + RR[1][0] + RR[1][1] +
              RR[1][2] + RR[1][3];
      AR[2] = X13 * V13[1] + X24 * V24[1] + MX13 * VHI[1] - Y13 * V13[0] - Y24 * V24[0] -
              MY13 * VHI[0] + RR[2][0] + RR[2][1] + RR[2][2] + RR[2][3];
      const TG XX1 = VC[0][0] * VC[0][0] + VC[0][1] * VC[0][1] + VC[0][2] * VC[0][2] +
                     VC[0][3] * VC[0][3];
      const TG YYv = VC[1][0] * VC[1][0] + VC[1][1] * VC[1][1] + VC[1][2] * VC[1][2] +
                     VC[1][3] * VC[1][3];
      const TG XY = VC[0][0] * VC[1][0] + VC[0][1] * VC[1][1] + VC[0][2] * VC[1][2] +
                    VC[0][3] * VC[1][3];
      const TG XZ1 = (VC[0][0] - VC[0][1] + VC[0][2] - VC[0][3]) * Z1;
      const TG YZ = (VC[1][0] - VC[1][1] + VC[1][2] - VC[1][3]) * Z1;
      const TG ZZ = FOUR * Z2;
      D[0] = YYv + ZZ + FOUR;
      D[1] = XX1 + ZZ + FOUR;
      D[2] = XX1 + YYv + FOUR;
      D[3] = -XY;
      D[4] = -XZ1;
      D[5] = -YZ;
      const TG ABC = D[0] * D[1] * D[2];
      const TG XXYZ2 = D[0] * D[5] * D[5];
      const TG YYXZ2 = D[1] * D[4] * D[4];
      const TG ZZXY2 = D[2] * D[3] * D[3];
      TG DETA = ABC + TWO * D[3] * D[4] * D[5] - XXYZ2 - YYXZ2 - ZZXY2;
      DETA = (DETA < EM20) ? ONE : M<TG>::div(ONE, DETA);
      o.DI[0] = M<TG>::div((ABC - XXYZ2) * DETA, D[0]);
      o.DI[1] = M<TG>::div((ABC - YYXZ2) * DETA, D[1]);
      o.DI[2] = M<TG>::div((ABC - ZZXY2) * DETA, D[2]);
      o.DI[3] = (D[4] * D[5] - D[3] * D[2]) * DETA;
      o.DI[4] = (D[5] * D[3] - D[4] * D[1]) * DETA;
      o.DI[5] = (D[3] * D[4] - D[5] * D[0]) * DETA;
      ALR[0] = o.DI[0] * AR[0] + o.DI[3] * AR[1] + o.DI[4] * AR[2];
      ALR[1] = o.DI[3] * AR[0] + o.DI[1] * AR[1] + o.DI[5] * AR[2];
      ALR[2] = o.DI[4] * AR[0] + o.DI[5] * AR[1] + o.DI[2] * AR[2];
      V13[0] = HALF * V13[0] + ALR[2] * Y13;
      V24[0] = HALF * V24[0] + ALR[2] * Y24;
      VHI[0] = FOURTH * VHI[0] + (ALR[2] * MY13 - Z1 * ALR[1]);
      V13[1] = HALF * V13[1] - ALR[2] * X13;
      V24[1] = HALF * V24[1] - ALR[2] * X24;
      VHI[1] = FOURTH * VHI[1] - (ALR[2] * MX13 - Z1 * ALR[0]);
      V13[2] = HALF * V13[2] - (Y13 * ALR[0] - X13 * ALR[1]);
      V24[2] = HALF * V24[2] - (Y24 * ALR[0] - X24 * ALR[1]);
      VHI[2] = FOURTH * VHI[2] + (MX13 * ALR[1] - MY13 * ALR[0]);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        o.VXYZ[d][0] = V13[d] + VHI[d];
        o.VXYZ[d][1] = V24[d] - VHI[d];
        o.VXYZ[d][2] = -V13[d] + VHI[d];
        o.VXYZ[d][3] = -V24[d] - VHI[d];
      }
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int c = 0; c < 4; ++c) RR[d][c] = RR[d][c] - ALR[d];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        o.RXYZ[0][c] = Q[0][c] * RR[0][c] + Q[1][c] * RR[1][c] + Q[2][c] * RR[2][c];
        o.RXYZ[1][c] = Q[3][c] * RR[0][c] + Q[4][c] * RR[1][c] + Q[5][c] * RR[2][c];
      }
    }
  }
  // characteristic length factors (default-REAL literals of CBACOOR)
  {
    const TG rx = XL2 + XL3 - XL4, ry = YL2 + YL3 - YL4;
    const TG sx = -XL2 + XL3 + XL4, sy = -YL2 + YL3 + YL4;
    const TG C1 = M<TG>::sqrt_(rx * rx + ry * ry);
    const TG C2 = M<TG>::sqrt_(sx * sx + sy * sy);
    TG S1 = FOURTH * (M<TG>::div(tmax(C1, C2), tmin(C1, C2)) - ONE);
    const TG FAC1 = tmin(HALF, S1) + ONE;
    TG FAC2 = M<TG>::div(FOUR * AREA, C1 * C2);
    FAC2 = TG(double(3.413f)) * tmax(ZERO, FAC2 - TG(double(0.7071f)));
    FAC2 = TG(double(0.78f)) + TG(double(0.22f)) * FAC2 * FAC2 * FAC2;
    const TG FACI = TWO * FAC1 * FAC2;
    o.FACN[0] = M<TG>::sqrt_(M<TG>::div(L24, LL));
    o.FACN[1] = M<TG>::sqrt_(M<TG>::div(L13, LL));
    S1 = M<TG>::sqrt_(FACI * (TG(k.four_over_3) + LM * AREA_I) * LL);
    S1 = tmax(S1, EM20);
    o.LL = M<TG>::div(AREA, S1);
  }
  o.AREA = AREA;
  if (offg < ZERO) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int c = 0; c < 4; ++c) o.VXYZ[d][c] = ZERO;
#pragma unroll
    for (int d = 0; d < 2; ++d)
#pragma unroll
      for (int c = 0; c < 4; ++c) o.RXYZ[d][c] = ZERO;
  }
  // CBADEFSH: the constant membrane shear rate (after CBACOOR's zeroing of deleted
  // elements' velocities)
  if (o.flat)
    o.VSH = o.Y24 * o.VXYZ[1][0] - o.Y13 * o.VXYZ[1][1] - o.X24 * o.VXYZ[0][0] +
            o.X13 * o.VXYZ[0][1];
  else
    o.VSH = o.Y24 * (o.VXYZ[1][0] - o.VXYZ[1][2]) + o.Y13 * (-o.VXYZ[1][1] + o.VXYZ[1][3]) -
            o.X24 * (o.VXYZ[0][0] - o.VXYZ[0][2]) + o.X13 * (o.VXYZ[0][1] - o.VXYZ[0][3]);
}

// Warped element, Gauss point ng, corner j: the in-plane terms of CBADEF's B
// matrices, C1, C2 ([B0] = <T> C), BC1, BC2 ([B1] = <T> B C) and the rotation terms
// V1 = (V1(1), V1(2)), V2 of the corner's frame seen from the Gauss point.
template <typename TG>
struct QbatCorner {
  TG C1, C2, BC1, BC2, V11, V12, V21, V22;
};

template <typename TG>
__device__ __forceinline__ QbatCorner<TG> qbat_corner(const QbatGeo<TG> &g, int ng, int j,
                                                      const TG TC[2][2], const TG TBC[2][2],
                                                      TG VKSI, TG VETA) {
  QbatCorner<TG> r;
  r.C1 = VKSI * TC[0][0] + VETA * TC[1][0];
  r.C2 = VKSI * TC[0][1] + VETA * TC[1][1];
  r.BC1 = VKSI * TBC[0][0] + VETA * TBC[1][0];
  r.BC2 = VKSI * TBC[0][1] + VETA * TBC[1][1];
  const TG q1 = g.VQG[0][ng], q2 = g.VQG[1][ng], q3 = g.VQG[2][ng];
  const TG q4 = g.VQG[3][ng], q5 = g.VQG[4][ng], q6 = g.VQG[5][ng];
  r.V12 = q1 * g.VQN[0][j] + q2 * g.VQN[1][j] + q3 * g.VQN[2][j];
  r.V11 = -(q1 * g.VQN[3][j] + q2 * g.VQN[4][j] + q3 * g.VQN[5][j]);
  r.V22 = q4 * g.VQN[0][j] + q5 * g.VQN[1][j] + q6 * g.VQN[2][j];
  r.V21 = -(q4 * g.VQN[3][j] + q5 * g.VQN[4][j] + q6 * g.VQN[5][j]);
  return r;
}

// Warped element, corner j: its ten transverse-shear entries BC(10j+1 .. 10j+10)
// (cbadef.F:580-665), formed from the two mid-edge tyings that meet at the corner.
template <typename TG>
__device__ __forceinline__ void qbat_corner_shear(const QbatGeo<TG> &g, int j, const TG W[4],
                                                  TG bc[10]) {
  // x-part edge, its sign and VASTN rows; y-part edge, sign and rows
  const int ex = (j == 0 || j == 1) ? 0 : 1;
  const int ey = (j == 0 || j == 3) ? 2 : 3;
  const bool nx = (j == 0 || j == 3), ny = (j == 0 || j == 1);
  const int rx = (j == 0 || j == 3) ? 0 : 2;
  const int ry = (j == 0 || j == 1) ? 0 : 2;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const TG cx = W[ex] * g.VNRM[d][ex];
    const TG cy = W[ey] * g.VNRM[d][ey];
    bc[2 * d] = nx ? -cx : cx;
    bc[2 * d + 1] = ny ? -cy : cy;
  }
  bc[6] = W[ex] * g.VASTN[rx][ex];
  bc[8] = W[ex] * g.VASTN[rx + 1][ex];
  bc[7] = W[ey] * g.VASTN[ry][ey];
  bc[9] = W[ey] * g.VASTN[ry + 1][ey];
}

// Flat element, Gauss point ng: CBADEF's BM(1:8) and BC(1:18).
template <typename TG>
__device__ __forceinline__ void qbat_flat_b(const QbatGeo<TG> &g, int ng, TG BM[8], TG BC[18],
                                            const ShConst &k) {
  const TG FOURTH = TG(0.25), PG = TG(qbat_pg());
  const TG *V = g.VCF;
  const TG vpg1 = (ng == 0 || ng == 3) ? -PG : PG;        // VPG(1,NG)
  const TG vpg2 = (ng == 0 || ng == 1) ? -PG : PG;        // VPG(2,NG)
  BM[0] = V[0] + g.HX[ng] * V[4];
  BM[1] = V[1] + g.HX[ng] * V[5];
  BM[2] = g.HX[ng] * FOURTH;
  BM[3] = -BM[2];
  BM[4] = V[2] + g.HY[ng] * V[4];
  BM[5] = V[3] + g.HY[ng] * V[5];
  BM[6] = g.HY[ng] * FOURTH;
  BM[7] = -BM[6];
  const TG A_1 = M<TG>::div(FOURTH, tmax(g.JAC[ng], TG(k.em20)));
  const TG C11 = (V[9] + V[11] * vpg1) * A_1;
  const TG C12 = -(V[7] + V[11] * vpg2) * A_1;
  const TG C21 = -(V[8] + V[10] * vpg1) * A_1;
  const TG C22 = (V[6] + V[10] * vpg2) * A_1;
  TG BETA1 = V[11] + V[7] * vpg2;
  TG KSI1 = V[11] + V[9] * vpg1;
  TG BETA2 = V[10] + V[6] * vpg2;
  TG KSI2 = V[10] + V[8] * vpg1;
  BC[0] = -C11 - C12;
  BC[1] = -C21 - C22;
  BC[2] = BETA1 * C11 + KSI1 * C12;
  BC[3] = BETA1 * C21 + KSI1 * C22;
  BC[4] = -BETA2 * C11 - KSI2 * C12;
  BC[5] = -BETA2 * C21 - KSI2 * C22;
  BC[6] = C11 - C12;
  BC[7] = C21 - C22;
  BC[8] = BETA1 * C11 - KSI1 * C12;
  BC[9] = BETA1 * C21 - KSI1 * C22;
  BC[10] = -BETA2 * C11 + KSI2 * C12;
  BC[11] = -BETA2 * C21 + KSI2 * C22;
  BETA1 = V[7] + V[11] * vpg2;
  KSI1 = V[9] + V[11] * vpg1;
  BETA2 = V[6] + V[10] * vpg2;
  KSI2 = V[8] + V[10] * vpg1;
  BC[12] = C11 * vpg2 + C12 * vpg1;
  BC[13] = C21 * vpg2 + C22 * vpg1;
  BC[14] = -BETA1 * C11 - KSI1 * C12;
  BC[15] = -BETA1 * C21 - KSI1 * C22;
  BC[16] = C11 * BETA2 + C12 * KSI2;
  BC[17] = C21 * BETA2 + C22 * KSI2;
}

// Warped element, Gauss point ng: THK (the frame's curvature term), TC = VJFI*VQ
// and TBC = TBI*TC of CBADEF.
template <typename TG>
__device__ __forceinline__ void qbat_warp_gp(const QbatGeo<TG> &g, int ng, const TG VKSI[4][4],
                                             const TG VETA[4][4], TG &THK, TG TC[2][2],
                                             TG TBC[2][2]) {
  TG TFN[3][2], TBI[2][2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    TFN[d][0] = VKSI[0][ng] * g.VQN[6 + d][0] + VKSI[1][ng] * g.VQN[6 + d][1] +
                VKSI[2][ng] * g.VQN[6 + d][2] + VKSI[3][ng] * g.VQN[6 + d][3];
    TFN[d][1] = VETA[0][ng] * g.VQN[6 + d][0] + VETA[1][ng] * g.VQN[6 + d][1] +
                VETA[2][ng] * g.VQN[6 + d][2] + VETA[3][ng] * g.VQN[6 + d][3];
  }
  const TG(&V)[6][4] = g.VJFI;
  TBI[1][1] = V[0][ng] * TFN[0][0] + V[1][ng] * TFN[1][0] + V[2][ng] * TFN[2][0];
  TBI[1][0] = V[3][ng] * TFN[0][0] + V[4][ng] * TFN[1][0] + V[5][ng] * TFN[2][0];
  TBI[0][1] = V[0][ng] * TFN[0][1] + V[1][ng] * TFN[1][1] + V[2][ng] * TFN[2][1];
  TBI[0][0] = V[3][ng] * TFN[0][1] + V[4][ng] * TFN[1][1] + V[5][ng] * TFN[2][1];
  THK = -(TBI[0][0] + TBI[1][1]);
  TBI[0][1] = -TBI[0][1];
  TBI[1][0] = -TBI[1][0];
  const TG(&Q)[6][4] = g.VQG;
  TC[0][0] = V[0][ng] * Q[0][ng] + V[1][ng] * Q[1][ng] + V[2][ng] * Q[2][ng];
  TC[1][0] = V[3][ng] * Q[0][ng] + V[4][ng] * Q[1][ng] + V[5][ng] * Q[2][ng];
  TC[0][1] = V[0][ng] * Q[3][ng] + V[1][ng] * Q[4][ng] + V[2][ng] * Q[5][ng];
  TC[1][1] = V[3][ng] * Q[3][ng] + V[4][ng] * Q[4][ng] + V[5][ng] * Q[5][ng];
  TBC[0][0] = TBI[0][0] * TC[0][0] + TBI[0][1] * TC[1][0];
  TBC[1][0] = TBI[1][0] * TC[0][0] + TBI[1][1] * TC[1][0];
  TBC[0][1] = TBI[0][0] * TC[0][1] + TBI[0][1] * TC[1][1];
  TBC[1][1] = TBI[1][0] * TC[0][1] + TBI[1][1] * TC[1][1];
}

// CBADEF strain rates VDEF(1,2,4..8) of Gauss point ng (VDEF(3) is CBADEFSH's).
template <typename TG>
__device__ __forceinline__ void qbat_strain_rates(const QbatGeo<TG> &g, int ng,
                                                  const TG VKSI[4][4], const TG VETA[4][4],
                                                  const ShConst &k, TG VDEF[8], TG TC[2][2],
                                                  TG &THK, TG TBC[2][2]) {
  VDEF[2] = g.VSH;
  if (g.flat) {
    TG BM[8], BC[18];
    qbat_flat_b<TG>(g, ng, BM, BC, k);
    const TG(&VX)[3][4] = g.VXYZ;
    const TG(&RX)[2][4] = g.RXYZ;
    VDEF[0] = BM[0] * VX[0][0] + BM[1] * VX[0][1] + BM[2] * VX[0][2];
    VDEF[1] = BM[4] * VX[1][0] + BM[5] * VX[1][1] + BM[6] * VX[1][2];
    VDEF[5] = BM[0] * RX[1][0] + BM[1] * RX[1][1] + BM[2] * RX[1][2];
    VDEF[6] = -(BM[4] * RX[0][0] + BM[5] * RX[0][1] + BM[6] * RX[0][2]);
    VDEF[7] = -(BM[0] * RX[0][0] + BM[1] * RX[0][1] + BM[2] * RX[0][2]) + BM[4] * RX[1][0] +
              BM[5] * RX[1][1] + BM[6] * RX[1][2];
    VDEF[3] = BC[0] * VX[2][0] + BC[6] * VX[2][1] + BC[12] * VX[2][2] + BC[2] * RX[0][0] +
              BC[8] * RX[0][1] + BC[14] * RX[0][3] + BC[4] * RX[1][0] + BC[10] * RX[1][1] +
              BC[16] * RX[1][3];
    VDEF[4] = BC[1] * VX[2][0] + BC[7] * VX[2][1] + BC[13] * VX[2][2] + BC[3] * RX[0][0] +
              BC[9] * RX[0][1] + BC[15] * RX[0][3] + BC[5] * RX[1][0] + BC[11] * RX[1][1] +
              BC[17] * RX[1][3];
    return;
  }
  qbat_warp_gp<TG>(g, ng, VKSI, VETA, THK, TC, TBC);
  const TG q1 = g.VQG[0][ng], q2 = g.VQG[1][ng], q3 = g.VQG[2][ng];
  const TG q4 = g.VQG[3][ng], q5 = g.VQG[4][ng], q6 = g.VQG[5][ng];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const QbatCorner<TG> r = qbat_corner<TG>(g, ng, j, TC, TBC, VKSI[j][ng], VETA[j][ng]);
    const TG VT1 = q1 * g.VXYZ[0][j] + q2 * g.VXYZ[1][j] + q3 * g.VXYZ[2][j];
    const TG VT2 = q4 * g.VXYZ[0][j] + q5 * g.VXYZ[1][j] + q6 * g.VXYZ[2][j];
    if (j == 0) {
      VDEF[0] = r.C1 * VT1;
      VDEF[1] = r.C2 * VT2;
      VDEF[5] = r.BC1 * VT1;
      VDEF[6] = r.BC2 * VT2;
      VDEF[7] = r.BC1 * VT2 + r.BC2 * VT1;
    } else {
      VDEF[0] = VDEF[0] + r.C1 * VT1;
      VDEF[1] = VDEF[1] + r.C2 * VT2;
      VDEF[5] = VDEF[5] + r.BC1 * VT1;
      VDEF[6] = VDEF[6] + r.BC2 * VT2;
      VDEF[7] = VDEF[7] + r.BC1 * VT2 + r.BC2 * VT1;
    }
    const TG RV1 = r.V11 * g.RXYZ[0][j] + r.V12 * g.RXYZ[1][j];
    const TG RV2 = r.V21 * g.RXYZ[0][j] + r.V22 * g.RXYZ[1][j];
    if (j < 3) {
      VDEF[5] = VDEF[5] + r.C1 * RV1;
      VDEF[6] = VDEF[6] + r.C2 * RV2;
      VDEF[7] = VDEF[7] + r.C1 * RV2 + r.C2 * RV1;
    } else {
      VDEF[5] = VDEF[5] + THK * VDEF[0] + r.C1 * RV1;
      VDEF[6] = VDEF[6] + THK * VDEF[1] + r.C2 * RV2;
      VDEF[7] = VDEF[7] + THK * VDEF[2] + r.C1 * RV2 + r.C2 * RV1;
    }
  }
  const TG W[4] = {VKSI[1][ng], VKSI[2][ng], VETA[3][ng], VETA[2][ng]};
  TG BCX = TG(0), BCY = TG(0);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    TG bc[10];
    qbat_corner_shear<TG>(g, j, W, bc);
    const TG(&VX)[3][4] = g.VXYZ;
    const TG(&RX)[2][4] = g.RXYZ;
    if (j == 0) {
      BCX = bc[0] * VX[0][0] + bc[2] * VX[1][0] + bc[4] * VX[2][0] + bc[6] * RX[0][0] +
            bc[8] * RX[1][0];
      BCY = bc[1] * VX[0][0] + bc[3] * VX[1][0] + bc[5] * VX[2][0] + bc[7] * RX[0][0] +
            bc[9] * RX[1][0];
    } else {
      BCX = BCX + bc[0] * VX[0][j] + bc[2] * VX[1][j] + bc[4] * VX[2][j] + bc[6] * RX[0][j] +
            bc[8] * RX[1][j];
      BCY = BCY + bc[1] * VX[0][j] + bc[3] * VX[1][j] + bc[5] * VX[2][j] + bc[7] * RX[0][j] +
            bc[9] * RX[1][j];
    }
  }
  VDEF[3] = TC[0][0] * BCX + TC[1][0] * BCY;
  VDEF[4] = TC[0][1] * BCX + TC[1][1] * BCY;
}

template <typename T, typename TG, typename IO, typename Emit>
__device__ __forceinline__ void qbat_chain(const TG xg[3][4], const TG vg[3][4],
                                           const TG vrg[3][4], T thk0, const IO &io,
                                           double dt1d, const ShMat &m, const T jc[3],
                                           const ShProp &p, const ShConst &k, QbatOut<T> &out,
                                           Emit emit) {
  const T ZERO = T(0), ONE = T(1), TWO = T(2), HALF = T(0.5), FOURTH = T(0.25);
  const T EM20 = T(k.em20);
  const T dt1 = T(dt1d);
  T offg = io.get(SC_OFFG);
  QbatGeo<TG> g;
  qbat_geometry<TG>(xg, vg, vrg, TG(dt1d), as_kin<TG>(offg), k, g);
  if (g.degenerate) {
    offg = ZERO;
    io.put(SC_OFFG, offg);
  }
  TG VKSI[4][4], VETA[4][4];
  qbat_shape_tables<TG>(VKSI, VETA);
  const T AREA = T(g.AREA);
  const T X13 = T(g.X13), X24 = T(g.X24), Y13 = T(g.Y13), Y24 = T(g.Y24);
  // ---- CNCOEF3, IGTYP = 1, law 2, ITHK = 0 (moduli from PM, resol_init.F:689-696) ----
  const T thk02 = thk0 * thk0, VOLG = thk0 * AREA;
  const T RHO = T(m.rho0), NU = T(m.nu), G = T(m.shear);
  const T A11 = T(m.a11), SSP = T(m.ssp);
  const T GS = G * T(p.fsh);
  const T AMU = (p.dn == 0.0) ? T(k.em03) : T(p.dn);      // IHBE = 11: EM3 when unset
  T off = tmin(ONE, tabs(offg));
  // ---- CBAENERS with the previous FOR ----
  T eint0 = io.get(SC_EINT), eint1 = io.get(SC_EINT + 1);
  const T VSH = T(g.VSH);
  eint0 = eint0 + io.get(SC_FOR + 2) * VSH * (off * AREA * thk0 * dt1 * HALF);
  const T asrate = (m.israte > 0) ? tmin(ONE, T(m.asrate_pm9) * dt1) : ONE;
  const T dtinv = M<T>::div(dt1, tmax(dt1 * dt1, EM20));
  T thk = io.get(SC_THK);
  T VF[3][4], VM[2][4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    VF[0][c] = ZERO; VF[1][c] = ZERO; VF[2][c] = ZERO;
    VM[0][c] = ZERO; VM[1][c] = ZERO;
  }
  T epsd_glob = ZERO, evis = ZERO, viscmx = ZERO;
  T forsum[5], momsum[3];   // left-to-right sums of the Gauss points' FORPG / MOMPG
#pragma unroll
  for (int ng = 0; ng < 4; ++ng) {
    const T CDET = T(g.JAC[ng]);
    const T VOL0 = thk0 * CDET;
    // ---- CBADEF strain rates, CBASTRA3 increments ----
    TG VDg[8], TCg[2][2], THKg = TG(0), TBCg[2][2];
    qbat_strain_rates<TG>(g, ng, VKSI, VETA, k, VDg, TCg, THKg, TBCg);
    T VDEF[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) VDEF[q] = T(VDg[q]);
    const T EXX = VDEF[0] * dt1, EYY = VDEF[1] * dt1, EXY = VDEF[2] * dt1;
    const T EYZ = VDEF[4] * dt1, EXZ = VDEF[3] * dt1;
    const T KXX = VDEF[5] * dt1, KYY = VDEF[6] * dt1, KXY = VDEF[7] * dt1;
    T FORL[5], MOML[3];
#pragma unroll
    for (int c = 0; c < 5; ++c) FORL[c] = io.get(QC_FORPG + 5 * ng + c);
#pragma unroll
    for (int c = 0; c < 3; ++c) MOML[c] = io.get(QC_MOMPG + 3 * ng + c);
    // ---- CBAENER with the Gauss point's previous resultants ----
    eint0 = eint0 - FORL[2] * EXY * (HALF * off * VOL0);
    // ---- the Gauss point's equivalent strain rate ----
    T epsd_pg;
    {
      const T eps_k2 = (KXX * KXX + KYY * KYY + KXX * KYY + FOURTH * (KXY * KXY)) *
                       T(k.one_over_9) * (thk * thk);
      const T eps_m2 =
          T(k.four_over_3) * (EXX * EXX + EYY * EYY + EXX * EYY + FOURTH * (EXY * EXY));
      epsd_pg = M<T>::sqrt_(eps_k2 + eps_m2) * dtinv;
      epsd_glob = epsd_glob + M<T>::div(epsd_pg, T(4));
    }
    // ---- MULAWC law 2 over the layers ----
    T degmb = FORL[0] * EXX + FORL[1] * EYY + FORL[2] * EXY + FORL[3] * EYZ + FORL[4] * EXZ;
    T degfx = MOML[0] * KXX + MOML[1] * KYY + MOML[2] * KXY;
    degmb = degmb * off;
    degfx = degfx * off;
    T thkn = thk;
#pragma unroll
    for (int c = 0; c < 5; ++c) FORL[c] = ZERO;
#pragma unroll
    for (int c = 0; c < 3; ++c) MOML[c] = ZERO;
    const T off_old = off;
    T epchk = ZERO;
    viscmx = ZERO;
#pragma unroll 1
    for (int it = 0; it < p.npt; ++it) {
      const T wf = T(p.wf[it]), wm = T(p.wm[it]);
      const T thklyl = wf * thk0;
      const T zt = T(p.z0[it]) * thk0;
      T deps[5];
      deps[0] = EXX + zt * KXX;
      deps[1] = EYY + zt * KYY;
      deps[2] = EXY + zt * KXY;
      deps[3] = EYZ;
      deps[4] = EXZ;
      const int cs = QC_SIG + 5 * (5 * ng + it);
      T s[5];
#pragma unroll
      for (int c = 0; c < 5; ++c) s[c] = io.get(cs + c);
      T pla = io.get(QC_PLA + 5 * ng + it), epsdp = io.get(QC_EPSDP + 5 * ng + it);
      T etse_unused, sigy_unused = ZERO;
      shell_law2_point<T>(m, jc, k, GS, dt1, asrate, epsd_pg, off_old, off, pla, epsdp, epchk,
                          thkn, thklyl, deps, s, etse_unused, sigy_unused, ONE);
      viscmx = tmax(T(p.dm), viscmx);
#pragma unroll
      for (int c = 0; c < 5; ++c) io.put(cs + c, s[c] * ONE);
      io.put(QC_PLA + 5 * ng + it, pla);
      io.put(QC_EPSDP + 5 * ng + it, epsdp);
#pragma unroll
      for (int c = 0; c < 5; ++c) FORL[c] = FORL[c] + wf * s[c];
#pragma unroll
      for (int c = 0; c < 3; ++c) MOML[c] = MOML[c] + wm * s[c];
    }
    thk = tmax(thkn, T(k.em30));
    {
      const T fact = T(k.onep414) * T(p.dm);
      const T visc = fact * SSP * M<T>::sqrt_(CDET) * dtinv * RHO;
      FORL[0] = FORL[0] + visc * (EXX + HALF * EYY);
      FORL[1] = FORL[1] + visc * (EYY + HALF * EXX);
      FORL[2] = FORL[2] + visc * EXY * T(k.third);
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) FORL[c] = FORL[c] * off * ONE;
#pragma unroll
    for (int c = 0; c < 3; ++c) MOML[c] = MOML[c] * off * ONE;
    degmb = degmb + FORL[0] * EXX + FORL[1] * EYY + FORL[2] * EXY + FORL[3] * EYZ + FORL[4] * EXZ;
    degfx = degfx + MOML[0] * KXX + MOML[1] * KYY + MOML[2] * KXY;
    {
      const T vol2 = HALF * VOL0;
      eint0 = eint0 + degmb * vol2;
      eint1 = eint1 + degfx * thk0 * vol2;
    }
    // ---- CBAENER with the new resultants ----
    eint0 = eint0 - FORL[2] * EXY * (HALF * off * VOL0);
    // ---- CBAVISC (AREA = CDET, THK = THK0) ----
    {
      T visc = T(k.onep414) * off * AMU * RHO * SSP * M<T>::sqrt_(CDET);
      const T GG = M<T>::div(HALF, ONE + NU);
      T FX = visc * (VDEF[0] + NU * VDEF[1]);
      T FY = visc * (VDEF[1] + NU * VDEF[0]);
      T FXY = visc * VDEF[2] * GG;
      FORL[0] = FORL[0] + FX;
      FORL[1] = FORL[1] + FY;
      FORL[2] = FORL[2] + FXY;
      T DV = CDET * thk0 * dt1;
      evis = evis + (FX * VDEF[0] + FY * VDEF[1]) * DV;
      visc = T(k.zep3) * thk0 * visc;
      FX = visc * (VDEF[5] + NU * VDEF[6]);
      FY = visc * (VDEF[6] + NU * VDEF[5]);
      FXY = visc * VDEF[7] * GG;
      MOML[0] = MOML[0] + FX;
      MOML[1] = MOML[1] + FY;
      MOML[2] = MOML[2] + FXY;
      DV = DV * thk0;
      evis = evis + (FX * VDEF[5] + FY * VDEF[6] + FXY * VDEF[7]) * DV;
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      io.put(QC_FORPG + 5 * ng + c, FORL[c]);
      forsum[c] = (ng == 0) ? FORL[c] : forsum[c] + FORL[c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      io.put(QC_MOMPG + 3 * ng + c, MOML[c]);
      momsum[c] = (ng == 0) ? MOML[c] : momsum[c] + MOML[c];
    }
    // ---- CBAFORI (VOL = VOL0, TH12 = THK0^2) ----
    {
      const T C2 = VOL0, C1 = thk02 * CDET;
      const T FF1 = FORL[0], FF2 = FORL[1], FF4 = FORL[4], FF5 = FORL[3];
      const T MM1 = MOML[0], MM2 = MOML[1], MM3 = MOML[2];
      if (g.flat) {
        TG BMg[8], BCg[18];
        qbat_flat_b<TG>(g, ng, BMg, BCg, k);
        T BM[8], BC[18];
#pragma unroll
        for (int q = 0; q < 8; ++q) BM[q] = T(BMg[q]);
#pragma unroll
        for (int q = 0; q < 18; ++q) BC[q] = T(BCg[q]);
        const T CM1 = C1 * (BM[6] * MM2 + BM[2] * MM3);
        const T CM2 = C1 * (BM[2] * MM1 + BM[6] * MM3);
        const T CC1 = C2 * (BC[14] * FF4 + BC[15] * FF5);
        const T CC2 = C2 * (BC[16] * FF4 + BC[17] * FF5);
        const T BC1 = BC[0] * FF4, BC3 = BC[2] * FF4, BC5 = BC[4] * FF4;
        VF[0][0] = VF[0][0] + C2 * (BM[0] * FF1);
        VF[1][0] = VF[1][0] + C2 * (BM[4] * FF2);
        VF[2][0] = VF[2][0] + C2 * (BC1 + BC[1] * FF5);
        VM[0][0] = VM[0][0] + C2 * (BC3 + BC[3] * FF5) - C1 * (BM[4] * MM2 + BM[0] * MM3);
        VM[1][0] = VM[1][0] + C2 * (BC5 + BC[5] * FF5) + C1 * (BM[0] * MM1 + BM[4] * MM3);
        VF[0][2] = VF[0][2] + C2 * (BM[2] * FF1);
        VF[1][2] = VF[1][2] + C2 * (BM[6] * FF2);
        VF[2][2] = VF[2][2] + C2 * (BC[12] * FF4 + BC[13] * FF5);
        VM[0][2] = VM[0][2] + CC1 - CM1;
        VM[1][2] = VM[1][2] + CC2 + CM2;
        VF[0][1] = VF[0][1] + C2 * (BM[1] * FF1);
        VF[1][1] = VF[1][1] + C2 * (BM[5] * FF2);
        VF[2][1] = VF[2][1] + C2 * (BC[6] * FF4 + BC[7] * FF5);
        VM[0][1] = VM[0][1] + C2 * (BC[8] * FF4 + BC[9] * FF5) - C1 * (BM[5] * MM2 + BM[1] * MM3);
        VM[1][1] = VM[1][1] + C2 * (BC[10] * FF4 + BC[11] * FF5) + C1 * (BM[1] * MM1 + BM[5] * MM3);
        VF[0][3] = -VF[0][2];
        VF[1][3] = -VF[1][2];
        VF[2][3] = -VF[2][2];
        VM[0][3] = VM[0][3] + CC1 + CM1;
        VM[1][3] = VM[1][3] + CC2 - CM2;
      } else {
        const T THK = T(THKg);
        const T BCX = T(TCg[0][0]) * FF4 + T(TCg[0][1]) * FF5;
        const T BCY = T(TCg[1][0]) * FF4 + T(TCg[1][1]) * FF5;
        const T q[6] = {T(g.VQG[0][ng]), T(g.VQG[1][ng]), T(g.VQG[2][ng]),
                        T(g.VQG[3][ng]), T(g.VQG[4][ng]), T(g.VQG[5][ng])};
        const TG W[4] = {VKSI[1][ng], VKSI[2][ng], VETA[3][ng], VETA[2][ng]};
        // CBADEFSH's constant-shear entries BM(9j+3), BM(9j+6)
        const T BMS[8] = {-X24, Y24, X13, -Y13, X24, -Y24, -X13, Y13};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const QbatCorner<TG> r = qbat_corner<TG>(g, ng, j, TCg, TBCg, VKSI[j][ng], VETA[j][ng]);
          const T c1 = T(r.C1), c2 = T(r.C2), bc1 = T(r.BC1), bc2 = T(r.BC2);
          TG bcg[10];
          qbat_corner_shear<TG>(g, j, W, bcg);
          T bc[10];
#pragma unroll
          for (int q2 = 0; q2 < 10; ++q2) bc[q2] = T(bcg[q2]);
          T BM[9], BMF[9];
          BM[2] = BMS[2 * j];
          BM[5] = BMS[2 * j + 1];
          BM[8] = ZERO;
          BM[0] = q[0] * c1;
          BMF[0] = THK * BM[0] + q[0] * bc1;
          BM[1] = q[3] * c2;
          BMF[1] = THK * BM[1] + q[3] * bc2;
          BMF[2] = THK * BM[2] + q[0] * bc2 + q[3] * bc1;
          BM[3] = q[1] * c1;
          BMF[3] = THK * BM[3] + q[1] * bc1;
          BM[4] = q[4] * c2;
          BMF[4] = THK * BM[4] + q[4] * bc2;
          BMF[5] = THK * BM[5] + q[1] * bc2 + q[4] * bc1;
          BM[6] = q[2] * c1;
          BMF[6] = THK * BM[6] + q[2] * bc1;
          BM[7] = q[5] * c2;
          BMF[7] = THK * BM[7] + q[5] * bc2;
          BMF[8] = THK * BM[8] + q[2] * bc2 + q[5] * bc1;
          const T v11 = T(r.V11), v12 = T(r.V12), v21 = T(r.V21), v22 = T(r.V22);
          const T BF[6] = {v11 * c1, v21 * c2, v11 * c2 + v21 * c1,
                           v12 * c1, v22 * c2, v12 * c2 + v22 * c1};
          VF[0][j] = VF[0][j] + C2 * (BM[0] * FF1 + BM[1] * FF2 + bc[0] * BCX + bc[1] * BCY) +
                     C1 * (BMF[0] * MM1 + BMF[1] * MM2 + BMF[2] * MM3);
          VM[0][j] = VM[0][j] + C2 * (bc[6] * BCX + bc[7] * BCY) +
                     C1 * (BF[0] * MM1 + BF[1] * MM2 + BF[2] * MM3);
          VF[1][j] = VF[1][j] + C2 * (BM[3] * FF1 + BM[4] * FF2 + bc[2] * BCX + bc[3] * BCY) +
                     C1 * (BMF[3] * MM1 + BMF[4] * MM2 + BMF[5] * MM3);
          VM[1][j] = VM[1][j] + C2 * (bc[8] * BCX + bc[9] * BCY) +
                     C1 * (BF[3] * MM1 + BF[4] * MM2 + BF[5] * MM3);
          VF[2][j] = VF[2][j] + C2 * (BM[6] * FF1 + BM[7] * FF2 + bc[4] * BCX + bc[5] * BCY) +
                     C1 * (BMF[6] * MM1 + BMF[7] * MM2 + BMF[8] * MM3);
        }
      }
    }
  }
  io.put(SC_THK, thk);
  io.put(SC_EPSD, ONE * epsd_glob + (ONE - ONE) * io.get(SC_EPSD));
  // ---- GBUF%FOR, GBUF%MOM: the Gauss-point means ----
  T FOR[5], MOM[3];
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    FOR[c] = FOURTH * forsum[c];
    io.put(SC_FOR + c, FOR[c]);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    MOM[c] = FOURTH * momsum[c];
    io.put(SC_MOM + c, MOM[c]);
  }
  // ---- CBAFORCT: the constant membrane shear's nodal forces ----
  {
    const T THOFF = VOLG * FOR[2] * off;
    const T SX1 = -THOFF * X24, SY1 = THOFF * Y24;
    const T SX2 = THOFF * X13, SY2 = -THOFF * Y13;
    VF[0][0] = VF[0][0] + SX1;
    VF[1][0] = VF[1][0] + SY1;
    VF[0][1] = VF[0][1] + SX2;
    VF[1][1] = VF[1][1] + SY2;
    if (!g.flat) {
      VF[0][2] = VF[0][2] - SX1;
      VF[1][2] = VF[1][2] - SY1;
      VF[0][3] = VF[0][3] - SX2;
      VF[1][3] = VF[1][3] - SY2;
    }
  }
  // ---- CBAENERS with the new FOR ----
  eint0 = eint0 + FOR[2] * VSH * (off * AREA * thk0 * dt1 * HALF);
  io.put(SC_EINT, eint0);
  io.put(SC_EINT + 1, eint1);
  // ---- CBAPROJ (ISROT = 0): back to the global frame ----
  T F[3][4], Mo[3][4];
  {
    const T e1[3] = {T(g.e1[0]), T(g.e1[1]), T(g.e1[2])};
    const T e2[3] = {T(g.e2[0]), T(g.e2[1]), T(g.e2[2])};
    const T e3[3] = {T(g.e3[0]), T(g.e3[1]), T(g.e3[2])};
    if (g.flat) {
      T FL[3][4], ML[2][4];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        FL[d][0] = VF[d][0] + VF[d][2];
        FL[d][1] = VF[d][1] + VF[d][3];
        FL[d][2] = -VF[d][0] + VF[d][2];
        FL[d][3] = -VF[d][1] + VF[d][3];
      }
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        ML[d][0] = VM[d][0] + VM[d][2];
        ML[d][1] = VM[d][1] + VM[d][3];
        ML[d][2] = -VM[d][0] + VM[d][2];
        ML[d][3] = -VM[d][1] + VM[d][3];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          F[d][c] = e1[d] * FL[0][c] + e2[d] * FL[1][c] + e3[d] * FL[2][c];
          Mo[d][c] = e1[d] * ML[0][c] + e2[d] * ML[1][c];
        }
    } else {
      T VC[3][4], DI[6], MM[3][4], AR[3], ALR[3];
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) VC[d][c] = T(g.VC[d][c]);
#pragma unroll
      for (int q = 0; q < 6; ++q) DI[q] = T(g.DI[q]);
      const T Z1 = VC[2][0];
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          MM[d][j] = T(g.VQN[d][j]) * VM[0][j] + T(g.VQN[d + 3][j]) * VM[1][j] + ZERO;
      AR[0] = -Z1 * (VF[1][0] - VF[1][1] + VF[1][2] - VF[1][3]) + VC[1][0] * VF[2][0] + MM[0][0] +
              VC[1][1] * VF[2][1] + MM[0][1] + VC[1][2] * VF[2][2] + MM[0][2] +
              VC[1][3] * VF[2][3] + MM[0][3];
      AR[1] = Z1 * (VF[0][0] - VF[0][1] + VF[0][2] - VF[0][3]) - VC[0][0] * VF[2][0] + MM[1][0] -
              VC[0][1] * VF[2][1] + MM[1][1] - VC[0][2] * VF[2][2] + MM[1][2] -
              VC[0][3] * VF[2][3] + MM[1][3];
      AR[2] = -VC[1][0] * VF[0][0] + VC[0][0] * VF[1][0] + MM[2][0] - VC[1][1] * VF[0][1] +
              VC[0][1] * VF[1][1] + MM[2][1] - VC[1][2] * VF[0][2] + VC[0][2] * VF[1][2] +
              MM[2][2] - VC[1][3] * VF[0][3] + VC[0][3] * VF[1][3] + MM[2][3];
      ALR[0] = DI[0] * AR[0] + DI[3] * AR[1] + DI[4] * AR[2];
      ALR[1] = DI[3] * AR[0] + DI[1] * AR[1] + DI[5] * AR[2];
      ALR[2] = DI[4] * AR[0] + DI[5] * AR[1] + DI[2] * AR[2];
      T C1 = Z1 * ALR[1];
      VF[0][0] = VF[0][0] - C1 + VC[1][0] * ALR[2];
      VF[0][1] = VF[0][1] + C1 + VC[1][1] * ALR[2];
      VF[0][2] = VF[0][2] - C1 + VC[1][2] * ALR[2];
      VF[0][3] = VF[0][3] + C1 + VC[1][3] * ALR[2];
      C1 = Z1 * ALR[0];
      VF[1][0] = VF[1][0] + C1 - VC[0][0] * ALR[2];
      VF[1][1] = VF[1][1] - C1 - VC[0][1] * ALR[2];
      VF[1][2] = VF[1][2] + C1 - VC[0][2] * ALR[2];
      VF[1][3] = VF[1][3] - C1 - VC[0][3] * ALR[2];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        VF[2][j] = VF[2][j] - VC[1][j] * ALR[0] + VC[0][j] * ALR[1];
        MM[0][j] = MM[0][j] - ALR[0];
        MM[1][j] = MM[1][j] - ALR[1];
        MM[2][j] = MM[2][j] - ALR[2];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          F[d][c] = e1[d] * VF[0][c] + e2[d] * VF[1][c] + e3[d] * VF[2][c];
          Mo[d][c] = e1[d] * MM[0][c] + e2[d] * MM[1][c] + e3[d] * MM[2][c];
        }
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      F[d][c] = F[d][c] * off;
      Mo[d][c] = Mo[d][c] * off;
    }
  // ---- CNDT3, NODADT = 0, IDTMINS = 0, IDTMIN(3) = 2, ZOFFSET = 0 (F_OSET = F_DTE = 1) ----
  viscmx = tmax(viscmx, AMU);
  viscmx = M<T>::sqrt_(ONE + viscmx * viscmx) - viscmx;
  const T ALDT = M<T>::div(T(g.LL) * viscmx, M<T>::sqrt_(ONE));
  const T DT = M<T>::div(T(p.dtfac) * ALDT, SSP);
  if (DT <= T(p.dtmin) && off >= ONE && offg != TWO && offg >= ZERO) off = ZERO;
  out.dt = DT;
  out.alive = (offg > ZERO && off != ZERO);
  out.ehou = evis;
  const T sti = M<T>::div(HALF * VOLG * A11 * off, tmax(ALDT * ALDT, EM20));
  // ---- CUPDTN3P: OFFG, deleted elements write zero, FSKY(7) = STI*FACN ----
  if (off < ONE) {
    offg = off;
    io.put(SC_OFFG, offg);
  }
  const bool zero_out = offg < ZERO;
  out.sti13 = zero_out ? ZERO : sti * T(g.FACN[0]);
  out.sti24 = zero_out ? ZERO : sti * T(g.FACN[1]);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (zero_out) emit(c, ZERO, ZERO, ZERO, ZERO, ZERO, ZERO);
    else emit(c, -F[0][c], -F[1][c], -F[2][c], -Mo[0][c], -Mo[1][c], -Mo[2][c]);
  }
}

}  // namespace orad

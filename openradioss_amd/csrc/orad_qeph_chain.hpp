// orad_qeph_chain.hpp -- one QEPH shell (Ishell=24) through the explicit force
// chain of CZFORC3 (engine/source/elements/shell/coquez/czforc3.F:436-808):
// IGTYP=1, Ismstr=2, ISROT=0, IORTH=0, law 2 at NPT >= 2 points, PARITH/ON slots.
//
// Stages and the routines they follow:
//   CZCORC1 (czcorc.F:37-618) with CLSKEW3 IREP=0 (cdkcoor3.F:303-368): local
//   frame from the diagonals, local corner coordinates, characteristic length,
//   projected relative velocities with the second-order rotation correction,
//   CZCORP5 (czcorp5.F:32-396) warped-element projection (flat elements skip it);
//   CNCOEF3B (cncoef3.F:29-293) thickness, moduli, Dn; CZDEF (czdef.F:29-209)
//   membrane/bending/shear rates and hourglass rates; CZSTRA3 increments; the
//   element strain rate (czforc3.F:554-561); MULAWC's law-2 point loop with the
//   QEPH factors ZCFAC and SIGY (mulawc.F90:509-569,2787-2794); CNDT3
//   (sh3n/coquedk/cndt3.F:84-315); CZFINTCE (czfintce.F); CZFINTN1 (czfintn.F:29-500)
//   physical hourglass with the 12-value HOURG history; CZPROJ1/CZPROJN
//   (czproj.F:32-185,880-1470) back to the global frame; CUPDTN3P
//   (coque/cupdtn3.F:488-700), whose nodal stiffness is STI*FACN per diagonal.
//
// T = double follows the reference's double-precision operation order (its file
// is compiled without FMA contraction).  T = float is the mixed form: the frame,
// projection and velocity stage (qeph_kinematics) runs in TG on coordinates and
// translational velocities taken relative to corner 1 in FP64; every quantity of
// that stage is invariant to the shift.
#pragma once
#include "orad_shell_chain.hpp"

namespace orad {

template <typename T>
struct QephOut {
  T sti13, sti24;  // FSKY(7) of corners 1,3 (STI*FACN(1)) and 2,4 (STI*FACN(2))
  T dt;            // element time step candidate
  T ehou;          // hourglass energy increment (PARTSAV(8))
  int alive;       // the candidate enters DT2T
};

// CZCORC1 + CZCORP5 + CZDEF results that the rest of the chain reads.  CZDEF runs
// in this stage because its transverse-shear and hourglass rates difference the
// out-of-plane velocities against the rotations, which nearly cancel in bending.
template <typename TG>
struct QephKin {
  TG e1[3], e2[3], e3[3];
  TG AREA, AREA_I, X13, X24, Y13, Y24, MX13, MX23, MX34, MY13, MY23, MY34, Z1;
  TG COREL[2][4];
  TG LL, FACN[2];
  TG VDEF[8], DHG[6];          // CZDEF: generalised strain rates and hourglass rates
  TG VQN[3][4], DI[6];
  bool plat;
  bool degenerate;             // CLSKEW3 found a zero-area frame and set OFFG = 0
};

template <typename TG>
__device__ __forceinline__ void qeph_kinematics(const TG xg[3][4], const TG vg[3][4],
                                                const TG vrg[3][4], TG dt1, TG offg, int npt,
                                                const ShConst &k, QephKin<TG> &o) {
  const TG ZERO = TG(0), ONE = TG(1), TWO = TG(2), FOUR = TG(4), HALF = TG(0.5),
           FOURTH = TG(0.25);
  const TG EM20 = TG(k.em20);
  // ---- CZCORC1 diagonals, CLSKEW3 IREP = 0 ----
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
  const TG off_loc = (tabs(offg) != ZERO) ? ONE : ZERO;
  {
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
  const TG AREA = FOURTH * DET;
  const TG AREA_I = tmax(M<TG>::div(off_loc, AREA), EM20);
  o.AREA = AREA;
  o.AREA_I = AREA_I;
  // ---- local coordinates (corner 1 at the origin, Z1 from the centroid) ----
  TG XL2, YL2, XL3, YL3, XL4, YL4, Z1;
  {
    TG L0[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) L0[d] = FOURTH * (xg[d][2] + xg[d][3] + xg[d][0] + xg[d][1]);
    TG XX = xg[0][1] - xg[0][0], YY = xg[1][1] - xg[1][0], ZZ = xg[2][1] - xg[2][0];
    XL2 = E1X * XX + E1Y * YY + E1Z * ZZ;
    YL2 = E2X * XX + E2Y * YY + E2Z * ZZ;
    XX = xg[0][0] - L0[0]; YY = xg[1][0] - L0[1]; ZZ = xg[2][0] - L0[2];
    Z1 = E3X * XX + E3Y * YY + E3Z * ZZ;
    XX = xg[0][2] - xg[0][0]; YY = xg[1][2] - xg[1][0]; ZZ = xg[2][2] - xg[2][0];
    XL3 = E1X * XX + E1Y * YY + E1Z * ZZ;
    YL3 = E2X * XX + E2Y * YY + E2Z * ZZ;
    XX = xg[0][3] - xg[0][0]; YY = xg[1][3] - xg[1][0]; ZZ = xg[2][3] - xg[2][0];
    XL4 = E1X * XX + E1Y * YY + E1Z * ZZ;
    YL4 = E2X * XX + E2Y * YY + E2Z * ZZ;
  }
  // Ismstr = 2 stores the local coordinates for a later switch to small strain
  // (OFFG = 2, reached only through /DT/SHELL/CST); they are not read here.
  TG(&COREL)[2][4] = o.COREL;
  {
    const TG lx = FOURTH * (XL2 + XL3 + XL4), ly = FOURTH * (YL2 + YL3 + YL4);
    COREL[0][0] = -lx; COREL[0][1] = XL2 - lx; COREL[0][2] = XL3 - lx; COREL[0][3] = XL4 - lx;
    COREL[1][0] = -ly; COREL[1][1] = YL2 - ly; COREL[1][2] = YL3 - ly; COREL[1][3] = YL4 - ly;
  }
  const TG X13 = (COREL[0][0] - COREL[0][2]) * HALF, X24 = (COREL[0][1] - COREL[0][3]) * HALF;
  const TG Y13 = (COREL[1][0] - COREL[1][2]) * HALF, Y24 = (COREL[1][1] - COREL[1][3]) * HALF;
  const TG MX13 = (COREL[0][0] + COREL[0][2]) * HALF;
  const TG MY13 = (COREL[1][0] + COREL[1][2]) * HALF;
  o.X13 = X13; o.X24 = X24; o.Y13 = Y13; o.Y24 = Y24; o.MX13 = MX13; o.MY13 = MY13;
  o.MX23 = (COREL[0][1] + COREL[0][2]) * HALF;
  o.MX34 = (COREL[0][2] + COREL[0][3]) * HALF;
  o.MY23 = (COREL[1][1] + COREL[1][2]) * HALF;
  o.MY34 = (COREL[1][2] + COREL[1][3]) * HALF;
  const TG L13 = X13 * X13 + Y13 * Y13, L24 = X24 * X24 + Y24 * Y24;
  TG HS;
  {
    const TG C1 = COREL[0][1] * COREL[1][3] - COREL[1][1] * COREL[0][3];
    const TG C2 = COREL[0][0] * COREL[1][2] - COREL[1][0] * COREL[0][2];
    HS = tmax(tabs(C1), tabs(C2)) * AREA_I;
  }
  // ---- characteristic length (FACDT = 5/4, IDT1SH = 0) and FACN ----
  TG LM;
  {
    const TG FACDT = TG(1.25);
    const TG rx = XL2 + XL3 - XL4, ry = YL2 + YL3 - YL4;
    const TG sx = -XL2 + XL3 + XL4, sy = -YL2 + YL3 + YL4;
    const TG C1 = M<TG>::sqrt_(rx * rx + ry * ry), C2 = M<TG>::sqrt_(sx * sx + sy * sy);
    TG S1 = FOURTH * (M<TG>::div(tmax(C1, C2), tmin(C1, C2)) - ONE);
    const TG FAC1 = tmin(HALF, S1) + ONE;
    TG FAC2 = M<TG>::div(FOUR * AREA, C1 * C2);
    // default-real literals of czcorc.F:393-394, widened
    FAC2 = TG((double)3.413f) * tmax(ZERO, FAC2 - TG((double)0.7071f));
    FAC2 = TG((double)0.78f) + TG((double)0.22f) * FAC2 * FAC2 * FAC2;
    const TG FACI = TWO * FAC1 * FAC2;
    TG LL = tmax(L13, L24);
    LM = HALF * (L13 + L24);
    o.FACN[0] = M<TG>::sqrt_(M<TG>::div(L24, LL));
    o.FACN[1] = M<TG>::sqrt_(M<TG>::div(L13, LL));
    S1 = M<TG>::sqrt_(FACI * (FACDT + HS) * LL);
    S1 = tmax(S1, TG(k.em10));
    o.LL = M<TG>::div(AREA, S1);
  }
  TG RL[2][4];                 // RLXYZ, projected below for warped elements
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    RL[0][c] = E1X * vrg[0][c] + E1Y * vrg[1][c] + E1Z * vrg[2][c];
    RL[1][c] = E2X * vrg[0][c] + E2Y * vrg[1][c] + E2Z * vrg[2][c];
  }
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
      V13[q] = (E[q][0] * G13[0] + E[q][1] * G13[1] + E[q][2] * G13[2]);
      V24[q] = (E[q][0] * G24[0] + E[q][1] * G24[1] + E[q][2] * G24[2]);
      VHI[q] = (E[q][0] * GHI[0] + E[q][1] * GHI[1] + E[q][2] * GHI[2]);
    }
  }
  // ---- second-order rigid-rotation correction (IMPL_S = 0) ----
  {
    const TG DT05 = HALF * dt1, DT025 = FOURTH * dt1;
    const TG EXZ = Y24 * V13[2] - Y13 * V24[2];
    const TG EYZ = -X24 * V13[2] + X13 * V24[2];
    const TG DDRY = DT05 * EXZ * AREA_I;
    const TG DDRX = DT05 * EYZ * AREA_I;
    const TG V13X = V13[0], V24X = V24[0], VHIX = VHI[0];
    const TG EM10 = TG(k.em10);
    TG DDRZ1, DDRZ2;
    if (tabs(X13 - X24) < EM10) DDRZ1 = ZERO;
    else DDRZ1 = M<TG>::div(DT025 * (V13[1] - V24[1]), X13 - X24);
    V13[0] = V13[0] - DDRY * V13[2] - DDRZ1 * V13[1];
    V24[0] = V24[0] - DDRY * V24[2] - DDRZ1 * V24[1];
    VHI[0] = VHI[0] - DDRY * VHI[2] - DDRZ1 * VHI[1];
    if (tabs(Y13 + Y24) < EM10) DDRZ2 = ZERO;
    else DDRZ2 = M<TG>::div(DT025 * (V13X + V24X), Y13 + Y24);
    V13[1] = V13[1] - DDRX * V13[2] - DDRZ2 * V13X;
    V24[1] = V24[1] - DDRX * V24[2] - DDRZ2 * V24X;
    VHI[1] = VHI[1] - DDRX * VHI[2] - DDRZ2 * VHIX;
  }
  // ---- CZCORP5: flatness test against LM = (L13+L24)/2, then the projection ----
  TG(&VQN)[3][4] = o.VQN;
  TG *DI = o.DI;
#pragma unroll
  for (int c = 0; c < 4; ++c) { VQN[0][c] = ZERO; VQN[1][c] = ZERO; VQN[2][c] = ZERO; }
#pragma unroll
  for (int q = 0; q < 6; ++q) DI[q] = ZERO;
  const TG Z2 = Z1 * Z1;
  o.plat = (Z2 < LM * TG(k.em08) || npt == 1);
  if (o.plat) {
    Z1 = ZERO;
  } else {
    const TG A_4 = AREA * FOURTH;
    TG SZ1 = MX13 * Y24 - MY13 * X24;
    TG SZ2 = A_4 + SZ1;
    TG SZs = Z2 * L24;
    TG SL = M<TG>::div(ONE, M<TG>::sqrt_(SZs + SZ2 * SZ2));
    VQN[0][0] = -Z1 * Y24;
    VQN[1][0] = Z1 * X24;
    VQN[2][0] = SZ2 * SL;
    VQN[0][2] = -VQN[0][0];
    VQN[1][2] = -VQN[1][0];
    VQN[0][0] = VQN[0][0] * SL;
    VQN[1][0] = VQN[1][0] * SL;
    SZ2 = A_4 - SZ1;
    SL = M<TG>::div(ONE, M<TG>::sqrt_(SZs + SZ2 * SZ2));
    VQN[0][2] = VQN[0][2] * SL;
    VQN[1][2] = VQN[1][2] * SL;
    VQN[2][2] = SZ2 * SL;
    SZ1 = MX13 * Y13 - MY13 * X13;
    SZ2 = A_4 + SZ1;
    SZs = Z2 * L13;
    SL = M<TG>::div(ONE, M<TG>::sqrt_(SZs + SZ2 * SZ2));
    VQN[0][1] = -Z1 * Y13;
    VQN[1][1] = Z1 * X13;
    VQN[2][1] = SZ2 * SL;
    VQN[0][3] = -VQN[0][1];
    VQN[1][3] = -VQN[1][1];
    VQN[0][1] = VQN[0][1] * SL;
    VQN[1][1] = VQN[1][1] * SL;
    SZ2 = A_4 - SZ1;
    SL = M<TG>::div(ONE, M<TG>::sqrt_(SZs + SZ2 * SZ2));
    VQN[0][3] = VQN[0][3] * SL;
    VQN[1][3] = VQN[1][3] * SL;
    VQN[2][3] = SZ2 * SL;
    TG RR[3][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      RR[0][c] = RL[0][c];
      RR[1][c] = RL[1][c];
      RR[2][c] = E3X * vrg[0][c] + E3Y * vrg[1][c] + E3Z * vrg[2][c];
    }
    TG AR[3], AD[4];
    AR[0] = -Z1 * VHI[1] + Y13 * V13[2] + Y24 * V24[2] + MY13 * VHI[2] + RR[0][0] + RR[0][1] +
            RR[0][2] + RR[0][3];
    AR[1] = Z1 * VHI[0] - X13 * V13[2] - X24 * V24[2] - MX13 * VHI[2] + RR[1][0] + RR[1][1] +
            RR[1][2] + RR[1][3];
    AR[2] = X13 * V13[1] + X24 * V24[1] + MX13 * VHI[1] - Y13 * V13[0] - Y24 * V24[0] -
            MY13 * VHI[0] + RR[2][0] + RR[2][1] + RR[2][2] + RR[2][3];
#pragma unroll
    for (int c = 0; c < 4; ++c)
      AD[c] = VQN[0][c] * RR[0][c] + VQN[1][c] * RR[1][c] + VQN[2][c] * RR[2][c];
    const TG XX = COREL[0][0] * COREL[0][0] + COREL[0][1] * COREL[0][1] +
                  COREL[0][2] * COREL[0][2] + COREL[0][3] * COREL[0][3];
    const TG YY = COREL[1][0] * COREL[1][0] + COREL[1][1] * COREL[1][1] +
                  COREL[1][2] * COREL[1][2] + COREL[1][3] * COREL[1][3];
    const TG XY = COREL[0][0] * COREL[1][0] + COREL[0][1] * COREL[1][1] +
                  COREL[0][2] * COREL[1][2] + COREL[0][3] * COREL[1][3];
    const TG XZ = (COREL[0][0] - COREL[0][1] + COREL[0][2] - COREL[0][3]) * Z1;
    const TG YZ = (COREL[1][0] - COREL[1][1] + COREL[1][2] - COREL[1][3]) * Z1;
    const TG ZZ = FOUR * Z2;
    TG BTB[6];
    BTB[0] = VQN[0][0] * VQN[0][0] + VQN[0][1] * VQN[0][1] + VQN[0][2] * VQN[0][2] + VQN[0][3] * VQN[0][3];
    BTB[1] = VQN[1][0] * VQN[1][0] + VQN[1][1] * VQN[1][1] + VQN[1][2] * VQN[1][2] + VQN[1][3] * VQN[1][3];
    BTB[2] = VQN[2][0] * VQN[2][0] + VQN[2][1] * VQN[2][1] + VQN[2][2] * VQN[2][2] + VQN[2][3] * VQN[2][3];
    BTB[3] = VQN[0][0] * VQN[1][0] + VQN[0][1] * VQN[1][1] + VQN[0][2] * VQN[1][2] + VQN[0][3] * VQN[1][3];
    BTB[4] = VQN[0][0] * VQN[2][0] + VQN[0][1] * VQN[2][1] + VQN[0][2] * VQN[2][2] + VQN[0][3] * VQN[2][3];
    BTB[5] = VQN[1][0] * VQN[2][0] + VQN[1][1] * VQN[2][1] + VQN[1][2] * VQN[2][2] + VQN[1][3] * VQN[2][3];
    TG D[6];
    D[0] = YY + ZZ + FOUR - BTB[0];
    D[1] = XX + ZZ + FOUR - BTB[1];
    D[2] = XX + YY + FOUR - BTB[2];
    D[3] = -XY - BTB[3];
    D[4] = -XZ - BTB[4];
    D[5] = -YZ - BTB[5];
    {  // A3INVDP_V, the symmetric 3x3 inverse (IRESP = 0)
      const TG ABC = D[0] * D[1] * D[2];
      const TG XXYZ2 = D[0] * D[5] * D[5];
      const TG YYXZ2 = D[1] * D[4] * D[4];
      const TG ZZXY2 = D[2] * D[3] * D[3];
      TG DETA = tabs(ABC + TWO * D[3] * D[4] * D[5] - XXYZ2 - YYXZ2 - ZZXY2);
      DETA = M<TG>::div(ONE, tmax(DETA, EM20));
      DI[0] = M<TG>::div((ABC - XXYZ2) * DETA, tmax(D[0], EM20));
      DI[1] = M<TG>::div((ABC - YYXZ2) * DETA, tmax(D[1], EM20));
      DI[2] = M<TG>::div((ABC - ZZXY2) * DETA, tmax(D[2], EM20));
      DI[3] = (D[4] * D[5] - D[3] * D[2]) * DETA;
      DI[4] = (D[5] * D[3] - D[4] * D[1]) * DETA;
      DI[5] = (D[3] * D[4] - D[5] * D[0]) * DETA;
    }
    TG DB[3][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      DB[0][c] = DI[0] * VQN[0][c] + DI[3] * VQN[1][c] + DI[4] * VQN[2][c];
      DB[1][c] = DI[3] * VQN[0][c] + DI[1] * VQN[1][c] + DI[5] * VQN[2][c];
      DB[2][c] = DI[4] * VQN[0][c] + DI[5] * VQN[1][c] + DI[2] * VQN[2][c];
    }
    TG DBAD[3], ALR[3], ALD[4];
#pragma unroll
    for (int q = 0; q < 3; ++q)
      DBAD[q] = DB[q][0] * AD[0] + DB[q][1] * AD[1] + DB[q][2] * AD[2] + DB[q][3] * AD[3];
    ALR[0] = DI[0] * AR[0] + DI[3] * AR[1] + DI[4] * AR[2] - DBAD[0];
    ALR[1] = DI[3] * AR[0] + DI[1] * AR[1] + DI[5] * AR[2] - DBAD[1];
    ALR[2] = DI[4] * AR[0] + DI[5] * AR[1] + DI[2] * AR[2] - DBAD[2];
#pragma unroll
    for (int c = 0; c < 4; ++c)
      ALD[c] = AD[c] + VQN[0][c] * DBAD[0] + VQN[1][c] * DBAD[1] + VQN[2][c] * DBAD[2] -
               DB[0][c] * AR[0] - DB[1][c] * AR[1] - DB[2][c] * AR[2];
    const TG C1 = TWO * ALR[2];
    V13[0] = V13[0] + C1 * Y13;
    V24[0] = V24[0] + C1 * Y24;
    VHI[0] = VHI[0] + FOUR * (ALR[2] * MY13 - Z1 * ALR[1]);
    V13[1] = V13[1] - C1 * X13;
    V24[1] = V24[1] - C1 * X24;
    VHI[1] = VHI[1] - FOUR * (ALR[2] * MX13 - Z1 * ALR[0]);
    V13[2] = V13[2] - TWO * (Y13 * ALR[0] - X13 * ALR[1]);
    V24[2] = V24[2] - TWO * (Y24 * ALR[0] - X24 * ALR[1]);
    VHI[2] = VHI[2] + FOUR * (MX13 * ALR[1] - MY13 * ALR[0]);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      RL[0][c] = RR[0][c] - ALR[0] - VQN[0][c] * ALD[c];
      RL[1][c] = RR[1][c] - ALR[1] - VQN[1][c] * ALD[c];
    }
  }
  o.Z1 = Z1;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    V13[q] = V13[q] * AREA_I;
    V24[q] = V24[q] * AREA_I;
    VHI[q] = VHI[q] * FOURTH;
  }
  // ---- CZDEF ----
  {
    TG *VDEF = o.VDEF, *DHG = o.DHG;
    const TG MX23 = o.MX23, MX34 = o.MX34, MY23 = o.MY23, MY34 = o.MY34;
    const TG AA = AREA_I;
    TG R13[2], R24[2], RSOM[2], RHI[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      R13[q] = (RL[q][0] - RL[q][2]) * AA;
      R24[q] = (RL[q][1] - RL[q][3]) * AA;
      RSOM[q] = (RL[q][3] + RL[q][2] + RL[q][0] + RL[q][1]) * AA;
      RHI[q] = (RL[q][0] - RL[q][1] + RL[q][2] - RL[q][3]) * FOURTH;
    }
    VDEF[0] = Y24 * V13[0] - Y13 * V24[0];
    VDEF[1] = -X24 * V13[1] + X13 * V24[1];
    const TG BXV2 = Y24 * V13[1] - Y13 * V24[1];
    const TG BYV1 = -X24 * V13[0] + X13 * V24[0];
    VDEF[2] = BXV2 + BYV1;
    VDEF[5] = Y24 * R13[1] - Y13 * R24[1];
    VDEF[6] = X24 * R13[0] - X13 * R24[0];
    const TG BXR1 = Y13 * R24[0] - Y24 * R13[0];
    const TG BYR2 = -X24 * R13[1] + X13 * R24[1];
    VDEF[7] = BXR1 + BYR2;
    const TG BCXY = AREA * FOURTH;
    const TG BCX = V13[2] - MY13 * R13[0] + MX13 * R13[1];
    const TG BCY = V24[2] + MY13 * R24[0] - MX13 * R24[1];
    VDEF[3] = Y24 * BCX - Y13 * BCY + BCXY * RSOM[1];
    VDEF[4] = X13 * BCY - X24 * BCX - BCXY * RSOM[0];
    DHG[0] = VHI[0] - MX13 * VDEF[0] - MY13 * BYV1;
    DHG[1] = VHI[1] - MX13 * BXV2 - MY13 * VDEF[1];
    DHG[2] = RHI[1] - MX13 * VDEF[5] - MY13 * BYR2;
    DHG[3] = -RHI[0] - MX13 * BXR1 - MY13 * VDEF[6];
    DHG[4] = (VHI[2] * FOUR - (MY13 * RSOM[0] - MY23 * (R13[0] + R24[0]) + MX23 * (R13[1] + R24[1]) -
                               MX13 * RSOM[1]) * AREA) * FOUR;
    DHG[5] = (VHI[2] * FOUR - (MY13 * RSOM[0] - MY34 * (R13[0] - R24[0]) + MX34 * (R13[1] - R24[1]) -
                               MX13 * RSOM[1]) * AREA) * FOUR;
    DHG[0] = DHG[0] + (Y24 * V13[2] - Y13 * V24[2]) * Z1;
    DHG[1] = DHG[1] + (-X24 * V13[2] + X13 * V24[2]) * Z1;
    const TG DETA1 = Z1 * FOUR * AA;
    VDEF[5] = VDEF[5] + (X13 * V13[0] - X24 * V24[0]) * DETA1;
    VDEF[6] = VDEF[6] + (Y13 * V13[1] - Y24 * V24[1]) * DETA1;
    VDEF[7] = VDEF[7] + (X13 * V13[1] - X24 * V24[1] + Y13 * V13[0] - Y24 * V24[0]) * DETA1;
    if (offg < ZERO) {
#pragma unroll
      for (int q = 0; q < 8; ++q) VDEF[q] = ZERO;
#pragma unroll
      for (int q = 0; q < 6; ++q) DHG[q] = ZERO;
    }
  }
}

// IO: get(col) / put(col, v) on this element's state columns (SC_NCOL_QEPH of them).
// emit(k, fx, fy, fz, mx, my, mz): slot values (FSKY(1:6), already negated) of corner k.
template <typename T, typename TG, typename IO, typename Emit>
__device__ __forceinline__ void qeph_chain(const TG xg[3][4], const TG vg[3][4],
                                           const TG vrg[3][4], T thk0, const IO &io,
                                           double dt1d, const ShMat &m, const T jc[3],
                                           const ShProp &p, const ShConst &k, QephOut<T> &out,
                                           Emit emit) {
  const T ZERO = T(0), ONE = T(1), TWO = T(2), THREE = T(3), FOUR = T(4), HALF = T(0.5),
          FOURTH = T(0.25);
  const T EM20 = T(k.em20);
  const T dt1 = T(dt1d);
  T offg = io.get(SC_OFFG);
  QephKin<TG> g;
  qeph_kinematics<TG>(xg, vg, vrg, TG(dt1d), as_kin<TG>(offg), p.npt, k, g);
  if (g.degenerate) {
    offg = ZERO;
    io.put(SC_OFFG, offg);
  }
  const T AREA = T(g.AREA), AREA_I = T(g.AREA_I);
  const T X13 = T(g.X13), X24 = T(g.X24), Y13 = T(g.Y13), Y24 = T(g.Y24);
  const T MX13 = T(g.MX13), MX23 = T(g.MX23), MX34 = T(g.MX34);
  const T MY13 = T(g.MY13), MY23 = T(g.MY23), MY34 = T(g.MY34), Z1 = T(g.Z1);
  // ---- CNCOEF3B, IGTYP = 1, law 2, ITHK = 0 (moduli from PM, resol_init.F:689-696) ----
  const T thk02 = thk0 * thk0, VOL0 = thk0 * AREA;
  const T RHO = T(m.rho0), G = T(m.shear);
  const T A11 = T(m.a11), A12 = T(m.a12), SSP = T(m.ssp);
  const T GSR = M<T>::sqrt_(tabs(G)), A11SR = M<T>::sqrt_(tabs(A11)),
          A12SR = M<T>::sqrt_(tabs(A12));
  const T SHF = T(p.fsh), SHFSR = M<T>::sqrt_(T(p.fsh));
  const T GS = G * SHF;
  const T AMU = T(p.dn);
  // ---- CZDEF (in the kinematics stage) ----
  T VDEF[8], DHG[6];
#pragma unroll
  for (int q = 0; q < 8; ++q) VDEF[q] = T(g.VDEF[q]);
#pragma unroll
  for (int q = 0; q < 6; ++q) DHG[q] = T(g.DHG[q]);
  T off = tmin(ONE, tabs(offg));
  // ---- CZSTRA3 ----
  const T EXX = VDEF[0] * dt1, EYY = VDEF[1] * dt1, EXY = VDEF[2] * dt1;
  const T EYZ = VDEF[4] * dt1, EXZ = VDEF[3] * dt1;
  const T KXX = VDEF[5] * dt1, KYY = VDEF[6] * dt1, KXY = VDEF[7] * dt1;
  // ---- element strain rate ----
  const T dtinv = M<T>::div(dt1, tmax(dt1 * dt1, EM20));
  const T thk = io.get(SC_THK);
  T epsd_pg;
  {
    const T eps_k2 = (KXX * KXX + KYY * KYY + KXX * KYY + FOURTH * (KXY * KXY)) *
                     T(k.one_over_9) * (thk * thk);
    const T eps_m2 = T(k.four_over_3) * (EXX * EXX + EYY * EYY + EXX * EYY + FOURTH * (EXY * EXY));
    epsd_pg = M<T>::sqrt_(eps_k2 + eps_m2) * dtinv;
    io.put(SC_EPSD, ONE * epsd_pg + (ONE - ONE) * io.get(SC_EPSD));
  }
  // ---- MULAWC, law 2, with ZCFAC (mean and minimum of ET/E) and SIGY (mean yield) ----
  T FOR[5], MOM[3];
#pragma unroll
  for (int c = 0; c < 5; ++c) FOR[c] = io.get(SC_FOR + c);
#pragma unroll
  for (int c = 0; c < 3; ++c) MOM[c] = io.get(SC_MOM + c);
  T degmb = FOR[0] * EXX + FOR[1] * EYY + FOR[2] * EXY + FOR[3] * EYZ + FOR[4] * EXZ;
  T degfx = MOM[0] * KXX + MOM[1] * KYY + MOM[2] * KXY;
  T thkn = thk;
#pragma unroll
  for (int c = 0; c < 5; ++c) FOR[c] = ZERO;
#pragma unroll
  for (int c = 0; c < 3; ++c) MOM[c] = ZERO;
  T sigy = ZERO, zcfac1 = ZERO, zcfac2 = ONE;
  const T off_old = off;
  T epchk = ZERO;
  T viscmx = ZERO;
  const T asrate = (m.israte > 0) ? tmin(ONE, T(m.asrate_pm9) * dt1) : ONE;
  const T fnpt = T(p.npt);
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
    T s[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) s[c] = io.get(SC_SIG + 5 * it + c);
    T pla = io.get(SC_PLA + it), epsdp = io.get(SC_EPSDP + it);
    T etse;
    shell_law2_point<T>(m, jc, k, GS, dt1, asrate, epsd_pg, off_old, off, pla, epsdp, epchk, thkn,
                        thklyl, deps, s, etse, sigy, fnpt);
    viscmx = tmax(T(p.dm), viscmx);
#pragma unroll
    for (int c = 0; c < 5; ++c) io.put(SC_SIG + 5 * it + c, s[c] * ONE);
    io.put(SC_PLA + it, pla);
    io.put(SC_EPSDP + it, epsdp);
#pragma unroll
    for (int c = 0; c < 5; ++c) FOR[c] = FOR[c] + wf * s[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) MOM[c] = MOM[c] + wm * s[c];
    zcfac1 = zcfac1 + M<T>::div(etse, fnpt);
    zcfac2 = tmin(etse, zcfac2);
  }
  io.put(SC_THK, tmax(thkn, T(k.em30)));
  {
    const T fact = T(k.onep414) * T(p.dm);
    const T visc = fact * SSP * M<T>::sqrt_(AREA) * dtinv * RHO;
    FOR[0] = FOR[0] + visc * (EXX + HALF * EYY);
    FOR[1] = FOR[1] + visc * (EYY + HALF * EXX);
    FOR[2] = FOR[2] + visc * EXY * T(k.third);
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) FOR[c] = FOR[c] * off * ONE;
#pragma unroll
  for (int c = 0; c < 3; ++c) MOM[c] = MOM[c] * off * ONE;
  degmb = degmb + FOR[0] * EXX + FOR[1] * EYY + FOR[2] * EXY + FOR[3] * EYZ + FOR[4] * EXZ;
  degfx = degfx + MOM[0] * KXX + MOM[1] * KYY + MOM[2] * KXY;
  T eint0, eint1;
  {
    const T vol2 = HALF * (AREA * thk0);
    eint0 = io.get(SC_EINT) + degmb * vol2;
    eint1 = io.get(SC_EINT + 1) + degfx * thk0 * vol2;
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) io.put(SC_FOR + c, FOR[c]);
#pragma unroll
  for (int c = 0; c < 3; ++c) io.put(SC_MOM + c, MOM[c]);
  // ---- CNDT3, NODADT = 0, IDTMINS = 0, IDTMIN(3) = 2, ZOFFSET = 0 (F_OSET = F_DTE = 1) ----
  viscmx = tmax(viscmx, AMU);
  viscmx = M<T>::sqrt_(ONE + viscmx * viscmx) - viscmx;
  T ALDT = M<T>::div(T(g.LL) * viscmx, M<T>::sqrt_(ONE));
  const T DT = M<T>::div(T(p.dtfac) * ALDT, SSP);
  if (DT <= T(p.dtmin) && off >= ONE && offg != TWO && offg >= ZERO) off = ZERO;
  out.dt = DT;
  out.alive = (offg > ZERO && off != ZERO);
  const T sti = M<T>::div(HALF * VOL0 * A11 * off, tmax(ALDT * ALDT, EM20));
  // ---- CZFINTCE ----
  T VF[3][4], VM[2][4];
  {
    const T THK = thk0, C1 = thk02, A_I = AREA_I;
    const T *VS = FOR, *MS = MOM;
    const T X13S8 = X13 * MS[2], X24S8 = X24 * MS[2], Y13S8 = Y13 * MS[2], Y24S8 = Y24 * MS[2];
    const T S1 = (MY34 * MX23 - MY23 * MX34) * THK;
    const T S42S = S1 * VS[4], S52S = S1 * VS[3];
    VF[0][0] = THK * (Y24 * VS[0] - X24 * VS[2]);
    VF[1][0] = THK * (-X24 * VS[1] + Y24 * VS[2]);
    VF[2][0] = THK * (-X24 * VS[3] + Y24 * VS[4]);
    VM[0][0] = C1 * (X24 * MS[1] - Y24S8) - MY13 * VF[2][0];
    VM[1][0] = C1 * (Y24 * MS[0] - X24S8) + MX13 * VF[2][0];
    VM[0][2] = -S52S;
    VM[1][2] = S42S;
    VF[0][1] = THK * (-Y13 * VS[0] + X13 * VS[2]);
    VF[1][1] = THK * (X13 * VS[1] - Y13 * VS[2]);
    VF[2][1] = THK * (X13 * VS[3] - Y13 * VS[4]);
    VM[0][1] = C1 * (-X13 * MS[1] + Y13S8) + MY13 * VF[2][1];
    VM[1][1] = C1 * (-Y13 * MS[0] + X13S8) - MX13 * VF[2][1];
    VM[0][3] = VM[0][2];
    VM[1][3] = VM[1][2];
    const T C2 = C1 * Z1 * FOUR * A_I;
    VF[0][0] = VF[0][0] + C2 * (X13 * MS[0] + Y13S8);
    VF[1][0] = VF[1][0] + C2 * (Y13 * MS[1] + X13S8);
    VF[0][1] = VF[0][1] - C2 * (X24 * MS[0] + Y24S8);
    VF[1][1] = VF[1][1] - C2 * (Y24 * MS[1] + X24S8);
    VF[0][2] = ZERO; VF[0][3] = ZERO; VF[1][2] = ZERO; VF[1][3] = ZERO;
    VF[2][2] = ZERO; VF[2][3] = ZERO;
  }
  // ---- CZFINTN1, MTN = 2: physical hourglass, HOURG(1:12) ----
  T TESY;
  {
    T vgl[13];
#pragma unroll
    for (int q = 1; q <= 5; ++q) vgl[q] = io.get(SC_HOUR + q - 1);
#pragma unroll
    for (int q = 6; q <= 12; ++q) vgl[q] = io.get(SC_HOURX + q - 6);
    const T THK = thk0, C1 = thk02, AA = AREA_I, FAC1 = T(p.cvis);
    const T C7 = T(k.four_over_3), COEFH = T(k.zep999), COEF = T(k.zep85), STIER = T(k.fivep333);
    const T TOLH = T(k.em18);
    T FBEND_V = T(k.threep464);
    const T COEF1 = T(25);   // NPT >= 1
    T FBEND;
    if (p.npt == 1) { FBEND = ZERO; FBEND_V = ZERO; } else { FBEND = T(k.one_over_12); }
    const T UNDOUZSR = M<T>::sqrt_(T(k.one_over_12));
    T dhg[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) dhg[q] = DHG[q] * dt1;
    T C3 = FOUR * AA;
    const T HXX = C3 * MY34, HYY = C3 * MX34, HXX_K = C3 * MY23, HYY_K = C3 * MX23;
    const T CXX = HXX * dhg[0], CYY = HYY * dhg[1], CXX_K = HXX_K * dhg[0], CYY_K = HYY_K * dhg[1];
    const T BXX = HXX * dhg[2], BYY = HYY * dhg[3], BXX_K = HXX_K * dhg[2], BYY_K = HYY_K * dhg[3];
    T C1M = A11 * FAC1, C2M = A12 * FAC1;
    const T C6 = C1 * FBEND;
    T SS1 = MY34 * vgl[1] + MY23 * vgl[7];
    T SS2 = MX23 * vgl[8] + MX34 * vgl[2];
    T SF1 = MY34 * vgl[3] + MY23 * vgl[9];
    T SF2 = -MX23 * vgl[10] - MX34 * vgl[4];
    T SC5 = MY34 * vgl[5] + MX34 * vgl[6];
    T SC6 = MY23 * vgl[11] + MX23 * vgl[12];
    const T C5 = HALF * off * THK * C7;
    const T ESX = SS1 * dhg[0] + SS2 * dhg[1];
    T ETMP1 = C5 * (ESX + FOURTH * (SC5 * dhg[4] + SC6 * dhg[5]));
    const T EMX = (SF1 * dhg[2] - SF2 * dhg[3]) * C6;
    T ETMP2 = C5 * EMX;
    T DG[13];
    DG[1] = C1M * CXX - C2M * CYY;
    DG[2] = C1M * CYY - C2M * CXX;
    DG[3] = C1M * BXX - C2M * BYY;
    DG[4] = C1M * BYY - C2M * BXX;
    DG[7] = C1M * CXX_K - C2M * CYY_K;
    DG[8] = C1M * CYY_K - C2M * CXX_K;
    DG[9] = C1M * BXX_K - C2M * BYY_K;
    DG[10] = C1M * BYY_K - C2M * BXX_K;
    T C2 = FAC1 * G * SHF * T(k.one_over_64);
    DG[5] = C2 * HXX * dhg[4];
    DG[6] = C2 * HYY * dhg[4];
    DG[11] = C2 * HXX_K * dhg[5];
    DG[12] = C2 * HYY_K * dhg[5];
#pragma unroll
    for (int q = 1; q <= 12; ++q) vgl[q] = vgl[q] + DG[q];
    eint0 = eint0 + ETMP1;
    eint1 = eint1 + ETMP2;
    if (sigy < T(k.zep9ep30)) {
      const T UFAC = tabs(tmin(zcfac1, zcfac2) - ONE);
      const T SIGY2 = sigy * sigy;
      T SVM = ZERO, SXY0 = ZERO;
      if (UFAC < TOLH) {
        const T *VS = FOR, *MS = MOM;
        SXY0 = VS[0] * VS[0] + VS[1] * VS[1] - VS[0] * VS[1] + THREE * VS[2] * VS[2];
        T MXY0 = MS[0] * MS[0] + MS[1] * MS[1] - MS[0] * MS[1] + THREE * MS[2] * MS[2];
        const T CNN = COEF, CMM = COEF * THK * T(k.one_over_16);
        const T CNNX = CNN * vgl[1], CNNY = CNN * vgl[2], CNNX_K = CNN * vgl[7], CNNY_K = CNN * vgl[8];
        const T CMMX = CMM * vgl[3], CMMY = CMM * vgl[4], CMMX_K = CMM * vgl[9], CMMY_K = CMM * vgl[10];
        SXY0 = SXY0 + CNNX * CNNX + CNNY * CNNY - CNNX * CNNY;
        MXY0 = MXY0 + CMMX * CMMX + CMMY * CMMY - CMMX * CMMY;
        SXY0 = SXY0 + CNNX_K * CNNX_K + CNNY_K * CNNY_K - CNNX_K * CNNY_K;
        MXY0 = MXY0 + CMMX_K * CMMX_K + CMMY_K * CMMY_K - CMMX_K * CMMY_K;
        SXY0 = SXY0 + tabs(CNNX * (TWO * CNNX_K - CNNY_K) + CNNY * (TWO * CNNY_K - CNNX_K));
        MXY0 = MXY0 + tabs(CMMX * (TWO * CMMX_K - CMMY_K) + CMMY * (TWO * CMMY_K - CMMX_K));
        SVM = SXY0 + COEF1 * MXY0;
      }
      if (UFAC >= TOLH || SVM > SIGY2) {
        T EH1 = tmin(M<T>::div(SXY0, tmax(SIGY2, TOLH)), ONE);
        EH1 = tmax(COEFH * EH1, (ONE - zcfac1));
        T EH2 = tmax(COEFH, (ONE - zcfac2));
        if (ESX < ZERO) EH1 = ZERO;
        if (EMX < ZERO) EH2 = ZERO;
        vgl[1] = vgl[1] - EH1 * DG[1];
        vgl[2] = vgl[2] - EH1 * DG[2];
        vgl[7] = vgl[7] - EH1 * DG[7];
        vgl[8] = vgl[8] - EH1 * DG[8];
        vgl[3] = vgl[3] - EH2 * DG[3];
        vgl[4] = vgl[4] - EH2 * DG[4];
        vgl[9] = vgl[9] - EH2 * DG[9];
        vgl[10] = vgl[10] - EH2 * DG[10];
      }
    }
    const T C8 = C7 * off;
    SS1 = (MY34 * vgl[1] + MY23 * vgl[7]) * C8;
    SS2 = (MX23 * vgl[8] + MX34 * vgl[2]) * C8;
    SF1 = (MY34 * vgl[3] + MY23 * vgl[9]) * C8;
    SF2 = -(MX23 * vgl[10] + MX34 * vgl[4]) * C8;
    const T HSURA = THK * AA;
    C2 = C8 * THK;
    SC5 = (MY34 * vgl[5] + MX34 * vgl[6]) * C2;
    SC6 = (MY23 * vgl[11] + MX23 * vgl[12]) * C2;
    T SS3 = SC5 + SC6;
    const T HVL = AMU * M<T>::sqrt_(RHO * AREA * FAC1) * off;
    const T SSV0 = MY23 * MY23, SSV1 = MY34 * MY34, SSV2 = MX23 * MX23, SSV3 = MX34 * MX34;
    const T HXX_V = STIER * (SSV1 + SSV0);
    const T HXY_V = -STIER * (MY34 * MX34 + MY23 * MX23);
    const T HYY_V = STIER * (SSV2 + SSV3);
    C2 = HVL * GSR * SHFSR * UNDOUZSR;
    const T CXZ_V = (SSV1 + SSV3) * C2;
    const T CYZ_V = (SSV2 + SSV0) * C2;
    const T AUX = AA * HVL;
    C1M = A11SR * AUX;
    C2M = A12SR * AUX;
    const T CXX_V = C1M * HXX_V, CYY_V = C1M * HYY_V, CXY_V = C2M * HXY_V;
    const T SS1_V = CXX_V * DHG[0] + CXY_V * DHG[1];
    const T SS2_V = CYY_V * DHG[1] + CXY_V * DHG[0];
    const T SF1_V = (CXX_V * DHG[2] + CXY_V * DHG[3]) * FBEND_V;
    const T SF2_V = (-CYY_V * DHG[3] - CXY_V * DHG[2]) * FBEND_V;
    const T SC5_V = CXZ_V * DHG[4] * HSURA;
    const T SC6_V = CYZ_V * DHG[5] * HSURA;
    const T SS3_V = SC5_V + SC6_V;
    SS1 = SS1 + SS1_V;
    SS2 = SS2 + SS2_V;
    SS3 = SS3 + SS3_V;
    SC5 = SC5 + SC5_V;
    SC6 = SC6 + SC6_V;
    SF1 = SF1 + SF1_V;
    SF2 = SF2 + SF2_V;
    const T Y13S = MY13 * SS3, X13S = MX13 * SS3;
    const T Y34S6 = MY34 * SC6, Y23S5 = MY23 * SC5, X23S5 = MX23 * SC5, X34S6 = MX34 * SC6;
    C2 = FOURTH * THK;
    const T B13 = (MY13 * X24 - MX13 * Y24) * HSURA;
    VF[0][0] = VF[0][0] + B13 * SS1;
    VF[0][2] = C2 * SS1;
    VF[1][0] = VF[1][0] + B13 * SS2;
    VF[1][2] = C2 * SS2;
    VF[2][2] = SS3;
    const T B24 = (MX13 * Y13 - MY13 * X13) * HSURA;
    VF[0][1] = VF[0][1] + B24 * SS1;
    VF[0][3] = -VF[0][2];
    VF[1][1] = VF[1][1] + B24 * SS2;
    VF[1][3] = -VF[1][2];
    VF[2][3] = -VF[2][2];
    C3 = C6 * B13;
    const T C4 = C6 * C2;
    VM[0][0] = VM[0][0] + C3 * SF2 + Y23S5 + Y34S6;
    VM[0][2] = VM[0][2] + C4 * SF2 - Y13S;
    VM[1][0] = VM[1][0] + C3 * SF1 - X23S5 - X34S6;
    VM[1][2] = VM[1][2] + C4 * SF1 + X13S;
    C3 = C6 * B24;
    VM[0][1] = VM[0][1] + C3 * SF2 + Y23S5 - Y34S6;
    VM[0][3] = VM[0][3] - C4 * SF2 - Y13S;
    VM[1][1] = VM[1][1] + C3 * SF1 - X23S5 + X34S6;
    VM[1][3] = VM[1][3] - C4 * SF1 + X13S;
    C2 = Z1 * HSURA;
    VF[2][0] = VF[2][0] + C2 * (SS1 * Y24 - SS2 * X24);
    VF[2][1] = VF[2][1] + C2 * (-SS1 * Y13 + SS2 * X13);
    const T ESY = ((SS1 - SS1_V) * dhg[0] + (SS2 - SS2_V) * dhg[1]) * THK +
                  FOURTH * ((SC5 - SC5_V) * dhg[4] + (SC6 - SC6_V) * dhg[5]);
    ETMP1 = HALF * ESY;
    const T EMY = (SF1 - SF1_V) * dhg[2] - (SF2 - SF2_V) * dhg[3];
    ETMP2 = HALF * C6 * EMY * THK;
    TESY = (SS1_V * dhg[0] + SS2_V * dhg[1]) * THK + (SF1_V * dhg[2] - SF2_V * dhg[3]) * THK * C6 +
           FOURTH * (SC5_V * dhg[4] + SC6_V * dhg[5]);
    eint0 = eint0 + ETMP1;
    eint1 = eint1 + ETMP2;
#pragma unroll
    for (int q = 1; q <= 5; ++q) io.put(SC_HOUR + q - 1, vgl[q]);
#pragma unroll
    for (int q = 6; q <= 12; ++q) io.put(SC_HOURX + q - 6, vgl[q]);
  }
  io.put(SC_EINT, eint0);
  io.put(SC_EINT + 1, eint1);
  out.ehou = TESY;
  // ---- CZPROJ1 -> CZPROJN (IFINI = 0): symmetric/antisymmetric parts to corners,
  // the warped-element projection, then the global frame ----
  T F[3][4], Mo[3][4];
  {
    const T e1[3] = {T(g.e1[0]), T(g.e1[1]), T(g.e1[2])};
    const T e2[3] = {T(g.e2[0]), T(g.e2[1]), T(g.e2[2])};
    const T e3[3] = {T(g.e3[0]), T(g.e3[1]), T(g.e3[2])};
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
    if (g.plat) {
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          F[d][c] = e1[d] * FL[0][c] + e2[d] * FL[1][c] + e3[d] * FL[2][c];
          Mo[d][c] = e1[d] * ML[0][c] + e2[d] * ML[1][c];
        }
    } else {
      T COREL[2][4], VQN[3][4], DI[6];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        COREL[0][c] = T(g.COREL[0][c]); COREL[1][c] = T(g.COREL[1][c]);
        VQN[0][c] = T(g.VQN[0][c]); VQN[1][c] = T(g.VQN[1][c]); VQN[2][c] = T(g.VQN[2][c]);
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) DI[q] = T(g.DI[q]);
      T DB[3][4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        DB[0][c] = DI[0] * VQN[0][c] + DI[3] * VQN[1][c] + DI[4] * VQN[2][c];
        DB[1][c] = DI[3] * VQN[0][c] + DI[1] * VQN[1][c] + DI[5] * VQN[2][c];
        DB[2][c] = DI[4] * VQN[0][c] + DI[5] * VQN[1][c] + DI[2] * VQN[2][c];
      }
      T AR[3], AD[4], DBAD[3], ALR[3], ALD[4], MM[3][4];
      AR[0] = -Z1 * (FL[1][0] - FL[1][1] + FL[1][2] - FL[1][3]) + COREL[1][0] * FL[2][0] +
              ML[0][0] + COREL[1][1] * FL[2][1] + ML[0][1] + COREL[1][2] * FL[2][2] + ML[0][2] +
              COREL[1][3] * FL[2][3] + ML[0][3];
      AR[1] = Z1 * (FL[0][0] - FL[0][1] + FL[0][2] - FL[0][3]) - COREL[0][0] * FL[2][0] +
              ML[1][0] - COREL[0][1] * FL[2][1] + ML[1][1] - COREL[0][2] * FL[2][2] + ML[1][2] -
              COREL[0][3] * FL[2][3] + ML[1][3];
      AR[2] = -COREL[1][0] * FL[0][0] + COREL[0][0] * FL[1][0] - COREL[1][1] * FL[0][1] +
              COREL[0][1] * FL[1][1] - COREL[1][2] * FL[0][2] + COREL[0][2] * FL[1][2] -
              COREL[1][3] * FL[0][3] + COREL[0][3] * FL[1][3];
#pragma unroll
      for (int c = 0; c < 4; ++c) AD[c] = VQN[0][c] * ML[0][c] + VQN[1][c] * ML[1][c];
#pragma unroll
      for (int q = 0; q < 3; ++q)
        DBAD[q] = DB[q][0] * AD[0] + DB[q][1] * AD[1] + DB[q][2] * AD[2] + DB[q][3] * AD[3];
      ALR[0] = DI[0] * AR[0] + DI[3] * AR[1] + DI[4] * AR[2] - DBAD[0];
      ALR[1] = DI[3] * AR[0] + DI[1] * AR[1] + DI[5] * AR[2] - DBAD[1];
      ALR[2] = DI[4] * AR[0] + DI[5] * AR[1] + DI[2] * AR[2] - DBAD[2];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        ALD[c] = AD[c] + VQN[0][c] * DBAD[0] + VQN[1][c] * DBAD[1] + VQN[2][c] * DBAD[2] -
                 DB[0][c] * AR[0] - DB[1][c] * AR[1] - DB[2][c] * AR[2];
      T C1 = Z1 * ALR[1];
      FL[0][0] = FL[0][0] - C1 + COREL[1][0] * ALR[2];
      FL[0][1] = FL[0][1] + C1 + COREL[1][1] * ALR[2];
      FL[0][2] = FL[0][2] - C1 + COREL[1][2] * ALR[2];
      FL[0][3] = FL[0][3] + C1 + COREL[1][3] * ALR[2];
      C1 = Z1 * ALR[0];
      FL[1][0] = FL[1][0] + C1 - COREL[0][0] * ALR[2];
      FL[1][1] = FL[1][1] - C1 - COREL[0][1] * ALR[2];
      FL[1][2] = FL[1][2] + C1 - COREL[0][2] * ALR[2];
      FL[1][3] = FL[1][3] - C1 - COREL[0][3] * ALR[2];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        FL[2][c] = FL[2][c] - COREL[1][c] * ALR[0] + COREL[0][c] * ALR[1];
        MM[0][c] = ML[0][c] - ALR[0] - VQN[0][c] * ALD[c];
        MM[1][c] = ML[1][c] - ALR[1] - VQN[1][c] * ALD[c];
        MM[2][c] = -ALR[2] - VQN[2][c] * ALD[c];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          F[d][c] = e1[d] * FL[0][c] + e2[d] * FL[1][c] + e3[d] * FL[2][c];
          Mo[d][c] = e1[d] * MM[0][c] + e2[d] * MM[1][c] + e3[d] * MM[2][c];
        }
    }
  }
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

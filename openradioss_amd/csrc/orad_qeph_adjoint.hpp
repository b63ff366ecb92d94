// orad_qeph_adjoint.hpp -- hand-written reverse (adjoint) pass of one QEPH shell
// (Ishell=24) through the CZFORC3 chain of orad_qeph_chain.hpp, FP64.
//
// qeph_vjp evaluates the chain once forward (values only, every branch taken on the
// values as the chain takes it) and then once backward, from the adjoints of the corner
// slots (forces and moments) and of the carried state (SIG and PLA of every point,
// HOURG(1:12)) to the adjoints of the corner coordinates, velocities and rotational
// velocities, of the incoming state, of the thickness and of A, B, n.  It is the exact
// derivative of the forward arithmetic for law 2 with radial return (IPLA = 0) and no
// strain-rate term (C = 0), including the hourglass limiter's dependence on ET/E and
// the mean yield stress of the points.  The characteristic length, FACN, the time
// step and STI carry no derivative.  The symmetric 3x3 inverse of CZCORP5
// (A3INVDP_V) is differentiated as the inverse it is (dS = -S dD S), for elements
// whose projection matrix has a positive determinant.  Stages, in reverse order:
// CUPDTN3P, CZPROJN, CZFINTN1, CZFINTCE, the law-2 point loop (MULAWC, SIGEPS02C,
// M2CPLR), CZSTRA3, then qeph_kin_rev: CZDEF, CZCORP5, the rigid-rotation correction
// and CZCORC1 with CLSKEW3 (see the forward chain for the routines' lines).
#pragma once
#include "orad_qeph_chain.hpp"

namespace orad {

namespace qadj {

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void cross3(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ void cross3_rev(const double a[3], const double b[3], const double lc[3],
                                           double la[3], double lb[3]) {
  double t[3];
  cross3(b, lc, t);
  la[0] += t[0]; la[1] += t[1]; la[2] += t[2];
  cross3(lc, a, t);
  lb[0] += t[0]; lb[1] += t[1]; lb[2] += t[2];
}

// The projection of CZCORP5 / CZPROJN on a 3x3 symmetric S = [[d0,d3,d4],[d3,d1,d5],
// [d4,d5,d2]] (the DI of A3INVDP_V) and the 3x4 VQN:  w = AR - VQN AD,  ALR = S w,
// ALD[c] = AD[c] - VQN[:,c] . ALR  (the chain's DB / DBAD / ALR / ALD steps).  Reverse:
// adds the adjoints of AR, AD, VQN and the full-matrix adjoint G of S (G_ij = dL/dS_ij).
__device__ inline void proj_rev(const double S[3][3], const double VQN[3][4], const double AR[3],
                                const double AD[4], const double lALR[3], const double lALD[4],
                                double lAR[3], double lAD[4], double lVQN[3][4], double G[3][3]) {
  double w[3], ALR[3];
  for (int p = 0; p < 3; ++p)
    w[p] = AR[p] - (VQN[p][0] * AD[0] + VQN[p][1] * AD[1] + VQN[p][2] * AD[2] + VQN[p][3] * AD[3]);
  for (int p = 0; p < 3; ++p) ALR[p] = S[p][0] * w[0] + S[p][1] * w[1] + S[p][2] * w[2];
  double lA[3];
  for (int p = 0; p < 3; ++p) {
    lA[p] = lALR[p];
    for (int c = 0; c < 4; ++c) {
      lA[p] -= VQN[p][c] * lALD[c];
      lVQN[p][c] -= lALD[c] * ALR[p];
    }
  }
  for (int c = 0; c < 4; ++c) lAD[c] += lALD[c];
  double lw[3];
  for (int p = 0; p < 3; ++p) {
    lw[p] = S[0][p] * lA[0] + S[1][p] * lA[1] + S[2][p] * lA[2];
    for (int q = 0; q < 3; ++q) G[p][q] += lA[p] * w[q];
  }
  for (int p = 0; p < 3; ++p) {
    lAR[p] += lw[p];
    for (int c = 0; c < 4; ++c) {
      lVQN[p][c] -= lw[p] * AD[c];
      lAD[c] -= VQN[p][c] * lw[p];
    }
  }
}

}  // namespace qadj

// Adjoints of the kinematics stage's outputs (QephKin) that the rest of the chain reads.
struct QKinAdj {
  double e1[3], e2[3], e3[3];
  double AREA, AREA_I, X13, X24, Y13, Y24, MX13, MX23, MX34, MY13, MY23, MY34, Z1;
  double COREL[2][4];
  double VDEF[8], DHG[6];
  double VQN[3][4];
  double G[3][3];       // full-matrix adjoint of the projection's S = DI
};

// Reverse of qeph_kinematics<double>: given L (the adjoints of its outputs), adds the
// adjoints of the corner coordinates, velocities and rotational velocities.
__device__ inline void qeph_kin_rev(const double xg[3][4], const double vg[3][4],
                                    const double vrg[3][4], double dt1, double offg_in, int npt,
                                    const ShConst &k, QKinAdj &L, double lx[3][4],
                                    double lv[3][4], double lvr[3][4]) {
  using namespace qadj;
  const double ZERO = 0.0, ONE = 1.0, TWO = 2.0, FOUR = 4.0, HALF = 0.5, FOURTH = 0.25;
  const double EM20 = k.em20, EM10 = k.em10;
  // ================= forward =================
  double Rv[3], Sv[3];
  for (int d = 0; d < 3; ++d) {
    Rv[d] = xg[d][1] + xg[d][2] - xg[d][0] - xg[d][3];
    Sv[d] = xg[d][2] + xg[d][3] - xg[d][0] - xg[d][1];
  }
  double e3r[3];
  cross3(Rv, Sv, e3r);
  const double DET0 = __builtin_sqrt(dot3(e3r, e3r));
  double offg = offg_in;
  if (DET0 < EM20 && offg != ZERO) offg = ZERO;
  const double off_loc = (tabs(offg) != ZERO) ? ONE : ZERO;
  const bool det_free = DET0 > EM20;
  const double DET = det_free ? DET0 : EM20;
  const double CC0 = off_loc / DET;
  const bool cc_free = CC0 > EM20;
  const double CC = cc_free ? CC0 : EM20;
  double E3[3];
  for (int i = 0; i < 3; ++i) E3[i] = e3r[i] * CC;
  const double C1C1 = dot3(Rv, Rv), C2C2 = dot3(Sv, Sv);
  double c2_1 = ZERO, c1_1 = ZERO;
  int cmode = 0;
  if (C1C1 != ZERO) {
    c2_1 = M<double>::sqrt_(C2C2 / tmax(EM20, C1C1)); c1_1 = ONE; cmode = 1;
  } else if (C2C2 != ZERO) {
    c2_1 = ONE; c1_1 = M<double>::sqrt_(C1C1 / tmax(EM20, C2C2)); cmode = 2;
  }
  double SxE3[3], E1u[3];
  cross3(Sv, E3, SxE3);
  for (int i = 0; i < 3; ++i) E1u[i] = Rv[i] * c2_1 + SxE3[i] * c1_1;
  const double C1n = M<double>::sqrt_(dot3(E1u, E1u));
  const bool e1_free = C1n > EM20;
  const double inv1 = (C1n != ZERO) ? ONE / tmax(EM20, C1n) : C1n;
  double E1[3], E2[3];
  for (int i = 0; i < 3; ++i) E1[i] = E1u[i] * inv1;
  cross3(E3, E1, E2);
  const double AREA = FOURTH * DET;
  const double AI0 = off_loc / AREA;
  const bool ai_free = AI0 > EM20;
  const double AREA_I = ai_free ? AI0 : EM20;
  double L0[3], dx[4][3];
  for (int d = 0; d < 3; ++d) L0[d] = FOURTH * (xg[d][2] + xg[d][3] + xg[d][0] + xg[d][1]);
  for (int c = 1; c < 4; ++c)
    for (int d = 0; d < 3; ++d) dx[c][d] = xg[d][c] - xg[d][0];
  for (int d = 0; d < 3; ++d) dx[0][d] = xg[d][0] - L0[d];
  const double XL2 = dot3(E1, dx[1]), YL2 = dot3(E2, dx[1]);
  double Z1 = dot3(E3, dx[0]);
  const double XL3 = dot3(E1, dx[2]), YL3 = dot3(E2, dx[2]);
  const double XL4 = dot3(E1, dx[3]), YL4 = dot3(E2, dx[3]);
  double CO[2][4];
  {
    const double lxm = FOURTH * (XL2 + XL3 + XL4), lym = FOURTH * (YL2 + YL3 + YL4);
    CO[0][0] = -lxm; CO[0][1] = XL2 - lxm; CO[0][2] = XL3 - lxm; CO[0][3] = XL4 - lxm;
    CO[1][0] = -lym; CO[1][1] = YL2 - lym; CO[1][2] = YL3 - lym; CO[1][3] = YL4 - lym;
  }
  const double X13 = (CO[0][0] - CO[0][2]) * HALF, X24 = (CO[0][1] - CO[0][3]) * HALF;
  const double Y13 = (CO[1][0] - CO[1][2]) * HALF, Y24 = (CO[1][1] - CO[1][3]) * HALF;
  const double MX13 = (CO[0][0] + CO[0][2]) * HALF, MY13 = (CO[1][0] + CO[1][2]) * HALF;
  const double MX23 = (CO[0][1] + CO[0][2]) * HALF, MX34 = (CO[0][2] + CO[0][3]) * HALF;
  const double MY23 = (CO[1][1] + CO[1][2]) * HALF, MY34 = (CO[1][2] + CO[1][3]) * HALF;
  const double L13 = X13 * X13 + Y13 * Y13, L24 = X24 * X24 + Y24 * Y24;
  const double LM = HALF * (L13 + L24);
  double RL[2][4];
  for (int c = 0; c < 4; ++c) {
    RL[0][c] = E1[0] * vrg[0][c] + E1[1] * vrg[1][c] + E1[2] * vrg[2][c];
    RL[1][c] = E2[0] * vrg[0][c] + E2[1] * vrg[1][c] + E2[2] * vrg[2][c];
  }
  double G13[3], G24[3], GHI[3];
  for (int d = 0; d < 3; ++d) {
    G13[d] = vg[d][0] - vg[d][2];
    G24[d] = vg[d][1] - vg[d][3];
    GHI[d] = vg[d][0] - vg[d][1] + vg[d][2] - vg[d][3];
  }
  const double *Eq[3] = {E1, E2, E3};
  double V13a[3], V24a[3], VHIa[3];   // before the rotation correction
  for (int q = 0; q < 3; ++q) {
    V13a[q] = dot3(Eq[q], G13);
    V24a[q] = dot3(Eq[q], G24);
    VHIa[q] = dot3(Eq[q], GHI);
  }
  const double DT05 = HALF * dt1, DT025 = FOURTH * dt1;
  const double EXZr = Y24 * V13a[2] - Y13 * V24a[2];
  const double EYZr = -X24 * V13a[2] + X13 * V24a[2];
  const double DDRY = DT05 * EXZr * AREA_I, DDRX = DT05 * EYZr * AREA_I;
  const bool z1_free = !(tabs(X13 - X24) < EM10);
  const double DDRZ1 = z1_free ? DT025 * (V13a[1] - V24a[1]) / (X13 - X24) : ZERO;
  const bool z2_free = !(tabs(Y13 + Y24) < EM10);
  const double DDRZ2 = z2_free ? DT025 * (V13a[0] + V24a[0]) / (Y13 + Y24) : ZERO;
  double V13b[3], V24b[3], VHIb[3];   // after the rotation correction
  V13b[0] = V13a[0] - DDRY * V13a[2] - DDRZ1 * V13a[1];
  V24b[0] = V24a[0] - DDRY * V24a[2] - DDRZ1 * V24a[1];
  VHIb[0] = VHIa[0] - DDRY * VHIa[2] - DDRZ1 * VHIa[1];
  V13b[1] = V13a[1] - DDRX * V13a[2] - DDRZ2 * V13a[0];
  V24b[1] = V24a[1] - DDRX * V24a[2] - DDRZ2 * V24a[0];
  VHIb[1] = VHIa[1] - DDRX * VHIa[2] - DDRZ2 * VHIa[0];
  V13b[2] = V13a[2]; V24b[2] = V24a[2]; VHIb[2] = VHIa[2];
  const double Z2 = Z1 * Z1;
  const bool plat = (Z2 < LM * k.em08 || npt == 1);
  // CZCORP5 (warped elements)
  double VQN[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
  double S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double RR2[4] = {0, 0, 0, 0}, ARp[3] = {0, 0, 0}, ADp[4] = {0, 0, 0, 0}, ALR[3] = {0, 0, 0};
  double ALD[4] = {0, 0, 0, 0};
  double SZ1a = 0, SZ2a = 0, SLa = 0, SZ2b = 0, SLb = 0, SZ1c = 0, SZ2c = 0, SLc = 0, SZ2d = 0,
         SLd = 0, Dm[6] = {0, 0, 0, 0, 0, 0}, detD = 0;
  bool inv_free = false;
  double V13c[3], V24c[3], VHIc[3], RLc[2][4];   // after the projection
  for (int q = 0; q < 3; ++q) { V13c[q] = V13b[q]; V24c[q] = V24b[q]; VHIc[q] = VHIb[q]; }
  for (int c = 0; c < 4; ++c) { RLc[0][c] = RL[0][c]; RLc[1][c] = RL[1][c]; }
  if (plat) {
    Z1 = ZERO;
  } else {
    const double A_4 = AREA * FOURTH;
    SZ1a = MX13 * Y24 - MY13 * X24;
    SZ2a = A_4 + SZ1a;
    const double SZsA = Z2 * L24;
    SLa = ONE / M<double>::sqrt_(SZsA + SZ2a * SZ2a);
    VQN[0][0] = -Z1 * Y24 * SLa; VQN[1][0] = Z1 * X24 * SLa; VQN[2][0] = SZ2a * SLa;
    SZ2b = A_4 - SZ1a;
    SLb = ONE / M<double>::sqrt_(SZsA + SZ2b * SZ2b);
    VQN[0][2] = Z1 * Y24 * SLb; VQN[1][2] = -Z1 * X24 * SLb; VQN[2][2] = SZ2b * SLb;
    SZ1c = MX13 * Y13 - MY13 * X13;
    SZ2c = A_4 + SZ1c;
    const double SZsC = Z2 * L13;
    SLc = ONE / M<double>::sqrt_(SZsC + SZ2c * SZ2c);
    VQN[0][1] = -Z1 * Y13 * SLc; VQN[1][1] = Z1 * X13 * SLc; VQN[2][1] = SZ2c * SLc;
    SZ2d = A_4 - SZ1c;
    SLd = ONE / M<double>::sqrt_(SZsC + SZ2d * SZ2d);
    VQN[0][3] = Z1 * Y13 * SLd; VQN[1][3] = -Z1 * X13 * SLd; VQN[2][3] = SZ2d * SLd;
    for (int c = 0; c < 4; ++c) RR2[c] = E3[0] * vrg[0][c] + E3[1] * vrg[1][c] + E3[2] * vrg[2][c];
    ARp[0] = -Z1 * VHIb[1] + Y13 * V13b[2] + Y24 * V24b[2] + MY13 * VHIb[2] + RL[0][0] + RL[0][1] +
             RL[0][2] + RL[0][3];
    ARp[1] = Z1 * VHIb[0] - X13 * V13b[2] - X24 * V24b[2] - MX13 * VHIb[2] + RL[1][0] + RL[1][1] +
             RL[1][2] + RL[1][3];
    ARp[2] = X13 * V13b[1] + X24 * V24b[1] + MX13 * VHIb[1] - Y13 * V13b[0] - Y24 * V24b[0] -
             MY13 * VHIb[0] + RR2[0] + RR2[1] + RR2[2] + RR2[3];
    for (int c = 0; c < 4; ++c)
      ADp[c] = VQN[0][c] * RL[0][c] + VQN[1][c] * RL[1][c] + VQN[2][c] * RR2[c];
    double XX = 0, YY = 0, XY = 0;
    for (int c = 0; c < 4; ++c) {
      XX += CO[0][c] * CO[0][c]; YY += CO[1][c] * CO[1][c]; XY += CO[0][c] * CO[1][c];
    }
    const double XZ = (CO[0][0] - CO[0][1] + CO[0][2] - CO[0][3]) * Z1;
    const double YZ = (CO[1][0] - CO[1][1] + CO[1][2] - CO[1][3]) * Z1;
    const double ZZ = FOUR * Z2;
    double BTB[6] = {0, 0, 0, 0, 0, 0};
    for (int c = 0; c < 4; ++c) {
      BTB[0] += VQN[0][c] * VQN[0][c]; BTB[1] += VQN[1][c] * VQN[1][c];
      BTB[2] += VQN[2][c] * VQN[2][c]; BTB[3] += VQN[0][c] * VQN[1][c];
      BTB[4] += VQN[0][c] * VQN[2][c]; BTB[5] += VQN[1][c] * VQN[2][c];
    }
    Dm[0] = YY + ZZ + FOUR - BTB[0]; Dm[1] = XX + ZZ + FOUR - BTB[1];
    Dm[2] = XX + YY + FOUR - BTB[2]; Dm[3] = -XY - BTB[3];
    Dm[4] = -XZ - BTB[4]; Dm[5] = -YZ - BTB[5];
    {
      const double ABC = Dm[0] * Dm[1] * Dm[2];
      const double XXYZ2 = Dm[0] * Dm[5] * Dm[5], YYXZ2 = Dm[1] * Dm[4] * Dm[4];
      const double ZZXY2 = Dm[2] * Dm[3] * Dm[3];
      detD = ABC + TWO * Dm[3] * Dm[4] * Dm[5] - XXYZ2 - YYXZ2 - ZZXY2;
      const double DETA = ONE / tmax(tabs(detD), EM20);
      const double DI0 = (ABC - XXYZ2) * DETA / tmax(Dm[0], EM20);
      const double DI1 = (ABC - YYXZ2) * DETA / tmax(Dm[1], EM20);
      const double DI2 = (ABC - ZZXY2) * DETA / tmax(Dm[2], EM20);
      const double DI3 = (Dm[4] * Dm[5] - Dm[3] * Dm[2]) * DETA;
      const double DI4 = (Dm[5] * Dm[3] - Dm[4] * Dm[1]) * DETA;
      const double DI5 = (Dm[3] * Dm[4] - Dm[5] * Dm[0]) * DETA;
      S[0][0] = DI0; S[1][1] = DI1; S[2][2] = DI2;
      S[0][1] = S[1][0] = DI3; S[0][2] = S[2][0] = DI4; S[1][2] = S[2][1] = DI5;
      inv_free = detD > EM20 && Dm[0] > EM20 && Dm[1] > EM20 && Dm[2] > EM20;
    }
    {
      double w[3];
      for (int p = 0; p < 3; ++p)
        w[p] = ARp[p] - (VQN[p][0] * ADp[0] + VQN[p][1] * ADp[1] + VQN[p][2] * ADp[2] +
                         VQN[p][3] * ADp[3]);
      for (int p = 0; p < 3; ++p) ALR[p] = S[p][0] * w[0] + S[p][1] * w[1] + S[p][2] * w[2];
      for (int c = 0; c < 4; ++c)
        ALD[c] = ADp[c] - (VQN[0][c] * ALR[0] + VQN[1][c] * ALR[1] + VQN[2][c] * ALR[2]);
    }
    const double C1p = TWO * ALR[2];
    V13c[0] = V13b[0] + C1p * Y13;
    V24c[0] = V24b[0] + C1p * Y24;
    VHIc[0] = VHIb[0] + FOUR * (ALR[2] * MY13 - Z1 * ALR[1]);
    V13c[1] = V13b[1] - C1p * X13;
    V24c[1] = V24b[1] - C1p * X24;
    VHIc[1] = VHIb[1] - FOUR * (ALR[2] * MX13 - Z1 * ALR[0]);
    V13c[2] = V13b[2] - TWO * (Y13 * ALR[0] - X13 * ALR[1]);
    V24c[2] = V24b[2] - TWO * (Y24 * ALR[0] - X24 * ALR[1]);
    VHIc[2] = VHIb[2] + FOUR * (MX13 * ALR[1] - MY13 * ALR[0]);
    for (int c = 0; c < 4; ++c) {
      RLc[0][c] = RL[0][c] - ALR[0] - VQN[0][c] * ALD[c];
      RLc[1][c] = RL[1][c] - ALR[1] - VQN[1][c] * ALD[c];
    }
  }
  double V13[3], V24[3], VHI[3];
  for (int q = 0; q < 3; ++q) {
    V13[q] = V13c[q] * AREA_I; V24[q] = V24c[q] * AREA_I; VHI[q] = VHIc[q] * FOURTH;
  }
  const double AA = AREA_I;
  double R13[2], R24[2], RSOM[2];
  for (int q = 0; q < 2; ++q) {
    R13[q] = (RLc[q][0] - RLc[q][2]) * AA;
    R24[q] = (RLc[q][1] - RLc[q][3]) * AA;
    RSOM[q] = (RLc[q][3] + RLc[q][2] + RLc[q][0] + RLc[q][1]) * AA;
  }
  const bool vdead = offg < ZERO;   // CZDEF's rates are zeroed for deleted elements

  // ================= reverse =================
  double lV13[3] = {0, 0, 0}, lV24[3] = {0, 0, 0}, lVHI[3] = {0, 0, 0};
  double lR13[2] = {0, 0}, lR24[2] = {0, 0}, lRSOM[2] = {0, 0}, lRHI[2] = {0, 0};
  double lX13 = L.X13, lX24 = L.X24, lY13 = L.Y13, lY24 = L.Y24;
  double lMX13 = L.MX13, lMY13 = L.MY13, lMX23 = L.MX23, lMX34 = L.MX34, lMY23 = L.MY23,
         lMY34 = L.MY34;
  double lAREA = L.AREA, lAA = L.AREA_I, lZ1 = plat ? 0.0 : L.Z1;
  // ---- CZDEF ----
  if (!vdead) {
    const double *lVD = L.VDEF, *lDH = L.DHG;
    // VDEF5..7 += (...) * DETA1, DETA1 = Z1 4 AA
    const double DETA1 = Z1 * FOUR * AA;
    const double t5 = X13 * V13[0] - X24 * V24[0];
    const double t6 = Y13 * V13[1] - Y24 * V24[1];
    const double t7 = X13 * V13[1] - X24 * V24[1] + Y13 * V13[0] - Y24 * V24[0];
    const double lDETA1 = lVD[5] * t5 + lVD[6] * t6 + lVD[7] * t7;
    lZ1 += lDETA1 * FOUR * AA;
    lAA += lDETA1 * Z1 * FOUR;
    {
      const double l5 = lVD[5] * DETA1, l6 = lVD[6] * DETA1, l7 = lVD[7] * DETA1;
      lX13 += l5 * V13[0] + l7 * V13[1]; lV13[0] += l5 * X13 + l7 * Y13;
      lX24 -= l5 * V24[0] + l7 * V24[1]; lV24[0] -= l5 * X24 + l7 * Y24;
      lY13 += l6 * V13[1] + l7 * V13[0]; lV13[1] += l6 * Y13 + l7 * X13;
      lY24 -= l6 * V24[1] + l7 * V24[0]; lV24[1] -= l6 * Y24 + l7 * X24;
    }
    // DHG0 += (Y24 V13_2 - Y13 V24_2) Z1; DHG1 += (-X24 V13_2 + X13 V24_2) Z1
    {
      const double a0 = Y24 * V13[2] - Y13 * V24[2], a1 = -X24 * V13[2] + X13 * V24[2];
      lZ1 += lDH[0] * a0 + lDH[1] * a1;
      const double l0 = lDH[0] * Z1, l1 = lDH[1] * Z1;
      lY24 += l0 * V13[2]; lV13[2] += l0 * Y24 - l1 * X24;
      lY13 -= l0 * V24[2]; lV24[2] += -l0 * Y13 + l1 * X13;
      lX24 -= l1 * V13[2]; lX13 += l1 * V24[2];
    }
    // DHG4 = (4 VHI2 - P4 AREA) 4, P4 = MY13 RSOM0 - MY23 (R13_0 + R24_0) + MX23 (R13_1 + R24_1) - MX13 RSOM1
    // DHG5 = (4 VHI2 - P5 AREA) 4, P5 = MY13 RSOM0 - MY34 (R13_0 - R24_0) + MX34 (R13_1 - R24_1) - MX13 RSOM1
    {
      const double P4 = MY13 * RSOM[0] - MY23 * (R13[0] + R24[0]) + MX23 * (R13[1] + R24[1]) -
                        MX13 * RSOM[1];
      const double P5 = MY13 * RSOM[0] - MY34 * (R13[0] - R24[0]) + MX34 * (R13[1] - R24[1]) -
                        MX13 * RSOM[1];
      lVHI[2] += 16.0 * (lDH[4] + lDH[5]);
      const double lP4 = -4.0 * lDH[4] * AREA, lP5 = -4.0 * lDH[5] * AREA;
      lAREA += -4.0 * (lDH[4] * P4 + lDH[5] * P5);
      lMY13 += (lP4 + lP5) * RSOM[0];
      lMX13 -= (lP4 + lP5) * RSOM[1];
      lRSOM[0] += (lP4 + lP5) * MY13;
      lRSOM[1] -= (lP4 + lP5) * MX13;
      lMY23 -= lP4 * (R13[0] + R24[0]);
      lMX23 += lP4 * (R13[1] + R24[1]);
      lMY34 -= lP5 * (R13[0] - R24[0]);
      lMX34 += lP5 * (R13[1] - R24[1]);
      lR13[0] += -lP4 * MY23 - lP5 * MY34;
      lR24[0] += -lP4 * MY23 + lP5 * MY34;
      lR13[1] += lP4 * MX23 + lP5 * MX34;
      lR24[1] += lP4 * MX23 - lP5 * MX34;
    }
    // VDEF0..7 and DHG0..3 (their base forms)
    const double BYV1 = -X24 * V13[0] + X13 * V24[0];
    const double BXV2 = Y24 * V13[1] - Y13 * V24[1];
    const double BYR2 = -X24 * R13[1] + X13 * R24[1];
    const double BXR1 = Y13 * R24[0] - Y24 * R13[0];
    const double VD0 = Y24 * V13[0] - Y13 * V24[0];
    const double VD1 = -X24 * V13[1] + X13 * V24[1];
    const double VD5 = Y24 * R13[1] - Y13 * R24[1];
    const double VD6 = X24 * R13[0] - X13 * R24[0];
    // DHG0 = VHI0 - MX13 VD0 - MY13 BYV1;  DHG1 = VHI1 - MX13 BXV2 - MY13 VD1
    // DHG2 = RHI1 - MX13 VD5 - MY13 BYR2;  DHG3 = -RHI0 - MX13 BXR1 - MY13 VD6
    lVHI[0] += lDH[0]; lVHI[1] += lDH[1];
    lRHI[1] += lDH[2]; lRHI[0] -= lDH[3];
    lMX13 -= lDH[0] * VD0 + lDH[1] * BXV2 + lDH[2] * VD5 + lDH[3] * BXR1;
    lMY13 -= lDH[0] * BYV1 + lDH[1] * VD1 + lDH[2] * BYR2 + lDH[3] * VD6;
    double lVD0 = lVD[0] - lDH[0] * MX13, lVD1 = lVD[1] - lDH[1] * MY13;
    double lBYV1 = lVD[2] - lDH[0] * MY13, lBXV2 = lVD[2] - lDH[1] * MX13;
    double lVD5 = lVD[5] - lDH[2] * MX13, lVD6 = lVD[6] - lDH[3] * MY13;
    double lBYR2 = lVD[7] - lDH[2] * MY13, lBXR1 = lVD[7] - lDH[3] * MX13;
    // VDEF3 = Y24 BCX - Y13 BCY + BCXY RSOM1; VDEF4 = X13 BCY - X24 BCX - BCXY RSOM0
    {
      const double BCXY = AREA * FOURTH;
      const double BCX = V13[2] - MY13 * R13[0] + MX13 * R13[1];
      const double BCY = V24[2] + MY13 * R24[0] - MX13 * R24[1];
      const double l3 = lVD[3], l4 = lVD[4];
      lY24 += l3 * BCX; lY13 -= l3 * BCY; lX13 += l4 * BCY; lX24 -= l4 * BCX;
      const double lBCX = l3 * Y24 - l4 * X24, lBCY = -l3 * Y13 + l4 * X13;
      lAREA += FOURTH * (l3 * RSOM[1] - l4 * RSOM[0]);
      lRSOM[1] += l3 * BCXY; lRSOM[0] -= l4 * BCXY;
      lV13[2] += lBCX; lMY13 += -lBCX * R13[0] + lBCY * R24[0];
      lMX13 += lBCX * R13[1] - lBCY * R24[1];
      lR13[0] -= lBCX * MY13; lR13[1] += lBCX * MX13;
      lV24[2] += lBCY; lR24[0] += lBCY * MY13; lR24[1] -= lBCY * MX13;
    }
    // VD0 = Y24 V13_0 - Y13 V24_0; VD1 = -X24 V13_1 + X13 V24_1
    lY24 += lVD0 * V13[0]; lV13[0] += lVD0 * Y24; lY13 -= lVD0 * V24[0]; lV24[0] -= lVD0 * Y13;
    lX24 -= lVD1 * V13[1]; lV13[1] -= lVD1 * X24; lX13 += lVD1 * V24[1]; lV24[1] += lVD1 * X13;
    // BXV2 = Y24 V13_1 - Y13 V24_1; BYV1 = -X24 V13_0 + X13 V24_0
    lY24 += lBXV2 * V13[1]; lV13[1] += lBXV2 * Y24; lY13 -= lBXV2 * V24[1]; lV24[1] -= lBXV2 * Y13;
    lX24 -= lBYV1 * V13[0]; lV13[0] -= lBYV1 * X24; lX13 += lBYV1 * V24[0]; lV24[0] += lBYV1 * X13;
    // VD5 = Y24 R13_1 - Y13 R24_1; VD6 = X24 R13_0 - X13 R24_0
    lY24 += lVD5 * R13[1]; lR13[1] += lVD5 * Y24; lY13 -= lVD5 * R24[1]; lR24[1] -= lVD5 * Y13;
    lX24 += lVD6 * R13[0]; lR13[0] += lVD6 * X24; lX13 -= lVD6 * R24[0]; lR24[0] -= lVD6 * X13;
    // BXR1 = Y13 R24_0 - Y24 R13_0; BYR2 = -X24 R13_1 + X13 R24_1
    lY13 += lBXR1 * R24[0]; lR24[0] += lBXR1 * Y13; lY24 -= lBXR1 * R13[0]; lR13[0] -= lBXR1 * Y24;
    lX24 -= lBYR2 * R13[1]; lR13[1] -= lBYR2 * X24; lX13 += lBYR2 * R24[1]; lR24[1] += lBYR2 * X13;
  }
  if (plat) lZ1 = 0.0;   // Z1 is set to zero for flat elements
  // R13, R24, RSOM scaled by AA; RHI by 1/4
  double lRLc[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
  for (int q = 0; q < 2; ++q) {
    lAA += lR13[q] * (RLc[q][0] - RLc[q][2]) + lR24[q] * (RLc[q][1] - RLc[q][3]) +
           lRSOM[q] * (RLc[q][3] + RLc[q][2] + RLc[q][0] + RLc[q][1]);
    const double a = lR13[q] * AA, b = lR24[q] * AA, s = lRSOM[q] * AA, h = lRHI[q] * FOURTH;
    lRLc[q][0] += a + s + h; lRLc[q][2] += -a + s + h;
    lRLc[q][1] += b + s - h; lRLc[q][3] += -b + s - h;
  }
  // V13 = V13c AA, V24 = V24c AA, VHI = VHIc / 4
  double lV13c[3], lV24c[3], lVHIc[3];
  for (int q = 0; q < 3; ++q) {
    lAA += lV13[q] * V13c[q] + lV24[q] * V24c[q];
    lV13c[q] = lV13[q] * AA; lV24c[q] = lV24[q] * AA; lVHIc[q] = lVHI[q] * FOURTH;
  }
  if (!ai_free) lAA = 0.0;
  // AREA_I = off_loc / AREA
  lAREA += -lAA * AREA_I / AREA;
  // ---- CZCORP5 reverse ----
  double lV13b[3], lV24b[3], lVHIb[3], lRL[2][4];
  for (int q = 0; q < 3; ++q) { lV13b[q] = lV13c[q]; lV24b[q] = lV24c[q]; lVHIb[q] = lVHIc[q]; }
  for (int c = 0; c < 4; ++c) { lRL[0][c] = lRLc[0][c]; lRL[1][c] = lRLc[1][c]; }
  double lE3[3] = {L.e3[0], L.e3[1], L.e3[2]};
  double lE1[3] = {L.e1[0], L.e1[1], L.e1[2]}, lE2[3] = {L.e2[0], L.e2[1], L.e2[2]};
  double lCO[2][4];
  for (int c = 0; c < 4; ++c) { lCO[0][c] = L.COREL[0][c]; lCO[1][c] = L.COREL[1][c]; }
  if (!plat) {
    double lALR[3] = {0, 0, 0}, lALD[4] = {0, 0, 0, 0};
    double lVQN[3][4];
    for (int q = 0; q < 3; ++q)
      for (int c = 0; c < 4; ++c) lVQN[q][c] = L.VQN[q][c];
    // RLc = RL - ALR - VQN ALD
    for (int c = 0; c < 4; ++c)
      for (int q = 0; q < 2; ++q) {
        lALR[q] -= lRLc[q][c];
        lVQN[q][c] -= lRLc[q][c] * ALD[c];
        lALD[c] -= lRLc[q][c] * VQN[q][c];
      }
    // V13c/V24c/VHIc updates
    {
      const double a13 = lV13c[0], a24 = lV24c[0], ah0 = lVHIc[0];
      const double b13 = lV13c[1], b24 = lV24c[1], bh1 = lVHIc[1];
      const double c13 = lV13c[2], c24 = lV24c[2], ch2 = lVHIc[2];
      // C1p = 2 ALR2: V13_0 += C1p Y13; V24_0 += C1p Y24; V13_1 -= C1p X13; V24_1 -= C1p X24
      const double lC1p = a13 * Y13 + a24 * Y24 - b13 * X13 - b24 * X24;
      lY13 += a13 * TWO * ALR[2]; lY24 += a24 * TWO * ALR[2];
      lX13 -= b13 * TWO * ALR[2]; lX24 -= b24 * TWO * ALR[2];
      lALR[2] += TWO * lC1p;
      // VHI0 += 4 (ALR2 MY13 - Z1 ALR1); VHI1 -= 4 (ALR2 MX13 - Z1 ALR0)
      lALR[2] += FOUR * (ah0 * MY13 - bh1 * MX13);
      lMY13 += FOUR * ah0 * ALR[2]; lMX13 -= FOUR * bh1 * ALR[2];
      lZ1 += FOUR * (-ah0 * ALR[1] + bh1 * ALR[0]);
      lALR[1] -= FOUR * ah0 * Z1; lALR[0] += FOUR * bh1 * Z1;
      // V13_2 -= 2 (Y13 ALR0 - X13 ALR1); V24_2 -= 2 (Y24 ALR0 - X24 ALR1)
      lY13 -= TWO * c13 * ALR[0]; lX13 += TWO * c13 * ALR[1];
      lY24 -= TWO * c24 * ALR[0]; lX24 += TWO * c24 * ALR[1];
      lALR[0] -= TWO * (c13 * Y13 + c24 * Y24);
      lALR[1] += TWO * (c13 * X13 + c24 * X24);
      // VHI2 += 4 (MX13 ALR1 - MY13 ALR0)
      lMX13 += FOUR * ch2 * ALR[1]; lMY13 -= FOUR * ch2 * ALR[0];
      lALR[1] += FOUR * ch2 * MX13; lALR[0] -= FOUR * ch2 * MY13;
    }
    double lARp[3] = {0, 0, 0}, lADp[4] = {0, 0, 0, 0};
    qadj::proj_rev(S, VQN, ARp, ADp, lALR, lALD, lARp, lADp, lVQN, L.G);
    // S = DI = inverse of D (dS = -S dD S) from the full-matrix adjoint L.G, which holds
    // CZPROJN's use as well
    double lDm[6] = {0, 0, 0, 0, 0, 0};
    if (inv_free) {
      double H[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double h = 0.0;
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) h += S[i][a] * L.G[a][b] * S[b][j];
          H[i][j] = -h;
        }
      lDm[0] = H[0][0]; lDm[1] = H[1][1]; lDm[2] = H[2][2];
      lDm[3] = H[0][1] + H[1][0]; lDm[4] = H[0][2] + H[2][0]; lDm[5] = H[1][2] + H[2][1];
    }
    // AD[c] = VQN0c RL0c + VQN1c RL1c + VQN2c RR2c
    double lRR2[4];
    for (int c = 0; c < 4; ++c) {
      lVQN[0][c] += lADp[c] * RL[0][c]; lVQN[1][c] += lADp[c] * RL[1][c];
      lVQN[2][c] += lADp[c] * RR2[c];
      lRL[0][c] += lADp[c] * VQN[0][c]; lRL[1][c] += lADp[c] * VQN[1][c];
      lRR2[c] = lADp[c] * VQN[2][c];
    }
    // AR0 = -Z1 VHI1 + Y13 V13_2 + Y24 V24_2 + MY13 VHI2 + sum RL0
    // AR1 = Z1 VHI0 - X13 V13_2 - X24 V24_2 - MX13 VHI2 + sum RL1
    // AR2 = X13 V13_1 + X24 V24_1 + MX13 VHI1 - Y13 V13_0 - Y24 V24_0 - MY13 VHI0 + sum RR2
    {
      const double a0 = lARp[0], a1 = lARp[1], a2 = lARp[2];
      lZ1 += -a0 * VHIb[1] + a1 * VHIb[0];
      lVHIb[1] += -a0 * Z1 + a2 * MX13; lVHIb[0] += a1 * Z1 - a2 * MY13;
      lY13 += a0 * V13b[2] - a2 * V13b[0]; lY24 += a0 * V24b[2] - a2 * V24b[0];
      lV13b[2] += a0 * Y13 - a1 * X13; lV24b[2] += a0 * Y24 - a1 * X24;
      lMY13 += a0 * VHIb[2] - a2 * VHIb[0]; lMX13 += -a1 * VHIb[2] + a2 * VHIb[1];
      lVHIb[2] += a0 * MY13 - a1 * MX13;
      lX13 += -a1 * V13b[2] + a2 * V13b[1]; lX24 += -a1 * V24b[2] + a2 * V24b[1];
      lV13b[1] += a2 * X13; lV24b[1] += a2 * X24;
      lV13b[0] -= a2 * Y13; lV24b[0] -= a2 * Y24;
      for (int c = 0; c < 4; ++c) { lRL[0][c] += a0; lRL[1][c] += a1; lRR2[c] += a2; }
    }
    // RR2 = E3 . vr
    for (int c = 0; c < 4; ++c)
      for (int d = 0; d < 3; ++d) {
        lE3[d] += lRR2[c] * vrg[d][c];
        lvr[d][c] += lRR2[c] * E3[d];
      }
    // D = (YY + ZZ + 4 - BTB0, XX + ZZ + 4 - BTB1, XX + YY + 4 - BTB2, -XY - BTB3, -XZ - BTB4, -YZ - BTB5)
    {
      const double lXX = lDm[1] + lDm[2], lYY = lDm[0] + lDm[2], lZZ = lDm[0] + lDm[1];
      const double lXY = -lDm[3], lXZ = -lDm[4], lYZ = -lDm[5];
      for (int c = 0; c < 4; ++c) {
        lCO[0][c] += 2.0 * lXX * CO[0][c] + lXY * CO[1][c];
        lCO[1][c] += 2.0 * lYY * CO[1][c] + lXY * CO[0][c];
        const double sg = (c & 1) ? -1.0 : 1.0;
        lCO[0][c] += sg * lXZ * Z1;
        lCO[1][c] += sg * lYZ * Z1;
        // BTB
        lVQN[0][c] += -2.0 * lDm[0] * VQN[0][c] - lDm[3] * VQN[1][c] - lDm[4] * VQN[2][c];
        lVQN[1][c] += -2.0 * lDm[1] * VQN[1][c] - lDm[3] * VQN[0][c] - lDm[5] * VQN[2][c];
        lVQN[2][c] += -2.0 * lDm[2] * VQN[2][c] - lDm[4] * VQN[0][c] - lDm[5] * VQN[1][c];
      }
      lZ1 += lXZ * (CO[0][0] - CO[0][1] + CO[0][2] - CO[0][3]) +
             lYZ * (CO[1][0] - CO[1][1] + CO[1][2] - CO[1][3]) + lZZ * 8.0 * Z1;
    }
    // VQN from Z1, X13.., MX13, MY13, AREA, L13 = X13^2 + Y13^2, L24
    {
      const double A_4 = AREA * FOURTH;
      double lA4 = 0.0, lL13 = 0.0, lL24 = 0.0, lSZ1a = 0.0, lSZ1c = 0.0;
      // pair (a, b) uses SZ1a, L24 and (-Z1 Y24, Z1 X24); pair (c, d) uses SZ1c, L13, (-Z1 Y13, Z1 X13)
      auto pair_rev = [&](double lq0a, double lq1a, double lq2a, double lq0b, double lq1b,
                          double lq2b, double SZ2p, double SLp, double SZ2m, double SLm,
                          double Xp, double Yp, double Lp, double &lXp, double &lYp,
                          double &lLp, double &lSZ1) {
        // first corner: VQN0 = -Z1 Yp SLp, VQN1 = Z1 Xp SLp, VQN2 = SZ2p SLp
        double lSLp = -lq0a * Z1 * Yp + lq1a * Z1 * Xp + lq2a * SZ2p;
        lZ1 += (-lq0a * Yp + lq1a * Xp) * SLp;
        lYp -= lq0a * Z1 * SLp; lXp += lq1a * Z1 * SLp;
        double lSZ2p = lq2a * SLp;
        // second corner: VQN0 = Z1 Yp SLm, VQN1 = -Z1 Xp SLm, VQN2 = SZ2m SLm
        double lSLm = lq0b * Z1 * Yp - lq1b * Z1 * Xp + lq2b * SZ2m;
        lZ1 += (lq0b * Yp - lq1b * Xp) * SLm;
        lYp += lq0b * Z1 * SLm; lXp -= lq1b * Z1 * SLm;
        double lSZ2m = lq2b * SLm;
        // SL = 1/sqrt(Z1^2 Lp + SZ2^2): dSL = -SL^3 / 2 d(...)
        const double gp = -0.5 * SLp * SLp * SLp * lSLp, gm = -0.5 * SLm * SLm * SLm * lSLm;
        lZ1 += (gp + gm) * 2.0 * Z1 * Lp;
        lLp += (gp + gm) * Z1 * Z1;
        lSZ2p += gp * 2.0 * SZ2p;
        lSZ2m += gm * 2.0 * SZ2m;
        // SZ2p = A_4 + SZ1, SZ2m = A_4 - SZ1
        lA4 += lSZ2p + lSZ2m;
        lSZ1 += lSZ2p - lSZ2m;
      };
      pair_rev(lVQN[0][0], lVQN[1][0], lVQN[2][0], lVQN[0][2], lVQN[1][2], lVQN[2][2], SZ2a, SLa,
               SZ2b, SLb, X24, Y24, L24, lX24, lY24, lL24, lSZ1a);
      pair_rev(lVQN[0][1], lVQN[1][1], lVQN[2][1], lVQN[0][3], lVQN[1][3], lVQN[2][3], SZ2c, SLc,
               SZ2d, SLd, X13, Y13, L13, lX13, lY13, lL13, lSZ1c);
      lAREA += lA4 * FOURTH;
      // SZ1a = MX13 Y24 - MY13 X24; SZ1c = MX13 Y13 - MY13 X13
      lMX13 += lSZ1a * Y24 + lSZ1c * Y13;
      lMY13 -= lSZ1a * X24 + lSZ1c * X13;
      lY24 += lSZ1a * MX13; lX24 -= lSZ1a * MY13;
      lY13 += lSZ1c * MX13; lX13 -= lSZ1c * MY13;
      lX13 += lL13 * 2.0 * X13; lY13 += lL13 * 2.0 * Y13;
      lX24 += lL24 * 2.0 * X24; lY24 += lL24 * 2.0 * Y24;
    }
  }
  // ---- rotation correction reverse ----
  double lV13a[3], lV24a[3], lVHIa[3];
  for (int q = 0; q < 3; ++q) { lV13a[q] = lV13b[q]; lV24a[q] = lV24b[q]; lVHIa[q] = lVHIb[q]; }
  {
    // row 1: V?_1 = V?a_1 - DDRX V?a_2 - DDRZ2 V?a_0
    const double l13 = lV13b[1], l24 = lV24b[1], lh = lVHIb[1];
    const double lDDRX = -(l13 * V13a[2] + l24 * V24a[2] + lh * VHIa[2]);
    const double lDDRZ2 = -(l13 * V13a[0] + l24 * V24a[0] + lh * VHIa[0]);
    lV13a[2] -= l13 * DDRX; lV24a[2] -= l24 * DDRX; lVHIa[2] -= lh * DDRX;
    lV13a[0] -= l13 * DDRZ2; lV24a[0] -= l24 * DDRZ2; lVHIa[0] -= lh * DDRZ2;
    // row 0: V?_0 = V?a_0 - DDRY V?a_2 - DDRZ1 V?a_1
    const double m13 = lV13b[0], m24 = lV24b[0], mh = lVHIb[0];
    const double lDDRY = -(m13 * V13a[2] + m24 * V24a[2] + mh * VHIa[2]);
    const double lDDRZ1 = -(m13 * V13a[1] + m24 * V24a[1] + mh * VHIa[1]);
    lV13a[2] -= m13 * DDRY; lV24a[2] -= m24 * DDRY; lVHIa[2] -= mh * DDRY;
    lV13a[1] -= m13 * DDRZ1; lV24a[1] -= m24 * DDRZ1; lVHIa[1] -= mh * DDRZ1;
    if (z2_free) {   // DDRZ2 = DT025 (V13a_0 + V24a_0) / (Y13 + Y24)
      const double den = Y13 + Y24;
      lV13a[0] += lDDRZ2 * DT025 / den; lV24a[0] += lDDRZ2 * DT025 / den;
      lY13 -= lDDRZ2 * DDRZ2 / den; lY24 -= lDDRZ2 * DDRZ2 / den;
    }
    if (z1_free) {   // DDRZ1 = DT025 (V13a_1 - V24a_1) / (X13 - X24)
      const double den = X13 - X24;
      lV13a[1] += lDDRZ1 * DT025 / den; lV24a[1] -= lDDRZ1 * DT025 / den;
      lX13 -= lDDRZ1 * DDRZ1 / den; lX24 += lDDRZ1 * DDRZ1 / den;
    }
    // DDRY = DT05 EXZr AREA_I, DDRX = DT05 EYZr AREA_I
    const double lEXZr = lDDRY * DT05 * AREA_I, lEYZr = lDDRX * DT05 * AREA_I;
    double lAA2 = lDDRY * DT05 * EXZr + lDDRX * DT05 * EYZr;
    if (!ai_free) lAA2 = 0.0;
    lAREA += -lAA2 * AREA_I / AREA;
    // EXZr = Y24 V13a_2 - Y13 V24a_2; EYZr = -X24 V13a_2 + X13 V24a_2
    lY24 += lEXZr * V13a[2]; lY13 -= lEXZr * V24a[2];
    lX24 -= lEYZr * V13a[2]; lX13 += lEYZr * V24a[2];
    lV13a[2] += lEXZr * Y24 - lEYZr * X24;
    lV24a[2] += -lEXZr * Y13 + lEYZr * X13;
  }
  // V?a[q] = Eq . G?
  double lG13[3] = {0, 0, 0}, lG24[3] = {0, 0, 0}, lGHI[3] = {0, 0, 0};
  double *lEq[3] = {lE1, lE2, lE3};
  for (int q = 0; q < 3; ++q)
    for (int d = 0; d < 3; ++d) {
      lEq[q][d] += lV13a[q] * G13[d] + lV24a[q] * G24[d] + lVHIa[q] * GHI[d];
      lG13[d] += lV13a[q] * Eq[q][d];
      lG24[d] += lV24a[q] * Eq[q][d];
      lGHI[d] += lVHIa[q] * Eq[q][d];
    }
  for (int d = 0; d < 3; ++d) {
    lv[d][0] += lG13[d] + lGHI[d];
    lv[d][2] += -lG13[d] + lGHI[d];
    lv[d][1] += lG24[d] - lGHI[d];
    lv[d][3] += -lG24[d] - lGHI[d];
  }
  // RL = E1 . vr, E2 . vr
  for (int c = 0; c < 4; ++c)
    for (int d = 0; d < 3; ++d) {
      lE1[d] += lRL[0][c] * vrg[d][c];
      lE2[d] += lRL[1][c] * vrg[d][c];
      lvr[d][c] += lRL[0][c] * E1[d] + lRL[1][c] * E2[d];
    }
  // ---- local coordinates: X13.. from COREL, COREL from XL/YL ----
  lCO[0][0] += HALF * (lX13 + lMX13); lCO[0][2] += HALF * (-lX13 + lMX13 + lMX23 + lMX34);
  lCO[0][1] += HALF * (lX24 + lMX23); lCO[0][3] += HALF * (-lX24 + lMX34);
  lCO[1][0] += HALF * (lY13 + lMY13); lCO[1][2] += HALF * (-lY13 + lMY13 + lMY23 + lMY34);
  lCO[1][1] += HALF * (lY24 + lMY23); lCO[1][3] += HALF * (-lY24 + lMY34);
  double lXL[4] = {0, 0, 0, 0}, lYL[4] = {0, 0, 0, 0};   // XL2..4 at 1..3
  {
    const double sx = lCO[0][0] + lCO[0][1] + lCO[0][2] + lCO[0][3];
    const double sy = lCO[1][0] + lCO[1][1] + lCO[1][2] + lCO[1][3];
    for (int c = 1; c < 4; ++c) {
      lXL[c] = lCO[0][c] - FOURTH * sx;
      lYL[c] = lCO[1][c] - FOURTH * sy;
    }
  }
  // XLc = E1 . dx[c], YLc = E2 . dx[c] (c = 1..3), Z1 = E3 . dx[0]
  double ldx[4][3];
  for (int d = 0; d < 3; ++d) {
    ldx[0][d] = lZ1 * E3[d];
    lE3[d] += lZ1 * dx[0][d];
  }
  for (int c = 1; c < 4; ++c)
    for (int d = 0; d < 3; ++d) {
      ldx[c][d] = lXL[c] * E1[d] + lYL[c] * E2[d];
      lE1[d] += lXL[c] * dx[c][d];
      lE2[d] += lYL[c] * dx[c][d];
    }
  for (int d = 0; d < 3; ++d) {
    // dx[0] = x0 - L0, L0 = mean of the corners; dx[c] = xc - x0
    const double l0 = ldx[0][d];
    lx[d][0] += l0 - FOURTH * l0;
    for (int c = 1; c < 4; ++c) lx[d][c] -= FOURTH * l0;
    for (int c = 1; c < 4; ++c) {
      lx[d][c] += ldx[c][d];
      lx[d][0] -= ldx[c][d];
    }
  }
  // AREA = DET / 4
  double lDET = det_free ? lAREA * FOURTH : 0.0;
  // ---- frame reverse: E2 = E3 x E1, E1 = E1u inv1, E1u = R c2_1 + (S x E3) c1_1, E3 = e3r CC ----
  cross3_rev(E3, E1, lE2, lE3, lE1);
  double lE1u[3] = {0, 0, 0};
  if (C1n != ZERO) {
    if (e1_free) {
      const double p = dot3(E1, lE1);
      for (int i = 0; i < 3; ++i) lE1u[i] = (lE1[i] - p * E1[i]) / C1n;
    } else {
      for (int i = 0; i < 3; ++i) lE1u[i] = lE1[i] * inv1;
    }
  }
  double lRv[3] = {0, 0, 0}, lSv[3] = {0, 0, 0};
  double lc2 = 0.0, lc1 = 0.0;
  {
    double lSxE3[3];
    for (int i = 0; i < 3; ++i) {
      lRv[i] += lE1u[i] * c2_1;
      lSxE3[i] = lE1u[i] * c1_1;
      lc2 += lE1u[i] * Rv[i];
      lc1 += lE1u[i] * SxE3[i];
    }
    cross3_rev(Sv, E3, lSxE3, lSv, lE3);
  }
  if (cmode == 1 && c2_1 > 0.0) {
    // c2_1 = sqrt(C2C2 / max(EM20, C1C1))
    const double den = tmax(EM20, C1C1);
    const double q = lc2 * 0.5 / c2_1;
    const double lC2C2 = q / den;
    const double lC1C1 = (C1C1 > EM20) ? -q * C2C2 / (den * den) : 0.0;
    for (int i = 0; i < 3; ++i) { lSv[i] += 2.0 * lC2C2 * Sv[i]; lRv[i] += 2.0 * lC1C1 * Rv[i]; }
  } else if (cmode == 2 && c1_1 > 0.0) {
    const double den = tmax(EM20, C2C2);
    const double q = lc1 * 0.5 / c1_1;
    const double lC1C1 = q / den;
    const double lC2C2 = (C2C2 > EM20) ? -q * C1C1 / (den * den) : 0.0;
    for (int i = 0; i < 3; ++i) { lSv[i] += 2.0 * lC2C2 * Sv[i]; lRv[i] += 2.0 * lC1C1 * Rv[i]; }
  }
  // E3 = e3r CC, CC = off_loc / DET (clamped), DET = |e3r| (clamped)
  double le3r[3];
  double lCC = 0.0;
  for (int i = 0; i < 3; ++i) { le3r[i] = lE3[i] * CC; lCC += lE3[i] * e3r[i]; }
  if (cc_free) lDET += -lCC * CC / DET;
  if (det_free && DET0 > 0.0)
    for (int i = 0; i < 3; ++i) le3r[i] += lDET * e3r[i] / DET0;
  cross3_rev(Rv, Sv, le3r, lRv, lSv);
  // R = x1 + x2 - x0 - x3, S = x2 + x3 - x0 - x1
  for (int d = 0; d < 3; ++d) {
    lx[d][1] += lRv[d] - lSv[d];
    lx[d][2] += lRv[d] + lSv[d];
    lx[d][0] += -lRv[d] - lSv[d];
    lx[d][3] += -lRv[d] + lSv[d];
  }
}

// Reverse pass of one QEPH element.  st(col) reads the pre-cycle state; lpost(j) is the
// adjoint of the post-cycle state entry j (SIG 0..24, PLA 25..29, HOURG(1:12) 30..41);
// lslot(c, i) the adjoint of corner c's slot component i (0..2 force, 3..5 moment).
// lpre(j, value) is called once for every entry j of the points in use and HOURG(1:12)
// with the adjoint of the pre-cycle state; lx/lv/lvr, g_thk and g_jc[3] are added to.
template <typename ST, typename LPOST, typename LSLOT, typename LPRE>
__device__ inline void qeph_vjp(const double xg[3][4], const double vg[3][4],
                                const double vrg[3][4], double thk0, ST st, double dt1,
                                const ShMat &m, const double jc[3], const ShProp &p,
                                const ShConst &k, LPOST lpost, LSLOT lslot, LPRE lpre,
                                double lx[3][4], double lv[3][4], double lvr[3][4],
                                double &g_thk, double g_jc[3]) {
  const double ZERO = 0.0, ONE = 1.0, TWO = 2.0, THREE = 3.0, FOUR = 4.0, HALF = 0.5,
               FOURTH = 0.25;
  const double EM20 = k.em20;
  // ================= forward =================
  double offg = st(SC_OFFG);
  QephKin<double> g;
  qeph_kinematics<double>(xg, vg, vrg, dt1, offg, p.npt, k, g);
  if (g.degenerate) offg = ZERO;
  const double AREA = g.AREA, AA = g.AREA_I;
  const double X13 = g.X13, X24 = g.X24, Y13 = g.Y13, Y24 = g.Y24;
  const double MX13 = g.MX13, MX23 = g.MX23, MX34 = g.MX34;
  const double MY13 = g.MY13, MY23 = g.MY23, MY34 = g.MY34, Z1 = g.Z1;
  const double THK = thk0, THK02 = thk0 * thk0;
  const double RHO = m.rho0, G = m.shear, YM = m.young, NU = m.nu;
  const double A11 = m.a11, A12 = m.a12, SSP = m.ssp;
  const double GSR = __builtin_sqrt(tabs(G)), A11SR = __builtin_sqrt(tabs(A11)),
               A12SR = __builtin_sqrt(tabs(A12));
  const double SHF = p.fsh, SHFSR = __builtin_sqrt(p.fsh), GS = G * SHF, AMU = p.dn;
  double off = tmin(ONE, tabs(offg));
  const double EXX = g.VDEF[0] * dt1, EYY = g.VDEF[1] * dt1, EXY = g.VDEF[2] * dt1;
  const double EYZ = g.VDEF[4] * dt1, EXZ = g.VDEF[3] * dt1;
  const double KXX = g.VDEF[5] * dt1, KYY = g.VDEF[6] * dt1, KXY = g.VDEF[7] * dt1;
  const double dtinv = dt1 / tmax(dt1 * dt1, EM20);
  // law 2 at every point (IPLA = 0, C = 0), with the means ZCFAC and SIGY
  const double a11 = YM / (ONE - NU * NU), a12 = a11 * NU;
  const double ca = jc[0], cb = jc[1], cn = jc[2], ymax = m.ymax;
  const double SMALL = k.em07;
  const double fnpt = double(p.npt);
  double FOR[5] = {ZERO, ZERO, ZERO, ZERO, ZERO}, MOM[3] = {ZERO, ZERO, ZERO};
  double off_pt[5];
  const double off_old = off;
  double epchk = ZERO, sigy = ZERO, zcfac1 = ZERO, zcfac2 = ONE, viscmx = ZERO;
  int zc2_at = -1;   // the point whose ETSE is the minimum ZCFAC2 (-1: none below 1)
#pragma unroll
  for (int it = 0; it < 5; ++it) {
    if (it >= p.npt) break;
    off_pt[it] = off;
    const double zt = p.z0[it] * thk0;
    const double d0 = EXX + zt * KXX, d1 = EYY + zt * KYY, d2 = EXY + zt * KXY;
    double s[5];
    s[0] = st(SC_SIG + 5 * it) + a11 * d0 + a12 * d1;
    s[1] = st(SC_SIG + 5 * it + 1) + a12 * d0 + a11 * d1;
    s[2] = st(SC_SIG + 5 * it + 2) + G * d2;
    s[3] = st(SC_SIG + 5 * it + 3) + GS * EYZ;
    s[4] = st(SC_SIG + 5 * it + 4) + GS * EXZ;
    double pla = st(SC_PLA + it);
    double yld = (pla == ZERO) ? ca : ca + cb * (ONE - ZERO) * M<double>::powe(pla, cn);
    yld = tmin(yld, ymax);
    const double svm = __builtin_sqrt(s[0] * s[0] + s[1] * s[1] - s[0] * s[1] + THREE * s[2] * s[2]);
    const double r = tmin(ONE, yld / (svm + k.em15));
    double etse = ONE;
    if (r < ONE) {
      s[0] = s[0] * r; s[1] = s[1] * r; s[2] = s[2] * r;
      pla = pla + off * tmax(ZERO, (svm - yld) / YM);
      epchk = tmax(pla, epchk);
      const double h = (yld >= ymax) ? ZERO : cn * cb * M<double>::powe(pla + SMALL, cn - ONE);
      etse = h / (h + YM);
    }
    sigy = sigy + yld / fnpt;
    if (off == off_old && off > ZERO) {
      if (off == ONE && epchk >= m.epmx) off = k.four_over_5;
      else if (off < ONE) off = off * k.four_over_5;
    }
    viscmx = tmax(p.dm, viscmx);
#pragma unroll
    for (int c = 0; c < 5; ++c) FOR[c] = FOR[c] + p.wf[it] * s[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) MOM[c] = MOM[c] + p.wm[it] * s[c];
    zcfac1 = zcfac1 + etse / fnpt;
    if (etse < zcfac2) { zcfac2 = etse; zc2_at = it; }
  }
  const double fact = k.onep414 * p.dm;
  const double SQA = __builtin_sqrt(AREA);
  const double visc = fact * SSP * SQA * dtinv * RHO;
  FOR[0] = FOR[0] + visc * (EXX + HALF * EYY);
  FOR[1] = FOR[1] + visc * (EYY + HALF * EXX);
  FOR[2] = FOR[2] + visc * EXY * k.third;
  const double off_f = off;   // the resultants are scaled by OFF before CNDT3
#pragma unroll
  for (int c = 0; c < 5; ++c) FOR[c] = FOR[c] * off;
#pragma unroll
  for (int c = 0; c < 3; ++c) MOM[c] = MOM[c] * off;
  // CNDT3's deletion
  viscmx = tmax(viscmx, AMU);
  viscmx = __builtin_sqrt(ONE + viscmx * viscmx) - viscmx;
  {
    const double ALDT = g.LL * viscmx / __builtin_sqrt(ONE);
    const double DT = p.dtfac * ALDT / SSP;
    if (DT <= p.dtmin && off >= ONE && offg != TWO && offg >= ZERO) off = ZERO;
  }
  // CZFINTCE
  double VF[3][4], VM[2][4];
  const double *VS = FOR, *MS = MOM;
  const double C1c = THK02, C2c = C1c * Z1 * FOUR * AA;
  const double S1m = (MY34 * MX23 - MY23 * MX34) * THK;
  {
    const double X13S8 = X13 * MS[2], X24S8 = X24 * MS[2], Y13S8 = Y13 * MS[2], Y24S8 = Y24 * MS[2];
    VF[0][0] = THK * (Y24 * VS[0] - X24 * VS[2]);
    VF[1][0] = THK * (-X24 * VS[1] + Y24 * VS[2]);
    VF[2][0] = THK * (-X24 * VS[3] + Y24 * VS[4]);
    VM[0][0] = C1c * (X24 * MS[1] - Y24S8) - MY13 * VF[2][0];
    VM[1][0] = C1c * (Y24 * MS[0] - X24S8) + MX13 * VF[2][0];
    VM[0][2] = -S1m * VS[3];
    VM[1][2] = S1m * VS[4];
    VF[0][1] = THK * (-Y13 * VS[0] + X13 * VS[2]);
    VF[1][1] = THK * (X13 * VS[1] - Y13 * VS[2]);
    VF[2][1] = THK * (X13 * VS[3] - Y13 * VS[4]);
    VM[0][1] = C1c * (-X13 * MS[1] + Y13S8) + MY13 * VF[2][1];
    VM[1][1] = C1c * (-Y13 * MS[0] + X13S8) - MX13 * VF[2][1];
    VM[0][3] = VM[0][2];
    VM[1][3] = VM[1][2];
    VF[0][0] = VF[0][0] + C2c * (X13 * MS[0] + Y13S8);
    VF[1][0] = VF[1][0] + C2c * (Y13 * MS[1] + X13S8);
    VF[0][1] = VF[0][1] - C2c * (X24 * MS[0] + Y24S8);
    VF[1][1] = VF[1][1] - C2c * (Y24 * MS[1] + X24S8);
  }
  const double VF20ce = VF[2][0], VF21ce = VF[2][1];
  // CZFINTN1
  const double FAC1 = p.cvis, COEFH = k.zep999, COEF = k.zep85, STIER = k.fivep333;
  const double C7 = k.four_over_3, TOLH = k.em18, COEF1 = 25.0;
  const double FBEND = (p.npt == 1) ? ZERO : k.one_over_12;
  const double FBV = (p.npt == 1) ? ZERO : k.threep464;
  const double UNDOUZSR = __builtin_sqrt(k.one_over_12);
  double vgl[13], DHG[6], dhg[6];
#pragma unroll
  for (int q = 1; q <= 5; ++q) vgl[q] = st(SC_HOUR + q - 1);
#pragma unroll
  for (int q = 6; q <= 12; ++q) vgl[q] = st(SC_HOURX + q - 6);
#pragma unroll
  for (int q = 0; q < 6; ++q) { DHG[q] = g.DHG[q]; dhg[q] = DHG[q] * dt1; }
  const double C3h = FOUR * AA;
  const double HXX = C3h * MY34, HYY = C3h * MX34, HXX_K = C3h * MY23, HYY_K = C3h * MX23;
  const double C1M = A11 * FAC1, C2M = A12 * FAC1, C2g = FAC1 * G * SHF * k.one_over_64;
  const double C6 = C1c * FBEND;
  const double ESX = (MY34 * vgl[1] + MY23 * vgl[7]) * dhg[0] + (MX23 * vgl[8] + MX34 * vgl[2]) * dhg[1];
  const double EMX = ((MY34 * vgl[3] + MY23 * vgl[9]) * dhg[2] -
                      (-MX23 * vgl[10] - MX34 * vgl[4]) * dhg[3]) * C6;
  double DG[13];
  {
    const double CXX = HXX * dhg[0], CYY = HYY * dhg[1], CXX_K = HXX_K * dhg[0], CYY_K = HYY_K * dhg[1];
    const double BXX = HXX * dhg[2], BYY = HYY * dhg[3], BXX_K = HXX_K * dhg[2], BYY_K = HYY_K * dhg[3];
    DG[1] = C1M * CXX - C2M * CYY;
    DG[2] = C1M * CYY - C2M * CXX;
    DG[3] = C1M * BXX - C2M * BYY;
    DG[4] = C1M * BYY - C2M * BXX;
    DG[7] = C1M * CXX_K - C2M * CYY_K;
    DG[8] = C1M * CYY_K - C2M * CXX_K;
    DG[9] = C1M * BXX_K - C2M * BYY_K;
    DG[10] = C1M * BYY_K - C2M * BXX_K;
    DG[5] = C2g * HXX * dhg[4];
    DG[6] = C2g * HYY * dhg[4];
    DG[11] = C2g * HXX_K * dhg[5];
    DG[12] = C2g * HYY_K * dhg[5];
  }
  double vglm[13];
#pragma unroll
  for (int q = 1; q <= 12; ++q) vglm[q] = vgl[q] + DG[q];
  // the hourglass limiter (ZCFAC, SIGY)
  bool lim = false, elastic = false, eh1_ratio = false, eh1_zc = false, eh2_zc = false;
  double EH1 = ZERO, EH2 = ZERO, SXY0 = ZERO, SIGY2 = ZERO, ratio = ZERO, rden = ONE;
  double CNN[4] = {ZERO, ZERO, ZERO, ZERO};
  double sgnA = ONE;
  if (sigy < k.zep9ep30) {
    const double UFAC = tabs(tmin(zcfac1, zcfac2) - ONE);
    SIGY2 = sigy * sigy;
    double SVM = ZERO;
    elastic = UFAC < TOLH;
    if (elastic) {
      SXY0 = VS[0] * VS[0] + VS[1] * VS[1] - VS[0] * VS[1] + THREE * VS[2] * VS[2];
      double MXY0 = MS[0] * MS[0] + MS[1] * MS[1] - MS[0] * MS[1] + THREE * MS[2] * MS[2];
      const double CMM = COEF * THK * k.one_over_16;
      CNN[0] = COEF * vglm[1]; CNN[1] = COEF * vglm[2]; CNN[2] = COEF * vglm[7]; CNN[3] = COEF * vglm[8];
      const double CMMX = CMM * vglm[3], CMMY = CMM * vglm[4], CMMX_K = CMM * vglm[9], CMMY_K = CMM * vglm[10];
      SXY0 = SXY0 + CNN[0] * CNN[0] + CNN[1] * CNN[1] - CNN[0] * CNN[1];
      MXY0 = MXY0 + CMMX * CMMX + CMMY * CMMY - CMMX * CMMY;
      SXY0 = SXY0 + CNN[2] * CNN[2] + CNN[3] * CNN[3] - CNN[2] * CNN[3];
      MXY0 = MXY0 + CMMX_K * CMMX_K + CMMY_K * CMMY_K - CMMX_K * CMMY_K;
      const double Aabs = CNN[0] * (TWO * CNN[2] - CNN[3]) + CNN[1] * (TWO * CNN[3] - CNN[2]);
      sgnA = (Aabs < ZERO) ? -ONE : ONE;
      SXY0 = SXY0 + tabs(Aabs);
      MXY0 = MXY0 + tabs(CMMX * (TWO * CMMX_K - CMMY_K) + CMMY * (TWO * CMMY_K - CMMX_K));
      SVM = SXY0 + COEF1 * MXY0;
    }
    if (!elastic || SVM > SIGY2) {
      lim = true;
      rden = tmax(SIGY2, TOLH);
      ratio = SXY0 / rden;
      const double EH1r = tmin(ratio, ONE);
      eh1_ratio = ratio < ONE;
      EH1 = tmax(COEFH * EH1r, (ONE - zcfac1));
      eh1_zc = !(COEFH * EH1r > (ONE - zcfac1));
      EH2 = tmax(COEFH, (ONE - zcfac2));
      eh2_zc = !(COEFH > (ONE - zcfac2));
      if (ESX < ZERO) EH1 = ZERO;
      if (EMX < ZERO) EH2 = ZERO;
    }
  }
  double vglf[13];
#pragma unroll
  for (int q = 1; q <= 12; ++q) vglf[q] = vglm[q];
  if (lim) {
    vglf[1] -= EH1 * DG[1]; vglf[2] -= EH1 * DG[2]; vglf[7] -= EH1 * DG[7]; vglf[8] -= EH1 * DG[8];
    vglf[3] -= EH2 * DG[3]; vglf[4] -= EH2 * DG[4]; vglf[9] -= EH2 * DG[9]; vglf[10] -= EH2 * DG[10];
  }
  const double C8 = C7 * off, C2h = C8 * THK, HSURA = THK * AA;
  const double HVL = AMU * __builtin_sqrt(RHO * AREA * FAC1) * off;
  const double SSV0 = MY23 * MY23, SSV1 = MY34 * MY34, SSV2 = MX23 * MX23, SSV3 = MX34 * MX34;
  const double HXX_V = STIER * (SSV1 + SSV0);
  const double HXY_V = -STIER * (MY34 * MX34 + MY23 * MX23);
  const double HYY_V = STIER * (SSV2 + SSV3);
  const double C2v = HVL * GSR * SHFSR * UNDOUZSR;
  const double CXZ_V = (SSV1 + SSV3) * C2v, CYZ_V = (SSV2 + SSV0) * C2v;
  const double AUX = AA * HVL, C1Mv = A11SR * AUX, C2Mv = A12SR * AUX;
  const double CXX_V = C1Mv * HXX_V, CYY_V = C1Mv * HYY_V, CXY_V = C2Mv * HXY_V;
  const double SS1 = (MY34 * vglf[1] + MY23 * vglf[7]) * C8 + (CXX_V * DHG[0] + CXY_V * DHG[1]);
  const double SS2 = (MX23 * vglf[8] + MX34 * vglf[2]) * C8 + (CYY_V * DHG[1] + CXY_V * DHG[0]);
  const double SF1 = (MY34 * vglf[3] + MY23 * vglf[9]) * C8 + (CXX_V * DHG[2] + CXY_V * DHG[3]) * FBV;
  const double SF2 = -(MX23 * vglf[10] + MX34 * vglf[4]) * C8 + (-CYY_V * DHG[3] - CXY_V * DHG[2]) * FBV;
  const double SC5 = (MY34 * vglf[5] + MX34 * vglf[6]) * C2h + CXZ_V * DHG[4] * HSURA;
  const double SC6 = (MY23 * vglf[11] + MX23 * vglf[12]) * C2h + CYZ_V * DHG[5] * HSURA;
  const double SS3 = SC5 + SC6;
  const double C2q = FOURTH * THK, C4 = C6 * C2q;
  const double B13 = (MY13 * X24 - MX13 * Y24) * HSURA, B24 = (MX13 * Y13 - MY13 * X13) * HSURA;
  const double C2z = Z1 * HSURA;
  VF[0][0] += B13 * SS1; VF[0][2] = C2q * SS1; VF[1][0] += B13 * SS2; VF[1][2] = C2q * SS2;
  VF[2][2] = SS3; VF[0][1] += B24 * SS1; VF[0][3] = -VF[0][2]; VF[1][1] += B24 * SS2;
  VF[1][3] = -VF[1][2]; VF[2][3] = -VF[2][2];
  {
    const double Y13S = MY13 * SS3, X13S = MX13 * SS3;
    const double Y34S6 = MY34 * SC6, Y23S5 = MY23 * SC5, X23S5 = MX23 * SC5, X34S6 = MX34 * SC6;
    double C3 = C6 * B13;
    VM[0][0] += C3 * SF2 + Y23S5 + Y34S6;
    VM[0][2] += C4 * SF2 - Y13S;
    VM[1][0] += C3 * SF1 - X23S5 - X34S6;
    VM[1][2] += C4 * SF1 + X13S;
    C3 = C6 * B24;
    VM[0][1] += C3 * SF2 + Y23S5 - Y34S6;
    VM[0][3] += -C4 * SF2 - Y13S;
    VM[1][1] += C3 * SF1 - X23S5 + X34S6;
    VM[1][3] += -C4 * SF1 + X13S;
  }
  VF[2][0] += C2z * (SS1 * Y24 - SS2 * X24);
  VF[2][1] += C2z * (-SS1 * Y13 + SS2 * X13);
  // CZPROJ1 / CZPROJN
  double FL[3][4], ML[2][4];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    FL[d][0] = VF[d][0] + VF[d][2]; FL[d][1] = VF[d][1] + VF[d][3];
    FL[d][2] = -VF[d][0] + VF[d][2]; FL[d][3] = -VF[d][1] + VF[d][3];
  }
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    ML[d][0] = VM[d][0] + VM[d][2]; ML[d][1] = VM[d][1] + VM[d][3];
    ML[d][2] = -VM[d][0] + VM[d][2]; ML[d][3] = -VM[d][1] + VM[d][3];
  }
  double S[3][3];
  S[0][0] = g.DI[0]; S[1][1] = g.DI[1]; S[2][2] = g.DI[2];
  S[0][1] = S[1][0] = g.DI[3]; S[0][2] = S[2][0] = g.DI[4]; S[1][2] = S[2][1] = g.DI[5];
  double FLf[3][4], MM[3][4];
  double AR[3] = {0, 0, 0}, AD[4] = {0, 0, 0, 0}, ALR[3] = {0, 0, 0}, ALD[4] = {0, 0, 0, 0};
  const double(&CO)[2][4] = g.COREL;
  const double(&VQN)[3][4] = g.VQN;
  if (g.plat) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      FLf[0][c] = FL[0][c]; FLf[1][c] = FL[1][c]; FLf[2][c] = FL[2][c];
      MM[0][c] = ML[0][c]; MM[1][c] = ML[1][c]; MM[2][c] = ZERO;
    }
  } else {
    AR[0] = -Z1 * (FL[1][0] - FL[1][1] + FL[1][2] - FL[1][3]);
    AR[1] = Z1 * (FL[0][0] - FL[0][1] + FL[0][2] - FL[0][3]);
    AR[2] = ZERO;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      AR[0] += CO[1][c] * FL[2][c] + ML[0][c];
      AR[1] += -CO[0][c] * FL[2][c] + ML[1][c];
      AR[2] += -CO[1][c] * FL[0][c] + CO[0][c] * FL[1][c];
      AD[c] = VQN[0][c] * ML[0][c] + VQN[1][c] * ML[1][c];
    }
    double w[3];
#pragma unroll
    for (int q = 0; q < 3; ++q)
      w[q] = AR[q] - (VQN[q][0] * AD[0] + VQN[q][1] * AD[1] + VQN[q][2] * AD[2] + VQN[q][3] * AD[3]);
#pragma unroll
    for (int q = 0; q < 3; ++q) ALR[q] = S[q][0] * w[0] + S[q][1] * w[1] + S[q][2] * w[2];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      ALD[c] = AD[c] - (VQN[0][c] * ALR[0] + VQN[1][c] * ALR[1] + VQN[2][c] * ALR[2]);
      const double sg = (c & 1) ? ONE : -ONE;
      FLf[0][c] = FL[0][c] + sg * Z1 * ALR[1] + CO[1][c] * ALR[2];
      FLf[1][c] = FL[1][c] - sg * Z1 * ALR[0] - CO[0][c] * ALR[2];
      FLf[2][c] = FL[2][c] - CO[1][c] * ALR[0] + CO[0][c] * ALR[1];
      MM[0][c] = ML[0][c] - ALR[0] - VQN[0][c] * ALD[c];
      MM[1][c] = ML[1][c] - ALR[1] - VQN[1][c] * ALD[c];
      MM[2][c] = -ALR[2] - VQN[2][c] * ALD[c];
    }
  }
  if (off < ONE) offg = off;
  const bool zero_out = offg < ZERO;

  // ================= reverse =================
  QKinAdj L{};
  double lthk0 = 0.0, lTHK02 = 0.0, lAREA = 0.0, lAA = 0.0;
  double lX13 = 0, lX24 = 0, lY13 = 0, lY24 = 0, lMX13 = 0, lMX23 = 0, lMX34 = 0;
  double lMY13 = 0, lMY23 = 0, lMY34 = 0, lZ1 = 0;
  double lFL[3][4] = {}, lML[2][4] = {};
  double lCO[2][4] = {}, lVQN[3][4] = {};
  double lE[3][3] = {};
  double lMM2[4] = {0, 0, 0, 0};
  if (!zero_out) {
    // slots emit -F and -M; F = sum_q e_q FL_q, M = sum_q e_q MM_q
    const double *e[3] = {g.e1, g.e2, g.e3};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double lf[3], lm[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) { lf[d] = -lslot(c, d); lm[d] = -lslot(c, 3 + d); }
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        lFL[q][c] = e[q][0] * lf[0] + e[q][1] * lf[1] + e[q][2] * lf[2];
        const double lmm = e[q][0] * lm[0] + e[q][1] * lm[1] + e[q][2] * lm[2];
#pragma unroll
        for (int d = 0; d < 3; ++d) lE[q][d] += lf[d] * FLf[q][c] + lm[d] * MM[q][c];
        if (q < 2) lML[q][c] = lmm;
        else lMM2[c] = lmm;
      }
    }
  }
  if (!g.plat && !zero_out) {
    double lALR[3] = {0, 0, 0}, lALD[4] = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double sg = (c & 1) ? ONE : -ONE;
      const double lf0 = lFL[0][c], lf1 = lFL[1][c], lf2 = lFL[2][c];
      const double lm0 = lML[0][c], lm1 = lML[1][c], lm2 = lMM2[c];
      // FLf0 = FL0 + sg Z1 ALR1 + CO1 ALR2; FLf1 = FL1 - sg Z1 ALR0 - CO0 ALR2;
      // FLf2 = FL2 - CO1 ALR0 + CO0 ALR1
      lALR[1] += lf0 * sg * Z1 + lf2 * CO[0][c];
      lALR[2] += lf0 * CO[1][c] - lf1 * CO[0][c];
      lALR[0] += -lf1 * sg * Z1 - lf2 * CO[1][c];
      lZ1 += sg * (lf0 * ALR[1] - lf1 * ALR[0]);
      lCO[1][c] += lf0 * ALR[2] - lf2 * ALR[0];
      lCO[0][c] += -lf1 * ALR[2] + lf2 * ALR[1];
      // MMq = MLq - ALRq - VQNq ALD (q < 2); MM2 = -ALR2 - VQN2 ALD
      lALR[0] -= lm0; lALR[1] -= lm1; lALR[2] -= lm2;
      lALD[c] -= lm0 * VQN[0][c] + lm1 * VQN[1][c] + lm2 * VQN[2][c];
      lVQN[0][c] -= lm0 * ALD[c]; lVQN[1][c] -= lm1 * ALD[c]; lVQN[2][c] -= lm2 * ALD[c];
    }
    double lAR[3] = {0, 0, 0}, lAD[4] = {0, 0, 0, 0};
    qadj::proj_rev(S, VQN, AR, AD, lALR, lALD, lAR, lAD, lVQN, L.G);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      // AD = VQN0 ML0 + VQN1 ML1
      lVQN[0][c] += lAD[c] * ML[0][c]; lVQN[1][c] += lAD[c] * ML[1][c];
      lML[0][c] += lAD[c] * VQN[0][c]; lML[1][c] += lAD[c] * VQN[1][c];
      // AR's sums
      const double sg = (c & 1) ? -ONE : ONE;
      lFL[1][c] += -lAR[0] * Z1 * sg + lAR[2] * CO[0][c];
      lFL[0][c] += lAR[1] * Z1 * sg - lAR[2] * CO[1][c];
      lFL[2][c] += lAR[0] * CO[1][c] - lAR[1] * CO[0][c];
      lCO[1][c] += lAR[0] * FL[2][c] - lAR[2] * FL[0][c];
      lCO[0][c] += -lAR[1] * FL[2][c] + lAR[2] * FL[1][c];
      lML[0][c] += lAR[0]; lML[1][c] += lAR[1];
    }
    lZ1 += -lAR[0] * (FL[1][0] - FL[1][1] + FL[1][2] - FL[1][3]) +
           lAR[1] * (FL[0][0] - FL[0][1] + FL[0][2] - FL[0][3]);
  }
  // FL, ML from VF, VM
  double lVF[3][4], lVM[2][4];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lVF[d][0] = lFL[d][0] - lFL[d][2]; lVF[d][2] = lFL[d][0] + lFL[d][2];
    lVF[d][1] = lFL[d][1] - lFL[d][3]; lVF[d][3] = lFL[d][1] + lFL[d][3];
  }
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    lVM[d][0] = lML[d][0] - lML[d][2]; lVM[d][2] = lML[d][0] + lML[d][2];
    lVM[d][1] = lML[d][1] - lML[d][3]; lVM[d][3] = lML[d][1] + lML[d][3];
  }
  // ---- CZFINTN1 reverse ----
  double lSS1 = 0, lSS2 = 0, lSS3 = 0, lSF1 = 0, lSF2 = 0, lSC5 = 0, lSC6 = 0;
  double lHSURA = 0, lC6 = 0, lC2q = 0;
  {
    // VF20 += C2z (SS1 Y24 - SS2 X24); VF21 += C2z (-SS1 Y13 + SS2 X13)
    const double a = lVF[2][0], b = lVF[2][1];
    const double lC2z = a * (SS1 * Y24 - SS2 * X24) + b * (-SS1 * Y13 + SS2 * X13);
    lSS1 += C2z * (a * Y24 - b * Y13);
    lSS2 += C2z * (-a * X24 + b * X13);
    lY24 += a * C2z * SS1; lX24 -= a * C2z * SS2; lY13 -= b * C2z * SS1; lX13 += b * C2z * SS2;
    lZ1 += lC2z * HSURA; lHSURA += lC2z * Z1;
    // moments
    const double C3a = C6 * B13, C3b = C6 * B24;
    lSF2 += C3a * lVM[0][0] + C4 * lVM[0][2] + C3b * lVM[0][1] - C4 * lVM[0][3];
    lSF1 += C3a * lVM[1][0] + C4 * lVM[1][2] + C3b * lVM[1][1] - C4 * lVM[1][3];
    const double lC3a = lVM[0][0] * SF2 + lVM[1][0] * SF1;
    const double lC3b = lVM[0][1] * SF2 + lVM[1][1] * SF1;
    const double lC4 = (lVM[0][2] - lVM[0][3]) * SF2 + (lVM[1][2] - lVM[1][3]) * SF1;
    const double lY23S5 = lVM[0][0] + lVM[0][1], lY34S6 = lVM[0][0] - lVM[0][1];
    const double lX23S5 = -lVM[1][0] - lVM[1][1], lX34S6 = -lVM[1][0] + lVM[1][1];
    const double lY13S = -lVM[0][2] - lVM[0][3], lX13S = lVM[1][2] + lVM[1][3];
    lMY13 += lY13S * SS3; lMX13 += lX13S * SS3; lSS3 += lY13S * MY13 + lX13S * MX13;
    lMY34 += lY34S6 * SC6; lMX34 += lX34S6 * SC6; lSC6 += lY34S6 * MY34 + lX34S6 * MX34;
    lMY23 += lY23S5 * SC5; lMX23 += lX23S5 * SC5; lSC5 += lY23S5 * MY23 + lX23S5 * MX23;
    lC6 += lC3a * B13 + lC3b * B24 + lC4 * C2q;
    double lB13 = lC3a * C6, lB24 = lC3b * C6;
    lC2q += lC4 * C6;
    // forces
    lSS1 += B13 * lVF[0][0] + C2q * (lVF[0][2] - lVF[0][3]) + B24 * lVF[0][1];
    lSS2 += B13 * lVF[1][0] + C2q * (lVF[1][2] - lVF[1][3]) + B24 * lVF[1][1];
    lSS3 += lVF[2][2] - lVF[2][3];
    lB13 += lVF[0][0] * SS1 + lVF[1][0] * SS2;
    lB24 += lVF[0][1] * SS1 + lVF[1][1] * SS2;
    lC2q += (lVF[0][2] - lVF[0][3]) * SS1 + (lVF[1][2] - lVF[1][3]) * SS2;
    // B13 = (MY13 X24 - MX13 Y24) HSURA; B24 = (MX13 Y13 - MY13 X13) HSURA
    lHSURA += lB13 * (MY13 * X24 - MX13 * Y24) + lB24 * (MX13 * Y13 - MY13 * X13);
    lMY13 += (lB13 * X24 - lB24 * X13) * HSURA;
    lMX13 += (-lB13 * Y24 + lB24 * Y13) * HSURA;
    lX24 += lB13 * HSURA * MY13; lY24 -= lB13 * HSURA * MX13;
    lY13 += lB24 * HSURA * MX13; lX13 -= lB24 * HSURA * MY13;
    lthk0 += lC2q * FOURTH;
    lSC5 += lSS3; lSC6 += lSS3;
  }
  double lDHG[6] = {0, 0, 0, 0, 0, 0};
  double lvglf[13];
#pragma unroll
  for (int q = 1; q <= 12; ++q) lvglf[q] = lpost(29 + q);
  {
    // viscous parts
    lDHG[0] += lSS1 * CXX_V + lSS2 * CXY_V;
    lDHG[1] += lSS1 * CXY_V + lSS2 * CYY_V;
    lDHG[2] += (lSF1 * CXX_V - lSF2 * CXY_V) * FBV;
    lDHG[3] += (lSF1 * CXY_V - lSF2 * CYY_V) * FBV;
    lDHG[4] += lSC5 * CXZ_V * HSURA;
    lDHG[5] += lSC6 * CYZ_V * HSURA;
    const double lCXX_V = lSS1 * DHG[0] + lSF1 * DHG[2] * FBV;
    const double lCYY_V = lSS2 * DHG[1] - lSF2 * DHG[3] * FBV;
    const double lCXY_V = lSS1 * DHG[1] + lSS2 * DHG[0] + (lSF1 * DHG[3] - lSF2 * DHG[2]) * FBV;
    const double lCXZ_V = lSC5 * DHG[4] * HSURA, lCYZ_V = lSC6 * DHG[5] * HSURA;
    lHSURA += lSC5 * CXZ_V * DHG[4] + lSC6 * CYZ_V * DHG[5];
    const double lAUX = (lCXX_V * HXX_V + lCYY_V * HYY_V) * A11SR + lCXY_V * HXY_V * A12SR;
    const double lHXX_V = lCXX_V * C1Mv, lHYY_V = lCYY_V * C1Mv, lHXY_V = lCXY_V * C2Mv;
    lAA += lAUX * HVL;
    double lHVL = lAUX * AA;
    const double lC2v = lCXZ_V * (SSV1 + SSV3) + lCYZ_V * (SSV2 + SSV0);
    const double lSSV0 = lCYZ_V * C2v + STIER * lHXX_V, lSSV1 = lCXZ_V * C2v + STIER * lHXX_V;
    const double lSSV2 = lCYZ_V * C2v + STIER * lHYY_V, lSSV3 = lCXZ_V * C2v + STIER * lHYY_V;
    lHVL += lC2v * GSR * SHFSR * UNDOUZSR;
    lMY34 += -STIER * lHXY_V * MX34 + TWO * lSSV1 * MY34;
    lMX34 += -STIER * lHXY_V * MY34 + TWO * lSSV3 * MX34;
    lMY23 += -STIER * lHXY_V * MX23 + TWO * lSSV0 * MY23;
    lMX23 += -STIER * lHXY_V * MY23 + TWO * lSSV2 * MX23;
    if (HVL != ZERO) lAREA += lHVL * HVL * HALF / AREA;
    // hourglass history parts
    lvglf[1] += lSS1 * MY34 * C8; lvglf[7] += lSS1 * MY23 * C8;
    lMY34 += lSS1 * vglf[1] * C8; lMY23 += lSS1 * vglf[7] * C8;
    lvglf[8] += lSS2 * MX23 * C8; lvglf[2] += lSS2 * MX34 * C8;
    lMX23 += lSS2 * vglf[8] * C8; lMX34 += lSS2 * vglf[2] * C8;
    lvglf[3] += lSF1 * MY34 * C8; lvglf[9] += lSF1 * MY23 * C8;
    lMY34 += lSF1 * vglf[3] * C8; lMY23 += lSF1 * vglf[9] * C8;
    lvglf[10] -= lSF2 * MX23 * C8; lvglf[4] -= lSF2 * MX34 * C8;
    lMX23 -= lSF2 * vglf[10] * C8; lMX34 -= lSF2 * vglf[4] * C8;
    lvglf[5] += lSC5 * MY34 * C2h; lvglf[6] += lSC5 * MX34 * C2h;
    lMY34 += lSC5 * vglf[5] * C2h; lMX34 += lSC5 * vglf[6] * C2h;
    lvglf[11] += lSC6 * MY23 * C2h; lvglf[12] += lSC6 * MX23 * C2h;
    lMY23 += lSC6 * vglf[11] * C2h; lMX23 += lSC6 * vglf[12] * C2h;
    const double lC2h = lSC5 * (MY34 * vglf[5] + MX34 * vglf[6]) +
                        lSC6 * (MY23 * vglf[11] + MX23 * vglf[12]);
    lthk0 += lC2h * C8;
    // HSURA = THK AA
    lthk0 += lHSURA * AA; lAA += lHSURA * THK;
  }
  // limiter and the history increment
  double lDG[13], lvglm[13];
  double lzcfac1 = 0.0, lzcfac2 = 0.0, lsigy = 0.0;
  double lFORf[5] = {0, 0, 0, 0, 0}, lMOMf[3] = {0, 0, 0};
#pragma unroll
  for (int q = 1; q <= 12; ++q) { lvglm[q] = lvglf[q]; lDG[q] = 0.0; }
  if (lim) {
    double lEH1 = 0.0, lEH2 = 0.0;
    const int q1[4] = {1, 2, 7, 8}, q2[4] = {3, 4, 9, 10};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      lEH1 -= lvglf[q1[i]] * DG[q1[i]]; lDG[q1[i]] -= EH1 * lvglf[q1[i]];
      lEH2 -= lvglf[q2[i]] * DG[q2[i]]; lDG[q2[i]] -= EH2 * lvglf[q2[i]];
    }
    if (!(ESX < ZERO)) {
      if (eh1_zc) {
        lzcfac1 -= lEH1;
      } else if (eh1_ratio) {
        const double lratio = COEFH * lEH1;
        double lSXY0 = lratio / rden;
        if (SIGY2 > TOLH) lsigy += -lratio * ratio / rden * TWO * sigy;
        if (elastic) {
          lFORf[0] += lSXY0 * (TWO * VS[0] - VS[1]);
          lFORf[1] += lSXY0 * (TWO * VS[1] - VS[0]);
          lFORf[2] += lSXY0 * 6.0 * VS[2];
          const double lc0 = lSXY0 * (TWO * CNN[0] - CNN[1] + sgnA * (TWO * CNN[2] - CNN[3]));
          const double lc1 = lSXY0 * (TWO * CNN[1] - CNN[0] + sgnA * (TWO * CNN[3] - CNN[2]));
          const double lc2 = lSXY0 * (TWO * CNN[2] - CNN[3] + sgnA * (TWO * CNN[0] - CNN[1]));
          const double lc3 = lSXY0 * (TWO * CNN[3] - CNN[2] + sgnA * (TWO * CNN[1] - CNN[0]));
          lvglm[1] += COEF * lc0; lvglm[2] += COEF * lc1;
          lvglm[7] += COEF * lc2; lvglm[8] += COEF * lc3;
        }
      }
    }
    if (!(EMX < ZERO) && eh2_zc) lzcfac2 -= lEH2;
  }
  // vglm = vgl_in + DG
#pragma unroll
  for (int q = 1; q <= 12; ++q) {
    lDG[q] += lvglm[q];
    lpre(29 + q, lvglm[q]);
  }
  {
    const double lCXX = C1M * lDG[1] - C2M * lDG[2], lCYY = -C2M * lDG[1] + C1M * lDG[2];
    const double lBXX = C1M * lDG[3] - C2M * lDG[4], lBYY = -C2M * lDG[3] + C1M * lDG[4];
    const double lCXX_K = C1M * lDG[7] - C2M * lDG[8], lCYY_K = -C2M * lDG[7] + C1M * lDG[8];
    const double lBXX_K = C1M * lDG[9] - C2M * lDG[10], lBYY_K = -C2M * lDG[9] + C1M * lDG[10];
    const double lHXX = lCXX * dhg[0] + lBXX * dhg[2] + C2g * lDG[5] * dhg[4];
    const double lHYY = lCYY * dhg[1] + lBYY * dhg[3] + C2g * lDG[6] * dhg[4];
    const double lHXX_K = lCXX_K * dhg[0] + lBXX_K * dhg[2] + C2g * lDG[11] * dhg[5];
    const double lHYY_K = lCYY_K * dhg[1] + lBYY_K * dhg[3] + C2g * lDG[12] * dhg[5];
    double ldhg[6];
    ldhg[0] = lCXX * HXX + lCXX_K * HXX_K;
    ldhg[1] = lCYY * HYY + lCYY_K * HYY_K;
    ldhg[2] = lBXX * HXX + lBXX_K * HXX_K;
    ldhg[3] = lBYY * HYY + lBYY_K * HYY_K;
    ldhg[4] = C2g * (lDG[5] * HXX + lDG[6] * HYY);
    ldhg[5] = C2g * (lDG[11] * HXX_K + lDG[12] * HYY_K);
#pragma unroll
    for (int q = 0; q < 6; ++q) lDHG[q] += ldhg[q] * dt1;
    const double lC3h = lHXX * MY34 + lHYY * MX34 + lHXX_K * MY23 + lHYY_K * MX23;
    lAA += FOUR * lC3h;
    lMY34 += lHXX * C3h; lMX34 += lHYY * C3h; lMY23 += lHXX_K * C3h; lMX23 += lHYY_K * C3h;
  }
  lTHK02 += lC6 * FBEND;
  // ---- CZFINTCE reverse ----
  double lVS[5] = {lFORf[0], lFORf[1], lFORf[2], lFORf[3], lFORf[4]};
  double lMS[3] = {lMOMf[0], lMOMf[1], lMOMf[2]};
  {
    const double lVM02t = lVM[0][2] + lVM[0][3], lVM12t = lVM[1][2] + lVM[1][3];
    const double lS1m = -lVM02t * VS[3] + lVM12t * VS[4];
    lVS[3] -= lVM02t * S1m; lVS[4] += lVM12t * S1m;
    lMY34 += lS1m * MX23 * THK; lMX23 += lS1m * MY34 * THK;
    lMY23 -= lS1m * MX34 * THK; lMX34 -= lS1m * MY23 * THK;
    lthk0 += lS1m * (MY34 * MX23 - MY23 * MX34);
    const double m00 = lVM[0][0], m10 = lVM[1][0], m01 = lVM[0][1], m11 = lVM[1][1];
    const double lF20 = lVF[2][0] - m00 * MY13 + m10 * MX13;
    const double lF21 = lVF[2][1] + m01 * MY13 - m11 * MX13;
    lMY13 += -m00 * VF20ce + m01 * VF21ce;
    lMX13 += m10 * VF20ce - m11 * VF21ce;
    double lC1c = m00 * (X24 * MS[1] - Y24 * MS[2]) + m10 * (Y24 * MS[0] - X24 * MS[2]) +
                  m01 * (-X13 * MS[1] + Y13 * MS[2]) + m11 * (-Y13 * MS[0] + X13 * MS[2]);
    lX24 += C1c * (m00 * MS[1] - m10 * MS[2]);
    lY24 += C1c * (-m00 * MS[2] + m10 * MS[0]);
    lX13 += C1c * (-m01 * MS[1] + m11 * MS[2]);
    lY13 += C1c * (m01 * MS[2] - m11 * MS[0]);
    lMS[0] += C1c * (m10 * Y24 - m11 * Y13);
    lMS[1] += C1c * (m00 * X24 - m01 * X13);
    lMS[2] += C1c * (-m00 * Y24 - m10 * X24 + m01 * Y13 + m11 * X13);
    const double a = lVF[0][0], b = lVF[1][0], c = lF20, d = lVF[0][1], e = lVF[1][1], f = lF21;
    lthk0 += a * (Y24 * VS[0] - X24 * VS[2]) + b * (-X24 * VS[1] + Y24 * VS[2]) +
             c * (-X24 * VS[3] + Y24 * VS[4]) + d * (-Y13 * VS[0] + X13 * VS[2]) +
             e * (X13 * VS[1] - Y13 * VS[2]) + f * (X13 * VS[3] - Y13 * VS[4]);
    lVS[0] += THK * (a * Y24 - d * Y13);
    lVS[1] += THK * (-b * X24 + e * X13);
    lVS[2] += THK * (-a * X24 + b * Y24 + d * X13 - e * Y13);
    lVS[3] += THK * (-c * X24 + f * X13);
    lVS[4] += THK * (c * Y24 - f * Y13);
    lY24 += THK * (a * VS[0] + b * VS[2] + c * VS[4]);
    lX24 -= THK * (a * VS[2] + b * VS[1] + c * VS[3]);
    lY13 -= THK * (d * VS[0] + e * VS[2] + f * VS[4]);
    lX13 += THK * (d * VS[2] + e * VS[1] + f * VS[3]);
    const double lC2c = a * (X13 * MS[0] + Y13 * MS[2]) + b * (Y13 * MS[1] + X13 * MS[2]) -
                        d * (X24 * MS[0] + Y24 * MS[2]) - e * (Y24 * MS[1] + X24 * MS[2]);
    lX13 += C2c * (a * MS[0] + b * MS[2]);
    lY13 += C2c * (a * MS[2] + b * MS[1]);
    lX24 -= C2c * (d * MS[0] + e * MS[2]);
    lY24 -= C2c * (d * MS[2] + e * MS[1]);
    lMS[0] += C2c * (a * X13 - d * X24);
    lMS[1] += C2c * (b * Y13 - e * Y24);
    lMS[2] += C2c * (a * Y13 + b * X13 - d * Y24 - e * X24);
    lC1c += lC2c * Z1 * FOUR * AA;
    lZ1 += lC2c * C1c * FOUR * AA;
    lAA += lC2c * C1c * Z1 * FOUR;
    lTHK02 += lC1c;
  }
  // ---- resultants: scaling by OFF and membrane damping ----
  double lFORs[5], lMOMs[3];
#pragma unroll
  for (int c = 0; c < 5; ++c) lFORs[c] = lVS[c] * off_f;
#pragma unroll
  for (int c = 0; c < 3; ++c) lMOMs[c] = lMS[c] * off_f;
  double lEXX = 0.0, lEYY = 0.0, lEXY = 0.0, lEYZ = 0.0, lEXZ = 0.0;
  double lKXX = 0.0, lKYY = 0.0, lKXY = 0.0;
  {
    const double lvisc = lFORs[0] * (EXX + HALF * EYY) + lFORs[1] * (EYY + HALF * EXX) +
                         lFORs[2] * EXY * k.third;
    lEXX += visc * (lFORs[0] + HALF * lFORs[1]);
    lEYY += visc * (lFORs[1] + HALF * lFORs[0]);
    lEXY += visc * k.third * lFORs[2];
    if (SQA > 0.0) lAREA += lvisc * fact * SSP * dtinv * RHO * HALF / SQA;
  }
  // ---- law 2 at every point, with ETSE (ZCFAC) and YLD (SIGY) ----
#pragma unroll
  for (int it = 0; it < 5; ++it) {
    if (it >= p.npt) break;
    const double offp = off_pt[it];
    const double z0 = p.z0[it], zt = z0 * thk0;
    const double d0 = EXX + zt * KXX, d1 = EYY + zt * KYY, d2 = EXY + zt * KXY;
    double s[5];
    s[0] = st(SC_SIG + 5 * it) + a11 * d0 + a12 * d1;
    s[1] = st(SC_SIG + 5 * it + 1) + a12 * d0 + a11 * d1;
    s[2] = st(SC_SIG + 5 * it + 2) + G * d2;
    s[3] = st(SC_SIG + 5 * it + 3) + GS * EYZ;
    s[4] = st(SC_SIG + 5 * it + 4) + GS * EXZ;
    const double pla = st(SC_PLA + it);
    double ppow = 0.0, yld;
    if (pla == ZERO) yld = ca;
    else { ppow = M<double>::powe(pla, cn); yld = ca + cb * ppow; }
    const bool capped = !(yld < ymax);
    yld = tmin(yld, ymax);
    const double q = s[0] * s[0] + s[1] * s[1] - s[0] * s[1] + THREE * s[2] * s[2];
    const double svm = __builtin_sqrt(q);
    const double den = svm + k.em15;
    const double r = tmin(ONE, yld / den);
    double ls[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) ls[c] = lpost(5 * it + c) + p.wf[it] * lFORs[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) ls[c] += p.wm[it] * lMOMs[c];
    double lpla_out = lpost(25 + it);
    double lstr[5] = {ls[0], ls[1], ls[2], ls[3], ls[4]};
    double lyld = lsigy / fnpt;
    if (r < ONE) {
      const double dpla = offp * tmax(ZERO, (svm - yld) / YM);
      if (!capped) {
        // ETSE = h / (h + E), h = n B (pla_out + SMALL)^(n - 1)
        const double letse = lzcfac1 / fnpt + ((zc2_at == it) ? lzcfac2 : 0.0);
        const double P = pla + dpla + SMALL;
        const double Pw = M<double>::powe(P, cn - ONE);
        const double h = cn * cb * Pw;
        const double lh = letse * YM / ((h + YM) * (h + YM));
        g_jc[1] += lh * cn * Pw;
        g_jc[2] += lh * (cb * Pw + h * M<double>::log_(P));
        lpla_out += lh * h * (cn - ONE) / P;
      }
      // s_out = r s_tr (c < 3); r = yld / (svm + em15); pla_out = pla + off (svm - yld)/E
      const double lr = ls[0] * s[0] + ls[1] * s[1] + ls[2] * s[2];
      lstr[0] = ls[0] * r; lstr[1] = ls[1] * r; lstr[2] = ls[2] * r;
      double lsvm = -lr * yld / (den * den);
      lyld += lr / den;
      if (!(ZERO > (svm - yld) / YM)) {   // tmax(0, .) keeps the second operand
        lsvm += lpla_out * offp / YM;
        lyld -= lpla_out * offp / YM;
      }
      if (svm > 0.0) {
        const double lq = lsvm * HALF / svm;
        lstr[0] += lq * (TWO * s[0] - s[1]);
        lstr[1] += lq * (TWO * s[1] - s[0]);
        lstr[2] += lq * (6.0 * s[2]);
      }
    }
    double lpla = lpla_out;
    if (!capped) {
      g_jc[0] += lyld;
      if (pla != ZERO) {
        g_jc[1] += lyld * ppow;
        const double lp = lyld * cb * ppow;
        g_jc[2] += lp * M<double>::log_(pla);
        lpla += lp * cn / pla;
      }
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) lpre(5 * it + c, lstr[c]);
    lpre(25 + it, lpla);
    const double ld0 = a11 * lstr[0] + a12 * lstr[1], ld1 = a12 * lstr[0] + a11 * lstr[1];
    const double ld2 = G * lstr[2];
    lEYZ += GS * lstr[3];
    lEXZ += GS * lstr[4];
    lEXX += ld0; lEYY += ld1; lEXY += ld2;
    lKXX += zt * ld0; lKYY += zt * ld1; lKXY += zt * ld2;
    lthk0 += z0 * (KXX * ld0 + KYY * ld1 + KXY * ld2);
  }
  // ---- CZSTRA3 (rates times DT1) and the kinematics ----
  L.VDEF[0] = lEXX * dt1; L.VDEF[1] = lEYY * dt1; L.VDEF[2] = lEXY * dt1;
  L.VDEF[3] = lEXZ * dt1; L.VDEF[4] = lEYZ * dt1;
  L.VDEF[5] = lKXX * dt1; L.VDEF[6] = lKYY * dt1; L.VDEF[7] = lKXY * dt1;
#pragma unroll
  for (int q = 0; q < 6; ++q) L.DHG[q] = lDHG[q];
#pragma unroll
  for (int d = 0; d < 3; ++d) { L.e1[d] = lE[0][d]; L.e2[d] = lE[1][d]; L.e3[d] = lE[2][d]; }
  L.AREA = lAREA; L.AREA_I = lAA;
  L.X13 = lX13; L.X24 = lX24; L.Y13 = lY13; L.Y24 = lY24;
  L.MX13 = lMX13; L.MX23 = lMX23; L.MX34 = lMX34; L.MY13 = lMY13; L.MY23 = lMY23; L.MY34 = lMY34;
  L.Z1 = lZ1;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    L.COREL[0][c] = lCO[0][c]; L.COREL[1][c] = lCO[1][c];
    L.VQN[0][c] = lVQN[0][c]; L.VQN[1][c] = lVQN[1][c]; L.VQN[2][c] = lVQN[2][c];
  }
  qeph_kin_rev(xg, vg, vrg, dt1, st(SC_OFFG), p.npt, k, L, lx, lv, lvr);
  g_thk += lthk0 + lTHK02 * TWO * thk0;
}

}  // namespace orad

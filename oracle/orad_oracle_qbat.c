/*
 * orad_oracle_qbat.c -- CPU restatement (FP64) of the OpenRadioss QBAT shell
 * (/PROP/SHELL Ishell=12 -> IHBE=11, CBAFORC3 with 2 x 2 Gauss points).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP QBAT kernels.  Never linked
 * into openradioss_amd/.
 *
 * Scope: IGTYP=1, law 2 (IPLA 0/1/2, rate term), Ismstr=2, ISROT=0 (no drilling
 * DOF), IREP=0, no pinching, NPT 2..5 through-thickness points, PARITH/ON slots,
 * IMPL_S=0.  Operation order follows the double-precision build so the result
 * can be compared bit for bit with the reference's routines compiled from source
 * and driven in CBAFORC3 order by oracle/ref/harness_qbat.F.  Literals that the
 * reference writes as default-REAL constants (the Gauss abscissa PG, the
 * characteristic-length factors of CBACOOR) are single-precision values widened
 * to double, as that build holds them.
 *
 * Parity pins (tests/test_oracle_qbat.py): every output of the golden groups
 * (forces, moments, per-corner STI, the whole 190-value state, the damping
 * energy, DT2T) bit for bit; closed forms (rigid motion gives zero forces and
 * strain rates, a flat element takes CBACOOR's flat branch and gives no moment
 * about its normal).
 */
#include "orad_oracle.h"
#include "orad_oracle_int.h"
#include <math.h>
#include <string.h>

#define ZERO 0.0
#define ONE 1.0
#define TWO 2.0
#define FOUR 4.0
#define HALF 0.5
#define FOURTH 0.25

static inline double dmin(double a, double b) { return (a < b) ? a : b; }
static inline double dmax(double a, double b) { return (a > b) ? a : b; }

void oracle_qbat_element(const double xg[3][4], const double vg[3][4], const double vrg[3][4],
                         double thke, orad_qbstate_t *st, double dt1,
                         const orad_shmat_t *m, const orad_shprop_t *p, orad_shout_t *out) {
  const shconst_t *K = oracle_shk();
  const double EM20 = K->EM20, EM10 = K->EM10;
  /* DATA PG/.577350269189626/ (cbacoor.F:158) and PARAMETER PG, PG1 (cbadef.F:73-74):
   * default-REAL literals */
  const double PG = (double)0.577350269189626f;
  const double PG1 = (double)(-0.577350269189626f);
  /* ---- CBACOOR (coqueba/cbacoor.F:37-1112), ISMSTR = 2, ISROT = 0, NPINCH = 0 ---- */
  double VKSI[4][4], VETA[4][4];                 /* [node][gauss point] */
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

  double RX = xg[0][1] + xg[0][2] - xg[0][0] - xg[0][3];
  double SX = xg[0][2] + xg[0][3] - xg[0][0] - xg[0][1];
  double RY = xg[1][1] + xg[1][2] - xg[1][0] - xg[1][3];
  double SY = xg[1][2] + xg[1][3] - xg[1][0] - xg[1][1];
  double RZ = xg[2][1] + xg[2][2] - xg[2][0] - xg[2][3];
  double SZ = xg[2][2] + xg[2][3] - xg[2][0] - xg[2][1];
  /* CLSKEW3 with IREP = 0 (sh3n/coquedk/cdkcoor3.F:303-380) */
  double E3X = RY * SZ - RZ * SY, E3Y = RZ * SX - RX * SZ, E3Z = RX * SY - RY * SX;
  double DETA1 = sqrt(E3X * E3X + E3Y * E3Y + E3Z * E3Z);
  if (DETA1 < EM20 && st->offg != ZERO) st->offg = ZERO;
  {
    double off_loc = (fabs(st->offg) != ZERO) ? ONE : ZERO;
    DETA1 = dmax(EM20, DETA1);
    double CC = off_loc / DETA1;
    CC = dmax(CC, EM20);
    E3X = E3X * CC; E3Y = E3Y * CC; E3Z = E3Z * CC;
  }
  double E1X, E1Y, E1Z;
  {
    double C1C1 = RX * RX + RY * RY + RZ * RZ;
    double C2C2 = SX * SX + SY * SY + SZ * SZ;
    double c2_1 = 0.0, c1_1 = 0.0;
    if (C1C1 != ZERO) {
      c2_1 = sqrt(C2C2 / dmax(EM20, C1C1));
      c1_1 = ONE;
    } else if (C2C2 != ZERO) {
      c2_1 = ONE;
      c1_1 = sqrt(C1C1 / dmax(EM20, C2C2));
    }
    E1X = RX * c2_1 + (SY * E3Z - SZ * E3Y) * c1_1;
    E1Y = RY * c2_1 + (SZ * E3X - SX * E3Z) * c1_1;
    E1Z = RZ * c2_1 + (SX * E3Y - SY * E3X) * c1_1;
    double C1 = sqrt(E1X * E1X + E1Y * E1Y + E1Z * E1Z);
    if (C1 != ZERO) C1 = ONE / dmax(EM20, C1);
    E1X = E1X * C1; E1Y = E1Y * C1; E1Z = E1Z * C1;
  }
  double E2X = E3Y * E1Z - E3Z * E1Y, E2Y = E3Z * E1X - E3X * E1Z, E2Z = E3X * E1Y - E3Y * E1X;
  /* VQ(:,j) = E_j */
  const double e1[3] = {E1X, E1Y, E1Z}, e2[3] = {E2X, E2Y, E2Z}, e3[3] = {E3X, E3Y, E3Z};
  double AREA = FOURTH * DETA1;
  const double AREA_I = ONE / AREA;
  double XL2, YL2, XL3, YL3, XL4, YL4, ZL1;
  {
    double L0[3];
    for (int d = 0; d < 3; ++d) L0[d] = FOURTH * (xg[d][2] + xg[d][3] + xg[d][0] + xg[d][1]);
    double XX = xg[0][1] - xg[0][0], YY = xg[1][1] - xg[1][0], ZZ = xg[2][1] - xg[2][0];
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
  /* ISMSTR = 2: the small-strain reference (SMSTR) is stored but read only after a
   * switch to small strain (|OFFG| = 2), which needs /DT/SHELL/CST. */
  double VC[3][4];                                  /* VCORE(:, node) */
  double X13T, X24T, Y13T, Y24T, L13, L24, LL, LM;
  {
    double LX = FOURTH * (XL2 + XL3 + XL4), LY = FOURTH * (YL2 + YL3 + YL4);
    VC[0][0] = -LX; VC[0][1] = XL2 - LX; VC[0][2] = XL3 - LX; VC[0][3] = XL4 - LX;
    VC[1][0] = -LY; VC[1][1] = YL2 - LY; VC[1][2] = YL3 - LY; VC[1][3] = YL4 - LY;
    VC[2][0] = VC[2][1] = VC[2][2] = VC[2][3] = ZERO;
    X13T = (VC[0][0] - VC[0][2]) * HALF;
    X24T = (VC[0][1] - VC[0][3]) * HALF;
    Y13T = (VC[1][0] - VC[1][2]) * HALF;
    Y24T = (VC[1][1] - VC[1][3]) * HALF;
    L13 = X13T * X13T + Y13T * Y13T;
    L24 = X24T * X24T + Y24T * Y24T;
    LL = dmax(L13, L24);
    double C1 = VC[0][1] * VC[1][3] - VC[1][1] * VC[0][3];
    double C2 = VC[0][0] * VC[1][2] - VC[1][0] * VC[0][2];
    LM = dmax(fabs(C1), fabs(C2));
  }
  /* relative velocities in the element frame */
  double V13[3], V24[3], VHI[3];
  {
    double G13[3], G24[3], GHI[3];
    for (int d = 0; d < 3; ++d) {
      G13[d] = vg[d][0] - vg[d][2];
      G24[d] = vg[d][1] - vg[d][3];
      GHI[d] = vg[d][0] - vg[d][1] + vg[d][2] - vg[d][3];
    }
    const double *E[3] = {e1, e2, e3};
    for (int k = 0; k < 3; ++k) {
      V13[k] = (E[k][0] * G13[0] + E[k][1] * G13[1] + E[k][2] * G13[2]);
      V24[k] = (E[k][0] * G24[0] + E[k][1] * G24[1] + E[k][2] * G24[2]);
      VHI[k] = (E[k][0] * GHI[0] + E[k][1] * GHI[1] + E[k][2] * GHI[2]);
    }
  }
  /* mid-step rotation correction (IMPL_S = 0) */
  {
    const double DT05 = HALF * dt1, DT025 = FOURTH * dt1;
    double EXZ = Y24T * V13[2] - Y13T * V24[2];
    double EYZ = -X24T * V13[2] + X13T * V24[2];
    double DDRY = DT05 * EXZ * AREA_I;
    double DDRX = DT05 * EYZ * AREA_I;
    double V13X = V13[0], V24X = V24[0], VHIX = VHI[0];
    double DDRZ1 = DT025 * (V13[1] - V24[1]) / (X13T - X24T);
    if (fabs(X13T - X24T) < EM10) DDRZ1 = ZERO;
    V13[0] = V13[0] - DDRY * V13[2] - DDRZ1 * V13[1];
    V24[0] = V24[0] - DDRY * V24[2] - DDRZ1 * V24[1];
    VHI[0] = VHI[0] - DDRY * VHI[2] - DDRZ1 * VHI[1];
    double DDRZ2 = DT025 * (V13X + V24X) / (Y13T + Y24T);
    if (fabs(Y13T + Y24T) < EM10) DDRZ2 = ZERO;
    V13[1] = V13[1] - DDRX * V13[2] - DDRZ2 * V13X;
    V24[1] = V24[1] - DDRX * V24[2] - DDRZ2 * V24X;
    VHI[1] = VHI[1] - DDRX * VHI[2] - DDRZ2 * VHIX;
  }
  const int flat = (ZL1 * ZL1 < K->EM12 * LL);    /* NPT > 1 here */
  double VXYZ[3][4], RXYZ[2][4];
  double VCF[12] = {0};                             /* flat: VCORE(EP,1..12) */
  double JAC[4] = {0}, HX[4] = {0}, HY[4] = {0};
  double VQN[9][4] = {{0}}, VQG[9][4] = {{0}}, VJFI[6][4] = {{0}}, VNRM[3][4] = {{0}};
  double VASTN[4][4] = {{0}}, DI[6] = {0};
  double RR[3][4];
  {
    const double *E[3] = {e1, e2, e3};
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 4; ++c)
        RR[k][c] = E[k][0] * vrg[0][c] + E[k][1] * vrg[1][c] + E[k][2] * vrg[2][c];
  }
  if (flat) {
    const double X13 = X13T, X24 = X24T, Y13 = Y13T, Y24 = Y24T;
    const double MX13 = (VC[0][0] + VC[0][2]) * HALF, MY13 = (VC[1][0] + VC[1][2]) * HALF;
    const double MX23 = (VC[0][1] + VC[0][2]) * HALF, MX34 = (VC[0][2] + VC[0][3]) * HALF;
    const double MY23 = (VC[1][1] + VC[1][2]) * HALF, MY34 = (VC[1][2] + VC[1][3]) * HALF;
    X13T = X13 * AREA_I; X24T = X24 * AREA_I; Y13T = Y13 * AREA_I; Y24T = Y24 * AREA_I;
    const double GAMA1 = -MX13 * Y24 + MY13 * X24;
    const double GAMA2 = MX13 * Y13 - MY13 * X13;
    for (int d = 0; d < 3; ++d) {
      VXYZ[d][0] = V13[d];
      VXYZ[d][1] = V24[d];
      VXYZ[d][2] = VHI[d];
      VXYZ[d][3] = ZERO;
    }
    for (int k = 0; k < 2; ++k) {
      RXYZ[k][0] = RR[k][0] - RR[k][2];
      RXYZ[k][1] = RR[k][1] - RR[k][3];
      RXYZ[k][2] = RR[k][0] - RR[k][1] + RR[k][2] - RR[k][3];
      RXYZ[k][3] = RR[k][0] + RR[k][1] + RR[k][2] + RR[k][3];
    }
    VCF[0] = Y24T; VCF[1] = -Y13T; VCF[2] = -X24T; VCF[3] = X13T;
    VCF[4] = GAMA1 * AREA_I; VCF[5] = GAMA2 * AREA_I;
    VCF[6] = MX23; VCF[7] = MY23; VCF[8] = MX34; VCF[9] = MY34; VCF[10] = MX13; VCF[11] = MY13;
    double J1 = (MX23 * MY13 - MX13 * MY23) * PG;
    double J2 = -(MX13 * MY34 - MX34 * MY13) * PG;
    double J0 = AREA * FOURTH;
    JAC[0] = fabs(J0 + J2 - J1);
    JAC[1] = fabs(J0 + J2 + J1);
    JAC[2] = fabs(J0 - J2 + J1);
    JAC[3] = fabs(J0 - J2 - J1);
    J1 = (MY23 - MY34) * PG;
    J2 = -(MY23 + MY34) * PG;
    HX[0] = J1 / JAC[0]; HX[1] = J2 / JAC[1]; HX[2] = -J1 / JAC[2]; HX[3] = -J2 / JAC[3];
    J1 = (MX34 - MX23) * PG;
    J2 = (MX34 + MX23) * PG;
    HY[0] = J1 / JAC[0]; HY[1] = J2 / JAC[1]; HY[2] = -J1 / JAC[2]; HY[3] = -J2 / JAC[3];
  } else {
    const double Z1 = ZL1, Z2 = Z1 * Z1;
    VC[2][0] = Z1; VC[2][1] = -Z1; VC[2][2] = Z1; VC[2][3] = -Z1;
    const double X13 = X13T, X24 = X24T, Y13 = Y13T, Y24 = Y24T;
    const double MX13 = (VC[0][0] + VC[0][2]) * HALF, MY13 = (VC[1][0] + VC[1][2]) * HALF;
    const double MX23 = (VC[0][1] + VC[0][2]) * HALF, MY23 = (VC[1][1] + VC[1][2]) * HALF;
    const double MX34 = (VC[0][2] + VC[0][3]) * HALF, MY34 = (VC[1][2] + VC[1][3]) * HALF;
    X13T = X13 * AREA_I; X24T = X24 * AREA_I; Y13T = Y13 * AREA_I; Y24T = Y24 * AREA_I;
    const double GAMA1 = -MX13 * Y24 + MY13 * X24;
    const double GAMA2 = MX13 * Y13 - MY13 * X13;
    /* [F] edges */
    const double X21 = MX23 - MX13, X34 = (VC[0][2] - VC[0][3]) * HALF;
    const double Y21 = MY23 - MY13, Y34 = (VC[1][2] - VC[1][3]) * HALF;
    const double Z21 = -Z1, Z34 = Z1;
    const double L12 = sqrt(X21 * X21 + Y21 * Y21 + Z2);
    const double L34 = sqrt(X34 * X34 + Y34 * Y34 + Z2);
    const double X41 = MX34 - MX13, X32 = (VC[0][2] - VC[0][1]) * HALF;
    const double Y41 = MY34 - MY13, Y32 = (VC[1][2] - VC[1][1]) * HALF;
    const double Z41 = -Z1, Z32 = Z1;
    /* nodal frames [QN] */
    double A_4 = AREA * FOURTH;
    double SL = ONE / dmax(L12, EM20);
    VQN[0][0] = X21 * SL; VQN[1][0] = Y21 * SL; VQN[2][0] = Z21 * SL;
    double SZ2 = A_4 - GAMA1;
    double SZv = Z2 * L24;
    SL = ONE / sqrt(SZv + SZ2 * SZ2);
    VQN[6][0] = -Z1 * Y24; VQN[7][0] = Z1 * X24; VQN[8][0] = SZ2 * SL;
    VQN[6][2] = -VQN[6][0]; VQN[7][2] = -VQN[7][0];
    VQN[6][0] = VQN[6][0] * SL; VQN[7][0] = VQN[7][0] * SL;
#define QCROSS(Q, n)                                                   \
    Q[3][n] = Q[7][n] * Q[2][n] - Q[8][n] * Q[1][n];                   \
    Q[4][n] = Q[8][n] * Q[0][n] - Q[6][n] * Q[2][n];                   \
    Q[5][n] = Q[6][n] * Q[1][n] - Q[7][n] * Q[0][n]
    QCROSS(VQN, 0);
    SL = ONE / dmax(L34, EM20);
    VQN[0][2] = X34 * SL; VQN[1][2] = Y34 * SL; VQN[2][2] = Z34 * SL;
    SZ2 = A_4 + GAMA1;
    SL = ONE / sqrt(SZv + SZ2 * SZ2);
    VQN[6][2] = VQN[6][2] * SL; VQN[7][2] = VQN[7][2] * SL; VQN[8][2] = SZ2 * SL;
    QCROSS(VQN, 2);
    VQN[0][1] = VQN[0][0]; VQN[1][1] = VQN[1][0]; VQN[2][1] = VQN[2][0];
    SZ2 = A_4 + GAMA2;
    SZv = Z2 * L13;
    SL = ONE / sqrt(SZv + SZ2 * SZ2);
    VQN[6][1] = -Z1 * Y13; VQN[7][1] = Z1 * X13; VQN[8][1] = SZ2 * SL;
    VQN[6][3] = -VQN[6][1]; VQN[7][3] = -VQN[7][1];
    VQN[6][1] = VQN[6][1] * SL; VQN[7][1] = VQN[7][1] * SL;
    QCROSS(VQN, 1);
    VQN[0][3] = VQN[0][2]; VQN[1][3] = VQN[1][2]; VQN[2][3] = VQN[2][2];
    SZ2 = A_4 - GAMA2;
    SL = ONE / sqrt(SZv + SZ2 * SZ2);
    VQN[6][3] = VQN[6][3] * SL; VQN[7][3] = VQN[7][3] * SL; VQN[8][3] = SZ2 * SL;
    QCROSS(VQN, 3);
    /* assumed-strain normals at mid-edges and their tying terms */
    {
      const int na[4] = {0, 3, 0, 1}, nb[4] = {1, 2, 3, 2};
      for (int j = 0; j < 4; ++j) {
        for (int d = 0; d < 3; ++d) VNRM[d][j] = VQN[6 + d][na[j]] + VQN[6 + d][nb[j]];
        double C1 = sqrt(VNRM[0][j] * VNRM[0][j] + VNRM[1][j] * VNRM[1][j] +
                         VNRM[2][j] * VNRM[2][j]);
        C1 = dmax(EM20, C1);
        for (int d = 0; d < 3; ++d) VNRM[d][j] = VNRM[d][j] / C1;
      }
    }
    VASTN[0][0] = ZERO; VASTN[1][0] = L12; VASTN[2][0] = VASTN[0][0]; VASTN[3][0] = VASTN[1][0];
    VASTN[0][1] = ZERO; VASTN[1][1] = L34; VASTN[2][1] = VASTN[0][1]; VASTN[3][1] = VASTN[1][1];
    VASTN[0][2] = -(X41 * VQN[3][0] + Y41 * VQN[4][0] + Z41 * VQN[5][0]);
    VASTN[1][2] = X41 * VQN[0][0] + Y41 * VQN[1][0] + Z41 * VQN[2][0];
    VASTN[2][2] = -(X41 * VQN[3][3] + Y41 * VQN[4][3] + Z41 * VQN[5][3]);
    VASTN[3][2] = X41 * VQN[0][3] + Y41 * VQN[1][3] + Z41 * VQN[2][3];
    VASTN[0][3] = -(X32 * VQN[3][1] + Y32 * VQN[4][1] + Z32 * VQN[5][1]);
    VASTN[1][3] = X32 * VQN[0][1] + Y32 * VQN[1][1] + Z32 * VQN[2][1];
    VASTN[2][3] = -(X32 * VQN[3][2] + Y32 * VQN[4][2] + Z32 * VQN[5][2]);
    VASTN[3][3] = X32 * VQN[0][2] + Y32 * VQN[1][2] + Z32 * VQN[2][2];
    /* Gauss-point frames [QG] and the inverse tangent maps VJFI */
    A_4 = A_4 / PG;
    const double JMX13 = MX13 * PG, JMY13 = MY13 * PG, JMZ13 = Z1 * PG;
    const double J2MYZ = JMZ13 * JMZ13;
#define QNORM(n)                                                                    \
    do {                                                                            \
      double s_ = sqrt(VQG[0][n] * VQG[0][n] + VQG[1][n] * VQG[1][n] +              \
                       VQG[2][n] * VQG[2][n]);                                      \
      if (s_ != ZERO) s_ = ONE / s_;                                                \
      VQG[0][n] = VQG[0][n] * s_; VQG[1][n] = VQG[1][n] * s_;                        \
      VQG[2][n] = VQG[2][n] * s_;                                                   \
    } while (0)
    /* NG = 1 */
    SZ2 = A_4 - GAMA1;
    SZv = Z2 * L24;
    SL = sqrt(SZv + SZ2 * SZ2);
    JAC[0] = SL * PG;
    SL = ONE / dmax(SL, EM20);
    VQG[6][0] = -Z1 * Y24; VQG[7][0] = Z1 * X24; VQG[8][0] = SZ2 * SL;
    VQG[6][2] = -VQG[6][0]; VQG[7][2] = -VQG[7][0];
    VQG[6][0] = VQG[6][0] * SL; VQG[7][0] = VQG[7][0] * SL;
    const double G1X1 = MX23 - JMX13, G1Y1 = MY23 - JMY13;
    const double C1g = sqrt(G1X1 * G1X1 + G1Y1 * G1Y1 + J2MYZ);
    const double G2X1 = MX34 - JMX13, G2Y1 = MY34 - JMY13;
    const double C2g = sqrt(G2X1 * G2X1 + G2Y1 * G2Y1 + J2MYZ);
    VJFI[0][0] = (G2Y1 * VQG[8][0] + JMZ13 * VQG[7][0]);
    VJFI[1][0] = (-JMZ13 * VQG[6][0] - G2X1 * VQG[8][0]);
    VJFI[2][0] = (G2X1 * VQG[7][0] - G2Y1 * VQG[6][0]);
    VQG[0][0] = G1X1 * C2g + VJFI[0][0] * C1g;
    VQG[1][0] = G1Y1 * C2g + VJFI[1][0] * C1g;
    VQG[2][0] = -JMZ13 * C2g + VJFI[2][0] * C1g;
    SL = SL / PG;
    VJFI[0][0] = VJFI[0][0] * SL; VJFI[1][0] = VJFI[1][0] * SL; VJFI[2][0] = VJFI[2][0] * SL;
    VJFI[3][0] = (-JMZ13 * VQG[7][0] - G1Y1 * VQG[8][0]) * SL;
    VJFI[4][0] = (G1X1 * VQG[8][0] + JMZ13 * VQG[6][0]) * SL;
    VJFI[5][0] = (G1Y1 * VQG[6][0] - G1X1 * VQG[7][0]) * SL;
    QNORM(0);
    QCROSS(VQG, 0);
    /* NG = 3 */
    SZ2 = A_4 + GAMA1;
    SL = sqrt(SZv + SZ2 * SZ2);
    JAC[2] = SL * PG;
    SL = ONE / dmax(SL, EM20);
    VQG[6][2] = VQG[6][2] * SL; VQG[7][2] = VQG[7][2] * SL; VQG[8][2] = SZ2 * SL;
    const double G1X3 = MX23 + JMX13, G1Y3 = MY23 + JMY13;
    const double J1g = sqrt(G1X3 * G1X3 + G1Y3 * G1Y3 + J2MYZ);
    const double G2X2 = MX34 + JMX13, G2Y2 = MY34 + JMY13;
    const double J2g = sqrt(G2X2 * G2X2 + G2Y2 * G2Y2 + J2MYZ);
    VJFI[0][2] = (G2Y2 * VQG[8][2] - JMZ13 * VQG[7][2]);
    VJFI[1][2] = (JMZ13 * VQG[6][2] - G2X2 * VQG[8][2]);
    VJFI[2][2] = (G2X2 * VQG[7][2] - G2Y2 * VQG[6][2]);
    VQG[0][2] = G1X3 * J2g + VJFI[0][2] * J1g;
    VQG[1][2] = G1Y3 * J2g + VJFI[1][2] * J1g;
    VQG[2][2] = JMZ13 * J2g + VJFI[2][2] * J1g;
    SL = SL / PG;
    VJFI[0][2] = VJFI[0][2] * SL; VJFI[1][2] = VJFI[1][2] * SL; VJFI[2][2] = VJFI[2][2] * SL;
    VJFI[3][2] = (JMZ13 * VQG[7][2] - G1Y3 * VQG[8][2]) * SL;
    VJFI[4][2] = (G1X3 * VQG[8][2] - JMZ13 * VQG[6][2]) * SL;
    VJFI[5][2] = (G1Y3 * VQG[6][2] - G1X3 * VQG[7][2]) * SL;
    QNORM(2);
    QCROSS(VQG, 2);
    /* NG = 2 */
    SZ2 = A_4 + GAMA2;
    SZv = Z2 * L13;
    SL = sqrt(SZv + SZ2 * SZ2);
    JAC[1] = SL * PG;
    SL = ONE / dmax(SL, EM20);
    VQG[6][1] = -Z1 * Y13; VQG[7][1] = Z1 * X13; VQG[8][1] = SZ2 * SL;
    VQG[6][3] = -VQG[6][1]; VQG[7][3] = -VQG[7][1];
    VQG[6][1] = VQG[6][1] * SL; VQG[7][1] = VQG[7][1] * SL;
    VJFI[0][1] = (G2Y2 * VQG[8][1] - JMZ13 * VQG[7][1]);
    VJFI[1][1] = (JMZ13 * VQG[6][1] - G2X2 * VQG[8][1]);
    VJFI[2][1] = (G2X2 * VQG[7][1] - G2Y2 * VQG[6][1]);
    VQG[0][1] = G1X1 * J2g + VJFI[0][1] * C1g;
    VQG[1][1] = G1Y1 * J2g + VJFI[1][1] * C1g;
    VQG[2][1] = -JMZ13 * J2g + VJFI[2][1] * C1g;
    SL = SL / PG;
    VJFI[0][1] = VJFI[0][1] * SL; VJFI[1][1] = VJFI[1][1] * SL; VJFI[2][1] = VJFI[2][1] * SL;
    VJFI[3][1] = (-JMZ13 * VQG[7][1] - G1Y1 * VQG[8][1]) * SL;
    VJFI[4][1] = (G1X1 * VQG[8][1] + JMZ13 * VQG[6][1]) * SL;
    VJFI[5][1] = (G1Y1 * VQG[6][1] - G1X1 * VQG[7][1]) * SL;
    QNORM(1);
    QCROSS(VQG, 1);
    /* NG = 4 */
    SZ2 = A_4 - GAMA2;
    SL = sqrt(SZv + SZ2 * SZ2);
    JAC[3] = SL * PG;
    SL = ONE / dmax(SL, EM20);
    VQG[6][3] = VQG[6][3] * SL; VQG[7][3] = VQG[7][3] * SL; VQG[8][3] = SZ2 * SL;
    VJFI[0][3] = (G2Y1 * VQG[8][3] + JMZ13 * VQG[7][3]);
    VJFI[1][3] = (-JMZ13 * VQG[6][3] - G2X1 * VQG[8][3]);
    VJFI[2][3] = (G2X1 * VQG[7][3] - G2Y1 * VQG[6][3]);
    VQG[0][3] = G1X3 * C2g + VJFI[0][3] * J1g;
    VQG[1][3] = G1Y3 * C2g + VJFI[1][3] * J1g;
    VQG[2][3] = JMZ13 * C2g + VJFI[2][3] * J1g;
    SL = SL / PG;
    VJFI[0][3] = VJFI[0][3] * SL; VJFI[1][3] = VJFI[1][3] * SL; VJFI[2][3] = VJFI[2][3] * SL;
    VJFI[3][3] = (JMZ13 * VQG[7][3] - G1Y3 * VQG[8][3]) * SL;
    VJFI[4][3] = (G1X3 * VQG[8][3] - JMZ13 * VQG[6][3]) * SL;
    VJFI[5][3] = (G1Y3 * VQG[6][3] - G1X3 * VQG[7][3]) * SL;
    QNORM(3);
    QCROSS(VQG, 3);
#undef QNORM
    AREA = JAC[0] + JAC[1] + JAC[2] + JAC[3];
    /* rigid-body projection of the nodal velocities (6 DOF -> 5 DOF) */
    {
      const double x13 = (VC[0][0] - VC[0][2]) * HALF, x24 = (VC[0][1] - VC[0][3]) * HALF;
      const double y13 = (VC[1][0] - VC[1][2]) * HALF, y24 = (VC[1][1] - VC[1][3]) * HALF;
      const double mx13 = (VC[0][0] + VC[0][2]) * HALF, my13 = (VC[1][0] + VC[1][2]) * HALF;
      double AR[3], ALR[3], D[6];
      AR[0] = -Z1 * VHI[1] + y13 * V13[2] + y24 * V24[2] + my13 * VHI[2] + RR[0][0] + RR[0][1] +
              RR[0][2] + RR[0][3];
      AR[1] = Z1 * VHI[0] - x13 * V13[2] - x24 * V24[2] - mx13 * VHI[2] + RR[1][0] + RR[1][1] +
              RR[1][2] + RR[1][3];
      AR[2] = x13 * V13[1] + x24 * V24[1] + mx13 * VHI[1] - y13 * V13[0] - y24 * V24[0] -
              my13 * VHI[0] + RR[2][0] + RR[2][1] + RR[2][2] + RR[2][3];
      double XX1 = VC[0][0] * VC[0][0] + VC[0][1] * VC[0][1] + VC[0][2] * VC[0][2] +
                   VC[0][3] * VC[0][3];
      double YYv = VC[1][0] * VC[1][0] + VC[1][1] * VC[1][1] + VC[1][2] * VC[1][2] +
                   VC[1][3] * VC[1][3];
      double XY = VC[0][0] * VC[1][0] + VC[0][1] * VC[1][1] + VC[0][2] * VC[1][2] +
                  VC[0][3] * VC[1][3];
      double XZ1 = (VC[0][0] - VC[0][1] + VC[0][2] - VC[0][3]) * Z1;
      double YZ = (VC[1][0] - VC[1][1] + VC[1][2] - VC[1][3]) * Z1;
      double ZZ = FOUR * Z2;
      D[0] = YYv + ZZ + 4.0;
      D[1] = XX1 + ZZ + 4.0;
      D[2] = XX1 + YYv + 4.0;
      D[3] = -XY;
      D[4] = -XZ1;
      D[5] = -YZ;
      double ABC = D[0] * D[1] * D[2];
      double XXYZ2 = D[0] * D[5] * D[5];
      double YYXZ2 = D[1] * D[4] * D[4];
      double ZZXY2 = D[2] * D[3] * D[3];
      double DETA = ABC + 2.0 * D[3] * D[4] * D[5] - XXYZ2 - YYXZ2 - ZZXY2;
      if (DETA < EM20) DETA = ONE;
      else DETA = ONE / DETA;
      DI[0] = (ABC - XXYZ2) * DETA / D[0];
      DI[1] = (ABC - YYXZ2) * DETA / D[1];
      DI[2] = (ABC - ZZXY2) * DETA / D[2];
      DI[3] = (D[4] * D[5] - D[3] * D[2]) * DETA;
      DI[4] = (D[5] * D[3] - D[4] * D[1]) * DETA;
      DI[5] = (D[3] * D[4] - D[5] * D[0]) * DETA;
      ALR[0] = DI[0] * AR[0] + DI[3] * AR[1] + DI[4] * AR[2];
      ALR[1] = DI[3] * AR[0] + DI[1] * AR[1] + DI[5] * AR[2];
      ALR[2] = DI[4] * AR[0] + DI[5] * AR[1] + DI[2] * AR[2];
      V13[0] = HALF * V13[0] + ALR[2] * y13;
      V24[0] = HALF * V24[0] + ALR[2] * y24;
      VHI[0] = FOURTH * VHI[0] + (ALR[2] * my13 - Z1 * ALR[1]);
      V13[1] = HALF * V13[1] - ALR[2] * x13;
      V24[1] = HALF * V24[1] - ALR[2] * x24;
      VHI[1] = FOURTH * VHI[1] - (ALR[2] * mx13 - Z1 * ALR[0]);
      V13[2] = HALF * V13[2] - (y13 * ALR[0] - x13 * ALR[1]);
      V24[2] = HALF * V24[2] - (y24 * ALR[0] - x24 * ALR[1]);
      VHI[2] = FOURTH * VHI[2] + (mx13 * ALR[1] - my13 * ALR[0]);
      for (int d = 0; d < 3; ++d) {
        VXYZ[d][0] = V13[d] + VHI[d];
        VXYZ[d][1] = V24[d] - VHI[d];
        VXYZ[d][2] = -V13[d] + VHI[d];
        VXYZ[d][3] = -V24[d] - VHI[d];
      }
      for (int d = 0; d < 3; ++d)
        for (int c = 0; c < 4; ++c) RR[d][c] = RR[d][c] - ALR[d];
      for (int c = 0; c < 4; ++c) {
        RXYZ[0][c] = VQN[0][c] * RR[0][c] + VQN[1][c] * RR[1][c] + VQN[2][c] * RR[2][c];
        RXYZ[1][c] = VQN[3][c] * RR[0][c] + VQN[4][c] * RR[1][c] + VQN[5][c] * RR[2][c];
      }
    }
  }
#undef QCROSS
  /* characteristic length factors */
  double FACN[2];
  {
    double rx = XL2 + XL3 - XL4, ry = YL2 + YL3 - YL4;
    double sx = -XL2 + XL3 + XL4, sy = -YL2 + YL3 + YL4;
    double C1 = sqrt(rx * rx + ry * ry);
    double C2 = sqrt(sx * sx + sy * sy);
    double S1 = FOURTH * (dmax(C1, C2) / dmin(C1, C2) - ONE);
    double FAC1 = dmin(HALF, S1) + ONE;
    double FAC2 = FOUR * AREA / (C1 * C2);
    FAC2 = (double)3.413f * dmax(ZERO, FAC2 - (double)0.7071f);
    FAC2 = (double)0.78f + (double)0.22f * FAC2 * FAC2 * FAC2;
    double FACI = TWO * FAC1 * FAC2;
    FACN[0] = sqrt(L24 / LL);
    FACN[1] = sqrt(L13 / LL);
    S1 = sqrt(FACI * (K->FOUR_OVER_3 + LM * AREA_I) * LL);
    S1 = dmax(S1, EM20);
    LL = AREA / S1;
  }
  double off = dmin(ONE, fabs(st->offg));
  if (st->offg < ZERO) {
    for (int d = 0; d < 3; ++d)
      for (int c = 0; c < 4; ++c) VXYZ[d][c] = ZERO;
    for (int d = 0; d < 2; ++d)
      for (int c = 0; c < 4; ++c) RXYZ[d][c] = ZERO;
  }
  /* ---- CNCOEF3 (sh3n/coquedk/cncoef3.F:295-440), IGTYP = 1, law 2, ITHK = 0 ---- */
  const double thk0 = thke, thk02 = thk0 * thk0, VOLG = thk0 * AREA;
  const double RHO = m->rho0, NU = m->nu, G = m->shear;
  const double A11 = m->a11, SSP = m->ssp;
  const double SHF = p->fsh, GS = G * SHF;
  const double AMU = (p->dn == ZERO) ? K->EM03 : p->dn;     /* IHBE = 11: EM3 */
  /* ---- CBADEFSH (cbadef.F:722-767): constant membrane shear ---- */
  double VDEF[8];
  double BMS[8] = {0, 0, 0, 0, 0, 0, 0, 0};   /* warped: BM(3,6,12,15,21,24,30,33) */
  if (flat) {
    VDEF[2] = Y24T * VXYZ[1][0] - Y13T * VXYZ[1][1] - X24T * VXYZ[0][0] + X13T * VXYZ[0][1];
  } else {
    BMS[0] = -X24T; BMS[1] = Y24T; BMS[2] = X13T; BMS[3] = -Y13T;
    BMS[4] = -BMS[0]; BMS[5] = -BMS[1]; BMS[6] = -BMS[2]; BMS[7] = -BMS[3];
    VDEF[2] = Y24T * (VXYZ[1][0] - VXYZ[1][2]) + Y13T * (-VXYZ[1][1] + VXYZ[1][3]) -
              X24T * (VXYZ[0][0] - VXYZ[0][2]) + X13T * (VXYZ[0][1] - VXYZ[0][3]);
  }
  /* ---- CBAENERS (cbaener.F:58-85) with the previous FOR ---- */
  {
    double THOFF = off * AREA * thk0 * dt1 * HALF;
    st->eint[0] = st->eint[0] + st->forc[2] * VDEF[2] * THOFF;
  }
  const double asrate = (m->israte > 0) ? dmin(ONE, m->asrate_pm9 * dt1) : ONE;
  const double dtinv = dt1 / dmax(dt1 * dt1, EM20);
  const int npt = p->npt;
  double z0[5], wf[5], wm[5];
  oracle_shell_tables(npt, z0, wf, wm);
  double VF[3][4], VM[2][4];
  memset(VF, 0, sizeof VF);
  memset(VM, 0, sizeof VM);
  double epsd_glob = ZERO, evis = ZERO, viscmx = ZERO;
  const double VPG[2][4] = {{PG1, PG, PG, PG1}, {PG1, PG1, PG, PG}};   /* VPG(1:2,NG) */
  for (int ng = 0; ng < 4; ++ng) {
    const double CDET = JAC[ng];
    const double VOL0 = thk0 * CDET;
    double BM[36] = {0}, BMF[36] = {0}, BF[24] = {0}, BC[40] = {0}, TC[2][2] = {{0}};
    /* ---- CBADEF (cbadef.F:28-716) ---- */
    if (flat) {
      const double *VCl = VCF;
      const double VX[12] = {VXYZ[0][0], VXYZ[1][0], VXYZ[2][0], VXYZ[0][1], VXYZ[1][1],
                             VXYZ[2][1], VXYZ[0][2], VXYZ[1][2], VXYZ[2][2], 0, 0, 0};
      const double RXl[8] = {RXYZ[0][0], RXYZ[1][0], RXYZ[0][1], RXYZ[1][1],
                             RXYZ[0][2], RXYZ[1][2], RXYZ[0][3], RXYZ[1][3]};
      BM[0] = VCl[0] + HX[ng] * VCl[4];
      BM[1] = VCl[1] + HX[ng] * VCl[5];
      BM[2] = HX[ng] * FOURTH;
      BM[3] = -BM[2];
      BM[4] = VCl[2] + HY[ng] * VCl[4];
      BM[5] = VCl[3] + HY[ng] * VCl[5];
      BM[6] = HY[ng] * FOURTH;
      BM[7] = -BM[6];
      VDEF[0] = BM[0] * VX[0] + BM[1] * VX[3] + BM[2] * VX[6];
      VDEF[1] = BM[4] * VX[1] + BM[5] * VX[4] + BM[6] * VX[7];
      VDEF[5] = BM[0] * RXl[1] + BM[1] * RXl[3] + BM[2] * RXl[5];
      VDEF[6] = -(BM[4] * RXl[0] + BM[5] * RXl[2] + BM[6] * RXl[4]);
      VDEF[7] = -(BM[0] * RXl[0] + BM[1] * RXl[2] + BM[2] * RXl[4]) + BM[4] * RXl[1] +
                BM[5] * RXl[3] + BM[6] * RXl[5];
      double A_1 = 0.25 / dmax(CDET, EM20);
      double C11 = (VCl[9] + VCl[11] * VPG[0][ng]) * A_1;
      double C12 = -(VCl[7] + VCl[11] * VPG[1][ng]) * A_1;
      double C21 = -(VCl[8] + VCl[10] * VPG[0][ng]) * A_1;
      double C22 = (VCl[6] + VCl[10] * VPG[1][ng]) * A_1;
      double BETA1 = VCl[11] + VCl[7] * VPG[1][ng];
      double KSI1 = VCl[11] + VCl[9] * VPG[0][ng];
      double BETA2 = VCl[10] + VCl[6] * VPG[1][ng];
      double KSI2 = VCl[10] + VCl[8] * VPG[0][ng];
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
      BETA1 = VCl[7] + VCl[11] * VPG[1][ng];
      KSI1 = VCl[9] + VCl[11] * VPG[0][ng];
      BETA2 = VCl[6] + VCl[10] * VPG[1][ng];
      KSI2 = VCl[8] + VCl[10] * VPG[0][ng];
      BC[12] = C11 * VPG[1][ng] + C12 * VPG[0][ng];
      BC[13] = C21 * VPG[1][ng] + C22 * VPG[0][ng];
      BC[14] = -BETA1 * C11 - KSI1 * C12;
      BC[15] = -BETA1 * C21 - KSI1 * C22;
      BC[16] = C11 * BETA2 + C12 * KSI2;
      BC[17] = C21 * BETA2 + C22 * KSI2;
      BC[18] = -BC[12];
      BC[19] = -BC[13];
      BC[20] = BC[14];
      BC[21] = BC[15];
      BC[22] = BC[16];
      BC[23] = BC[17];
      VDEF[3] = BC[0] * VX[2] + BC[6] * VX[5] + BC[12] * VX[8] + BC[2] * RXl[0] +
                BC[8] * RXl[2] + BC[14] * RXl[6] + BC[4] * RXl[1] + BC[10] * RXl[3] +
                BC[16] * RXl[7];
      VDEF[4] = BC[1] * VX[2] + BC[7] * VX[5] + BC[13] * VX[8] + BC[3] * RXl[0] +
                BC[9] * RXl[2] + BC[15] * RXl[6] + BC[5] * RXl[1] + BC[11] * RXl[3] +
                BC[17] * RXl[7];
    } else {
      double TFN[3][2], TBI[2][2], TBC[2][2];
      for (int d = 0; d < 3; ++d) {
        TFN[d][0] = VKSI[0][ng] * VQN[6 + d][0] + VKSI[1][ng] * VQN[6 + d][1] +
                    VKSI[2][ng] * VQN[6 + d][2] + VKSI[3][ng] * VQN[6 + d][3];
        TFN[d][1] = VETA[0][ng] * VQN[6 + d][0] + VETA[1][ng] * VQN[6 + d][1] +
                    VETA[2][ng] * VQN[6 + d][2] + VETA[3][ng] * VQN[6 + d][3];
      }
      TBI[1][1] = VJFI[0][ng] * TFN[0][0] + VJFI[1][ng] * TFN[1][0] + VJFI[2][ng] * TFN[2][0];
      TBI[1][0] = VJFI[3][ng] * TFN[0][0] + VJFI[4][ng] * TFN[1][0] + VJFI[5][ng] * TFN[2][0];
      TBI[0][1] = VJFI[0][ng] * TFN[0][1] + VJFI[1][ng] * TFN[1][1] + VJFI[2][ng] * TFN[2][1];
      TBI[0][0] = VJFI[3][ng] * TFN[0][1] + VJFI[4][ng] * TFN[1][1] + VJFI[5][ng] * TFN[2][1];
      const double THK = -(TBI[0][0] + TBI[1][1]);
      TBI[0][1] = -TBI[0][1];
      TBI[1][0] = -TBI[1][0];
      /* TC = VJFI * VQ (VQ(:,1) = VQG(1:3), VQ(:,2) = VQG(4:6)) */
      TC[0][0] = VJFI[0][ng] * VQG[0][ng] + VJFI[1][ng] * VQG[1][ng] + VJFI[2][ng] * VQG[2][ng];
      TC[1][0] = VJFI[3][ng] * VQG[0][ng] + VJFI[4][ng] * VQG[1][ng] + VJFI[5][ng] * VQG[2][ng];
      TC[0][1] = VJFI[0][ng] * VQG[3][ng] + VJFI[1][ng] * VQG[4][ng] + VJFI[2][ng] * VQG[5][ng];
      TC[1][1] = VJFI[3][ng] * VQG[3][ng] + VJFI[4][ng] * VQG[4][ng] + VJFI[5][ng] * VQG[5][ng];
      TBC[0][0] = TBI[0][0] * TC[0][0] + TBI[0][1] * TC[1][0];
      TBC[1][0] = TBI[1][0] * TC[0][0] + TBI[1][1] * TC[1][0];
      TBC[0][1] = TBI[0][0] * TC[0][1] + TBI[0][1] * TC[1][1];
      TBC[1][1] = TBI[1][0] * TC[0][1] + TBI[1][1] * TC[1][1];
      const double q1 = VQG[0][ng], q2 = VQG[1][ng], q3 = VQG[2][ng];
      const double q4 = VQG[3][ng], q5 = VQG[4][ng], q6 = VQG[5][ng];
      for (int j = 0; j < 4; ++j) {
        const int b = 9 * j;
        double VT1 = q1 * VXYZ[0][j] + q2 * VXYZ[1][j] + q3 * VXYZ[2][j];
        double VT2 = q4 * VXYZ[0][j] + q5 * VXYZ[1][j] + q6 * VXYZ[2][j];
        double C1 = VKSI[j][ng] * TC[0][0] + VETA[j][ng] * TC[1][0];
        double C2 = VKSI[j][ng] * TC[0][1] + VETA[j][ng] * TC[1][1];
        double BC1 = VKSI[j][ng] * TBC[0][0] + VETA[j][ng] * TBC[1][0];
        double BC2 = VKSI[j][ng] * TBC[0][1] + VETA[j][ng] * TBC[1][1];
        BM[b + 2] = BMS[2 * j];
        BM[b + 5] = BMS[2 * j + 1];
        BM[b + 8] = ZERO;
        BM[b + 0] = q1 * C1;
        BMF[b + 0] = THK * BM[b + 0] + q1 * BC1;
        BM[b + 1] = q4 * C2;
        BMF[b + 1] = THK * BM[b + 1] + q4 * BC2;
        BMF[b + 2] = THK * BM[b + 2] + q1 * BC2 + q4 * BC1;
        BM[b + 3] = q2 * C1;
        BMF[b + 3] = THK * BM[b + 3] + q2 * BC1;
        BM[b + 4] = q5 * C2;
        BMF[b + 4] = THK * BM[b + 4] + q5 * BC2;
        BMF[b + 5] = THK * BM[b + 5] + q2 * BC2 + q5 * BC1;
        BM[b + 6] = q3 * C1;
        BMF[b + 6] = THK * BM[b + 6] + q3 * BC1;
        BM[b + 7] = q6 * C2;
        BMF[b + 7] = THK * BM[b + 7] + q6 * BC2;
        BMF[b + 8] = THK * BM[b + 8] + q3 * BC2 + q6 * BC1;
        if (j == 0) {
          VDEF[0] = C1 * VT1;
          VDEF[1] = C2 * VT2;
          VDEF[5] = BC1 * VT1;
          VDEF[6] = BC2 * VT2;
          VDEF[7] = BC1 * VT2 + BC2 * VT1;
        } else {
          VDEF[0] = VDEF[0] + C1 * VT1;
          VDEF[1] = VDEF[1] + C2 * VT2;
          VDEF[5] = VDEF[5] + BC1 * VT1;
          VDEF[6] = VDEF[6] + BC2 * VT2;
          VDEF[7] = VDEF[7] + BC1 * VT2 + BC2 * VT1;
        }
        double V12 = q1 * VQN[0][j] + q2 * VQN[1][j] + q3 * VQN[2][j];
        double V11 = -(q1 * VQN[3][j] + q2 * VQN[4][j] + q3 * VQN[5][j]);
        double V22 = q4 * VQN[0][j] + q5 * VQN[1][j] + q6 * VQN[2][j];
        double V21 = -(q4 * VQN[3][j] + q5 * VQN[4][j] + q6 * VQN[5][j]);
        double RV1 = V11 * RXYZ[0][j] + V12 * RXYZ[1][j];
        double RV2 = V21 * RXYZ[0][j] + V22 * RXYZ[1][j];
        const int f = 6 * j;
        BF[f + 0] = V11 * C1;
        BF[f + 1] = V21 * C2;
        BF[f + 2] = V11 * C2 + V21 * C1;
        BF[f + 3] = V12 * C1;
        BF[f + 4] = V22 * C2;
        BF[f + 5] = V12 * C2 + V22 * C1;
        if (j < 3) {
          VDEF[5] = VDEF[5] + C1 * RV1;
          VDEF[6] = VDEF[6] + C2 * RV2;
          VDEF[7] = VDEF[7] + C1 * RV2 + C2 * RV1;
        } else {
          VDEF[5] = VDEF[5] + THK * VDEF[0] + C1 * RV1;
          VDEF[6] = VDEF[6] + THK * VDEF[1] + C2 * RV2;
          VDEF[7] = VDEF[7] + THK * VDEF[2] + C1 * RV2 + C2 * RV1;
        }
      }
      /* transverse shear (assumed strain at mid-edges) */
      const double W1 = VKSI[1][ng], W2 = VKSI[2][ng], W3 = VETA[3][ng], W4 = VETA[2][ng];
      double C1;
      C1 = W1 * VNRM[0][0]; BC[0] = -C1; BC[10] = C1;
      C1 = W1 * VNRM[1][0]; BC[2] = -C1; BC[12] = C1;
      C1 = W1 * VNRM[2][0]; BC[4] = -C1; BC[14] = C1;
      BC[6] = W1 * VASTN[0][0]; BC[8] = W1 * VASTN[1][0];
      BC[16] = W1 * VASTN[2][0]; BC[18] = W1 * VASTN[3][0];
      C1 = W2 * VNRM[0][1]; BC[30] = -C1; BC[20] = C1;
      C1 = W2 * VNRM[1][1]; BC[32] = -C1; BC[22] = C1;
      C1 = W2 * VNRM[2][1]; BC[34] = -C1; BC[24] = C1;
      BC[36] = W2 * VASTN[0][1]; BC[38] = W2 * VASTN[1][1];
      BC[26] = W2 * VASTN[2][1]; BC[28] = W2 * VASTN[3][1];
      C1 = W3 * VNRM[0][2]; BC[1] = -C1; BC[31] = C1;
      C1 = W3 * VNRM[1][2]; BC[3] = -C1; BC[33] = C1;
      C1 = W3 * VNRM[2][2]; BC[5] = -C1; BC[35] = C1;
      BC[7] = W3 * VASTN[0][2]; BC[9] = W3 * VASTN[1][2];
      BC[37] = W3 * VASTN[2][2]; BC[39] = W3 * VASTN[3][2];
      C1 = W4 * VNRM[0][3]; BC[11] = -C1; BC[21] = C1;
      C1 = W4 * VNRM[1][3]; BC[13] = -C1; BC[23] = C1;
      C1 = W4 * VNRM[2][3]; BC[15] = -C1; BC[25] = C1;
      BC[17] = W4 * VASTN[0][3]; BC[19] = W4 * VASTN[1][3];
      BC[27] = W4 * VASTN[2][3]; BC[29] = W4 * VASTN[3][3];
      double BCX = BC[0] * VXYZ[0][0] + BC[2] * VXYZ[1][0] + BC[4] * VXYZ[2][0] +
                   BC[6] * RXYZ[0][0] + BC[8] * RXYZ[1][0];
      double BCY = BC[1] * VXYZ[0][0] + BC[3] * VXYZ[1][0] + BC[5] * VXYZ[2][0] +
                   BC[7] * RXYZ[0][0] + BC[9] * RXYZ[1][0];
      for (int j = 1; j < 4; ++j) {
        const int k = 10 * j;
        BCX = BCX + BC[k] * VXYZ[0][j] + BC[k + 2] * VXYZ[1][j] + BC[k + 4] * VXYZ[2][j] +
              BC[k + 6] * RXYZ[0][j] + BC[k + 8] * RXYZ[1][j];
        BCY = BCY + BC[k + 1] * VXYZ[0][j] + BC[k + 3] * VXYZ[1][j] + BC[k + 5] * VXYZ[2][j] +
              BC[k + 7] * RXYZ[0][j] + BC[k + 9] * RXYZ[1][j];
      }
      VDEF[3] = TC[0][0] * BCX + TC[1][0] * BCY;
      VDEF[4] = TC[0][1] * BCX + TC[1][1] * BCY;
    }
    /* ---- CBASTRA3 (cbastra3.F:29-100), ISTRAIN = 0 ---- */
    const double EXX = VDEF[0] * dt1, EYY = VDEF[1] * dt1, EXY = VDEF[2] * dt1;
    const double EYZ = VDEF[4] * dt1, EXZ = VDEF[3] * dt1;
    const double KXX = VDEF[5] * dt1, KYY = VDEF[6] * dt1, KXY = VDEF[7] * dt1;
    double FORL[5], MOML[3];
    for (int c = 0; c < 5; ++c) FORL[c] = st->forpg[ng][c];
    for (int c = 0; c < 3; ++c) MOML[c] = st->mompg[ng][c];
    /* ---- CBAENER (cbaener.F:24-50) with the Gauss point's previous resultants ---- */
    {
      double THOFF = HALF * off * VOL0;
      st->eint[0] = st->eint[0] - FORL[2] * EXY * THOFF;
    }
    /* ---- the Gauss point's equivalent strain rate (cbaforc3.F:759-767) ---- */
    double epsd_pg;
    {
      double eps_k2 = (KXX * KXX + KYY * KYY + KXX * KYY + FOURTH * (KXY * KXY)) *
                      K->ONE_OVER_9 * (st->thk * st->thk);
      double eps_m2 = K->FOUR_OVER_3 * (EXX * EXX + EYY * EYY + EXX * EYY + FOURTH * (EXY * EXY));
      epsd_pg = sqrt(eps_k2 + eps_m2) * dtinv;
      epsd_glob = epsd_glob + epsd_pg / 4.0;
    }
    /* ---- CMAIN3 -> MULAWC law 2 for NPG > 1 (mulawc.F90:530-560,2968-3020) ---- */
    double degmb = FORL[0] * EXX + FORL[1] * EYY + FORL[2] * EXY + FORL[3] * EYZ + FORL[4] * EXZ;
    double degfx = MOML[0] * KXX + MOML[1] * KYY + MOML[2] * KXY;
    degmb = degmb * off;
    degfx = degfx * off;
    double thkn = st->thk;
    for (int c = 0; c < 5; ++c) FORL[c] = ZERO;
    for (int c = 0; c < 3; ++c) MOML[c] = ZERO;
    double sigy = ZERO;
    const double off_old = off;
    double epchk = ZERO;
    viscmx = ZERO;
    for (int it = 0; it < npt; ++it) {
      double thklyl = wf[it] * thk0;
      double wmc = wm[it];
      double zt = z0[it] * thk0;
      double deps[5];
      deps[0] = EXX + zt * KXX;
      deps[1] = EYY + zt * KYY;
      deps[2] = EXY + zt * KXY;
      deps[3] = EYZ;
      deps[4] = EXZ;
      double sign[5], etse = ONE;
      oracle_sigeps02c_point(m, GS, dt1, asrate, epsd_pg, off_old, &off, &st->pla[ng][it],
                             &st->epsdp[ng][it], &epchk, &sigy, &thkn, thklyl, npt, deps,
                             st->sig[ng][it], sign, &etse);
      viscmx = dmax(p->dm, viscmx);
      for (int c = 0; c < 5; ++c) st->sig[ng][it][c] = sign[c] * ONE;
      for (int c = 0; c < 5; ++c) FORL[c] = FORL[c] + wf[it] * sign[c];
      for (int c = 0; c < 3; ++c) MOML[c] = MOML[c] + wmc * sign[c];
    }
    st->thk = dmax(thkn, K->EM30);
    {
      double fact = K->ONEP414 * p->dm;
      double visc = fact * SSP * sqrt(CDET) * dtinv * RHO;
      FORL[0] = FORL[0] + visc * (EXX + HALF * EYY);
      FORL[1] = FORL[1] + visc * (EYY + HALF * EXX);
      FORL[2] = FORL[2] + visc * EXY * K->THIRD;
    }
    for (int c = 0; c < 5; ++c) FORL[c] = FORL[c] * off * ONE;
    for (int c = 0; c < 3; ++c) MOML[c] = MOML[c] * off * ONE;
    degmb = degmb + FORL[0] * EXX + FORL[1] * EYY + FORL[2] * EXY + FORL[3] * EYZ + FORL[4] * EXZ;
    degfx = degfx + MOML[0] * KXX + MOML[1] * KYY + MOML[2] * KXY;
    {
      double vol2 = HALF * VOL0;
      st->eint[0] = st->eint[0] + degmb * vol2;
      st->eint[1] = st->eint[1] + degfx * thk0 * vol2;
    }
    /* ---- CBAENER with the new resultants ---- */
    {
      double THOFF = HALF * off * VOL0;
      st->eint[0] = st->eint[0] - FORL[2] * EXY * THOFF;
    }
    /* ---- CBAVISC (cbavisc.F:24-95), AREA = CDET, THK = THK0 ---- */
    {
      double visc = K->ONEP414 * off * AMU * RHO * SSP * sqrt(CDET);
      double GG = HALF / (ONE + NU);
      double FX = visc * (VDEF[0] + NU * VDEF[1]);
      double FY = visc * (VDEF[1] + NU * VDEF[0]);
      double FXY = visc * VDEF[2] * GG;
      FORL[0] = FORL[0] + FX;
      FORL[1] = FORL[1] + FY;
      FORL[2] = FORL[2] + FXY;
      double DV = CDET * thk0 * dt1;
      evis = evis + (FX * VDEF[0] + FY * VDEF[1]) * DV;
      visc = K->ZEP3 * thk0 * visc;
      FX = visc * (VDEF[5] + NU * VDEF[6]);
      FY = visc * (VDEF[6] + NU * VDEF[5]);
      FXY = visc * VDEF[7] * GG;
      MOML[0] = MOML[0] + FX;
      MOML[1] = MOML[1] + FY;
      MOML[2] = MOML[2] + FXY;
      DV = DV * thk0;
      evis = evis + (FX * VDEF[5] + FY * VDEF[6] + FXY * VDEF[7]) * DV;
    }
    /* ---- CBAFORI (cbafori.F:24-300), VOL = VOL0, TH12 = THK0^2 ---- */
    {
      const double C2 = VOL0, C1 = thk02 * CDET;
      const double FF1 = FORL[0], FF2 = FORL[1], FF4 = FORL[4], FF5 = FORL[3];
      const double MM1 = MOML[0], MM2 = MOML[1], MM3 = MOML[2];
      if (flat) {
        double CM1 = C1 * (BM[6] * MM2 + BM[2] * MM3);
        double CM2 = C1 * (BM[2] * MM1 + BM[6] * MM3);
        double CC1 = C2 * (BC[14] * FF4 + BC[15] * FF5);
        double CC2 = C2 * (BC[16] * FF4 + BC[17] * FF5);
        double BC1 = BC[0] * FF4, BC3 = BC[2] * FF4, BC5 = BC[4] * FF4;
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
        double BCX = TC[0][0] * FF4 + TC[0][1] * FF5;
        double BCY = TC[1][0] * FF4 + TC[1][1] * FF5;
        for (int j = 0; j < 4; ++j) {
          const int b = 9 * j, c = 10 * j, f = 6 * j;
          VF[0][j] = VF[0][j] + C2 * (BM[b] * FF1 + BM[b + 1] * FF2 + BC[c] * BCX + BC[c + 1] * BCY) +
                     C1 * (BMF[b] * MM1 + BMF[b + 1] * MM2 + BMF[b + 2] * MM3);
          VM[0][j] = VM[0][j] + C2 * (BC[c + 6] * BCX + BC[c + 7] * BCY) +
                     C1 * (BF[f] * MM1 + BF[f + 1] * MM2 + BF[f + 2] * MM3);
          VF[1][j] = VF[1][j] +
                     C2 * (BM[b + 3] * FF1 + BM[b + 4] * FF2 + BC[c + 2] * BCX + BC[c + 3] * BCY) +
                     C1 * (BMF[b + 3] * MM1 + BMF[b + 4] * MM2 + BMF[b + 5] * MM3);
          VM[1][j] = VM[1][j] + C2 * (BC[c + 8] * BCX + BC[c + 9] * BCY) +
                     C1 * (BF[f + 3] * MM1 + BF[f + 4] * MM2 + BF[f + 5] * MM3);
          VF[2][j] = VF[2][j] +
                     C2 * (BM[b + 6] * FF1 + BM[b + 7] * FF2 + BC[c + 4] * BCX + BC[c + 5] * BCY) +
                     C1 * (BMF[b + 6] * MM1 + BMF[b + 7] * MM2 + BMF[b + 8] * MM3);
        }
      }
    }
    for (int c = 0; c < 5; ++c) st->forpg[ng][c] = FORL[c];
    for (int c = 0; c < 3; ++c) st->mompg[ng][c] = MOML[c];
  }
  st->epsd = ONE * epsd_glob + (ONE - ONE) * st->epsd;
  for (int c = 0; c < 5; ++c)
    st->forc[c] = FOURTH * (st->forpg[0][c] + st->forpg[1][c] + st->forpg[2][c] + st->forpg[3][c]);
  for (int c = 0; c < 3; ++c)
    st->mom[c] = FOURTH * (st->mompg[0][c] + st->mompg[1][c] + st->mompg[2][c] + st->mompg[3][c]);
  /* ---- CBAFORCT (cbafori.F:305-350): constant membrane shear forces ---- */
  {
    double THOFF = VOLG * st->forc[2] * off;
    double SX1 = -THOFF * X24T, SY1 = THOFF * Y24T;
    double SX2 = THOFF * X13T, SY2 = -THOFF * Y13T;
    VF[0][0] = VF[0][0] + SX1;
    VF[1][0] = VF[1][0] + SY1;
    VF[0][1] = VF[0][1] + SX2;
    VF[1][1] = VF[1][1] + SY2;
    if (!flat) {
      VF[0][2] = VF[0][2] - SX1;
      VF[1][2] = VF[1][2] - SY1;
      VF[0][3] = VF[0][3] - SX2;
      VF[1][3] = VF[1][3] - SY2;
    }
  }
  /* ---- CBAENERS with the new FOR ---- */
  {
    double THOFF = off * AREA * thk0 * dt1 * HALF;
    st->eint[0] = st->eint[0] + st->forc[2] * VDEF[2] * THOFF;
  }
  /* ---- CBAPROJ (cbaproj.F:24-330), ISROT = 0 ---- */
  double F[3][4], M[3][4];
  if (flat) {
    double FL[3][4], ML[2][4];
    for (int d = 0; d < 3; ++d) {
      FL[d][0] = VF[d][0] + VF[d][2];
      FL[d][1] = VF[d][1] + VF[d][3];
      FL[d][2] = -VF[d][0] + VF[d][2];
      FL[d][3] = -VF[d][1] + VF[d][3];
    }
    for (int d = 0; d < 2; ++d) {
      ML[d][0] = VM[d][0] + VM[d][2];
      ML[d][1] = VM[d][1] + VM[d][3];
      ML[d][2] = -VM[d][0] + VM[d][2];
      ML[d][3] = -VM[d][1] + VM[d][3];
    }
    for (int c = 0; c < 4; ++c)
      for (int d = 0; d < 3; ++d) {
        F[d][c] = e1[d] * FL[0][c] + e2[d] * FL[1][c] + e3[d] * FL[2][c];
        M[d][c] = e1[d] * ML[0][c] + e2[d] * ML[1][c];
      }
  } else {
    double MM[3][4], AR[3], ALR[3];
    const double Z1 = VC[2][0];
    for (int j = 0; j < 4; ++j)
      for (int d = 0; d < 3; ++d)
        MM[d][j] = VQN[d][j] * VM[0][j] + VQN[d + 3][j] * VM[1][j] + ZERO;
    AR[0] = -Z1 * (VF[1][0] - VF[1][1] + VF[1][2] - VF[1][3]) + VC[1][0] * VF[2][0] + MM[0][0] +
            VC[1][1] * VF[2][1] + MM[0][1] + VC[1][2] * VF[2][2] + MM[0][2] + VC[1][3] * VF[2][3] +
            MM[0][3];
    AR[1] = Z1 * (VF[0][0] - VF[0][1] + VF[0][2] - VF[0][3]) - VC[0][0] * VF[2][0] + MM[1][0] -
            VC[0][1] * VF[2][1] + MM[1][1] - VC[0][2] * VF[2][2] + MM[1][2] - VC[0][3] * VF[2][3] +
            MM[1][3];
    AR[2] = -VC[1][0] * VF[0][0] + VC[0][0] * VF[1][0] + MM[2][0] - VC[1][1] * VF[0][1] +
            VC[0][1] * VF[1][1] + MM[2][1] - VC[1][2] * VF[0][2] + VC[0][2] * VF[1][2] + MM[2][2] -
            VC[1][3] * VF[0][3] + VC[0][3] * VF[1][3] + MM[2][3];
    ALR[0] = DI[0] * AR[0] + DI[3] * AR[1] + DI[4] * AR[2];
    ALR[1] = DI[3] * AR[0] + DI[1] * AR[1] + DI[5] * AR[2];
    ALR[2] = DI[4] * AR[0] + DI[5] * AR[1] + DI[2] * AR[2];
    double C1 = Z1 * ALR[1];
    VF[0][0] = VF[0][0] - C1 + VC[1][0] * ALR[2];
    VF[0][1] = VF[0][1] + C1 + VC[1][1] * ALR[2];
    VF[0][2] = VF[0][2] - C1 + VC[1][2] * ALR[2];
    VF[0][3] = VF[0][3] + C1 + VC[1][3] * ALR[2];
    C1 = Z1 * ALR[0];
    VF[1][0] = VF[1][0] + C1 - VC[0][0] * ALR[2];
    VF[1][1] = VF[1][1] - C1 - VC[0][1] * ALR[2];
    VF[1][2] = VF[1][2] + C1 - VC[0][2] * ALR[2];
    VF[1][3] = VF[1][3] - C1 - VC[0][3] * ALR[2];
    for (int j = 0; j < 4; ++j) {
      VF[2][j] = VF[2][j] - VC[1][j] * ALR[0] + VC[0][j] * ALR[1];
      MM[0][j] = MM[0][j] - ALR[0];
      MM[1][j] = MM[1][j] - ALR[1];
      MM[2][j] = MM[2][j] - ALR[2];
    }
    for (int c = 0; c < 4; ++c)
      for (int d = 0; d < 3; ++d) {
        F[d][c] = e1[d] * VF[0][c] + e2[d] * VF[1][c] + e3[d] * VF[2][c];
        M[d][c] = e1[d] * MM[0][c] + e2[d] * MM[1][c] + e3[d] * MM[2][c];
      }
  }
  for (int d = 0; d < 3; ++d)
    for (int c = 0; c < 4; ++c) {
      F[d][c] = F[d][c] * off;
      M[d][c] = M[d][c] * off;
    }
  /* ---- CNDT3 (sh3n/coquedk/cndt3.F:84-300), NODADT = 0, IDTMINS = 0 ---- */
  double ALDT = LL;
  viscmx = dmax(viscmx, AMU);
  viscmx = sqrt(ONE + viscmx * viscmx) - viscmx;
  ALDT = ALDT * viscmx / sqrt(ONE);                 /* ALPE = 1 */
  double F_OSET = ONE + K->ZEP1 * TWO * fabs(ZERO) / thk0;   /* ZOFFSET = 0 */
  double F_DTE = ONE / sqrt(F_OSET);
  double DT = p->dtfac * F_DTE * ALDT / SSP;
  if (DT <= p->dtmin && off >= ONE && st->offg != TWO && st->offg >= ZERO) off = ZERO;
  out->dt = DT;
  out->alive = (st->offg > ZERO && off != ZERO);
  double sti, stir;
  {
    double DIVM = dmax(ALDT * ALDT, EM20);
    sti = HALF * F_OSET * VOLG * A11 * off / DIVM;
    stir = ZERO;
  }
  /* ---- CUPDTN3P (coque/cupdtn3.F:488-700) ---- */
  if (off < ONE) st->offg = off;
  if (st->offg < ZERO) {
    for (int d = 0; d < 3; ++d)
      for (int k = 0; k < 4; ++k) { F[d][k] = ZERO; M[d][k] = ZERO; }
    sti = ZERO; stir = ZERO;
  }
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < 4; ++k) { out->f[d][k] = -F[d][k]; out->m[d][k] = -M[d][k]; }
  out->sti = sti;
  out->stir = stir;
  out->stic[0] = sti * FACN[0];
  out->stic[1] = sti * FACN[1];
  out->stic[2] = sti * FACN[0];
  out->stic[3] = sti * FACN[1];
  out->ehou = evis;
}

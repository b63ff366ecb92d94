/*
 * orad_oracle.c -- CPU restatement (FP64) of the OpenRadioss explicit brick path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the CPU
 * baseline timed by bench.py.  Never linked into openradioss_amd/.
 *
 * Operation order follows the reference's double-precision build so that the
 * result can be compared bit-for-bit with the reference's own leaf kernels
 * compiled from source (oracle/_ref).  Paths are relative to the reference root.
 *
 * Pinning (tests/test_oracle_*.py):
 *   - every element-chain stage against oracle/_ref (reference SRCOOR3, SDERI3,
 *     SDLEN3, SDEFO3, SRHO3, M2LAW, MQVISCB, SHVIS3, SFINT3, SRROTA3 driven by
 *     oracle/ref/harness_brick.F90);
 *   - closed forms: unit-cube volume and shape derivatives, rigid motion gives
 *     zero strain rate and zero force, uniaxial radial return, rho*V = rho0*V0;
 *   - CSR ordering against a brute-force sort of (node, user id, corner).
 */
#include "orad_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Constants with the exact bit patterns of common_source/modules/constant_mod.F
 * in the double-precision build (checked against the reference by a test). */
#define ZERO        0.0
#define ONE         1.0
#define TWO         2.0
#define THREE       3.0
#define FOUR        4.0
#define HALF        0.5
#define HUNDRED     100.0
#define FOURTH      0x1.0p-2
#define ONE_OVER_64 0x1.0p-6
#define THIRD       0x1.5555555555555p-2  /* ONE/THREE */
#define TWO_THIRD   0x1.5555555555555p-1  /* TWO/THREE */
#define ONEP333     0x1.553f7ced91687p+0  /* ONEP33 + THREEEM3 */
#define EM15        0x1.203af9ee75616p-50
#define EM20        0x1.79ca10c924223p-67
#define ONEP1       0x1.199999999999ap+0

static inline double dmin(double a, double b) { return (a < b) ? a : b; }
static inline double dmax(double a, double b) { return (a > b) ? a : b; }
/* Fortran MIN/MAX with the first argument kept on ties/NaN is irrelevant here. */

/* SREPISO3 (engine/source/elements/solid/solide/srepiso3.F:60-92) */
static void srepiso3(const double x[8], const double y[8], const double z[8],
                     double r[3], double s[3], double t[3]) {
  double X17 = x[6] - x[0], X28 = x[7] - x[1], X35 = x[4] - x[2], X46 = x[5] - x[3];
  double Y17 = y[6] - y[0], Y28 = y[7] - y[1], Y35 = y[4] - y[2], Y46 = y[5] - y[3];
  double Z17 = z[6] - z[0], Z28 = z[7] - z[1], Z35 = z[4] - z[2], Z46 = z[5] - z[3];
  double A17 = X17 + X46, A28 = X28 + X35;
  double B17 = Y17 + Y46, B28 = Y28 + Y35;
  double C17 = Z17 + Z46, C28 = Z28 + Z35;
  r[0] = X17 + X28 - X35 - X46;
  r[1] = Y17 + Y28 - Y35 - Y46;
  r[2] = Z17 + Z28 - Z35 - Z46;
  s[0] = A17 + A28; s[1] = B17 + B28; s[2] = C17 + C28;
  t[0] = A17 - A28; t[1] = B17 - B28; t[2] = C17 - C28;
}

/* SORTHO3 (engine/source/elements/solid/solide/sortho3.F:70-160), NITER=3.
 * Output e[i][j] = component j of unit vector e_(i+1); R(:,i) = e_(i+1). */
static void sortho3(const double r[3], const double s[3], const double t[3], double e[3][3]) {
  double aa, bb, U[3], V[3], W[3], e1[3], e2[3], e3[3];
  aa = sqrt(r[0]*r[0] + r[1]*r[1] + r[2]*r[2]); if (aa != ZERO) aa = ONE / aa;
  U[0] = r[0]*aa; U[1] = r[1]*aa; U[2] = r[2]*aa;
  aa = sqrt(s[0]*s[0] + s[1]*s[1] + s[2]*s[2]); if (aa != ZERO) aa = ONE / aa;
  V[0] = s[0]*aa; V[1] = s[1]*aa; V[2] = s[2]*aa;
  aa = sqrt(t[0]*t[0] + t[1]*t[1] + t[2]*t[2]); if (aa != ZERO) aa = ONE / aa;
  W[0] = t[0]*aa; W[1] = t[1]*aa; W[2] = t[2]*aa;
  for (int n = 0; n < 3; ++n) {
    e1[0] = V[1]*W[2] - V[2]*W[1] + U[0];
    e1[1] = V[2]*W[0] - V[0]*W[2] + U[1];
    e1[2] = V[0]*W[1] - V[1]*W[0] + U[2];
    e2[0] = W[1]*U[2] - W[2]*U[1] + V[0];
    e2[1] = W[2]*U[0] - W[0]*U[2] + V[1];
    e2[2] = W[0]*U[1] - W[1]*U[0] + V[2];
    e3[0] = U[1]*V[2] - U[2]*V[1] + W[0];
    e3[1] = U[2]*V[0] - U[0]*V[2] + W[1];
    e3[2] = U[0]*V[1] - U[1]*V[0] + W[2];
    bb = sqrt(e1[0]*e1[0] + e1[1]*e1[1] + e1[2]*e1[2]); if (bb != ZERO) bb = ONE / bb;
    U[0] = e1[0]*bb; U[1] = e1[1]*bb; U[2] = e1[2]*bb;
    bb = sqrt(e2[0]*e2[0] + e2[1]*e2[1] + e2[2]*e2[2]); if (bb != ZERO) bb = ONE / bb;
    V[0] = e2[0]*bb; V[1] = e2[1]*bb; V[2] = e2[2]*bb;
    bb = sqrt(e3[0]*e3[0] + e3[1]*e3[1] + e3[2]*e3[2]); if (bb != ZERO) bb = ONE / bb;
    W[0] = e3[0]*bb; W[1] = e3[1]*bb; W[2] = e3[2]*bb;
  }
  e1[0] = U[0]; e1[1] = U[1]; e1[2] = U[2];
  e3[0] = e1[1]*V[2] - e1[2]*V[1];
  e3[1] = e1[2]*V[0] - e1[0]*V[2];
  e3[2] = e1[0]*V[1] - e1[1]*V[0];
  aa = sqrt(e3[0]*e3[0] + e3[1]*e3[1] + e3[2]*e3[2]); if (aa != ZERO) aa = ONE / aa;
  e3[0] = e3[0]*aa; e3[1] = e3[1]*aa; e3[2] = e3[2]*aa;
  e2[0] = e3[1]*e1[2] - e3[2]*e1[1];
  e2[1] = e3[2]*e1[0] - e3[0]*e1[2];
  e2[2] = e3[0]*e1[1] - e3[1]*e1[0];
  for (int j = 0; j < 3; ++j) { e[0][j] = e1[j]; e[1][j] = e2[j]; e[2][j] = e3[j]; }
}

/* SLEN (engine/source/elements/solid/solide/slen.F:67-87): one face's E*G-F^2. */
static double slen(double x1, double x2, double x3, double x4,
                   double y1, double y2, double y3, double y4,
                   double z1, double z2, double z3, double z4) {
  double X13 = x3 - x1, X24 = x4 - x2, Y13 = y3 - y1, Y24 = y4 - y2;
  double Z13 = z3 - z1, Z24 = z4 - z2;
  double FS1 = X13 - X24, FT1 = X13 + X24, FS2 = Y13 - Y24, FT2 = Y13 + Y24;
  double FS3 = Z13 - Z24, FT3 = Z13 + Z24;
  double E = FS1*FS1 + FS2*FS2 + FS3*FS3;
  double F = FS1*FT1 + FS2*FT2 + FS3*FT3;
  double G = FT1*FT1 + FT2*FT2 + FT3*FT3;
  return E*G - F*F;
}

double oracle_brick_volume(const double xg[3][8]) {
  const double *x = xg[0], *y = xg[1], *z = xg[2];
  double X17 = x[6]-x[0], X28 = x[7]-x[1], X35 = x[4]-x[2], X46 = x[5]-x[3];
  double Y17 = y[6]-y[0], Y28 = y[7]-y[1], Y35 = y[4]-y[2], Y46 = y[5]-y[3];
  double Z17 = z[6]-z[0], Z28 = z[7]-z[1], Z35 = z[4]-z[2], Z46 = z[5]-z[3];
  double J1 = X17+X28-X35-X46, J2 = Y17+Y28-Y35-Y46, J3 = Z17+Z28-Z35-Z46;
  double x1746 = X17+X46, x2835 = X28+X35, y1746 = Y17+Y46, y2835 = Y28+Y35;
  double z1746 = Z17+Z46, z2835 = Z28+Z35;
  double J4 = x1746+x2835, J5 = y1746+y2835, J6 = z1746+z2835;
  double J7 = x1746-x2835, J8 = y1746-y2835, J9 = z1746-z2835;
  double J5968 = J5*J9 - J6*J8, J6749 = J6*J7 - J4*J9, J4857 = J4*J8 - J5*J7;
  return ONE_OVER_64*(J1*J5968 + J2*J6749 + J3*J4857);
}

void oracle_brick_element(const double xg[3][8], const double vg[3][8],
                          orad_bstate_t *st, double dt1,
                          const orad_law2_t *mat, const orad_sprop_t *prop,
                          orad_bout_t *out) {
  double xd[8], yd[8], zd[8], vx[8], vy[8], vz[8];
  double R[3][3]; /* R[i][j] = R(i+1,j+1) of the reference */
  double e[3][3];
  double r[3], s[3], t[3];
  double off, offg = st->offg;
  int mstop = 0;

  /* ---- SRCOOR3 (srcoor3.F:128-176 gather, 297-317 frame, 325-447 rotate, 642-684 vel) ---- */
  for (int k = 0; k < 8; ++k) {
    xd[k] = xg[0][k]; yd[k] = xg[1][k]; zd[k] = xg[2][k];
    vx[k] = vg[0][k]; vy[k] = vg[1][k]; vz[k] = vg[2][k];
  }
  srepiso3(xd, yd, zd, r, s, t);
  sortho3(r, s, t, e);
  /* SORTHO3(RX..TZ, R11,R12,R13,R21,R22,R23,R31,R32,R33) binds E1X=R11, E2X=R12, E3X=R13 ... */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = e[j][i];
  /* Ismstr=4, JLAG>0, |OFFG|<=1: X' = R^T X in double precision. */
  for (int k = 0; k < 8; ++k) {
    double XDL = R[0][0]*xd[k] + R[1][0]*yd[k] + R[2][0]*zd[k];
    double YDL = R[0][1]*xd[k] + R[1][1]*yd[k] + R[2][1]*zd[k];
    double ZDL = R[0][2]*xd[k] + R[1][2]*yd[k] + R[2][2]*zd[k];
    xd[k] = XDL; yd[k] = YDL; zd[k] = ZDL;
  }
  off = fabs(offg);
  /* SRROTA3 (srrota3.F:60-108) velocities to the convected frame */
  for (int k = 0; k < 8; ++k) {
    double X = R[0][0]*vx[k] + R[1][0]*vy[k] + R[2][0]*vz[k];
    double Y = R[0][1]*vx[k] + R[1][1]*vy[k] + R[2][1]*vz[k];
    double Z = R[0][2]*vx[k] + R[1][2]*vy[k] + R[2][2]*vz[k];
    vx[k] = X; vy[k] = Y; vz[k] = Z;
  }
  if (offg < ZERO)
    for (int k = 0; k < 8; ++k) { vx[k] = ZERO; vy[k] = ZERO; vz[k] = ZERO; }

  /* ---- SDERI3 (sderi3.F:104-149 Jacobian, 152-155 volume, 265-300 inverse, 302-335 PxHy) ---- */
  double X17 = xd[6]-xd[0], X28 = xd[7]-xd[1], X35 = xd[4]-xd[2], X46 = xd[5]-xd[3];
  double Y17 = yd[6]-yd[0], Y28 = yd[7]-yd[1], Y35 = yd[4]-yd[2], Y46 = yd[5]-yd[3];
  double Z17 = zd[6]-zd[0], Z28 = zd[7]-zd[1], Z35 = zd[4]-zd[2], Z46 = zd[5]-zd[3];
  double JAC1 = X17+X28-X35-X46, JAC2 = Y17+Y28-Y35-Y46, JAC3 = Z17+Z28-Z35-Z46;
  double X_17_46 = X17+X46, X_28_35 = X28+X35;
  double Y_17_46 = Y17+Y46, Y_28_35 = Y28+Y35;
  double Z_17_46 = Z17+Z46, Z_28_35 = Z28+Z35;
  double JAC4 = X_17_46+X_28_35, JAC5 = Y_17_46+Y_28_35, JAC6 = Z_17_46+Z_28_35;
  double JAC7 = X_17_46-X_28_35, JAC8 = Y_17_46-Y_28_35, JAC9 = Z_17_46-Z_28_35;
  double JAC_59_68 = JAC5*JAC9 - JAC6*JAC8;
  double JAC_67_49 = JAC6*JAC7 - JAC4*JAC9;
  double JAC_48_57 = JAC4*JAC8 - JAC5*JAC7;
  double det = ONE_OVER_64*(JAC1*JAC_59_68 + JAC2*JAC_67_49 + JAC3*JAC_48_57);
  /* SCHKJABT3 (solide4/schkjabt3.F), INEG_V=0: non-positive volume stops the run. */
  if (off == ZERO) {
    det = ONE;
  } else if (offg > ONE) {
    /* small-strain switch state: not reachable with Ismstr=4 defaults */
  } else if (det <= ZERO) {
    det = ONE; off = ZERO; offg = ZERO; mstop = 1;
  }
  double voln = det;
  double DETT = ONE_OVER_64/det;
  double JACI1 = DETT*JAC_59_68, JACI4 = DETT*JAC_67_49, JACI7 = DETT*JAC_48_57;
  double JACI2 = DETT*(JAC3*JAC8 - JAC2*JAC9);
  double JACI5 = DETT*(JAC1*JAC9 - JAC3*JAC7);
  double JACI8 = DETT*(JAC2*JAC7 - JAC1*JAC8);
  double JACI3 = DETT*(JAC2*JAC6 - JAC3*JAC5);
  double JACI6 = DETT*(JAC3*JAC4 - JAC1*JAC6);
  double JACI9 = DETT*(JAC1*JAC5 - JAC2*JAC4);
  double JAC12 = JACI1+JACI2, JAC45 = JACI4+JACI5, JAC78 = JACI7+JACI8;
  double PX[4], PY[4], PZ[4];
  PX[0] = -JAC12-JACI3; PY[0] = -JAC45-JACI6; PZ[0] = -JAC78-JACI9;
  PX[1] = -JAC12+JACI3; PY[1] = -JAC45+JACI6; PZ[1] = -JAC78+JACI9;
  JAC12 = JACI1-JACI2; JAC45 = JACI4-JACI5; JAC78 = JACI7-JACI8;
  PX[2] = JAC12+JACI3; PY[2] = JAC45+JACI6; PZ[2] = JAC78+JACI9;
  PX[3] = JAC12-JACI3; PY[3] = JAC45-JACI6; PZ[3] = JAC78-JACI9;
  double PH[3][4]; /* PH[m][a] = PX(a+1)H(m+1) */
  {
    static const double hm[3][8] = {
      { 1,-1, 1,-1, 1,-1, 1,-1},
      { 1, 1,-1,-1,-1,-1, 1, 1},
      { 1,-1,-1, 1,-1, 1, 1,-1}};
    for (int m = 0; m < 3; ++m) {
      /* HX=(X1-X2+X3-X4+X5-X6+X7-X8) with the mode's signs, left to right */
      double HX = xd[0], HY = yd[0], HZ = zd[0];
      for (int k = 1; k < 8; ++k) {
        if (hm[m][k] > 0) { HX = HX + xd[k]; HY = HY + yd[k]; HZ = HZ + zd[k]; }
        else              { HX = HX - xd[k]; HY = HY - yd[k]; HZ = HZ - zd[k]; }
      }
      for (int a = 0; a < 4; ++a) PH[m][a] = PX[a]*HX + PY[a]*HY + PZ[a]*HZ;
    }
  }

  /* ---- SDLEN3 / SLEN (sdlen3.F:109-167, slen.F:67-87) ---- */
  double aream = EM20, ar;
#define FACE(a,b,c,d) slen(xd[a],xd[b],xd[c],xd[d], yd[a],yd[b],yd[c],yd[d], zd[a],zd[b],zd[c],zd[d])
  ar = FACE(0,1,2,3); aream = dmax(ar, aream);
  ar = FACE(4,5,6,7); aream = dmax(ar, aream);
  ar = FACE(0,1,5,4); aream = dmax(ar, aream);
  ar = FACE(1,2,6,5); aream = dmax(ar, aream);
  ar = FACE(2,3,7,6); aream = dmax(ar, aream);
  ar = FACE(3,0,4,7); aream = dmax(ar, aream);
#undef FACE
  double deltax = FOUR*voln*ONE/sqrt(aream);

  /* ---- SDEFO3 (sdefo3.F:114-138 gradients, JCVT/=0 branch: W=0 + mid-step correction) ---- */
  double VX17 = vx[0]-vx[6], VX28 = vx[1]-vx[7], VX35 = vx[2]-vx[4], VX46 = vx[3]-vx[5];
  double VY17 = vy[0]-vy[6], VY28 = vy[1]-vy[7], VY35 = vy[2]-vy[4], VY46 = vy[3]-vy[5];
  double VZ17 = vz[0]-vz[6], VZ28 = vz[1]-vz[7], VZ35 = vz[2]-vz[4], VZ46 = vz[3]-vz[5];
  double DXX = PX[0]*VX17+PX[1]*VX28+PX[2]*VX35+PX[3]*VX46;
  double DYY = PY[0]*VY17+PY[1]*VY28+PY[2]*VY35+PY[3]*VY46;
  double DZZ = PZ[0]*VZ17+PZ[1]*VZ28+PZ[2]*VZ35+PZ[3]*VZ46;
  double DXY = PY[0]*VX17+PY[1]*VX28+PY[2]*VX35+PY[3]*VX46;
  double DXZ = PZ[0]*VX17+PZ[1]*VX28+PZ[2]*VX35+PZ[3]*VX46;
  double DYX = PX[0]*VY17+PX[1]*VY28+PX[2]*VY35+PX[3]*VY46;
  double DYZ = PZ[0]*VY17+PZ[1]*VY28+PZ[2]*VY35+PZ[3]*VY46;
  double DZX = PX[0]*VZ17+PX[1]*VZ28+PX[2]*VZ35+PX[3]*VZ46;
  double DZY = PY[0]*VZ17+PY[1]*VZ28+PY[2]*VZ35+PY[3]*VZ46;
  double DT1D2 = HALF*dt1;
  double D4, D5, D6;
  {
    double EXX = DXX, EYY = DYY, EZZ = DZZ, EXY = DXY, EYX = DYX;
    double EXZ = DXZ, EZX = DZX, EYZ = DYZ, EZY = DZY, AAA;
    DXX = DXX - DT1D2*(EXX*EXX + EYX*EYX + EZX*EZX);
    DYY = DYY - DT1D2*(EYY*EYY + EZY*EZY + EXY*EXY);
    DZZ = DZZ - DT1D2*(EZZ*EZZ + EXZ*EXZ + EYZ*EYZ);
    AAA = DT1D2*(EXX*EXY + EYX*EYY + EZX*EZY);
    DXY = DXY - AAA; DYX = DYX - AAA; D4 = DXY + DYX;
    AAA = DT1D2*(EYY*EYZ + EZY*EZZ + EXY*EXZ);
    DYZ = DYZ - AAA; DZY = DZY - AAA; D5 = DYZ + DZY;
    AAA = DT1D2*(EZZ*EZX + EXZ*EXX + EYZ*EYX);
    DXZ = DXZ - AAA; DZX = DZX - AAA; D6 = DXZ + DZX;
  }

  /* ---- SRHO3 (srho3.F, Lagrangian branch) ---- */
  const double rho0 = mat->rho0;
  double volo = st->vol0;
  if (offg == ZERO && voln == ONE) voln = volo;
  double dvol = voln - (rho0/st->rho)*volo;
  double rho = rho0*(volo/voln);
  double eint = st->eint*volo;

  /* ---- SRMALLA3 (srmall3.F): old stress ---- */
  double sold[6];
  for (int i = 0; i < 6; ++i) sold[i] = st->sig[i];

  /* ---- MMAIN law-2 prologue (mmain.F90:678-737) ---- */
  double vol_avg = voln - HALF*dvol;
  double amu = rho/rho0 - ONE;
  double rhoref = rho;

  /* ---- M2LAW (materials/mat/mat002/m2law.F:162-483) ---- */
  double G = mat->shear*off;
  double CA = mat->ca, SIGMX = mat->sigmx;
  double sig[6];
  for (int i = 0; i < 6; ++i) sig[i] = st->sig[i];
  double P = -THIRD*(sig[0] + sig[1] + sig[2]);
  double DAV = -THIRD*(DXX + DYY + DZZ);
  double G1 = dt1*G, G2 = TWO*G1;
  double SSP = sqrt((ONEP333*G + mat->bulk)/rho0);
  sig[0] = sig[0] + P + G2*(DXX + DAV);
  sig[1] = sig[1] + P + G2*(DYY + DAV);
  sig[2] = sig[2] + P + G2*(DZZ + DAV);
  sig[3] = sig[3] + G1*D4;
  sig[4] = sig[4] + G1*D5;
  sig[5] = sig[5] + G1*D6;
  double AJ2 = HALF*(sig[0]*sig[0] + sig[1]*sig[1] + sig[2]*sig[2])
             + sig[3]*sig[3] + sig[4]*sig[4] + sig[5]*sig[5];
  AJ2 = sqrt(THREE*AJ2);
  /* MSTRAIN_RATE (mat_share/mstrain_rate.F), IDEV = VP-2 */
  double EPSDOT;
  if (mat->vp - 2 == 0) {
    double E4 = HALF*D4, E5 = HALF*D5, E6 = HALF*D6;
    double EPSP = DXX*DXX + DYY*DYY + DZZ*DZZ + TWO*(E4*E4 + E5*E5 + E6*E6);
    EPSDOT = sqrt(EPSP);
  } else {
    double dav = (DXX + DYY + DZZ)*THIRD;
    double E1 = DXX - dav, E2 = DYY - dav, E3 = DZZ - dav;
    double E4 = HALF*D4, E5 = HALF*D5, E6 = HALF*D6;
    double EPSP = HALF*(E1*E1 + E2*E2 + E3*E3) + E4*E4 + E5*E5 + E6*E6;
    EPSDOT = sqrt(THREE*EPSP)/1.5;
  }
  double epsd = (mat->israte == 0) ? EPSDOT : st->epsd; /* rate filtering only with C>0 */
  if (mat->israte > 0) {
    double asrate = dmin(ONE, 0x1.5af1d78b58c40p+66*dt1); /* min(1, EP20*dt1) */
    epsd = asrate*EPSDOT + (ONE - asrate)*st->epsd;
  }
  double EPD = ONE;
  if (mat->cc != ZERO) {
    if (mat->vp == 1) { EPD = dmax(epsd, mat->epdr); EPD = log(EPD/mat->epdr); }
    else              { EPD = dmax(epsd, EM15);      EPD = log(EPD/mat->epdr); }
    double MT = dmax(EM15, mat->mexp);
    EPD = dmax(ZERO, EPD);
    EPD = (ONE + mat->cc*EPD)*(ONE - pow(ZERO, MT));
    if (mat->icc == 1) SIGMX = mat->sigmx*EPD;
  } else {
    double MT = dmax(EM15, mat->mexp);
    EPD = ONE - pow(ZERO, MT);
    if (mat->icc == 1) SIGMX = mat->sigmx*EPD;
  }
  double EPXE = st->pla, AK, QH;
  if (mat->cn == ONE) {
    AK = CA + mat->cb*EPXE;
    QH = mat->cb*EPD;
  } else if (EPXE > ZERO) {
    AK = CA + mat->cb*pow(EPXE, mat->cn);
    if (mat->cn > ONE) QH = (mat->cb*mat->cn*pow(EPXE, mat->cn - ONE))*EPD;
    else               QH = (mat->cb*mat->cn/pow(EPXE, ONE - mat->cn))*EPD;
  } else {
    AK = CA; QH = ZERO;
  }
  AK = AK*EPD;
  if (SIGMX < AK) { AK = SIGMX; QH = ZERO; }
  double SIGY = AK;
  if (EPXE > mat->epmx) { AK = ZERO; QH = ZERO; }
  double SCALE = dmin(ONE, AK/dmax(AJ2, EM15));
  double DPLA = (ONE - SCALE)*AJ2/dmax(THREE*G + QH, EM15);
  AK = AK + (ONE - mat->fisokin)*DPLA*QH;
  SCALE = dmin(ONE, AK/dmax(AJ2, EM15));
  for (int i = 0; i < 6; ++i) sig[i] = SCALE*sig[i];
  EPXE = EPXE + DPLA;

  /* ---- MQVISCB (mat_share/mqviscb.F:139-263, 382-416) ---- */
  double DD = -DXX - DYY - DZZ;
  double AD = ZERO, AL = ZERO;
  double CX = SSP + sqrt(ZERO); /* VD2 = 0 for Lagrangian bricks */
  /* MQVISCB's VOL dummy is M2LAW's VNEW (current volume); its VOL0 is V0. */
  if (off == ONE) {
    if (voln > ZERO) AL = pow(voln, ONE/THREE);
    AD = dmax(ZERO, DD);
  }
  double QA = ONE*prop->qa, QB = ONE*prop->qb;
  double QAA = QA*QA*AD;
  double QX = QB*SSP + AL*QAA + ONE*TWO*ZERO/dmax(EM20, rho*deltax)
            + (ZERO + ONE*ZERO)/dmax(EM20, rhoref*deltax);
  double qnew = rho*AD*AL*(QAA*AL + QB*SSP);
  double SSP_EQ = dmax(EM20, QX + sqrt(QX*QX + CX*CX));
  double DTX = deltax/SSP_EQ;
  double STI = ZERO;
  if (!(off == ZERO || offg < ZERO)) {
    double TIDT = ONE/DTX;
    double TRHO = rho*TIDT;
    double TVOL = voln*TIDT;
    STI = TRHO*TVOL;
  }
  DTX = prop->dtfac*DTX;
  double dtcand = 1.0e300;
  if (voln > ZERO && off != ZERO && offg >= ZERO) dtcand = DTX;

  /* ---- M2LAW epilogue: pressure, energy (m2law.F:390-483) ---- */
  double DTA = HALF*dt1;
  double dmg = st->dmg;
  if (EPXE > mat->epmx && dmg == ZERO) dmg = ONE;
  double PNEW = mat->bulk*amu;
  sig[0] = (sig[0] - PNEW)*off;
  sig[1] = (sig[1] - PNEW)*off;
  sig[2] = (sig[2] - PNEW)*off;
  sig[3] = sig[3]*off; sig[4] = sig[4]*off; sig[5] = sig[5]*off;
  {
    double E1 = DXX*(sold[0] + sig[0]);
    double E2 = DYY*(sold[1] + sig[1]);
    double E3 = DZZ*(sold[2] + sig[2]);
    double E4 = D4*(sold[3] + sig[3]);
    double E5 = D5*(sold[4] + sig[4]);
    double E6 = D6*(sold[5] + sig[5]);
    double EINC = vol_avg*(E1 + E2 + E3 + E4 + E5 + E6)*DTA - HALF*dvol*(st->qvis + qnew);
    eint = (eint + EINC*off)/dmax(EM15, volo);
  }

  /* ---- SHVIS3 (shvis3.F:163-238 coefficients, 317-411 JHBE>=1, 414-428 energy) ---- */
  double CAQ = FOURTH*off*prop->hvis;
  double FCL = CAQ*rho*pow(voln, TWO_THIRD);
  double FCQ = FCL*CAQ*HUNDRED;
  FCL = FCL*SSP;
  double Gm[3][8];
  Gm[0][0] =  ONE-PH[0][0]; Gm[0][1] = -ONE-PH[0][1]; Gm[0][2] =  ONE-PH[0][2]; Gm[0][3] = -ONE-PH[0][3];
  Gm[0][4] =  ONE+PH[0][2]; Gm[0][5] = -ONE+PH[0][3]; Gm[0][6] =  ONE+PH[0][0]; Gm[0][7] = -ONE+PH[0][1];
  Gm[1][0] =  ONE-PH[1][0]; Gm[1][1] =  ONE-PH[1][1]; Gm[1][2] = -ONE-PH[1][2]; Gm[1][3] = -ONE-PH[1][3];
  Gm[1][4] = -ONE+PH[1][2]; Gm[1][5] = -ONE+PH[1][3]; Gm[1][6] =  ONE+PH[1][0]; Gm[1][7] =  ONE+PH[1][1];
  Gm[2][0] =  ONE-PH[2][0]; Gm[2][1] = -ONE-PH[2][1]; Gm[2][2] = -ONE-PH[2][2]; Gm[2][3] =  ONE-PH[2][3];
  Gm[2][4] = -ONE+PH[2][2]; Gm[2][5] =  ONE+PH[2][3]; Gm[2][6] =  ONE+PH[2][0]; Gm[2][7] = -ONE+PH[2][1];
  double HG[3][4], HH[3][4]; /* [dir][mode] */
  const double *vv[3] = {vx, vy, vz};
  for (int d = 0; d < 3; ++d) {
    const double *w = vv[d];
    for (int m = 0; m < 3; ++m)
      HG[d][m] = Gm[m][0]*w[0] + Gm[m][1]*w[1] + Gm[m][2]*w[2] + Gm[m][3]*w[3]
               + Gm[m][4]*w[4] + Gm[m][5]*w[5] + Gm[m][6]*w[6] + Gm[m][7]*w[7];
    HG[d][3] = w[0] - w[1] + w[2] - w[3] - w[4] + w[5] - w[6] + w[7];
    for (int m = 0; m < 4; ++m) HH[d][m] = HG[d][m]*(FCL + fabs(HG[d][m])*FCQ);
  }
  static const double s4[8] = {-1, 1, -1, 1, 1, -1, 1, -1}; /* sign of HX4 term per node */
  double F[3][8];
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < 8; ++k) {
      double h4 = (s4[k] > 0) ? HH[d][3] : -HH[d][3];
      F[d][k] = -Gm[0][k]*HH[d][0] - Gm[1][k]*HH[d][1] - Gm[2][k]*HH[d][2] + h4;
    }
  double EHOU = dt1*(HH[2][0]*HG[2][0] + HH[2][1]*HG[2][1] + HH[2][2]*HG[2][2] + HH[2][3]*HG[2][3]
                   + HH[0][0]*HG[0][0] + HH[0][1]*HG[0][1] + HH[0][2]*HG[0][2] + HH[0][3]*HG[0][3]
                   + HH[1][0]*HG[1][0] + HH[1][1]*HG[1][1] + HH[1][2]*HG[1][2] + HH[1][3]*HG[1][3]);

  /* ---- SFINT3 (sfint3.F:257-322), SVIS = 0 for law 2 ---- */
  {
    double S1 = (sig[0] + ZERO - qnew)*voln;
    double S2 = (sig[1] + ZERO - qnew)*voln;
    double S3 = (sig[2] + ZERO - qnew)*voln;
    double S4 = (sig[3] + ZERO)*voln;
    double S5 = (sig[4] + ZERO)*voln;
    double S6 = (sig[5] + ZERO)*voln;
    static const int pa[4] = {0, 1, 2, 3}, pb[4] = {6, 7, 4, 5};
    for (int a = 0; a < 4; ++a) {
      double FINT;
      FINT = S1*PX[a] + S4*PY[a] + S6*PZ[a];
      F[0][pa[a]] = F[0][pa[a]] - FINT; F[0][pb[a]] = F[0][pb[a]] + FINT;
      FINT = S2*PY[a] + S4*PX[a] + S5*PZ[a];
      F[1][pa[a]] = F[1][pa[a]] - FINT; F[1][pb[a]] = F[1][pb[a]] + FINT;
      FINT = S3*PZ[a] + S6*PX[a] + S5*PY[a];
      F[2][pa[a]] = F[2][pa[a]] - FINT; F[2][pb[a]] = F[2][pb[a]] + FINT;
    }
  }

  /* ---- SRROTA3 with R transposed (sforc3.F:1604-1614): F_global = R F_local ---- */
  for (int k = 0; k < 8; ++k) {
    double X = R[0][0]*F[0][k] + R[0][1]*F[1][k] + R[0][2]*F[2][k];
    double Y = R[1][0]*F[0][k] + R[1][1]*F[1][k] + R[1][2]*F[2][k];
    double Z = R[2][0]*F[0][k] + R[2][1]*F[1][k] + R[2][2]*F[2][k];
    F[0][k] = X; F[1][k] = Y; F[2][k] = Z;
  }

  /* ---- SCUMU3P (scumu3p.F:105-172,196-198) ---- */
  if (offg < ZERO)
    for (int d = 0; d < 3; ++d) for (int k = 0; k < 8; ++k) F[d][k] = ZERO;
  STI = FOURTH*STI;

  for (int d = 0; d < 3; ++d) for (int k = 0; k < 8; ++k) out->f[d][k] = F[d][k];
  out->sti = STI; out->dtx = dtcand; out->ehou = EHOU; out->voln = voln; out->mstop = mstop;

  for (int i = 0; i < 6; ++i) st->sig[i] = sig[i];
  st->eint = eint; st->rho = rho; st->pla = EPXE; st->qvis = qnew;
  st->offg = offg; st->epsd = epsd; st->dmg = dmg;
  (void)SIGY;
}

/* ---- FILLCNE (starter/source/spmd/domdec2.F:2100-2150) + W_PON (w_pon.F:550-573)
 *      + GET_FSKY_ADDRESS (get_fsky_address.F90:90-102) ---- */
typedef struct { int32_t uid, idx; } uid_pair;
static int cmp_uid(const void *a, const void *b) {
  const uid_pair *p = a, *q = b;
  if (p->uid != q->uid) return (p->uid < q->uid) ? -1 : 1;
  return (p->idx < q->idx) ? -1 : (p->idx > q->idx);
}
int oracle_build_pon(int numnod, int numels, const int32_t *ixs,
                     const int32_t *uid, int32_t *adsky, int32_t *iads) {
  int32_t *cnt = calloc((size_t)numnod + 1, sizeof(int32_t));
  uid_pair *ord = malloc(sizeof(uid_pair)*(size_t)(numels > 0 ? numels : 1));
  if (!cnt || !ord) { free(cnt); free(ord); return -1; }
  for (int e = 0; e < numels; ++e)
    for (int k = 0; k < 8; ++k) {
      int n = ixs[8*e + k];
      if (n < 0 || n >= numnod) { free(cnt); free(ord); return -2; }
      cnt[n]++;
    }
  adsky[0] = 0;
  for (int n = 0; n < numnod; ++n) adsky[n + 1] = adsky[n] + cnt[n];
  for (int n = 0; n < numnod; ++n) cnt[n] = adsky[n];
  /* MY_ORDERS: elements in increasing user id (stable) */
  for (int e = 0; e < numels; ++e) { ord[e].uid = uid[e]; ord[e].idx = e; }
  qsort(ord, (size_t)numels, sizeof(uid_pair), cmp_uid);
  /* CNE(ADSKY(N)) = I in that order; each occurrence of node N in element I
   * takes the first corner not yet given a slot (GET_FSKY_ADDRESS bitmask). */
  for (int j = 0; j < numels; ++j) {
    int e = ord[j].idx;
    for (int k = 0; k < 8; ++k) {
      int n = ixs[8*e + k];
      iads[8*e + k] = cnt[n]++;
    }
  }
  free(cnt); free(ord);
  return 0;
}

/* ---- Whole cycle in RESOL order ---- */
void oracle_cycle(orad_omodel_t *m) {
  const int ne = m->numels, nn = m->numnod;
  double dt2t = 1.0e300;
  int mstop = 0;
  double *ehou = malloc(sizeof(double)*(size_t)(ne > 0 ? ne : 1));
  /* resol.F:2639-2640 */
  m->dt1 = m->dt2;
  double dt1 = m->dt1;
#ifdef _OPENMP
  int nth = m->nthreads > 0 ? m->nthreads : omp_get_max_threads();
#endif
  /* FORINT/SFORC3 over elements (forint.F:325), slots are disjoint. */
#pragma omp parallel for num_threads(nth) schedule(static) reduction(min:dt2t) reduction(|:mstop)
  for (int e = 0; e < ne; ++e) {
    double xg[3][8], vg[3][8];
    for (int k = 0; k < 8; ++k) {
      int n = m->ixs[8*e + k];
      for (int d = 0; d < 3; ++d) { xg[d][k] = m->x[3*n + d]; vg[d][k] = m->v[3*n + d]; }
    }
    orad_bout_t o;
    oracle_brick_element((const double (*)[8])xg, (const double (*)[8])vg, &m->st[e], dt1,
                         &m->mat, &m->prop, &o);
    for (int k = 0; k < 8; ++k) {
      double *f = m->fsky + 4*(size_t)m->iads[8*e + k];
      f[0] = o.f[0][k]; f[1] = o.f[1][k]; f[2] = o.f[2][k]; f[3] = o.sti;
    }
    if (o.dtx < dt2t) dt2t = o.dtx;
    mstop |= o.mstop;
    ehou[e] = o.ehou;
  }
  /* Hourglass energy (SHVIS3 -> PARTSAV(8,.)), summed in element order. */
  for (int e = 0; e < ne; ++e) m->ehour += ehou[e];
  free(ehou);
  /* /INTER/TYPE7 (i7mainf.F), its skyline entries FSKYI sorted by node */
  double *isk = NULL;
  int32_t *iord = NULL, *ifirst = NULL;
  if (m->contact)
    oracle_int7_cycle(m->contact, nn, m->x, m->v, m->ms, dt1, &dt2t, &isk, &iord, &ifirst);
  /* ASSPAR4 (asspar4.F:157-175) ordered per-node gather, then the node's interface
   * entries in skyline order (asspar4.F:560-583) */
#pragma omp parallel for num_threads(nth) schedule(static)
  for (int n = 0; n < nn; ++n) {
    double ax = 0.0, ay = 0.0, az = 0.0, sf = 0.0;
    for (int k = m->adsky[n]; k < m->adsky[n + 1]; ++k) {
      const double *f = m->fsky + 4*(size_t)k;
      ax = ax + f[0]; ay = ay + f[1]; az = az + f[2]; sf = sf + f[3];
    }
    if (iord)
      for (int k = ifirst[n]; k < ifirst[n + 1]; ++k) {
        const double *f = isk + 4*(size_t)iord[k];
        ax = ax + f[0]; ay = ay + f[1]; az = az + f[2]; sf = sf + f[3];
      }
    m->a[3*n] = ax; m->a[3*n + 1] = ay; m->a[3*n + 2] = az; m->stifn[n] = sf;
  }
  free(isk); free(iord); free(ifirst);
  orad_onodes_t nd = {0};
  nd.numnod = nn; nd.rot = 0;
  nd.x = m->x; nd.v = m->v; nd.d = m->d; nd.a = m->a; nd.ms = m->ms; nd.icodt = m->icodt;
  nd.fv_dir = m->fv_dir; nd.fv_val = m->fv_val; nd.fv_mem = m->fv_mem;
  nd.tt = m->tt; nd.dt1 = m->dt1; nd.dt2 = m->dt2; nd.dt12 = m->dt12; nd.dt2old = m->dt2old;
  nd.dtmx = m->dtmx; nd.wfext = m->wfext; nd.ncycle = m->ncycle; nd.nthreads = m->nthreads;
  oracle_nodal_update(&nd, dt2t);
  m->tt = nd.tt; m->dt1 = nd.dt1; m->dt2 = nd.dt2; m->dt12 = nd.dt12; m->dt2old = nd.dt2old;
  m->wfext = nd.wfext; m->ncycle = nd.ncycle;
  m->encin = nd.encin; m->enrot = nd.enrot;
  m->mstop |= mstop;
}

/* Nodal part of the cycle (declared in orad_oracle.h).  The FIXVEL loop runs over the
 * imposed nodes in increasing node order, the order of the reference's entries. */
void oracle_nodal_update(orad_onodes_t *m, double dt2t) {
  const int nn = m->numnod;
#ifdef _OPENMP
  int nth = m->nthreads > 0 ? m->nthreads : omp_get_max_threads();
#endif
  /* resol.F:2639-2640 */
  m->dt1 = m->dt2;
  const double dt1 = m->dt1;
  /* resol.F:5857-5861 */
  double dt2 = 1.0e6;                         /* DT2 = EP06 */
  if (dt2t < dt2) dt2 = dt2t;
  /* resol.F:6080 writes 1.1*DT2OLD with a default-REAL literal */
  dt2 = dmin(dmin(dt2, (double)1.1f*m->dt2old), m->dtmx);
  /* resol.F:6224-6226 */
  m->dt2old = dt2;
  m->dt2 = dt2;
  m->dt12 = HALF*(dt1 + dt2);
  const double dt12 = m->dt12;
  /* ACCELE (accele.F:64-110) and BCS10 -> BCS1V (bcs1.F:46-76), node by node */
#pragma omp parallel for num_threads(nth) schedule(static)
  for (int n = 0; n < nn; ++n) {
    double *a = m->a + 3*n, *v = m->v + 3*n;
    if (m->ms[n] > ZERO) {
      double RTMP = ONE/m->ms[n];
      a[0] = a[0]*RTMP; a[1] = a[1]*RTMP; a[2] = a[2]*RTMP;
    } else {
      a[0] = ZERO; a[1] = ZERO; a[2] = ZERO;
    }
    int c = m->icodt[n];
    if (c & 4) { v[0] = ZERO; a[0] = ZERO; }
    if (c & 2) { v[1] = ZERO; a[1] = ZERO; }
    if (c & 1) { v[2] = ZERO; a[2] = ZERO; }
    if (m->rot) {
      double *ar = m->ar + 3*n, *vr = m->vr + 3*n;
      if (m->inr[n] > ZERO) {
        double RTMP = ONE/m->inr[n];
        ar[0] = ar[0]*RTMP; ar[1] = ar[1]*RTMP; ar[2] = ar[2]*RTMP;
      } else {
        ar[0] = ZERO; ar[1] = ZERO; ar[2] = ZERO;
      }
      int r = m->icodr[n];
      if (r & 4) { vr[0] = ZERO; ar[0] = ZERO; }
      if (r & 2) { vr[1] = ZERO; ar[1] = ZERO; }
      if (r & 1) { vr[2] = ZERO; ar[2] = ZERO; }
    }
  }
  /* FIXVEL, constant curve, global direction (fixvel.F:333-374): the imposed component's
   * acceleration becomes (VEL - V)/DT12; its work DW = MS/4 (A DT12 + 2V)(A - AOLD) enters
   * WFEXT weighted DT1 now and DT2 at the next cycle (fixvel.F:333-335, 834-836). */
  if (m->fv_dir) {
    double wf = m->wfext;
    for (int n = 0; n < nn; ++n) {
      if (m->fv_dir[n] <= 0) continue;
      const int j = m->fv_dir[n] - 1;
      double *a = m->a + 3*n, *v = m->v + 3*n;
      double dwm = m->fv_mem ? m->fv_mem[n] : ZERO;
      wf = wf + ONE*dwm;
      const double yc = (m->fv_val[n] - v[j]) / dt12;
      const double aold = a[j];
      a[j] = yc;
      const double dw = FOURTH*m->ms[n]*ONE*ONE*(a[j]*dt12 + TWO*v[j])*(a[j] - aold);
      wf = wf + dt1*dw;
      if (m->fv_mem) m->fv_mem[n] = ZERO + m->dt2*dw;
    }
    m->wfext = wf;
  }
  /* ECRIT (output/ecrit.F:192-226, 303-325), called from SORTIE_MAIN between FIXVEL and
   * VELOCITY: kinetic energies at the listing's velocity V + DT1/2 A, in node order */
  {
    const double DT05 = HALF*dt1;
    double ek = ZERO, er = ZERO;
    for (int n = 0; n < nn; ++n) {
      const double *a = m->a + 3*n, *v = m->v + 3*n;
      const double VX = v[0] + DT05*a[0], VY = v[1] + DT05*a[1], VZ = v[2] + DT05*a[2];
      ek = ek + (VX*VX + VY*VY + VZ*VZ)*HALF*m->ms[n];
      if (m->rot) {
        const double *ar = m->ar + 3*n, *vr = m->vr + 3*n;
        const double WX = vr[0] + DT05*ar[0], WY = vr[1] + DT05*ar[1], WZ = vr[2] + DT05*ar[2];
        er = er + (WX*WX + WY*WY + WZ*WZ)*HALF*m->inr[n];
      }
    }
    m->encin = ek;
    m->enrot = er;
  }
  /* TT (resol.F:8256-8266), VELOCITY (velocity.F:56-83), DEPLA (displacement.F:88-101) */
  m->tt = m->tt + dt2;
#pragma omp parallel for num_threads(nth) schedule(static)
  for (int n = 0; n < nn; ++n) {
    double *a = m->a + 3*n, *v = m->v + 3*n;
    for (int d = 0; d < 3; ++d) v[d] = v[d] + dt12*a[d];
    if (m->rot) {
      double *ar = m->ar + 3*n, *vr = m->vr + 3*n;
      for (int d = 0; d < 3; ++d) vr[d] = vr[d] + dt12*ar[d];
    }
    for (int d = 0; d < 3; ++d) {
      double VDT = dt2*v[d];
      m->d[3*n + d] = m->d[3*n + d] + VDT;
      m->x[3*n + d] = m->x[3*n + d] + VDT;
    }
  }
  m->ncycle += 1;
}

void oracle_energies(const orad_omodel_t *m, double *eint, double *ekin) {
  double ei = 0.0, ek = 0.0;
  for (int e = 0; e < m->numels; ++e) ei += m->st[e].eint*m->st[e].vol0;
  for (int n = 0; n < m->numnod; ++n) {
    const double *v = m->v + 3*n;
    ek += HALF*m->ms[n]*(v[0]*v[0] + v[1]*v[1] + v[2]*v[2]);
  }
  *eint = ei; *ekin = ek;
}

void oracle_brick_group(int n, const double *xg, const double *vg, orad_bstate_t *st,
                        double dt1, const orad_law2_t *mat, const orad_sprop_t *prop,
                        orad_bout_t *out) {
  for (int e = 0; e < n; ++e)
    oracle_brick_element((const double(*)[8])(xg + 24 * (size_t)e),
                         (const double(*)[8])(vg + 24 * (size_t)e), st + e, dt1, mat, prop,
                         out + e);
}

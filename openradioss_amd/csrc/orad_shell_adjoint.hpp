// orad_shell_adjoint.hpp -- hand-written reverse (adjoint) pass of one
// Belytschko-Tsay shell through the CFORC3 chain of orad_shell_chain.hpp, FP64.
//
// shell_vjp_bt evaluates the chain once forward (values only, the branches taken on
// the values, as the chain takes them) and then once backward: from the adjoints of
// the corner slots (forces and moments) and of the state carried to the next cycle
// (SIG and PLA of every point, the incremental hourglass forces HOUR(1:3)) it forms
// the adjoints of the corner coordinates, velocities and rotational velocities, of
// the state that came in, of the thickness and of A, B, n.  It is the exact
// derivative of the arithmetic the forward chain performs for law 2 with radial
// return (IPLA = 0) and no strain-rate term (C = 0); element deletion and the time
// step are branch decisions and carry no derivative.  Stages, in reverse order:
// CFINT3 (cfint3.F:165-253), CHVIS3 (chvis3.F:120-330), MULAWC / SIGEPS02C / M2CPLR
// radial return (mulawc.F90:530-3046, m2cplr.F:95-229), CSTRA3, CCOEF3, CCURV3,
// CDEFO3, CDERI3, CNVEC3 (see the forward chain for each routine's lines).
#pragma once
#include "orad_shell_chain.hpp"

namespace orad {

struct V3 { double x, y, z; };
__device__ __forceinline__ V3 v3(double x, double y, double z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// the chain's cross products, component for component
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// adjoint of c = a x b: la += b x lc, lb += lc x a
__device__ __forceinline__ void cross_rev(V3 a, V3 b, V3 lc, V3 &la, V3 &lb) {
  la = la + cross(b, lc);
  lb = lb + cross(lc, a);
}
// adjoint of u = w / max(|w|, em20): lw = (lu - u (u . lu)) / |w| (unclipped)
__device__ __forceinline__ V3 unit_rev(V3 u, double nw, bool clipped, double inv, V3 lu) {
  if (clipped) return inv * lu;
  return (1.0 / nw) * (lu - dot(u, lu) * u);
}

// Adjoints of the kinematics stage's outputs.
struct KinAdj {
  V3 E1, E2, E3;
  double PX1, PY1, PX2, PY2, AREA, VHX, VHY;
  double VX[4], VY[4], VZ[4], RX[4], RY[4];
  double EXX, EYY, EXY, EYZ, EXZ, KXX, KYY, KXY;   // of the unscaled rates
};

// Reverse of shell_kinematics<double>: given the adjoints of its outputs, adds the
// adjoints of the corner coordinates (lx), velocities (lv) and rotational
// velocities (lvr).  Recomputes the forward values it needs.
__device__ inline void shell_kinematics_rev(const double xg[3][4], const double vg[3][4],
                                            const double vrg[3][4], bool dead, double dt1,
                                            double EM20, KinAdj &L, double lx[3][4],
                                            double lv[3][4], double lvr[3][4]) {
  auto P = [&](int c) { return v3(xg[0][c], xg[1][c], xg[2][c]); };
  const V3 d21 = P(1) - P(0), d32 = P(2) - P(1), d34 = P(2) - P(3), d41 = P(3) - P(0);
  const V3 d31 = P(2) - P(0);
  // ---- CNVEC3 forward ----
  const V3 e1 = d21 + d34, e2 = d32 + d41;
  const V3 e3 = cross(e1, e2);
  const double n3 = __builtin_sqrt(dot(e3, e3));
  const bool c3 = !(n3 > EM20);
  const double s3 = 1.0 / (c3 ? EM20 : n3);
  const V3 E3 = s3 * e3;
  const double S1 = dot(e1, e1), S2 = dot(e2, e2);
  const double SUMA = __builtin_sqrt(S1 / S2);
  const V3 w = cross(e2, E3);
  const V3 e1p = e1 + SUMA * w;
  const double n1 = __builtin_sqrt(dot(e1p, e1p));
  const bool c1 = !(n1 > EM20);
  const double s1 = 1.0 / (c1 ? EM20 : n1);
  const V3 E1 = s1 * e1p;
  const V3 E2 = cross(E3, E1);
  // ---- CDERI3 forward ----
  const double X2 = dot(E1, d21), Y2 = dot(E2, d21), Y3 = dot(E2, d31), X3 = dot(E1, d31);
  const double X4 = dot(E1, d41), Y4 = dot(E2, d41);
  const double PX1 = 0.5 * (Y2 - Y4), PY1 = 0.5 * (X4 - X2), PX2 = 0.5 * Y3, PY2 = -0.5 * X3;
  const double AREA0 = 2.0 * (PY2 * PX1 - PY1 * PX2);
  const bool cA = !(AREA0 > EM20);
  const double AREA = cA ? EM20 : AREA0;
  // ---- CDEFO3 / CCURV3 forward values used by the reverse ----
  V3 vc[4], wc[4];
  double VX[4], VY[4], VZ[4], RX[4], RY[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    vc[c] = dead ? v3(0.0, 0.0, 0.0) : v3(vg[0][c], vg[1][c], vg[2][c]);
    wc[c] = dead ? v3(0.0, 0.0, 0.0) : v3(vrg[0][c], vrg[1][c], vrg[2][c]);
    VX[c] = dot(E1, vc[c]); VY[c] = dot(E2, vc[c]); VZ[c] = dot(E3, vc[c]);
    RX[c] = dot(E1, wc[c]); RY[c] = dot(E2, wc[c]);
  }
  const double VZ13 = VZ[0] - VZ[2], VZ24 = VZ[1] - VZ[3];
  const double DT1V4 = 0.25 * dt1;
  const double TMP2A = PY2 + PY1;
  const bool cTA = !(tabs(TMP2A) > EM20);
  const double TMP3A = tsign(cTA ? EM20 : tabs(TMP2A), TMP2A);
  const double DA = VZ13 - VZ24;
  const double TMP1A = DT1V4 * DA * DA / TMP3A;
  const double VX13 = VX[0] - VX[2] - TMP1A, VX24 = VX[1] - VX[3] + TMP1A;
  const double TMP1B = PX2 - PX1;
  const bool cTB = !(tabs(TMP1B) > EM20);
  const double TMP3B = tsign(cTB ? EM20 : tabs(TMP1B), TMP1B);
  const double DB = VZ13 + VZ24;
  const double TMP2B = DT1V4 * DB * DB / TMP3B;
  const double VY13 = VY[0] - VY[2] + TMP2B, VY24 = VY[1] - VY[3] + TMP2B;
  const double RX13 = RX[0] - RX[2], RX24 = RX[1] - RX[3];
  const double RXAV = RX[0] + RX[1] + RX[2] + RX[3];
  const double RY13 = RY[0] - RY[2], RY24 = RY[1] - RY[3];
  const double RYAV = RY[0] + RY[1] + RY[2] + RY[3];

  // ---- reverse: CCURV3 ----
  double lPX1 = L.PX1, lPY1 = L.PY1, lPX2 = L.PX2, lPY2 = L.PY2, lAREA = L.AREA;
  double lRX[4], lRY[4], lVX[4], lVY[4], lVZ[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    lRX[c] = L.RX[c]; lRY[c] = L.RY[c]; lVX[c] = L.VX[c]; lVY[c] = L.VY[c]; lVZ[c] = L.VZ[c];
  }
  // EXZ += RYAV * (AREA/4); EYZ -= RXAV * (AREA/4)
  const double lRYAV = L.EXZ * (0.25 * AREA), lRXAV = -L.EYZ * (0.25 * AREA);
  lAREA += 0.25 * (L.EXZ * RYAV - L.EYZ * RXAV);
  // KXY = PY1 RY13 + PY2 RY24 - (PX1 RX13 + PX2 RX24); KXX = PX1 RY13 + PX2 RY24;
  // KYY = -PY1 RX13 - PY2 RX24
  const double lKXX = L.KXX, lKYY = L.KYY, lKXY = L.KXY;
  const double lRY13 = lKXX * PX1 + lKXY * PY1, lRY24 = lKXX * PX2 + lKXY * PY2;
  const double lRX13 = -lKYY * PY1 - lKXY * PX1, lRX24 = -lKYY * PY2 - lKXY * PX2;
  lPX1 += lKXX * RY13 - lKXY * RX13;
  lPX2 += lKXX * RY24 - lKXY * RX24;
  lPY1 += lKXY * RY13 - lKYY * RX13;
  lPY2 += lKXY * RY24 - lKYY * RX24;
  lRY[0] += lRY13 + lRYAV; lRY[2] += -lRY13 + lRYAV;
  lRY[1] += lRY24 + lRYAV; lRY[3] += -lRY24 + lRYAV;
  lRX[0] += lRX13 + lRXAV; lRX[2] += -lRX13 + lRXAV;
  lRX[1] += lRX24 + lRXAV; lRX[3] += -lRX24 + lRXAV;
  // ---- reverse: CDEFO3 ----
  // EYY = PY1 VY13 + PY2 VY24; EXY = PY1 VX13 + PY2 VX24 + PX1 VY13 + PX2 VY24;
  // EXX = PX1 VX13 + PX2 VX24
  const double lEXX = L.EXX, lEYY = L.EYY, lEXY = L.EXY;
  const double lVY13 = lEYY * PY1 + lEXY * PX1, lVY24 = lEYY * PY2 + lEXY * PX2;
  const double lVX13 = lEXX * PX1 + lEXY * PY1, lVX24 = lEXX * PX2 + lEXY * PY2;
  lPY1 += lEYY * VY13 + lEXY * VX13;
  lPY2 += lEYY * VY24 + lEXY * VX24;
  lPX1 += lEXY * VY13 + lEXX * VX13;
  lPX2 += lEXY * VY24 + lEXX * VX24;
  // VY13 = VY1 - VY3 + TMP2B, VY24 = VY2 - VY4 + TMP2B
  lVY[0] += lVY13; lVY[2] -= lVY13; lVY[1] += lVY24; lVY[3] -= lVY24;
  const double lTMP2B = lVY13 + lVY24;
  // TMP2B = DT1V4 DB^2 / TMP3B
  const double lDB = lTMP2B * DT1V4 * 2.0 * DB / TMP3B;
  if (!cTB) {
    // TMP3B = TMP1B (sign(|t|, t) = t): d/dTMP1B of DT1V4 DB^2 / TMP1B
    const double lT = -lTMP2B * TMP2B / TMP3B;
    lPX2 += lT; lPX1 -= lT;
  }
  // VX13 = VX1 - VX3 - TMP1A, VX24 = VX2 - VX4 + TMP1A
  lVX[0] += lVX13; lVX[2] -= lVX13; lVX[1] += lVX24; lVX[3] -= lVX24;
  const double lTMP1A = -lVX13 + lVX24;
  const double lDA = lTMP1A * DT1V4 * 2.0 * DA / TMP3A;
  if (!cTA) {
    const double lT = -lTMP1A * TMP1A / TMP3A;
    lPY2 += lT; lPY1 += lT;
  }
  // EXZ = PX1 VZ13 + PX2 VZ24; EYZ = PY1 VZ13 + PY2 VZ24 (before the curvature terms)
  const double lEXZ = L.EXZ, lEYZ = L.EYZ;
  double lVZ13 = lEXZ * PX1 + lEYZ * PY1 + lDA + lDB;
  double lVZ24 = lEXZ * PX2 + lEYZ * PY2 - lDA + lDB;
  lPX1 += lEXZ * VZ13; lPX2 += lEXZ * VZ24; lPY1 += lEYZ * VZ13; lPY2 += lEYZ * VZ24;
  lVZ[0] += lVZ13; lVZ[2] -= lVZ13; lVZ[1] += lVZ24; lVZ[3] -= lVZ24;
  // projections VX = E1.v, VY = E2.v, VZ = E3.v, RX = E1.vr, RY = E2.vr
  V3 lE1 = L.E1, lE2 = L.E2, lE3 = L.E3;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    lE1 = lE1 + lVX[c] * vc[c] + lRX[c] * wc[c];
    lE2 = lE2 + lVY[c] * vc[c] + lRY[c] * wc[c];
    lE3 = lE3 + lVZ[c] * vc[c];
    if (!dead) {
      const V3 g = lVX[c] * E1 + lVY[c] * E2 + lVZ[c] * E3;
      const V3 r = lRX[c] * E1 + lRY[c] * E2;
      lv[0][c] += g.x; lv[1][c] += g.y; lv[2][c] += g.z;
      lvr[0][c] += r.x; lvr[1][c] += r.y; lvr[2][c] += r.z;
    }
  }
  // ---- reverse: CDERI3 ----
  // VHX = (-X2 + X3 - X4)/AREA, VHY = (-Y2 + Y3 - Y4)/AREA
  const double VHX = (-X2 + X3 - X4) / AREA, VHY = (-Y2 + Y3 - Y4) / AREA;
  double lX2 = -L.VHX / AREA, lX3 = L.VHX / AREA, lX4 = -L.VHX / AREA;
  double lY2 = -L.VHY / AREA, lY3 = L.VHY / AREA, lY4 = -L.VHY / AREA;
  lAREA += -(L.VHX * VHX + L.VHY * VHY) / AREA;
  if (!cA) {
    // AREA = 2 (PY2 PX1 - PY1 PX2)
    lPY2 += 2.0 * lAREA * PX1; lPX1 += 2.0 * lAREA * PY2;
    lPY1 -= 2.0 * lAREA * PX2; lPX2 -= 2.0 * lAREA * PY1;
  }
  // PX1 = (Y2 - Y4)/2, PY1 = (X4 - X2)/2, PX2 = Y3/2, PY2 = -X3/2
  lY2 += 0.5 * lPX1; lY4 -= 0.5 * lPX1;
  lX4 += 0.5 * lPY1; lX2 -= 0.5 * lPY1;
  lY3 += 0.5 * lPX2;
  lX3 -= 0.5 * lPY2;
  lE1 = lE1 + lX2 * d21 + lX3 * d31 + lX4 * d41;
  lE2 = lE2 + lY2 * d21 + lY3 * d31 + lY4 * d41;
  V3 ld21 = lX2 * E1 + lY2 * E2, ld31 = lX3 * E1 + lY3 * E2, ld41 = lX4 * E1 + lY4 * E2;
  // ---- reverse: CNVEC3 ----
  // E2 = E3 x E1
  cross_rev(E3, E1, lE2, lE3, lE1);
  // E1 = e1p / |e1p|
  const V3 le1p = unit_rev(E1, n1, c1, s1, lE1);
  // e1p = e1 + SUMA w, w = e2 x E3, SUMA = sqrt(S1/S2)
  V3 le1 = le1p;
  const V3 lw = SUMA * le1p;
  const double lSUMA = dot(le1p, w);
  V3 le2 = v3(0.0, 0.0, 0.0);
  cross_rev(e2, E3, lw, le2, lE3);
  if (SUMA > 0.0) {
    const double q = lSUMA * 0.5 / SUMA;
    const double lS1 = q / S2, lS2 = -q * S1 / (S2 * S2);
    le1 = le1 + (2.0 * lS1) * e1;
    le2 = le2 + (2.0 * lS2) * e2;
  }
  // E3 = e3 / |e3|, e3 = e1 x e2
  const V3 le3 = unit_rev(E3, n3, c3, s3, lE3);
  cross_rev(e1, e2, le3, le1, le2);
  // e1 = d21 + d34, e2 = d32 + d41
  ld21 = ld21 + le1;
  const V3 ld34 = le1, ld32 = le2;
  ld41 = ld41 + le2;
  // corners: d21 = P1 - P0, d32 = P2 - P1, d34 = P2 - P3, d41 = P3 - P0, d31 = P2 - P0
  const V3 lP0 = v3(0.0, 0.0, 0.0) - ld21 - ld41 - ld31;
  const V3 lP1 = ld21 - ld32;
  const V3 lP2 = ld32 + ld34 + ld31;
  const V3 lP3 = ld41 - ld34;
  const V3 lPc[4] = {lP0, lP1, lP2, lP3};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    lx[0][c] += lPc[c].x; lx[1][c] += lPc[c].y; lx[2][c] += lPc[c].z;
  }
}

// Reverse pass of one BT element.  st(col) reads the pre-cycle state; lpost(j) is
// the adjoint of the post-cycle state entry j (adjoint numbering of
// orad_adjoint_args.hpp: SIG 0..24, PLA 25..29, HOUR(1:3) 30..32); lslot(c, i) the
// adjoint of corner c's slot component i (0..2 force, 3..5 moment).  lpre(j, value)
// is called once for every entry j of the points in use and HOUR(1:3) with the
// adjoint of the pre-cycle state; lx/lv/lvr, g_thk and g_jc[3] are added to.
template <typename ST, typename LPOST, typename LSLOT, typename LPRE>
__device__ inline void shell_vjp_bt(const double xg[3][4], const double vg[3][4],
                                    const double vrg[3][4], double thk0, bool tri, ST st,
                                    double dt1, const ShMat &m, const double jc[3],
                                    const ShProp &p, const ShConst &k, LPOST lpost,
                                    LSLOT lslot, LPRE lpre, double lx[3][4], double lv[3][4],
                                    double lvr[3][4], double &g_thk, double g_jc[3]) {
  const double ZERO = 0.0, ONE = 1.0, TWO = 2.0, THREE = 3.0, HALF = 0.5, FOURTH = 0.25;
  const double EM20 = k.em20;
  // ---- forward: CCOOR3, kinematics, CCOEF3, CSTRA3 ----
  double offg = st(SC_OFFG);
  double off = tmin(ONE, tabs(offg));
  const bool dead = offg < ZERO;
  ShellKin<double> g;
  shell_kinematics<double>(xg, vg, vrg, dead, dt1, EM20, g);
  const double AREA = g.AREA, PX1 = g.PX1, PY1 = g.PY1, PX2 = g.PX2, PY2 = g.PY2;
  const double VHX = g.VHX, VHY = g.VHY;
  const V3 E1 = v3(g.E1X, g.E1Y, g.E1Z), E2 = v3(g.E2X, g.E2Y, g.E2Z), E3 = v3(g.E3X, g.E3Y, g.E3Z);
  const double THK02 = thk0 * thk0;
  const double RHO = m.rho0, YM = m.young, NU = m.nu, G = m.shear, SSP = m.ssp;
  const double H1 = p.hm, H2 = p.hf, H3 = p.hr;
  const double FAC1TMP = TWO * (ONE + NU) * THK02;
  const double ISH = p.ish, FSH = p.fsh;
  const double SDEN = FSH * AREA + FAC1TMP;
  const double SHF = FSH * ((ONE - ISH) + ISH * FAC1TMP / SDEN);
  const double GS = G * SHF;
  const double FAC1 = dt1 / AREA;
  const double EXX = g.EXX * FAC1, EYY = g.EYY * FAC1, EXY = g.EXY * FAC1;
  const double EYZ = g.EYZ * FAC1, EXZ = g.EXZ * FAC1;
  const double KXX = g.KXX * FAC1, KYY = g.KYY * FAC1, KXY = g.KXY * FAC1;
  const double dtinv = dt1 / tmax(dt1 * dt1, EM20);
  // ---- forward: law 2 at every point (IPLA = 0, C = 0) ----
  const double a11 = YM / (ONE - NU * NU), a12 = a11 * NU;
  const double ca = jc[0], cb = jc[1], cn = jc[2], ymax = m.ymax;
  double FOR[5] = {ZERO, ZERO, ZERO, ZERO, ZERO}, MOM[3] = {ZERO, ZERO, ZERO};
  double off_pt[5];
  const double off_old = off;
  double epchk = ZERO;
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
    if (r < ONE) {
      s[0] = s[0] * r; s[1] = s[1] * r; s[2] = s[2] * r;
      pla = pla + off * tmax(ZERO, (svm - yld) / YM);
      epchk = tmax(pla, epchk);
    }
    if (off == off_old && off > ZERO) {
      if (off == ONE && epchk >= m.epmx) off = k.four_over_5;
      else if (off < ONE) off = off * k.four_over_5;
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) FOR[c] = FOR[c] + p.wf[it] * s[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) MOM[c] = MOM[c] + p.wm[it] * s[c];
  }
  const double fact = k.onep414 * p.dm;
  const double SQA = __builtin_sqrt(AREA);
  const double visc = fact * SSP * SQA * dtinv * RHO;
  FOR[0] = FOR[0] + visc * (EXX + HALF * EYY);
  FOR[1] = FOR[1] + visc * (EYY + HALF * EXX);
  FOR[2] = FOR[2] + visc * EXY * k.third;
#pragma unroll
  for (int c = 0; c < 5; ++c) FOR[c] = FOR[c] * off;
#pragma unroll
  for (int c = 0; c < 3; ++c) MOM[c] = MOM[c] * off;
  // ---- forward: CHVIS3 ----
  const double GAMA1 = off * (ONE - PX1 * VHX - PY1 * VHY);
  const double GAMA3 = off * (ONE + PX1 * VHX + PY1 * VHY);
  const double GAMA2 = off * (-ONE - PX2 * VHX - PY2 * VHY);
  const double GAMA4 = off * (-ONE + PX2 * VHX + PY2 * VHY);
  const double GAMA[4] = {GAMA1, GAMA2, GAMA3, GAMA4};
  const double SHFPR3 = SHF / (THREE * (ONE + NU));
  const double R1 = FOURTH * RHO * 100.0;
  const double A1 = R1 * (HALF * H1);
  const double SRS = __builtin_sqrt(SHFPR3);
  const double A3 = R1 * (HALF * H2) * SRS;
  const double A5 = (HALF * H3) * R1 * k.zep072169;
  const double R0E = FOURTH * YM * HALF;
  const double A7 = H1 * R0E;
  const double A8 = H2 * R0E * SHFPR3;
  const double T2A = THK02 * AREA;
  const double TSA = __builtin_sqrt(T2A);
  const double TD = thk0 * dt1;
  const double B1 = PX1 * PX1 + PY1 * PY1, B2 = PX2 * PX2 + PY2 * PY2;
  double H1Q = A1 * TSA, H2Q = A3 * THK02, H3Q = A5 * T2A, HH1 = A7 * TD;
  double HH2 = A8 * THK02 * TD / (B1 + B2);
  if (tri) { H1Q = H2Q = H3Q = HH1 = HH2 = ZERO; }
  double VX[4], VY[4], VZ[4], RX[4], RY[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    VX[c] = g.VX[c]; VY[c] = g.VY[c]; VZ[c] = g.VZ[c]; RX[c] = g.RX[c]; RY[c] = g.RY[c];
  }
  const double HG1 = VX[0] * GAMA1 + VX[1] * GAMA2 + VX[2] * GAMA3 + VX[3] * GAMA4;
  const double HG2 = VY[0] * GAMA1 + VY[1] * GAMA2 + VY[2] * GAMA3 + VY[3] * GAMA4;
  const double HG3 = VZ[0] * GAMA1 + VZ[1] * GAMA2 + VZ[2] * GAMA3 + VZ[3] * GAMA4;
  const double HG4 = RX[0] - RX[1] + RX[2] - RX[3];
  const double HG5 = RY[0] - RY[1] + RY[2] - RY[3];
  const double hour0 = st(SC_HOUR) + HG1 * HH1;
  const double hour1 = st(SC_HOUR + 1) + HG2 * HH1;
  const double hour2 = st(SC_HOUR + 2) + HG3 * HH2;
  const double HOUR1A = hour0 + HG1 * (H1Q * tabs(HG1));
  const double HOUR2A = hour1 + HG2 * (H1Q * tabs(HG2));
  const double HOUR3A = hour2 + HG3 * (H2Q * tabs(HG3));
  const double hour3 = HG4 * (H3Q * tabs(HG4));
  const double hour4 = HG5 * (H3Q * tabs(HG5));
  // ---- forward: CUPDT3P (which elements emit forces) and CFINT3's combinations ----
  {
    // time-step deletion (CDT3) only lowers OFF after the forces are formed
    if (off < ONE) offg = off;
  }
  const bool zero_out = offg < ZERO;
  const double F1A = FOR[0] * thk0, F2A = FOR[1] * thk0, F3A = FOR[2] * thk0;
  const double F4A = FOR[3] * thk0, F5A = FOR[4] * thk0;
  const double F12 = F1A * PX2 + F3A * PY2, F22 = F2A * PY2 + F3A * PX2, F32 = F5A * PX2 + F4A * PY2;
  const double F11 = F1A * PX1 + F3A * PY1, F21 = F2A * PY1 + F3A * PX1, F31 = F5A * PX1 + F4A * PY1;
  const double G1[3] = {F11 + HOUR1A * GAMA1, F12 + HOUR1A * GAMA2, HOUR1A * GAMA3 - F11};
  const double G2[3] = {F21 + HOUR2A * GAMA1, F22 + HOUR2A * GAMA2, HOUR2A * GAMA3 - F21};
  const double G3[3] = {F31 + HOUR3A * GAMA1, F32 + HOUR3A * GAMA2, HOUR3A * GAMA3 - F31};
  const double M1A = MOM[0] * THK02, M2A = MOM[1] * THK02, M3A = MOM[2] * THK02;
  const double M4 = F4A * AREA * FOURTH, M5 = F5A * AREA * FOURTH;
  const double M11 = -M2A * PY1 - M3A * PX1, M21 = M1A * PX1 + M3A * PY1;
  const double M12 = -M2A * PY2 - M3A * PX2, M22 = M1A * PX2 + M3A * PY2;
  const double Bs[4] = {hour3 * off, -hour3 * off, hour3 * off, -hour3 * off};
  const double Cs[4] = {hour4 * off, -hour4 * off, hour4 * off, -hour4 * off};
  const double Q1[4] = {M11 - M4 + Bs[0], M12 - M4 + Bs[1], -M11 - M4 + Bs[2], -M12 - M4 + Bs[3]};
  const double Q2[4] = {M21 + M5 + Cs[0], M22 + M5 + Cs[1], -M21 + M5 + Cs[2], -M22 + M5 + Cs[3]};

  // ================= reverse =================
  KinAdj L{};
  double lthk0 = 0.0, lTHK02 = 0.0, lAREA = 0.0, lSHF = 0.0;
  double lPX1 = 0.0, lPY1 = 0.0, lPX2 = 0.0, lPY2 = 0.0, lVHX = 0.0, lVHY = 0.0;
  V3 lE1 = v3(0, 0, 0), lE2 = v3(0, 0, 0), lE3 = v3(0, 0, 0);
  double lFOR[5] = {0, 0, 0, 0, 0}, lMOM[3] = {0, 0, 0};
  double lHOUR1A = 0.0, lHOUR2A = 0.0, lHOUR3A = 0.0, lhour3 = 0.0, lhour4 = 0.0;
  double lGAMA[4] = {0, 0, 0, 0};
  if (!zero_out) {
    // slots emit -F and -M; corner 4's force is minus the sum of the other three
    double lf[3][4], lm[3][4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        lf[i][c] = -lslot(c, i);
        lm[i][c] = -lslot(c, 3 + i);
      }
    double lFc[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int c = 0; c < 3; ++c) lFc[i][c] = lf[i][c] - lf[i][3];
    // Fc[i][c] = E1_i G1[c] + E2_i G2[c] + E3_i G3[c]
    double lG1[3], lG2[3], lG3[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const V3 l = v3(lFc[0][c], lFc[1][c], lFc[2][c]);
      lG1[c] = dot(E1, l); lG2[c] = dot(E2, l); lG3[c] = dot(E3, l);
      lE1 = lE1 + G1[c] * l; lE2 = lE2 + G2[c] * l; lE3 = lE3 + G3[c] * l;
    }
    // moments m_i = E1_i Q1[c] + E2_i Q2[c]
    double lQ1[4], lQ2[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const V3 l = v3(lm[0][c], lm[1][c], lm[2][c]);
      lQ1[c] = dot(E1, l); lQ2[c] = dot(E2, l);
      lE1 = lE1 + Q1[c] * l; lE2 = lE2 + Q2[c] * l;
    }
    const double lM11 = lQ1[0] - lQ1[2], lM12 = lQ1[1] - lQ1[3];
    const double lM4 = -(lQ1[0] + lQ1[1] + lQ1[2] + lQ1[3]);
    const double lM21 = lQ2[0] - lQ2[2], lM22 = lQ2[1] - lQ2[3];
    const double lM5 = lQ2[0] + lQ2[1] + lQ2[2] + lQ2[3];
    lhour3 += off * (lQ1[0] - lQ1[1] + lQ1[2] - lQ1[3]);
    lhour4 += off * (lQ2[0] - lQ2[1] + lQ2[2] - lQ2[3]);
    // M11 = -M2A PY1 - M3A PX1; M21 = M1A PX1 + M3A PY1; M12, M22 likewise with PX2, PY2
    const double lM1A = lM21 * PX1 + lM22 * PX2;
    const double lM2A = -lM11 * PY1 - lM12 * PY2;
    const double lM3A = -lM11 * PX1 + lM21 * PY1 - lM12 * PX2 + lM22 * PY2;
    lPY1 += -lM11 * M2A + lM21 * M3A;
    lPX1 += -lM11 * M3A + lM21 * M1A;
    lPY2 += -lM12 * M2A + lM22 * M3A;
    lPX2 += -lM12 * M3A + lM22 * M1A;
    lMOM[0] += lM1A * THK02; lMOM[1] += lM2A * THK02; lMOM[2] += lM3A * THK02;
    lTHK02 += lM1A * MOM[0] + lM2A * MOM[1] + lM3A * MOM[2];
    // M4 = F4A AREA / 4, M5 = F5A AREA / 4
    double lF4A = lM4 * AREA * FOURTH, lF5A = lM5 * AREA * FOURTH;
    lAREA += FOURTH * (lM4 * F4A + lM5 * F5A);
    // G = F combinations + hourglass
    const double lF11 = lG1[0] - lG1[2], lF12 = lG1[1];
    const double lF21 = lG2[0] - lG2[2], lF22 = lG2[1];
    const double lF31 = lG3[0] - lG3[2], lF32 = lG3[1];
    lHOUR1A += lG1[0] * GAMA1 + lG1[1] * GAMA2 + lG1[2] * GAMA3;
    lHOUR2A += lG2[0] * GAMA1 + lG2[1] * GAMA2 + lG2[2] * GAMA3;
    lHOUR3A += lG3[0] * GAMA1 + lG3[1] * GAMA2 + lG3[2] * GAMA3;
    lGAMA[0] += lG1[0] * HOUR1A + lG2[0] * HOUR2A + lG3[0] * HOUR3A;
    lGAMA[1] += lG1[1] * HOUR1A + lG2[1] * HOUR2A + lG3[1] * HOUR3A;
    lGAMA[2] += lG1[2] * HOUR1A + lG2[2] * HOUR2A + lG3[2] * HOUR3A;
    // F11 = F1A PX1 + F3A PY1; F21 = F2A PY1 + F3A PX1; F31 = F5A PX1 + F4A PY1; same at 2
    const double lF1A = lF11 * PX1 + lF12 * PX2;
    const double lF2A = lF21 * PY1 + lF22 * PY2;
    const double lF3A = lF11 * PY1 + lF21 * PX1 + lF12 * PY2 + lF22 * PX2;
    lF4A += lF31 * PY1 + lF32 * PY2;
    lF5A += lF31 * PX1 + lF32 * PX2;
    lPX1 += lF11 * F1A + lF21 * F3A + lF31 * F5A;
    lPY1 += lF11 * F3A + lF21 * F2A + lF31 * F4A;
    lPX2 += lF12 * F1A + lF22 * F3A + lF32 * F5A;
    lPY2 += lF12 * F3A + lF22 * F2A + lF32 * F4A;
    // F?A = FOR * thk0
    const double lFA[5] = {lF1A, lF2A, lF3A, lF4A, lF5A};
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      lFOR[c] += lFA[c] * thk0;
      lthk0 += lFA[c] * FOR[c];
    }
  }
  // ---- reverse: CHVIS3 ----
  double lHG1 = 0.0, lHG2 = 0.0, lHG3 = 0.0, lHG4 = 0.0, lHG5 = 0.0;
  double lH1Q = 0.0, lH2Q = 0.0, lH3Q = 0.0, lHH1 = 0.0, lHH2 = 0.0;
  // hour3/4 = HG (H3Q |HG|)
  lHG4 += lhour3 * 2.0 * H3Q * tabs(HG4); lH3Q += lhour3 * HG4 * tabs(HG4);
  lHG5 += lhour4 * 2.0 * H3Q * tabs(HG5); lH3Q += lhour4 * HG5 * tabs(HG5);
  // HOURnA = hour_n + HG (HnQ |HG|); hour_n = HOUR_in + HG HHn (the carried state)
  const double lhour0 = lHOUR1A + lpost(30), lhour1 = lHOUR2A + lpost(31);
  const double lhour2 = lHOUR3A + lpost(32);
  lHG1 += lHOUR1A * 2.0 * H1Q * tabs(HG1) + lhour0 * HH1;
  lHG2 += lHOUR2A * 2.0 * H1Q * tabs(HG2) + lhour1 * HH1;
  lHG3 += lHOUR3A * 2.0 * H2Q * tabs(HG3) + lhour2 * HH2;
  lH1Q += lHOUR1A * HG1 * tabs(HG1) + lHOUR2A * HG2 * tabs(HG2);
  lH2Q += lHOUR3A * HG3 * tabs(HG3);
  lHH1 += lhour0 * HG1 + lhour1 * HG2;
  lHH2 += lhour2 * HG3;
  lpre(30, lhour0);
  lpre(31, lhour1);
  lpre(32, lhour2);
  // HG4 = RX1 - RX2 + RX3 - RX4, HG5 likewise; HG1..3 = sum V? GAMA
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double sg = (c & 1) ? -1.0 : 1.0;
    L.RX[c] += sg * lHG4;
    L.RY[c] += sg * lHG5;
    L.VX[c] += lHG1 * GAMA[c];
    L.VY[c] += lHG2 * GAMA[c];
    L.VZ[c] += lHG3 * GAMA[c];
    lGAMA[c] += lHG1 * VX[c] + lHG2 * VY[c] + lHG3 * VZ[c];
  }
  // GAMA1 = off (1 - PX1 VHX - PY1 VHY), GAMA3 = off (1 + ...), GAMA2/4 with PX2, PY2
  {
    const double l13 = off * (lGAMA[2] - lGAMA[0]), l24 = off * (lGAMA[3] - lGAMA[1]);
    lPX1 += l13 * VHX; lPY1 += l13 * VHY;
    lPX2 += l24 * VHX; lPY2 += l24 * VHY;
    lVHX += l13 * PX1 + l24 * PX2;
    lVHY += l13 * PY1 + l24 * PY2;
  }
  if (!tri) {
    // H1Q = A1 sqrt(T2A); H3Q = A5 T2A; T2A = THK02 AREA
    const double lT2A = lH1Q * A1 * HALF / TSA + lH3Q * A5;
    lTHK02 += lT2A * AREA;
    lAREA += lT2A * THK02;
    // H2Q = A3 THK02, A3 ~ sqrt(SHFPR3); HH1 = A7 thk0 dt1;
    // HH2 = A8 THK02 TD / (B1 + B2), A8 ~ SHFPR3, TD = thk0 dt1
    lTHK02 += lH2Q * A3;
    double lSHFPR3 = lH2Q * THK02 * A3 * HALF / SHFPR3;
    lthk0 += lHH1 * A7 * dt1;
    const double BB = B1 + B2;
    lSHFPR3 += lHH2 * HH2 / SHFPR3;
    lTHK02 += lHH2 * A8 * TD / BB;
    lthk0 += lHH2 * A8 * THK02 * dt1 / BB;
    const double lBB = -lHH2 * HH2 / BB;
    lPX1 += lBB * TWO * PX1; lPY1 += lBB * TWO * PY1;
    lPX2 += lBB * TWO * PX2; lPY2 += lBB * TWO * PY2;
    lSHF += lSHFPR3 / (THREE * (ONE + NU));
  }
  // ---- reverse: FOR/MOM scaling and membrane damping ----
  double lFORs[5], lMOMs[3];
#pragma unroll
  for (int c = 0; c < 5; ++c) lFORs[c] = lFOR[c] * off;
#pragma unroll
  for (int c = 0; c < 3; ++c) lMOMs[c] = lMOM[c] * off;
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
  // ---- reverse: law 2 at every point ----
  double lGS = 0.0;
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
    // adjoints of the point's outputs: carried state and the stress resultants
    double ls[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) ls[c] = lpost(5 * it + c) + p.wf[it] * lFORs[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) ls[c] += p.wm[it] * lMOMs[c];
    const double lpla_out = lpost(25 + it);
    double lstr[5] = {ls[0], ls[1], ls[2], ls[3], ls[4]};
    double lpla = lpla_out, lyld = 0.0;
    if (r < ONE) {
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
    if (!capped) {
      if (pla == ZERO) {
        g_jc[0] += lyld;
      } else {
        g_jc[0] += lyld;
        g_jc[1] += lyld * ppow;
        const double lp = lyld * cb * ppow;
        g_jc[2] += lp * M<double>::log_(pla);
        lpla += lp * cn / pla;
      }
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) lpre(5 * it + c, lstr[c]);
    lpre(25 + it, lpla);
    // trial stress: s = S_in + C deps, deps = E + zt K
    const double ld0 = a11 * lstr[0] + a12 * lstr[1], ld1 = a12 * lstr[0] + a11 * lstr[1];
    const double ld2 = G * lstr[2];
    lGS += lstr[3] * EYZ + lstr[4] * EXZ;
    lEYZ += GS * lstr[3];
    lEXZ += GS * lstr[4];
    lEXX += ld0; lEYY += ld1; lEXY += ld2;
    lKXX += zt * ld0; lKYY += zt * ld1; lKXY += zt * ld2;
    lthk0 += z0 * (KXX * ld0 + KYY * ld1 + KXY * ld2);
  }
  // ---- reverse: CCOEF3 (GS = G SHF) ----
  lSHF += lGS * G;
  if (ISH != 0.0) {
    // SHF = FSH ISH FAC1TMP / (FSH AREA + FAC1TMP) + const; FAC1TMP = 2 (1 + NU) THK02
    const double t = FSH * ISH;
    const double lF = lSHF * t * (ONE / SDEN - FAC1TMP / (SDEN * SDEN));
    lAREA += -lSHF * t * FAC1TMP * FSH / (SDEN * SDEN);
    lTHK02 += lF * TWO * (ONE + NU);
  }
  lthk0 += lTHK02 * TWO * thk0;
  // ---- reverse: CSTRA3 (rates times DT1 / AREA) ----
  L.EXX = lEXX * FAC1; L.EYY = lEYY * FAC1; L.EXY = lEXY * FAC1;
  L.EYZ = lEYZ * FAC1; L.EXZ = lEXZ * FAC1;
  L.KXX = lKXX * FAC1; L.KYY = lKYY * FAC1; L.KXY = lKXY * FAC1;
  const double lFAC1 = lEXX * g.EXX + lEYY * g.EYY + lEXY * g.EXY + lEYZ * g.EYZ + lEXZ * g.EXZ +
                       lKXX * g.KXX + lKYY * g.KYY + lKXY * g.KXY;
  lAREA += -lFAC1 * FAC1 / AREA;
  // ---- reverse: kinematics ----
  L.E1 = lE1; L.E2 = lE2; L.E3 = lE3;
  L.PX1 = lPX1; L.PY1 = lPY1; L.PX2 = lPX2; L.PY2 = lPY2;
  L.AREA = lAREA; L.VHX = lVHX; L.VHY = lVHY;
  shell_kinematics_rev(xg, vg, vrg, dead, dt1, EM20, L, lx, lv, lvr);
  g_thk += lthk0;
}

}  // namespace orad

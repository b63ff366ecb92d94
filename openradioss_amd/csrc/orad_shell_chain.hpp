// orad_shell_chain.hpp -- one Belytschko-Tsay shell through the explicit force
// chain of CFORC3 (engine/source/elements/shell/coque/cforc3.F:398-746) for the
// default 4-node shell: IHBE=1, Ismstr=2, ISHFRAM=0, law 2 at NPT points
// through the thickness, viscous/elastic hourglass (CHVIS3), PARITH/ON slots.
//
// Stages and the routines they follow:
//   CCOOR3 (ccoor3.F:78-143) gather and OFF; CNVEC3 (cnvec3.F:79-128) convected
//   orthonormal frame; CDERI3 (cderi3.F:89-188) local coordinates, PX/PY, area,
//   hourglass vectors; CCOEF3 (ccoef3.F:76-160) thickness, shear factor;
//   CDLEN3 (cdlen3.F:70-121) characteristic length; CDEFO3 (cdefo3.F:79-127)
//   membrane and transverse-shear rates with the BT warping correction;
//   CCURV3 (ccurv3.F:64-93) curvature rates; CSTRA3 (cstra3.F:100-110) scaling
//   to increments; element strain rate (cforc3.F:535-546); MULAWC's law-2
//   point loop (mulawc.F90:530-3046) with SIGEPS02C/M2CPLR
//   (sigeps02c.F:90-226, m2cplr.F:95-443); CHVIS3 (chvis3.F:120-330);
//   CDT3 (cdt3.F:110-229); CFINT3 (cfint3.F:165-253); CUPDT3P (cupdt3.F:1000-1067).
//
// T = double follows the reference's double-precision operation order (the
// file instantiating it is compiled without FMA contraction).  T = float is the
// mixed form: the caller passes coordinates and translational velocities
// relative to corner 1, formed in FP64; every quantity of the chain is
// invariant to that shift (frame, rates, hourglass modes), and the rotational
// velocities are used as they are.
#pragma once
#include "orad_common.hpp"
#include "orad_math.hpp"

namespace orad {

template <typename T>
struct ShellOut {
  T sti;     // FSKY(7,·)
  T dt;      // element time step candidate
  T ehou;    // hourglass energy increment
  int alive; // the candidate enters DT2T
};

template <typename T> __device__ __forceinline__ T tsign(T a, T b) {
  return (b >= T(0)) ? (a < T(0) ? -a : a) : (a < T(0) ? a : -a);
}
template <typename T> __device__ __forceinline__ T tabs(T a) { return a < T(0) ? -a : a; }

// SIGEPS02C + M2CPLR for one through-thickness point (FISOKIN = 0, IFORM = 0,
// TSTAR = 0).  off: element OFF, lowered by the ductile-rupture test; thkn: the
// element thickness updated point by point; etse: ET/E of the point (1 when
// elastic, m2cplr.F:99,227,282,414); sigy: SIGY, to which the point adds its
// yield stress / NPT (sigeps02c.F:187).  The QEPH chain reads etse and sigy; in
// the Belytschko-Tsay chain they are unused and compile away.
template <typename T>
__device__ __forceinline__ void shell_law2_point(const ShMat &m, const T jc[3], const ShConst &k,
                                                 T gs, T dt1,
                                                 T asrate, T epsd_pg, T off_old, T &off, T &pla,
                                                 T &epsdp, T &epchk, T &thkn, T thklyl,
                                                 const T deps[5], T s[5], T &etse, T &sigy,
                                                 T fnpt) {
  const T ZERO = T(0), ONE = T(1), TWO = T(2), THREE = T(3), HALF = T(0.5);
  const T young = T(m.young), g = T(m.shear), nu = T(m.nu);
  const T a11 = M<T>::div(young, ONE - nu * nu);
  const T a12 = a11 * nu;
  const T ca0 = jc[0], cb0 = jc[1], cn = jc[2], ymax0 = T(m.ymax);
  const T cc = T(m.cc);
  const T EM20 = T(k.em20);
  const T epdr = tmax(T(m.epdr) * dt1, EM20);
  T epsdot;
  if (m.vp == 1) {
    epsdot = epsdp * dt1;
  } else if (m.vp == 2) {
    epsdp = asrate * epsd_pg + (ONE - asrate) * epsdp;
    epsdot = epsdp * dt1;
  } else {
    const T dtinv = M<T>::div(dt1, tmax(dt1 * dt1, EM20));
    const T ex = deps[0] * dtinv, ey = deps[1] * dtinv, exy = deps[2] * dtinv;
    const T dav = (ex + ey) * T(k.third);
    const T d1 = ex - dav, d2 = ey - dav, d3 = -dav, d4 = HALF * exy;
    epsdot = HALF * (d1 * d1 + d2 * d2 + d3 * d3) + d4 * d4;
    epsdot = M<T>::div(M<T>::sqrt_(THREE * epsdot), T(k.three_half));
    if (m.israte > 0) epsdot = asrate * epsdot + (ONE - asrate) * epsdp;
    epsdp = epsdot;
    epsdot = epsdot * dt1;
  }
  // ---- M2CPLR ----
  T ezz = ZERO, h = ZERO, ca = ca0, cb = cb0, ymax = ymax0, dpla = ZERO;
  etse = ONE;
  s[0] = s[0] + a11 * deps[0] + a12 * deps[1];
  s[1] = s[1] + a12 * deps[0] + a11 * deps[1];
  s[2] = s[2] + g * deps[2];
  s[3] = s[3] + gs * deps[3];
  s[4] = s[4] + gs * deps[4];
  if (cc != ZERO) {
    T epsp = epsdot;
    if (m.israte == 0 && m.vp == 2) epsp = tmax(tmax(tabs(deps[0]), tabs(deps[1])), HALF * tabs(deps[2]));
    epsp = tmax(epsp, epdr);
    const T logep = M<T>::log_(M<T>::div(epsp, epdr));
    T q = (ONE + cc * logep);
    q = tmax(q, EM20);
    ca = ca * q;
    cb = cb * q;
    if (m.icc == 1) ymax = ymax * q;
  }
  T yld;
  if (pla == ZERO) yld = ca;
  else yld = ca + cb * (ONE - ZERO) * M<T>::powe(pla, cn);
  yld = tmin(yld, ymax);
  const T SMALL = T(k.em07);
  if (m.ipla == 0) {
    const T svm = M<T>::sqrt_(s[0] * s[0] + s[1] * s[1] - s[0] * s[1] + THREE * s[2] * s[2]);
    const T r = tmin(ONE, M<T>::div(yld, svm + T(k.em15)));
    if (r < ONE) {
      s[0] = s[0] * r;
      s[1] = s[1] * r;
      s[2] = s[2] * r;
      dpla = off * tmax(ZERO, M<T>::div(svm - yld, young));
      const T s1 = HALF * (s[0] + s[1]);
      ezz = M<T>::div(dpla * s1, yld);
      pla = pla + dpla;
      epchk = tmax(pla, epchk);
      if (yld >= ymax) h = ZERO;
      else h = cn * cb * M<T>::powe(pla + SMALL, cn - ONE);
      etse = M<T>::div(h, h + young);
    }
  } else if (m.ipla == 1) {
    const T s1 = s[0] + s[1], s2 = s[0] - s[1], s3 = s[2];
    const T A = T(0.25) * s1 * s1;
    const T B = T(k.three_over_4) * s2 * s2 + THREE * s3 * s3;
    const T svm = M<T>::sqrt_(A + B);
    if (svm > yld && off == ONE) {
      const T nu1 = M<T>::div(ONE, ONE - nu), nu2 = M<T>::div(ONE, ONE + nu);
      if (yld >= ymax) h = ZERO;
      else h = cn * cb * M<T>::powe(pla + SMALL, cn - ONE);
      T dpla_j = M<T>::div(svm - yld, THREE * g + h);
      etse = M<T>::div(h, h + young);
      const T anu1 = A * nu1, bnu2 = THREE * B * nu2, h2 = TWO * h;
      T dpla_i = ZERO, dr = ZERO, p = ZERO, q = ZERO;
      for (int it = 0; it < 3; ++it) {
        dpla_i = dpla_j;
        const T pla_i = pla + dpla_i;
        dpla = dpla_j;
        T yld_i;
        if (pla_i == ZERO) yld_i = tmin(ymax, ca);
        else yld_i = tmin(ymax, ca + cb * M<T>::powe(pla_i, cn));
        dr = M<T>::div(HALF * young * dpla_i, yld_i);
        p = M<T>::div(ONE, ONE + dr * nu1);
        q = M<T>::div(ONE, ONE + THREE * dr * nu2);
        const T p2 = p * p, q2 = q * q;
        const T f = A * p2 + B * q2 - yld_i * yld_i;
        const T df = M<T>::div(-(anu1 * p2 * p + bnu2 * q2 * q) * (young - dr * h2), yld_i) - h2 * yld_i;
        if (dpla_i > ZERO) dpla_j = tmax(ZERO, dpla_i - M<T>::div(f, df));
        else dpla_j = ZERO;
      }
      pla = pla + dpla_i;
      epchk = tmax(pla, epchk);
      const T t1 = (s[0] + s[1]) * p;
      const T t2 = (s[0] - s[1]) * q;
      s[0] = HALF * (t1 + t2);
      s[1] = HALF * (t1 - t2);
      s[2] = s[2] * q;
      ezz = M<T>::div(dr * t1, young);
    }
  } else {
    const T svm2 = s[0] * s[0] + s[1] * s[1] - s[0] * s[1] + THREE * s[2] * s[2];
    T svm = M<T>::sqrt_(svm2);
    const T yld2 = yld * yld;
    if (svm2 > yld2 && off == ONE) {
      if (yld >= ymax) h = ZERO;
      else h = cn * cb * M<T>::powe(pla + SMALL, cn - ONE);
      etse = M<T>::div(h, h + young);
      const T aa = M<T>::div(svm2 - yld2, T(5) * svm2 + THREE * (-s[0] * s[1] + s[2] * s[2]));
      const T s1 = (ONE - TWO * aa) * s[0] + aa * s[1];
      const T s2 = aa * s[0] + (ONE - TWO * aa) * s[1];
      const T s3 = (ONE - THREE * aa) * s[2];
      s[0] = s1; s[1] = s2; s[2] = s3;
      dpla = M<T>::div(off * (svm - yld), THREE * g + h);
      pla = pla + dpla;
      yld = yld + h * dpla;
      svm = M<T>::sqrt_(s[0] * s[0] + s[1] * s[1] - s[0] * s[1] + THREE * s[2] * s[2]);
      const T r = tmin(ONE, M<T>::div(yld, tmax(EM20, svm)));
      s[0] = s[0] * r; s[1] = s[1] * r; s[2] = s[2] * r;
      ezz = M<T>::div(dpla * HALF * (s[0] + s[1]), yld);
    }
  }
  // ---- back in SIGEPS02C ----
  if (m.vp == 1) {
    const T epsdot_p = M<T>::div(dpla, tmax(EM20, dt1));
    epsdp = asrate * epsdot_p + (ONE - asrate) * epsdp;
  }
  sigy = sigy + M<T>::div(yld, fnpt);
  if (off == off_old && off > ZERO) {
    if (off == ONE && epchk >= T(m.epmx)) off = T(k.four_over_5);
    else if (off < ONE) off = off * T(k.four_over_5);
  }
  ezz = -(deps[0] + deps[1]) * nu - (ONE - TWO * nu) * ezz;
  ezz = M<T>::div(ezz, ONE - nu);
  thkn = thkn + ezz * thklyl * off;
}

// Local frame and kinematics of one element in precision TG: CNVEC3 (ISHFRAM = 0),
// CDERI3, CDEFO3 (IHBE <= 1) and CCURV3, up to the unscaled strain and curvature
// rates.  The mixed form runs this stage in FP64 from FP64 relative coordinates:
// an FP32 frame is off by about 1e-7 rad, which leaks that fraction of the large
// out-of-plane relative velocity of a bending element into its membrane rates,
// with a sign that persists from cycle to cycle.
template <typename TG>
struct ShellKin {
  TG E1X, E1Y, E1Z, E2X, E2Y, E2Z, E3X, E3Y, E3Z;
  TG X2, Y2, X3, Y3, X4, Y4, PX1, PY1, PX2, PY2, AREA, VHX, VHY;
  TG VX[4], VY[4], VZ[4], RX[4], RY[4];
  TG EXX, EYY, EXY, EYZ, EXZ, KXX, KYY, KXY;
};

template <typename TG>
__device__ __forceinline__ void shell_kinematics(const TG xg[3][4], const TG vg[3][4],
                                                 const TG vrg[3][4], bool dead, TG dt1, TG EM20,
                                                 ShellKin<TG> &o) {
  const TG ZERO = TG(0), TWO = TG(2), HALF = TG(0.5), FOURTH = TG(0.25);
  TG &E1X = o.E1X, &E1Y = o.E1Y, &E1Z = o.E1Z, &E2X = o.E2X, &E2Y = o.E2Y, &E2Z = o.E2Z;
  TG &E3X = o.E3X, &E3Y = o.E3Y, &E3Z = o.E3Z;
  TG &X2 = o.X2, &Y2 = o.Y2, &X3 = o.X3, &Y3 = o.Y3, &X4 = o.X4, &Y4 = o.Y4;
  TG &PX1 = o.PX1, &PY1 = o.PY1, &PX2 = o.PX2, &PY2 = o.PY2, &AREA = o.AREA;
  TG &VHX = o.VHX, &VHY = o.VHY;
  TG *VX = o.VX, *VY = o.VY, *VZ = o.VZ, *RX = o.RX, *RY = o.RY;
  TG &EXX = o.EXX, &EYY = o.EYY, &EXY = o.EXY, &EYZ = o.EYZ, &EXZ = o.EXZ;
  TG &KXX = o.KXX, &KYY = o.KYY, &KXY = o.KXY;
  const TG ONE = TG(1);
  // ---- CNVEC3, ISHFRAM = 0 ----
  const TG X21 = xg[0][1] - xg[0][0], X32 = xg[0][2] - xg[0][1], X34 = xg[0][2] - xg[0][3],
          X41 = xg[0][3] - xg[0][0];
  const TG Y21 = xg[1][1] - xg[1][0], Y32 = xg[1][2] - xg[1][1], Y34 = xg[1][2] - xg[1][3],
          Y41 = xg[1][3] - xg[1][0];
  const TG Z21 = xg[2][1] - xg[2][0], Z32 = xg[2][2] - xg[2][1], Z34 = xg[2][2] - xg[2][3],
          Z41 = xg[2][3] - xg[2][0];
  E1X = (X21 + X34); E1Y = (Y21 + Y34); E1Z = (Z21 + Z34);
  E2X = (X32 + X41); E2Y = (Y32 + Y41); E2Z = (Z32 + Z41);
  E3X = E1Y * E2Z - E1Z * E2Y;
  E3Y = E1Z * E2X - E1X * E2Z;
  E3Z = E1X * E2Y - E1Y * E2X;
  TG SUMA = E3X * E3X + E3Y * E3Y + E3Z * E3Z;
  SUMA = M<TG>::div(ONE, tmax(M<TG>::sqrt_(SUMA), EM20));
  E3X = E3X * SUMA; E3Y = E3Y * SUMA; E3Z = E3Z * SUMA;
  {
    const TG S1 = E1X * E1X + E1Y * E1Y + E1Z * E1Z;
    const TG S2 = E2X * E2X + E2Y * E2Y + E2Z * E2Z;
    SUMA = M<TG>::sqrt_(M<TG>::div(S1, S2));
  }
  E1X = E1X + (E2Y * E3Z - E2Z * E3Y) * SUMA;
  E1Y = E1Y + (E2Z * E3X - E2X * E3Z) * SUMA;
  E1Z = E1Z + (E2X * E3Y - E2Y * E3X) * SUMA;
  SUMA = E1X * E1X + E1Y * E1Y + E1Z * E1Z;
  SUMA = M<TG>::div(ONE, tmax(M<TG>::sqrt_(SUMA), EM20));
  E1X = E1X * SUMA; E1Y = E1Y * SUMA; E1Z = E1Z * SUMA;
  E2X = E3Y * E1Z - E3Z * E1Y;
  E2Y = E3Z * E1X - E3X * E1Z;
  E2Z = E3X * E1Y - E3Y * E1X;
  // ---- CDERI3 ----
  const TG X31GA = xg[0][2] - xg[0][0], Y31GA = xg[1][2] - xg[1][0], Z31GA = xg[2][2] - xg[2][0];
  X2 = E1X * X21 + E1Y * Y21 + E1Z * Z21;
  Y2 = E2X * X21 + E2Y * Y21 + E2Z * Z21;
  Y3 = E2X * X31GA + E2Y * Y31GA + E2Z * Z31GA;
  X3 = E1X * X31GA + E1Y * Y31GA + E1Z * Z31GA;
  X4 = E1X * X41 + E1Y * Y41 + E1Z * Z41;
  Y4 = E2X * X41 + E2Y * Y41 + E2Z * Z41;
  PX1 = HALF * (Y2 - Y4);
  PY1 = HALF * (X4 - X2);
  PX2 = HALF * Y3;
  PY2 = -HALF * X3;
  AREA = tmax(TWO * (PY2 * PX1 - PY1 * PX2), EM20);
  VHX = M<TG>::div(-X2 + X3 - X4, AREA);
  VHY = M<TG>::div(-Y2 + Y3 - Y4, AREA);
  // ---- CDEFO3, IHBE <= 1 ----
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const TG a = dead ? ZERO : vg[0][c], b = dead ? ZERO : vg[1][c], d = dead ? ZERO : vg[2][c];
    VX[c] = E1X * a + E1Y * b + E1Z * d;
    VY[c] = E2X * a + E2Y * b + E2Z * d;
    VZ[c] = E3X * a + E3Y * b + E3Z * d;
  }
  const TG VZ13 = VZ[0] - VZ[2], VZ24 = VZ[1] - VZ[3];
  EYZ = PY1 * VZ13 + PY2 * VZ24;
  EXZ = PX1 * VZ13 + PX2 * VZ24;
  {
    const TG DT1V4 = FOURTH * dt1;
    const TG TMP2A = PY2 + PY1;
    const TG TMP3A = tsign(tmax(tabs(TMP2A), EM20), TMP2A);
    const TG TMP1A = M<TG>::div(DT1V4 * (VZ13 - VZ24) * (VZ13 - VZ24), TMP3A);
    TG VX13 = VX[0] - VX[2], VX24 = VX[1] - VX[3];
    VX13 = VX13 - TMP1A;
    VX24 = VX24 + TMP1A;
    EXX = PX1 * VX13 + PX2 * VX24;
    EXY = PY1 * VX13 + PY2 * VX24;
    const TG TMP1B = PX2 - PX1;
    const TG TMP3B = tsign(tmax(tabs(TMP1B), EM20), TMP1B);
    const TG TMP2B = M<TG>::div(DT1V4 * (VZ13 + VZ24) * (VZ13 + VZ24), TMP3B);
    TG VY13 = VY[0] - VY[2], VY24 = VY[1] - VY[3];
    VY13 = VY13 + TMP2B;
    VY24 = VY24 + TMP2B;
    EXY = EXY + PX1 * VY13 + PX2 * VY24;
    EYY = PY1 * VY13 + PY2 * VY24;
  }
  // ---- CCURV3 ----
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const TG a = dead ? ZERO : vrg[0][c], b = dead ? ZERO : vrg[1][c], d = dead ? ZERO : vrg[2][c];
    RX[c] = E1X * a + E1Y * b + E1Z * d;
    RY[c] = E2X * a + E2Y * b + E2Z * d;
  }
  {
    const TG RX13TA = RX[0] - RX[2];
    const TG RXAVTA = RX[0] + RX[1] + RX[2] + RX[3];
    const TG RX24TA = RX[1] - RX[3];
    KYY = -PY1 * RX13TA - PY2 * RX24TA;
    KXY = PX1 * RX13TA + PX2 * RX24TA;
    const TG RY13TA = RY[0] - RY[2];
    const TG RYAVTA = RY[0] + RY[1] + RY[2] + RY[3];
    const TG RY24TA = RY[1] - RY[3];
    KXX = PX1 * RY13TA + PX2 * RY24TA;
    KXY = PY1 * RY13TA + PY2 * RY24TA - KXY;
    EXZ = EXZ + RYAVTA * (FOURTH * AREA);
    EYZ = EYZ - RXAVTA * (FOURTH * AREA);
  }
  (void)VHX; (void)VHY; (void)ZERO; (void)TWO;
}

// IO: get(col) / put(col, v) on this element's state columns.
// emit(k, fx, fy, fz, mx, my, mz): slot values (FSKY(1:6), already negated) of corner k.
template <typename T, typename TG, typename IO, typename Emit>
__device__ __forceinline__ void shell_chain(const TG xg[3][4], const TG vg[3][4],
                                            const TG vrg[3][4], T thk0, bool tri, const IO &io,
                                            double dt1d, const ShMat &m, const T jc[3],
                                            const ShProp &p, const ShConst &k,
                                            ShellOut<T> &out, Emit emit) {
  const T ZERO = T(0), ONE = T(1), TWO = T(2), THREE = T(3), HALF = T(0.5), FOURTH = T(0.25);
  const T EM20 = T(k.em20);
  const T dt1 = T(dt1d);
  // ---- CCOOR3 ----
  T offg = io.get(SC_OFFG);
  T off = tmin(ONE, tabs(offg));
  const bool dead = offg < ZERO;
  ShellKin<TG> g;
  shell_kinematics<TG>(xg, vg, vrg, dead, TG(dt1d), TG(k.em20), g);
  const T E1X = T(g.E1X), E1Y = T(g.E1Y), E1Z = T(g.E1Z);
  const T E2X = T(g.E2X), E2Y = T(g.E2Y), E2Z = T(g.E2Z);
  const T E3X = T(g.E3X), E3Y = T(g.E3Y), E3Z = T(g.E3Z);
  const T X2 = T(g.X2), Y2 = T(g.Y2), X3 = T(g.X3), Y3 = T(g.Y3), X4 = T(g.X4), Y4 = T(g.Y4);
  const T PX1 = T(g.PX1), PY1 = T(g.PY1), PX2 = T(g.PX2), PY2 = T(g.PY2);
  const T AREA = T(g.AREA), VHX = T(g.VHX), VHY = T(g.VHY);
  T VX[4], VY[4], VZ[4], RX[4], RY[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    VX[c] = T(g.VX[c]); VY[c] = T(g.VY[c]); VZ[c] = T(g.VZ[c]);
    RX[c] = T(g.RX[c]); RY[c] = T(g.RY[c]);
  }
  T EXX = T(g.EXX), EYY = T(g.EYY), EXY = T(g.EXY), EYZ = T(g.EYZ), EXZ = T(g.EXZ);
  T KXX = T(g.KXX), KYY = T(g.KYY), KXY = T(g.KXY);
  // ---- CCOEF3, ITHK = 0 ----
  const T VOL0 = thk0 * AREA;
  const T THK02 = thk0 * thk0;
  const T RHO = T(m.rho0), YM = T(m.young), NU = T(m.nu), G = T(m.shear), SSP = T(m.ssp);
  const T H1 = T(p.hm), H2 = T(p.hf), H3 = T(p.hr);
  T SHF;
  {
    const T FAC1TMP = TWO * (ONE + NU) * THK02;
    const T ISH = T(p.ish), FSH = T(p.fsh);
    SHF = FSH * ((ONE - ISH) + M<T>::div(ISH * FAC1TMP, FSH * AREA + FAC1TMP));
  }
  const T GS = G * SHF;
  // ---- CDLEN3, IHBE = 1 ----
  T ALDT;
  {
    const T AL1 = X2 * X2 + Y2 * Y2;
    const T AL2 = (X3 - X2) * (X3 - X2) + (Y3 - Y2) * (Y3 - Y2);
    const T AL6 = X3 * X3 + Y3 * Y3;
    const T AL3 = (X4 - X3) * (X4 - X3) + (Y4 - Y3) * (Y4 - Y3);
    const T AL4 = X4 * X4 + Y4 * Y4;
    const T AL5 = (X4 - X2) * (X4 - X2) + (Y4 - Y2) * (Y4 - Y2);
    T ALMIN = tmin(tmin(AL1, AL2), AL4);
    const T ALQUAD = tmin(tmin(AL3, AL5), AL6);
    if (AL3 != ZERO) ALMIN = tmin(ALMIN, ALQUAD);
    const T DTDYN = M<T>::div(AREA * AREA, tmax(tmax(AL5, AL6), EM20));
    ALDT = tmax(DTDYN, ALMIN);
    const T DTHOUR = M<T>::div(HALF * (ALMIN + ALDT), tmax(H1, H2));
    if (DTHOUR < ALDT) ALDT = DTHOUR;
    ALDT = M<T>::sqrt_(ALDT);
  }
  // ---- CSTRA3 ----
  {
    const T FAC1 = M<T>::div(dt1, AREA);
    EXX = EXX * FAC1; EYY = EYY * FAC1; EXY = EXY * FAC1;
    EYZ = EYZ * FAC1; EXZ = EXZ * FAC1;
    KXX = KXX * FAC1; KYY = KYY * FAC1; KXY = KXY * FAC1;
  }
  // ---- element strain rate ----
  const T dtinv = M<T>::div(dt1, tmax(dt1 * dt1, EM20));
  T thk = io.get(SC_THK);
  T epsd_pg;
  {
    const T eps_k2 = (KXX * KXX + KYY * KYY + KXX * KYY + FOURTH * (KXY * KXY)) *
                     T(k.one_over_9) * (thk * thk);
    const T eps_m2 = T(k.four_over_3) * (EXX * EXX + EYY * EYY + EXX * EYY + FOURTH * (EXY * EXY));
    epsd_pg = M<T>::sqrt_(eps_k2 + eps_m2) * dtinv;
    io.put(SC_EPSD, ONE * epsd_pg + (ONE - ONE) * io.get(SC_EPSD));
  }
  // ---- MULAWC, law 2 ----
  T FOR[5], MOM[3];
#pragma unroll
  for (int c = 0; c < 5; ++c) FOR[c] = io.get(SC_FOR + c);
#pragma unroll
  for (int c = 0; c < 3; ++c) MOM[c] = io.get(SC_MOM + c);
  T degmb = FOR[0] * EXX + FOR[1] * EYY + FOR[2] * EXY + FOR[3] * EYZ + FOR[4] * EXZ;
  T degfx = MOM[0] * KXX + MOM[1] * KYY + MOM[2] * KXY;
  const T vol0m = AREA * thk0;
  T thkn = thk;
#pragma unroll
  for (int c = 0; c < 5; ++c) FOR[c] = ZERO;
#pragma unroll
  for (int c = 0; c < 3; ++c) MOM[c] = ZERO;
  const T off_old = off;
  T epchk = ZERO;
  T viscmx = ZERO;
  const T asrate = (m.israte > 0) ? tmin(ONE, T(m.asrate_pm9) * dt1) : ONE;
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
    T etse_unused, sigy_unused = ZERO;
    shell_law2_point<T>(m, jc, k, GS, dt1, asrate, epsd_pg, off_old, off, pla, epsdp, epchk, thkn,
                        thklyl, deps, s, etse_unused, sigy_unused, ONE);
    viscmx = tmax(T(p.dm), viscmx);
#pragma unroll
    for (int c = 0; c < 5; ++c) io.put(SC_SIG + 5 * it + c, s[c] * ONE);
    io.put(SC_PLA + it, pla);
    io.put(SC_EPSDP + it, epsdp);
#pragma unroll
    for (int c = 0; c < 5; ++c) FOR[c] = FOR[c] + wf * s[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) MOM[c] = MOM[c] + wm * s[c];
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
  {
    const T vol2 = HALF * vol0m;
    io.put(SC_EINT, io.get(SC_EINT) + degmb * vol2);
    io.put(SC_EINT + 1, io.get(SC_EINT + 1) + degfx * thk0 * vol2);
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) io.put(SC_FOR + c, FOR[c]);
#pragma unroll
  for (int c = 0; c < 3; ++c) io.put(SC_MOM + c, MOM[c]);
  viscmx = M<T>::sqrt_(ONE + viscmx * viscmx) - viscmx;
  // ---- CHVIS3, ISMSTR = 2, IHBE = 1 ----
  T GAMA1, GAMA2, GAMA3, GAMA4;
  {
    const T PX1V = PX1 * VHX, PX2V = PX2 * VHX, PY1V = PY1 * VHY, PY2V = PY2 * VHY;
    GAMA1 = off * (ONE - PX1V - PY1V);
    GAMA3 = off * (ONE + PX1V + PY1V);
    GAMA2 = off * (-ONE - PX2V - PY2V);
    GAMA4 = off * (-ONE + PX2V + PY2V);
  }
  T H1Q, H1L, H2Q, H2L, H3Q, H3L, HH1, HH2;
  {
    const T HVISC = HALF, HELAS = HALF, HVLIN = ZERO;
    const T SRH = ZERO;                     // GEO(18..20), never set for TYPE1
    const T SR2D2 = T(k.sr2d2), ZEP072169 = T(k.zep072169);
    const T SHFPR3 = M<T>::div(SHF, THREE * (ONE + NU));
    const T HVISH1 = HVISC * H1, HVISH2 = HVISC * H2;
    T R0 = FOURTH * RHO;
    const T R1 = R0 * T(100);
    R0 = R0 * HVLIN;
    const T A1 = R1 * HVISH1;
    const T A2 = R0 * SR2D2 * SRH;
    const T SRSHFPR3 = M<T>::sqrt_(SHFPR3);
    const T A3 = R1 * HVISH2 * SRSHFPR3;
    const T A4 = R0 * SR2D2 * SRH * SRSHFPR3;
    T HH3 = HELAS * H3;
    const T A5 = HH3 * R1 * ZEP072169;
    HH3 = SR2D2 * SRH;
    const T A6 = HH3 * R0 * ZEP072169;
    R0 = FOURTH * YM * HELAS;
    const T A7 = H1 * R0;
    const T A8 = H2 * R0 * SHFPR3;
    const T T2A = THK02 * AREA;
    const T TSA = M<T>::sqrt_(T2A);
    H1Q = A1 * TSA; H1L = A2 * SSP * TSA;
    H2Q = A3 * THK02; H2L = A4 * SSP * THK02;
    H3Q = A5 * T2A; H3L = A6 * SSP * T2A;
    const T TD = thk0 * dt1;
    HH1 = A7 * TD;
    const T B1 = PX1 * PX1 + PY1 * PY1;
    const T B2 = PX2 * PX2 + PY2 * PY2;
    HH2 = M<T>::div(A8 * THK02 * TD, B1 + B2);
    if (tri) { H1Q = H1L = H2Q = H2L = H3Q = H3L = HH1 = HH2 = ZERO; }
  }
  T hour[5];
#pragma unroll
  for (int c = 0; c < 3; ++c) hour[c] = io.get(SC_HOUR + c);
  T HG1 = VX[0] * GAMA1 + VX[1] * GAMA2 + VX[2] * GAMA3 + VX[3] * GAMA4;
  T HG2 = VY[0] * GAMA1 + VY[1] * GAMA2 + VY[2] * GAMA3 + VY[3] * GAMA4;
  hour[0] = hour[0] + HG1 * HH1;
  hour[1] = hour[1] + HG2 * HH1;
  const T HOUR1A = hour[0] + HG1 * (H1L + H1Q * tabs(HG1));
  const T H11 = HOUR1A * GAMA1, H12 = HOUR1A * GAMA2, H13 = HOUR1A * GAMA3;
  const T HOUR2A = hour[1] + HG2 * (H1L + H1Q * tabs(HG2));
  const T H21 = HOUR2A * GAMA1, H22 = HOUR2A * GAMA2, H23 = HOUR2A * GAMA3;
  T EHOU = HOUR1A * HG1 + HOUR2A * HG2;
  HG1 = VZ[0] * GAMA1 + VZ[1] * GAMA2 + VZ[2] * GAMA3 + VZ[3] * GAMA4;
  hour[2] = hour[2] + HG1 * HH2;
  const T HOUR3A = hour[2] + HG1 * (H2L + H2Q * tabs(HG1));
  const T H31 = HOUR3A * GAMA1, H32 = HOUR3A * GAMA2, H33 = HOUR3A * GAMA3;
  EHOU = EHOU + HOUR3A * HG1;
  HG1 = RX[0] - RX[1] + RX[2] - RX[3];
  HG2 = RY[0] - RY[1] + RY[2] - RY[3];
  hour[3] = HG1 * (H3L + H3Q * tabs(HG1));
  hour[4] = HG2 * (H3L + H3Q * tabs(HG2));
  EHOU = EHOU + hour[3] * HG1 + hour[4] * HG2;
  EHOU = dt1 * EHOU * off;
#pragma unroll
  for (int c = 0; c < 5; ++c) io.put(SC_HOUR + c, hour[c]);
  const T B11 = hour[3] * off, B12 = -hour[3] * off, B13 = hour[3] * off, B14 = -hour[3] * off;
  const T B21 = hour[4] * off, B22 = -hour[4] * off, B23 = hour[4] * off, B24 = -hour[4] * off;
  // ---- CDT3, IDTMIN(3) = 2 ----
  ALDT = M<T>::div(ALDT * viscmx, M<T>::sqrt_(ONE));
  const T DT = M<T>::div(T(p.dtfac) * ALDT, SSP);
  if (!(DT > T(p.dtmin) || off < ONE || offg < ZERO)) off = ZERO;
  out.dt = DT;
  out.alive = (offg > ZERO && off != ZERO);
  T sti = M<T>::div(HALF * VOL0 * YM, tmax(ALDT * ALDT, EM20));
  sti = T(k.zep81) * sti * off;
  out.ehou = EHOU;
  // ---- CUPDT3P: OFFG, deleted elements write zero ----
  if (off < ONE) offg = off;
  if (off < ONE) io.put(SC_OFFG, offg);
  const bool zero_out = offg < ZERO;
  out.sti = zero_out ? ZERO : sti;
  // ---- CFINT3, ISIGI = 0: forces then moments, one corner at a time ----
  const T F1A = FOR[0] * thk0, F2A = FOR[1] * thk0, F3A = FOR[2] * thk0;
  const T F4A = FOR[3] * thk0, F5A = FOR[4] * thk0;
  T M4 = F4A * AREA, M5 = F5A * AREA;
  const T F12 = F1A * PX2 + F3A * PY2;
  const T F22 = F2A * PY2 + F3A * PX2;
  const T F32 = F5A * PX2 + F4A * PY2;
  const T F11 = F1A * PX1 + F3A * PY1;
  const T F21 = F2A * PY1 + F3A * PX1;
  const T F31 = F5A * PX1 + F4A * PY1;
  const T G11 = F11 + H11, G13 = H13 - F11, G21 = F21 + H21;
  const T G23 = H23 - F21, G31 = F31 + H31, G33 = H33 - F31;
  const T G12 = F12 + H12, G22 = F22 + H22, G32 = F32 + H32;
  T Fc[3][3];
  Fc[0][0] = E1X * G11 + E2X * G21 + E3X * G31;
  Fc[0][1] = E1X * G12 + E2X * G22 + E3X * G32;
  Fc[0][2] = E1X * G13 + E2X * G23 + E3X * G33;
  Fc[1][0] = E1Y * G11 + E2Y * G21 + E3Y * G31;
  Fc[1][1] = E1Y * G12 + E2Y * G22 + E3Y * G32;
  Fc[1][2] = E1Y * G13 + E2Y * G23 + E3Y * G33;
  Fc[2][0] = E1Z * G11 + E2Z * G21 + E3Z * G31;
  Fc[2][1] = E1Z * G12 + E2Z * G22 + E3Z * G32;
  Fc[2][2] = E1Z * G13 + E2Z * G23 + E3Z * G33;
  const T M1A = MOM[0] * THK02, M2A = MOM[1] * THK02, M3A = MOM[2] * THK02;
  M4 = M4 * FOURTH;
  M5 = M5 * FOURTH;
  const T M11 = -M2A * PY1 - M3A * PX1;
  const T M21 = M1A * PX1 + M3A * PY1;
  const T M12 = -M2A * PY2 - M3A * PX2;
  const T M22 = M1A * PX2 + M3A * PY2;
  const T Q1[4] = {M11 - M4 + B11, M12 - M4 + B12, -M11 - M4 + B13, -M12 - M4 + B14};
  const T Q2[4] = {M21 + M5 + B21, M22 + M5 + B22, -M21 + M5 + B23, -M22 + M5 + B24};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    T fx, fy, fz;
    if (c < 3) {
      fx = Fc[0][c]; fy = Fc[1][c]; fz = Fc[2][c];
    } else {
      fx = -Fc[0][0] - Fc[0][1] - Fc[0][2];
      fy = -Fc[1][0] - Fc[1][1] - Fc[1][2];
      fz = -Fc[2][0] - Fc[2][1] - Fc[2][2];
    }
    const T mx = E1X * Q1[c] + E2X * Q2[c];
    const T my = E1Y * Q1[c] + E2Y * Q2[c];
    const T mz = E1Z * Q1[c] + E2Z * Q2[c];
    if (zero_out) emit(c, ZERO, ZERO, ZERO, ZERO, ZERO, ZERO);
    else emit(c, -fx, -fy, -fz, -mx, -my, -mz);
  }
}

}  // namespace orad

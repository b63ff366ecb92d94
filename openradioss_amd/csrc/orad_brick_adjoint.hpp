// orad_brick_adjoint.hpp -- hand-written reverse (adjoint) pass of one 8-node brick
// through the SFORC3 chain of orad_brick_chain.hpp (FP64 form).
//
// brick_vjp evaluates the chain once forward (values only, every branch taken on the
// values as the chain takes it) and then once backward: from the adjoints of the
// eight corner slots (global forces) and of the carried state that reaches the forces
// (SIG(6), PLA), it forms the adjoints of the corner coordinates and velocities, of
// the incoming SIG and PLA, and of the hardening curve's A, B, n.  It is the exact
// derivative of the arithmetic the forward chain performs for law 2 without the
// strain-rate term (C = 0: the rate factor EPD is 1).  EINT, RHO and QVIS reach only
// the energy and are not adjoint state (RHO is recomputed from the volume); the
// element time step, the characteristic length and STI carry no derivative.
// Stages, in reverse order: SRROTA3 and SCUMU3P (sforc3.F:1604-1614,
// scumu3p.F:105-308), SFINT3 (sfint3.F:257-322), SHVIS3 (shvis3.F:163-428), MQVISCB
// (mqviscb.F:139-263), M2LAW (m2law.F:162-426), SRHO3 / MMAIN (mmain.F90:678-737),
// SDEFO3 (sdefo3.F:114-194), SDERI3 (sderi3.F:104-335), SRCOOR3 with SREPISO3 and
// SORTHO3 (srcoor3.F:415-678, srepiso3.F:60-92, sortho3.F:70-160).
#pragma once
#include "orad_brick_chain.hpp"

namespace orad {

namespace badj {

struct V3 { double v[3]; };
__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void cross3(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
// adjoint of c = a x b: la += b x lc, lb += lc x a
__device__ __forceinline__ void cross3_rev(const double a[3], const double b[3], const double lc[3],
                                           double la[3], double lb[3]) {
  double t[3];
  cross3(b, lc, t);
  la[0] += t[0]; la[1] += t[1]; la[2] += t[2];
  cross3(lc, a, t);
  lb[0] += t[0]; lb[1] += t[1]; lb[2] += t[2];
}
// u = w * rnorm(|w|^2), rnorm(q) = 1/sqrt(q) or 0 at 0 (sortho3.F): adjoint of w
__device__ __forceinline__ void unit_rev(const double w[3], const double u[3], const double lu[3],
                                         double lw[3]) {
  const double n = __builtin_sqrt(dot3(w, w));
  if (n == 0.0) return;
  const double p = dot3(u, lu);
  for (int i = 0; i < 3; ++i) lw[i] += (lu[i] - p * u[i]) / n;
}

// SORTHO3's vectors after n iterations (n = 0: the normalised r, s, t), and with
// want_e the unnormalised vectors of iteration n (n < 3) that produce stage n + 1.
__device__ inline void sortho_stage(const double r0[3], const double s0[3], const double t0[3],
                                    int n, double U[3], double V[3], double W[3], bool want_e,
                                    double E1[3], double E2[3], double E3[3]) {
  double aa = M<double>::rnorm(dot3(r0, r0));
  for (int i = 0; i < 3; ++i) U[i] = r0[i] * aa;
  aa = M<double>::rnorm(dot3(s0, s0));
  for (int i = 0; i < 3; ++i) V[i] = s0[i] * aa;
  aa = M<double>::rnorm(dot3(t0, t0));
  for (int i = 0; i < 3; ++i) W[i] = t0[i] * aa;
  const int last = want_e ? n + 1 : n;
  for (int it = 0; it < last; ++it) {
    double c[3], a1[3], a2[3], a3[3];
    cross3(V, W, c);
    for (int i = 0; i < 3; ++i) a1[i] = c[i] + U[i];
    cross3(W, U, c);
    for (int i = 0; i < 3; ++i) a2[i] = c[i] + V[i];
    cross3(U, V, c);
    for (int i = 0; i < 3; ++i) a3[i] = c[i] + W[i];
    if (want_e && it == n) {
      for (int i = 0; i < 3; ++i) { E1[i] = a1[i]; E2[i] = a2[i]; E3[i] = a3[i]; }
      return;
    }
    double bb = M<double>::rnorm(dot3(a1, a1));
    for (int i = 0; i < 3; ++i) U[i] = a1[i] * bb;
    bb = M<double>::rnorm(dot3(a2, a2));
    for (int i = 0; i < 3; ++i) V[i] = a2[i] * bb;
    bb = M<double>::rnorm(dot3(a3, a3));
    for (int i = 0; i < 3; ++i) W[i] = a3[i] * bb;
  }
}

}  // namespace badj

// corner(k, xk, vk) loads corner k's coordinates and velocity (global frame, FP64);
// it is called in the forward pass and again in the reverse, so that they need not be
// held in registers in between.
// st_eint, st_rho, st_qvis, st_offg, st_epsd: the incoming values of the state
// without an adjoint.  lpost(j): adjoint of the outgoing SIG(j) (j < 6) and PLA (j = 6);
// lslot(k, i): adjoint of corner k's slot component i; lpre(j, v) is called once for
// j = 0..6 with the adjoint of the incoming SIG(j) / PLA; lcorner(k, lx, lv) once per
// corner with the adjoints of its coordinates and velocity; g_jc is added to.
template <typename CORNER, typename LPOST, typename LSLOT, typename LPRE, typename LCORNER>
__device__ inline void brick_vjp(CORNER corner, const double sig_in[6],
                                 double st_pla, double st_rho, double st_offg, double vol0d,
                                 double dt1, const Law2 &mat, const double jc[3],
                                 const SProp &prop, LPOST lpost, LSLOT lslot, LPRE lpre,
                                 LCORNER lcorner, double g_jc[3]) {
  using namespace badj;
  const double ZERO = 0.0, ONE = 1.0, TWO = 2.0, THREE = 3.0, FOUR = 4.0, HALF = 0.5;
  const double THIRD = K<double>::THIRD, ONEP333 = K<double>::ONEP333;
  const double EM15 = K<double>::EM15, ONE_OVER_64 = K<double>::ONE_OVER_64;
  (void)st_rho;
  double offg = st_offg;
  // ================= forward =================
  double r0[3], s0[3], t0[3];
  double R[3][3], e1[3], e3u[3], e3[3], e2[3], U3[3], V3v[3];
  double xl[3][8], vl[3][8];
  double off = (offg < ZERO) ? -offg : offg;
  const bool vdead = offg < ZERO;
  {
  double x[3][8], v[3][8];
  for (int k = 0; k < 8; ++k) {
    double xk[3], vk[3];
    corner(k, xk, vk);
    for (int d = 0; d < 3; ++d) { x[d][k] = xk[d]; v[d][k] = vk[d]; }
  }
  // ---- SREPISO3 ----
  for (int d = 0; d < 3; ++d) {
    const double *q = x[d];
    const double A17 = (q[6] - q[0]) + (q[5] - q[3]), A28 = (q[7] - q[1]) + (q[4] - q[2]);
    r0[d] = (q[6] - q[0]) + (q[7] - q[1]) - (q[4] - q[2]) - (q[5] - q[3]);
    s0[d] = A17 + A28;
    t0[d] = A17 - A28;
  }
  // ---- SORTHO3 (its iterations are recomputed in the reverse, not kept) ----
  {
    double W3[3], d0[3], d1[3], d2[3];
    sortho_stage(r0, s0, t0, 3, U3, V3v, W3, false, d0, d1, d2);
  }
  for (int i = 0; i < 3; ++i) e1[i] = U3[i];
  cross3(e1, V3v, e3u);
  {
    const double aa = M<double>::rnorm(dot3(e3u, e3u));
    for (int i = 0; i < 3; ++i) e3[i] = e3u[i] * aa;
  }
  cross3(e3, e1, e2);
  for (int i = 0; i < 3; ++i) { R[i][0] = e1[i]; R[i][1] = e2[i]; R[i][2] = e3[i]; }
  // ---- local coordinates and velocities ----
  for (int k = 0; k < 8; ++k)
    for (int j = 0; j < 3; ++j) {
      xl[j][k] = R[0][j] * x[0][k] + R[1][j] * x[1][k] + R[2][j] * x[2][k];
      vl[j][k] = vdead ? ZERO : R[0][j] * v[0][k] + R[1][j] * v[1][k] + R[2][j] * v[2][k];
    }
  }
  // ---- SDERI3 ----
  double J[9];
  {
    for (int d = 0; d < 3; ++d) {
      const double *q = xl[d];
      const double X17 = q[6] - q[0], X28 = q[7] - q[1], X35 = q[4] - q[2], X46 = q[5] - q[3];
      J[d] = X17 + X28 - X35 - X46;
      J[3 + d] = (X17 + X46) + (X28 + X35);
      J[6 + d] = (X17 + X46) - (X28 + X35);
    }
  }
  // cofactor-based inverse: JACI(a,b) = DETT * C[a][b]
  double C[9];
  C[0] = J[4] * J[8] - J[5] * J[7];
  C[1] = J[2] * J[7] - J[1] * J[8];
  C[2] = J[1] * J[5] - J[2] * J[4];
  C[3] = J[5] * J[6] - J[3] * J[8];
  C[4] = J[0] * J[8] - J[2] * J[6];
  C[5] = J[2] * J[3] - J[0] * J[5];
  C[6] = J[3] * J[7] - J[4] * J[6];
  C[7] = J[1] * J[6] - J[0] * J[7];
  C[8] = J[0] * J[4] - J[1] * J[3];
  double det = ONE_OVER_64 * (J[0] * C[0] + J[1] * C[3] + J[2] * C[6]);
  bool det_free = true;        // det follows the geometry (not replaced by 1)
  if (off == ZERO) {
    det = ONE; det_free = false;
  } else if (offg > ONE) {
  } else if (det <= ZERO) {
    det = ONE; off = ZERO; offg = ZERO; det_free = false;
  }
  double voln = det;
  const double DETT = ONE_OVER_64 / det;
  double JI[9];
  for (int q = 0; q < 9; ++q) JI[q] = DETT * C[q];
  double PX[4], PY[4], PZ[4];
  {
    const double *a = JI;
    PX[0] = -(a[0] + a[1]) - a[2]; PX[1] = -(a[0] + a[1]) + a[2];
    PX[2] = (a[0] - a[1]) + a[2];  PX[3] = (a[0] - a[1]) - a[2];
    PY[0] = -(a[3] + a[4]) - a[5]; PY[1] = -(a[3] + a[4]) + a[5];
    PY[2] = (a[3] - a[4]) + a[5];  PY[3] = (a[3] - a[4]) - a[5];
    PZ[0] = -(a[6] + a[7]) - a[8]; PZ[1] = -(a[6] + a[7]) + a[8];
    PZ[2] = (a[6] - a[7]) + a[8];  PZ[3] = (a[6] - a[7]) - a[8];
  }
  double Hm[3][3];   // Hm[m][d] = sum_k gamma_m[k] xl[d][k]
  double PH[3][4];
  for (int m = 0; m < 3; ++m) {
    for (int d = 0; d < 3; ++d) {
      double h = 0.0;
      for (int k = 0; k < 8; ++k) h += gamma_sign(m, k) * xl[d][k];
      Hm[m][d] = h;
    }
    for (int a = 0; a < 4; ++a) PH[m][a] = PX[a] * Hm[m][0] + PY[a] * Hm[m][1] + PZ[a] * Hm[m][2];
  }
  // ---- SDEFO3 ----
  const int pa[4] = {0, 1, 2, 3}, pb[4] = {6, 7, 4, 5};   // pairs 1-7, 2-8, 3-5, 4-6
  double Vd[3][4];   // Vd[d][a] = vl[d][pa] - vl[d][pb]
  for (int d = 0; d < 3; ++d)
    for (int a = 0; a < 4; ++a) Vd[d][a] = vl[d][pa[a]] - vl[d][pb[a]];
  const double *Pd[3] = {PX, PY, PZ};
  double Dg[3][3];   // Dg[i][j] = sum_a P_j[a] Vd[i][a]  (D_ij = dv_i/dx_j)
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      Dg[i][j] = Pd[j][0] * Vd[i][0] + Pd[j][1] * Vd[i][1] + Pd[j][2] * Vd[i][2] + Pd[j][3] * Vd[i][3];
  const double DT1D2 = HALF * dt1;
  // corrected rates: Dii -= DT1D2 sum_k D_ki^2; off-diagonal pair (i,j): both -= DT1D2 sum_k D_ki D_kj
  double Dc[3][3];
  for (int i = 0; i < 3; ++i) {
    double sq = 0.0;
    for (int k = 0; k < 3; ++k) sq += Dg[k][i] * Dg[k][i];
    Dc[i][i] = Dg[i][i] - DT1D2 * sq;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j) {
      double aaa = 0.0;
      for (int k = 0; k < 3; ++k) aaa += Dg[k][i] * Dg[k][j];
      Dc[i][j] = Dg[i][j] - DT1D2 * aaa;
      Dc[j][i] = Dg[j][i] - DT1D2 * aaa;
    }
  const double DXX = Dc[0][0], DYY = Dc[1][1], DZZ = Dc[2][2];
  const double D4 = Dc[0][1] + Dc[1][0], D5 = Dc[1][2] + Dc[2][1], D6 = Dc[0][2] + Dc[2][0];
  // ---- SRHO3 / MMAIN ----
  const double rho0 = mat.rho0, vol0 = vol0d;
  bool voln_free = det_free;
  if (offg == ZERO && voln == ONE) { voln = vol0; voln_free = false; }
  const double rho = rho0 * (vol0 / voln);
  const double amu = rho / rho0 - ONE;
  // ---- M2LAW ----
  const double G = mat.shear * off, bulk = mat.bulk;
  const double P = -THIRD * (sig_in[0] + sig_in[1] + sig_in[2]);
  const double DAV = -THIRD * (DXX + DYY + DZZ);
  const double G1 = dt1 * G, G2 = TWO * G1;
  const double SSP = M<double>::sqrt_((ONEP333 * G + bulk) / rho0);
  double str[6];
  str[0] = sig_in[0] + P + G2 * (DXX + DAV);
  str[1] = sig_in[1] + P + G2 * (DYY + DAV);
  str[2] = sig_in[2] + P + G2 * (DZZ + DAV);
  str[3] = sig_in[3] + G1 * D4;
  str[4] = sig_in[4] + G1 * D5;
  str[5] = sig_in[5] + G1 * D6;
  const double AJ2q = THREE * (HALF * (str[0] * str[0] + str[1] * str[1] + str[2] * str[2]) +
                               str[3] * str[3] + str[4] * str[4] + str[5] * str[5]);
  const double AJ2 = M<double>::sqrt_(AJ2q);
  const double EPD = ONE;
  const double SIGMX = (mat.icc == 1) ? mat.sigmx * EPD : mat.sigmx;
  const double CA = jc[0], CB = jc[1], CN = jc[2];
  const double EPXE = st_pla;
  double AK, QH, ppow = 0.0, qpow = 0.0;
  int hform;   // 0: CN = 1, 1: CN > 1 power, 2: CN < 1 power, 3: EPXE = 0
  if (CN == ONE) {
    AK = CA + CB * EPXE; QH = CB * EPD; hform = 0;
  } else if (EPXE > ZERO) {
    ppow = M<double>::pow_(EPXE, CN);
    AK = CA + CB * ppow;
    if (CN > ONE) { qpow = M<double>::pow_(EPXE, CN - ONE); QH = (CB * CN * qpow) * EPD; hform = 1; }
    else { qpow = M<double>::pow_(EPXE, ONE - CN); QH = (CB * CN) / qpow * EPD; hform = 2; }
  } else {
    AK = CA; QH = ZERO; hform = 3;
  }
  AK = AK * EPD;
  bool ak_free = true, qh_free = true;
  if (SIGMX < AK) { AK = SIGMX; QH = ZERO; ak_free = false; qh_free = false; }
  if (EPXE > mat.epmx) { AK = ZERO; QH = ZERO; ak_free = false; qh_free = false; }
  const bool aj_free = AJ2 > EM15;
  const double AJ2m = aj_free ? AJ2 : EM15;
  const double q1 = AK / AJ2m;
  const bool s1_free = !(ONE < q1);
  const double SCALE1 = s1_free ? q1 : ONE;
  const double den = THREE * G + QH;
  const bool den_free = den > EM15;
  const double denm = den_free ? den : EM15;
  const double DPLA = ((ONE - SCALE1) * AJ2) / denm;
  const double AK2 = AK + ONE * DPLA * QH;
  const double q2 = AK2 / AJ2m;
  const bool s2_free = !(ONE < q2);
  const double SCALE2 = s2_free ? q2 : ONE;
  const double PNEW = bulk * amu;
  double sgo[6];   // outgoing stresses
  for (int i = 0; i < 3; ++i) sgo[i] = (SCALE2 * str[i] - PNEW) * off;
  for (int i = 3; i < 6; ++i) sgo[i] = SCALE2 * str[i] * off;
  // ---- MQVISCB ----
  const double DD = -DXX - DYY - DZZ;
  double AD = ZERO, AL = ZERO;
  bool al_free = false, ad_free = false;
  if (off == ONE) {
    if (voln > ZERO) { AL = M<double>::cbrt_pow(voln); al_free = true; }
    AD = tmax(ZERO, DD);
    ad_free = !(ZERO > DD);
  }
  const double QA = prop.qa, QB = prop.qb;
  const double QAA = QA * QA * AD;
  const double qnew = rho * AD * AL * (QAA * AL + QB * SSP);
  // ---- SHVIS3 ----
  const double CAQ = 0.25 * off * prop.hvis;
  const double V23 = M<double>::two_third_pow(voln);
  const double FCL0 = CAQ * rho * V23;
  const double FCQ = FCL0 * CAQ * 100.0;
  const double FCL = FCL0 * SSP;
  double HG[3][4], HH[3][4];
  for (int d = 0; d < 3; ++d) {
    for (int m = 0; m < 3; ++m) {
      double h = 0.0;
      for (int k = 0; k < 8; ++k) h += G_of<double>(m, k, PH) * vl[d][k];
      HG[d][m] = h;
    }
    HG[d][3] = vl[d][0] - vl[d][1] + vl[d][2] - vl[d][3] - vl[d][4] + vl[d][5] - vl[d][6] + vl[d][7];
    for (int m = 0; m < 4; ++m) HH[d][m] = HG[d][m] * (FCL + fabs(HG[d][m]) * FCQ);
  }
  // ---- SFINT3 ----
  const double Sd[6] = {(sgo[0] - qnew) * voln, (sgo[1] - qnew) * voln, (sgo[2] - qnew) * voln,
                        sgo[3] * voln, sgo[4] * voln, sgo[5] * voln};
  // FI[0][a] = S1 PX + S4 PY + S6 PZ; FI[1][a] = S2 PY + S4 PX + S5 PZ; FI[2][a] = S3 PZ + S6 PX + S5 PY
  const bool zero_f = offg < ZERO;
  const int s4[8] = {-1, 1, -1, 1, 1, -1, 1, -1};
  const int ak[8] = {0, 1, 2, 3, 2, 3, 0, 1};
  // local corner force Fl (SHVIS3 hourglass force -/+ SFINT3 pair term)
  auto fl_of = [&](int d, int k) {
    const int a = ak[k];
    double FI;
    if (d == 0) FI = Sd[0] * PX[a] + Sd[3] * PY[a] + Sd[5] * PZ[a];
    else if (d == 1) FI = Sd[1] * PY[a] + Sd[3] * PX[a] + Sd[4] * PZ[a];
    else FI = Sd[2] * PZ[a] + Sd[5] * PX[a] + Sd[4] * PY[a];
    const double hg = -G_of<double>(0, k, PH) * HH[d][0] - G_of<double>(1, k, PH) * HH[d][1] -
                      G_of<double>(2, k, PH) * HH[d][2] + s4[k] * HH[d][3];
    return (k < 4) ? hg - FI : hg + FI;
  };

  // ================= reverse =================
  double lR[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double lHH[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
  double lPH[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
  double lPX[4] = {0, 0, 0, 0}, lPY[4] = {0, 0, 0, 0}, lPZ[4] = {0, 0, 0, 0};
  double lS[6] = {0, 0, 0, 0, 0, 0};
  if (!zero_f) {
    for (int k = 0; k < 8; ++k) {
      double lF[3] = {lslot(k, 0), lslot(k, 1), lslot(k, 2)};
      double lFl[3];
      for (int j = 0; j < 3; ++j) {
        lFl[j] = R[0][j] * lF[0] + R[1][j] * lF[1] + R[2][j] * lF[2];
        const double flj = fl_of(j, k);
        for (int i = 0; i < 3; ++i) lR[i][j] += lF[i] * flj;
      }
      const int a = ak[k];
      const double sg = (k < 4) ? -1.0 : 1.0;   // Fl = hg -/+ FI
      for (int d = 0; d < 3; ++d) {
        const double l = lFl[d];
        // hg = -sum_m G(m,k) HH[d][m] + s4 HH[d][3]
        for (int m = 0; m < 3; ++m) {
          lHH[d][m] -= G_of<double>(m, k, PH) * l;
          // G(m,k) = gamma -/+ PH[m][a]
          lPH[m][a] += ((k < 4) ? 1.0 : -1.0) * HH[d][m] * l;
        }
        lHH[d][3] += s4[k] * l;
        const double lFI = sg * l;
        if (d == 0) {
          lS[0] += lFI * PX[a]; lS[3] += lFI * PY[a]; lS[5] += lFI * PZ[a];
          lPX[a] += lFI * Sd[0]; lPY[a] += lFI * Sd[3]; lPZ[a] += lFI * Sd[5];
        } else if (d == 1) {
          lS[1] += lFI * PY[a]; lS[3] += lFI * PX[a]; lS[4] += lFI * PZ[a];
          lPY[a] += lFI * Sd[1]; lPX[a] += lFI * Sd[3]; lPZ[a] += lFI * Sd[4];
        } else {
          lS[2] += lFI * PZ[a]; lS[5] += lFI * PX[a]; lS[4] += lFI * PY[a];
          lPZ[a] += lFI * Sd[2]; lPX[a] += lFI * Sd[5]; lPY[a] += lFI * Sd[4];
        }
      }
    }
  }
  // ---- SFINT3: S_i = (sgo_i - qnew) voln (i < 3), sgo_i voln ----
  double lsgo[6], lqnew = 0.0, lvoln = 0.0;
  for (int i = 0; i < 6; ++i) {
    lsgo[i] = lS[i] * voln + lpost(i);
    lvoln += lS[i] * ((i < 3) ? (sgo[i] - qnew) : sgo[i]);
    if (i < 3) lqnew -= lS[i] * voln;
  }
  // ---- SHVIS3 ----
  // The local velocities' adjoint is kept compact until the corners are emitted:
  // lhg[d][m] = adjoint of HG[d][m] (HG = sum_k G(m,k) vl[d][k], G(3,k) = g4[k]) and
  // lVd[d][a] = adjoint of Vd[d][a] = vl[d][pa] - vl[d][pb] (SDEFO3, below).
  double lhg[3][4], lVd[3][4];
  for (int d = 0; d < 3; ++d)
    for (int a = 0; a < 4; ++a) lVd[d][a] = 0.0;
  double lFCL = 0.0, lFCQ = 0.0;
  for (int d = 0; d < 3; ++d) {
    for (int m = 0; m < 4; ++m) {
      const double h = HG[d][m], ah = fabs(h);
      const double l = lHH[d][m];
      const double lh = l * (FCL + 2.0 * ah * FCQ);
      lFCL += l * h;
      lFCQ += l * h * ah;
      lhg[d][m] = lh;
      // G(m,k) = gamma -/+ PH[m][a(k)]: corners a and pb[a] share PH[m][a] with signs
      // -, +, so HG picks up -PH[m][a] Vd[d][a]
      if (m < 3)
        for (int a = 0; a < 4; ++a) lPH[m][a] -= Vd[d][a] * lh;
    }
  }
  auto lvl_of = [&](int d, int k) {
    const int g4[8] = {1, -1, 1, -1, -1, 1, -1, 1};
    double l = G_of<double>(0, k, PH) * lhg[d][0] + G_of<double>(1, k, PH) * lhg[d][1] +
               G_of<double>(2, k, PH) * lhg[d][2] + g4[k] * lhg[d][3];
    return (k < 4) ? l + lVd[d][k] : l - lVd[d][ak[k]];
  };
  // FCL = CAQ rho V23 SSP, FCQ = CAQ rho V23 CAQ 100
  const double lFCL0 = lFCL * SSP + lFCQ * CAQ * 100.0;
  double lrho = lFCL0 * CAQ * V23;
  const double lV23 = lFCL0 * CAQ * rho;
  if (voln > 0.0) lvoln += lV23 * (2.0 / 3.0) * V23 / voln;
  // ---- MQVISCB: qnew = rho AD AL (QA^2 AD AL + QB SSP) ----
  double lDdiag_q = 0.0;   // adjoint of each corrected diagonal rate through AD
  {
    const double lq = lqnew, Bq = QAA * AL + QB * SSP;
    lrho += lq * AD * AL * Bq;
    const double lAD = lq * rho * AL * Bq + lq * rho * AD * AL * (QA * QA * AL);
    const double lAL = lq * rho * AD * Bq + lq * rho * AD * AL * QAA;
    if (al_free) lvoln += lAL * AL / (3.0 * voln);
    if (ad_free) lDdiag_q = -lAD;   // AD = DD = -(DXX + DYY + DZZ)
  }
  // ---- M2LAW epilogue and radial return ----
  double lstr[6] = {0, 0, 0, 0, 0, 0};
  double lSCALE2 = 0.0, lPNEW = 0.0;
  for (int i = 0; i < 6; ++i) {
    const double l = lsgo[i] * off;
    lstr[i] += SCALE2 * l;
    lSCALE2 += str[i] * l;
    if (i < 3) lPNEW -= l;
  }
  double lAJ2 = 0.0, lAK2 = 0.0;
  if (s2_free) {
    lAK2 += lSCALE2 / AJ2m;
    if (aj_free) lAJ2 -= lSCALE2 * q2 / AJ2m;
  }
  // AK2 = AK + DPLA QH
  double lAK = lAK2, lDPLA = lAK2 * QH, lQH = lAK2 * DPLA;
  // EPXE_out = EPXE + DPLA
  lDPLA += lpost(6);
  double lEPXE = lpost(6);
  // DPLA = (1 - SCALE1) AJ2 / denm
  double lSCALE1 = -lDPLA * AJ2 / denm;
  lAJ2 += lDPLA * (ONE - SCALE1) / denm;
  if (den_free) lQH -= lDPLA * DPLA / denm;
  if (s1_free) {
    lAK += lSCALE1 / AJ2m;
    if (aj_free) lAJ2 -= lSCALE1 * q1 / AJ2m;
  }
  // hardening (EPD = 1)
  if (!ak_free) lAK = 0.0;
  if (!qh_free) lQH = 0.0;
  switch (hform) {
    case 0:   // AK = CA + CB EPXE, QH = CB
      g_jc[0] += lAK; g_jc[1] += lAK * EPXE + lQH; lEPXE += lAK * CB;
      break;
    case 1:   // AK = CA + CB p^CN; QH = CB CN p^(CN-1)
    case 2: { // QH = CB CN / p^(1-CN) = CB CN p^(CN-1)
      const double lp = M<double>::log_(EPXE);
      const double qh = CB * CN * ((hform == 1) ? qpow : 1.0 / qpow);   // CB CN p^(CN-1)
      g_jc[0] += lAK;
      g_jc[1] += lAK * ppow;
      g_jc[2] += lAK * CB * ppow * lp;
      lEPXE += lAK * CB * CN * ppow / EPXE;
      // QH = CB CN p^(CN-1)
      const double pw = (hform == 1) ? qpow : 1.0 / qpow;               // p^(CN-1)
      g_jc[1] += lQH * CN * pw;
      g_jc[2] += lQH * (CB * pw + qh * lp);
      lEPXE += lQH * qh * (CN - ONE) / EPXE;
      break;
    }
    default:  // AK = CA, QH = 0
      g_jc[0] += lAK;
      break;
  }
  // AJ2 = sqrt(3 (1/2 sum s_i^2 (i<3) + sum s_i^2 (i>=3)))
  if (AJ2 > 0.0) {
    const double lq = lAJ2 * HALF / AJ2 * THREE;
    for (int i = 0; i < 3; ++i) lstr[i] += lq * str[i];
    for (int i = 3; i < 6; ++i) lstr[i] += lq * 2.0 * str[i];
  }
  // trial stress
  double lsig_in[6];
  double lD[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};   // adjoints of Dc
  {
    const double lP = lstr[0] + lstr[1] + lstr[2];
    double lDAV = 0.0;
    for (int i = 0; i < 3; ++i) {
      lsig_in[i] = lstr[i] - THIRD * lP;
      lD[i][i] += G2 * lstr[i];
      lDAV += G2 * lstr[i];
    }
    for (int i = 0; i < 3; ++i) lD[i][i] += -THIRD * lDAV;
    for (int i = 3; i < 6; ++i) lsig_in[i] = lstr[i];
    // D4 = Dc01 + Dc10, D5 = Dc12 + Dc21, D6 = Dc02 + Dc20
    lD[0][1] += G1 * lstr[3]; lD[1][0] += G1 * lstr[3];
    lD[1][2] += G1 * lstr[4]; lD[2][1] += G1 * lstr[4];
    lD[0][2] += G1 * lstr[5]; lD[2][0] += G1 * lstr[5];
  }
  for (int i = 0; i < 3; ++i) lD[i][i] += lDdiag_q;
  for (int j = 0; j < 6; ++j) lpre(j, lsig_in[j]);
  lpre(6, lEPXE);
  // PNEW = bulk amu, amu = rho / rho0 - 1, rho = rho0 vol0 / voln
  lrho += lPNEW * bulk / rho0;
  if (voln_free) lvoln -= lrho * rho / voln;
  // ---- SDEFO3 ----
  double lDg[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < 3; ++i) {
    lDg[i][i] += lD[i][i];
    for (int k = 0; k < 3; ++k) lDg[k][i] -= lD[i][i] * DT1D2 * 2.0 * Dg[k][i];
  }
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j) {
      lDg[i][j] += lD[i][j];
      lDg[j][i] += lD[j][i];
      const double laaa = -DT1D2 * (lD[i][j] + lD[j][i]);
      for (int k = 0; k < 3; ++k) {
        lDg[k][i] += laaa * Dg[k][j];
        lDg[k][j] += laaa * Dg[k][i];
      }
    }
  double *lPd[3] = {lPX, lPY, lPZ};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int a = 0; a < 4; ++a) {
        lPd[j][a] += lDg[i][j] * Vd[i][a];
        const double lvd = lDg[i][j] * Pd[j][a];
        lVd[i][a] += lvd;
      }
  // ---- SDERI3 ----
  double lxl[3][8];
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < 8; ++k) lxl[d][k] = 0.0;
  for (int m = 0; m < 3; ++m) {
    double lHm[3] = {0, 0, 0};
    for (int a = 0; a < 4; ++a) {
      const double l = lPH[m][a];
      lPX[a] += l * Hm[m][0]; lPY[a] += l * Hm[m][1]; lPZ[a] += l * Hm[m][2];
      lHm[0] += l * PX[a]; lHm[1] += l * PY[a]; lHm[2] += l * PZ[a];
    }
    for (int d = 0; d < 3; ++d)
      for (int k = 0; k < 8; ++k) lxl[d][k] += gamma_sign(m, k) * lHm[d];
  }
  double lJI[9];
  lJI[0] = -lPX[0] - lPX[1] + lPX[2] + lPX[3];
  lJI[1] = -lPX[0] - lPX[1] - lPX[2] - lPX[3];
  lJI[2] = -lPX[0] + lPX[1] + lPX[2] - lPX[3];
  lJI[3] = -lPY[0] - lPY[1] + lPY[2] + lPY[3];
  lJI[4] = -lPY[0] - lPY[1] - lPY[2] - lPY[3];
  lJI[5] = -lPY[0] + lPY[1] + lPY[2] - lPY[3];
  lJI[6] = -lPZ[0] - lPZ[1] + lPZ[2] + lPZ[3];
  lJI[7] = -lPZ[0] - lPZ[1] - lPZ[2] - lPZ[3];
  lJI[8] = -lPZ[0] + lPZ[1] + lPZ[2] - lPZ[3];
  double lC[9], lDETT = 0.0;
  for (int q = 0; q < 9; ++q) {
    lC[q] = DETT * lJI[q];
    lDETT += lJI[q] * C[q];
  }
  double ldet = det_free ? -lDETT * DETT / det : 0.0;
  if (det_free && voln_free) ldet += lvoln;
  // det = (J0 C0 + J1 C3 + J2 C6) / 64
  double lJ[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (det_free) {
    const double l = ldet * ONE_OVER_64;
    lJ[0] += l * C[0]; lJ[1] += l * C[3]; lJ[2] += l * C[6];
    lC[0] += l * J[0]; lC[3] += l * J[1]; lC[6] += l * J[2];
  }
  // cofactors
  lJ[4] += lC[0] * J[8]; lJ[8] += lC[0] * J[4]; lJ[5] -= lC[0] * J[7]; lJ[7] -= lC[0] * J[5];
  lJ[2] += lC[1] * J[7]; lJ[7] += lC[1] * J[2]; lJ[1] -= lC[1] * J[8]; lJ[8] -= lC[1] * J[1];
  lJ[1] += lC[2] * J[5]; lJ[5] += lC[2] * J[1]; lJ[2] -= lC[2] * J[4]; lJ[4] -= lC[2] * J[2];
  lJ[5] += lC[3] * J[6]; lJ[6] += lC[3] * J[5]; lJ[3] -= lC[3] * J[8]; lJ[8] -= lC[3] * J[3];
  lJ[0] += lC[4] * J[8]; lJ[8] += lC[4] * J[0]; lJ[2] -= lC[4] * J[6]; lJ[6] -= lC[4] * J[2];
  lJ[2] += lC[5] * J[3]; lJ[3] += lC[5] * J[2]; lJ[0] -= lC[5] * J[5]; lJ[5] -= lC[5] * J[0];
  lJ[3] += lC[6] * J[7]; lJ[7] += lC[6] * J[3]; lJ[4] -= lC[6] * J[6]; lJ[6] -= lC[6] * J[4];
  lJ[1] += lC[7] * J[6]; lJ[6] += lC[7] * J[1]; lJ[0] -= lC[7] * J[7]; lJ[7] -= lC[7] * J[0];
  lJ[0] += lC[8] * J[4]; lJ[4] += lC[8] * J[0]; lJ[1] -= lC[8] * J[3]; lJ[3] -= lC[8] * J[1];
  // J[d] = X17 + X28 - X35 - X46; J[3+d] = (X17+X46) + (X28+X35); J[6+d] = (X17+X46) - (X28+X35)
  for (int d = 0; d < 3; ++d) {
    const double l17 = lJ[d] + lJ[3 + d] + lJ[6 + d];
    const double l46 = -lJ[d] + lJ[3 + d] + lJ[6 + d];
    const double l28 = lJ[d] + lJ[3 + d] - lJ[6 + d];
    const double l35 = -lJ[d] + lJ[3 + d] - lJ[6 + d];
    lxl[d][6] += l17; lxl[d][0] -= l17;
    lxl[d][7] += l28; lxl[d][1] -= l28;
    lxl[d][4] += l35; lxl[d][2] -= l35;
    lxl[d][5] += l46; lxl[d][3] -= l46;
  }
  // ---- local frame: xl[j][k] = sum_i R[i][j] x[i][k], vl likewise: the frame's adjoint ----
  for (int k = 0; k < 8; ++k) {
    double xk[3], vk[3];
    corner(k, xk, vk);
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) {
        lR[i][j] += lxl[j][k] * xk[i];
        if (!vdead) lR[i][j] += lvl_of(j, k) * vk[i];
      }
  }
  // ---- SORTHO3 reverse ----
  double le1[3], le2[3], le3[3];
  for (int i = 0; i < 3; ++i) { le1[i] = lR[i][0]; le2[i] = lR[i][1]; le3[i] = lR[i][2]; }
  // e2 = e3 x e1
  cross3_rev(e3, e1, le2, le3, le1);
  // e3 = e3u rnorm, e3u = e1 x V3
  double le3u[3] = {0, 0, 0};
  unit_rev(e3u, e3, le3, le3u);
  double lU[3] = {le1[0], le1[1], le1[2]}, lV[3] = {0, 0, 0}, lW[3] = {0, 0, 0};
  cross3_rev(e1, V3v, le3u, lU, lV);
  for (int n = 2; n >= 0; --n) {
    // stage n vectors and the unnormalised E of iteration n, recomputed
    double Un[3], Vn[3], Wn[3], E1n[3], E2n[3], E3n[3];
    sortho_stage(r0, s0, t0, n, Un, Vn, Wn, true, E1n, E2n, E3n);
    double U1[3], V1[3], W1[3];
    {
      double bb = M<double>::rnorm(dot3(E1n, E1n));
      for (int i = 0; i < 3; ++i) U1[i] = E1n[i] * bb;
      bb = M<double>::rnorm(dot3(E2n, E2n));
      for (int i = 0; i < 3; ++i) V1[i] = E2n[i] * bb;
      bb = M<double>::rnorm(dot3(E3n, E3n));
      for (int i = 0; i < 3; ++i) W1[i] = E3n[i] * bb;
    }
    double lE1[3] = {0, 0, 0}, lE2[3] = {0, 0, 0}, lE3[3] = {0, 0, 0};
    unit_rev(E1n, U1, lU, lE1);
    unit_rev(E2n, V1, lV, lE2);
    unit_rev(E3n, W1, lW, lE3);
    double nU[3] = {lE1[0], lE1[1], lE1[2]}, nV[3] = {lE2[0], lE2[1], lE2[2]};
    double nW[3] = {lE3[0], lE3[1], lE3[2]};
    cross3_rev(Vn, Wn, lE1, nV, nW);   // E1 = V x W + U
    cross3_rev(Wn, Un, lE2, nW, nU);   // E2 = W x U + V
    cross3_rev(Un, Vn, lE3, nU, nV);   // E3 = U x V + W
    for (int i = 0; i < 3; ++i) { lU[i] = nU[i]; lV[i] = nV[i]; lW[i] = nW[i]; }
  }
  double lr[3] = {0, 0, 0}, ls[3] = {0, 0, 0}, lt[3] = {0, 0, 0};
  {
    double U0[3], V0[3], W0[3], d0[3], d1[3], d2[3];
    sortho_stage(r0, s0, t0, 0, U0, V0, W0, false, d0, d1, d2);
    unit_rev(r0, U0, lU, lr);
    unit_rev(s0, V0, lV, ls);
    unit_rev(t0, W0, lW, lt);
  }
  // ---- corner adjoints: the local frame's x = R xl, v = R vl, plus SREPISO3's
  // r = 17 + 28 - 35 - 46; s = 17 + 46 + 28 + 35; t = 17 + 46 - 28 - 35 (pairs 1-7,
  // 2-8, 3-5, 4-6 with the higher corner positive) ----
  double l17[3], l28[3], l35[3], l46[3];
  for (int d = 0; d < 3; ++d) {
    l17[d] = lr[d] + ls[d] + lt[d];
    l28[d] = lr[d] + ls[d] - lt[d];
    l35[d] = -lr[d] + ls[d] - lt[d];
    l46[d] = -lr[d] + ls[d] + lt[d];
  }
  for (int k = 0; k < 8; ++k) {
    double lxk[3], lvk[3];
    for (int i = 0; i < 3; ++i) {
      lxk[i] = R[i][0] * lxl[0][k] + R[i][1] * lxl[1][k] + R[i][2] * lxl[2][k];
      lvk[i] = vdead ? 0.0 : R[i][0] * lvl_of(0, k) + R[i][1] * lvl_of(1, k) + R[i][2] * lvl_of(2, k);
      double sr = 0.0;
      switch (k) {
        case 0: sr = -l17[i]; break;
        case 6: sr = l17[i]; break;
        case 1: sr = -l28[i]; break;
        case 7: sr = l28[i]; break;
        case 2: sr = -l35[i]; break;
        case 4: sr = l35[i]; break;
        case 3: sr = -l46[i]; break;
        default: sr = l46[i]; break;   // corner 5
      }
      lxk[i] += sr;
    }
    lcorner(k, lxk, lvk);
  }
}

}  // namespace orad

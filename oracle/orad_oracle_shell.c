/*
 * orad_oracle_shell.c -- CPU restatement (FP64) of the OpenRadioss explicit
 * Belytschko-Tsay shell path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP shell kernels and the CPU
 * baseline timed by bench.py.  Never linked into openradioss_amd/.
 *
 * Default 4-node shell of /PROP/SHELL (TYPE1) with law 2 in the explicit
 * engine: Ishell=1 -> IHBE=1 (starter/source/modules/defaults_mod.F90:164,
 * hm_read_prop01.F:305-310), Ismstr=2 (:232), Iplas default -> IPLA=0 radial
 * return (cgrtails.F:823-829; IPLA=1 Newton and IPLA=2 are kept too),
 * Ithick=0 (:832), ISHFRAM=0, NPT through-thickness points with the COQINI
 * tables, /DT/SHELL absent -> IDTMIN(3)=2 (engine/source/input/lectur.F:
 * 430-434), membrane damping DM=0 for law 2 (set_elgroup_param.F:81-103),
 * HVISC=HELAS=1/2, HVLIN=0 (engine/source/engine/radioss2.F:634-636), GEO(18..20)
 * never set for TYPE1 shells (SRH1..3 = 0), ISIGI=0, NODADT=0, IDT1SH=0.
 * Operation order follows the double-precision build (my_real = double) so
 * the result can be compared bit for bit with the reference's own routines
 * compiled from source (oracle/ref/harness_shell.F).
 *
 * Parity pins (tests/test_oracle_shell.py): every stage against oracle/_ref
 * on random groups; closed forms (rigid translation and rotation give zero
 * strain rates and forces; uniform membrane stretch gives uniform stress);
 * the constants against the reference's constant_mod values.
 */
#include "orad_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ZERO 0.0
#define ONE 1.0
#define TWO 2.0
#define THREE 3.0
#define HALF 0.5
#define FOURTH 0.25
#define HUNDRED 100.0
#define TEN 10.0

/* common_source/modules/constant_mod.F, re-derived from their definitions
 * (each a double operation on exact or previously rounded values). */
typedef struct {
  double EM01, EM02, EM03, EM04, EM05, EM06, EM07, EM15, EM20, EM30, EP20;
  double ZEP1, ZEP01, ZEP03, ZEP07, ZEP09, ZEP19, ZEP81, ZEP4;
  double TWOEM3, FOUREM3, FOUREM5, SIXEM5, NINEEM6, ZEP072169;
  double ONEP4, ONEP41, ONEP414;
  double ONE_OVER_9, ONE_OVER_12, FOUR_OVER_3, FOUR_OVER_5, FIVE_OVER_6, THIRD;
  double THREE_OVER_4, THREE_HALF, SR2D2;
} shconst_t;

static shconst_t K;
static int k_ready = 0;

static void init_constants(void) {
  if (k_ready) return;
  volatile double ten = TEN;   /* keep each step a run-time IEEE operation */
  double EP02 = HUNDRED, EP03 = EP02 * ten, EP04 = EP03 * ten, EP05 = EP04 * ten;
  double EP06 = EP05 * ten, EP07 = EP06 * ten, EP08 = EP07 * ten, EP09 = EP08 * ten;
  double EP10 = EP09 * ten, EP11 = EP10 * ten, EP12 = EP11 * ten, EP13 = EP12 * ten;
  double EP14 = EP13 * ten, EP15 = EP14 * ten, EP16 = EP15 * ten, EP17 = EP16 * ten;
  double EP18 = EP17 * ten, EP19 = EP18 * ten, EP20 = EP19 * ten;
  double EP30 = EP20 * EP10;
  K.EP20 = EP20;
  K.EM01 = ONE / ten;
  K.EM02 = ONE / EP02;
  K.EM03 = ONE / EP03;
  K.EM04 = ONE / EP04;
  K.EM05 = ONE / EP05;
  K.EM06 = ONE / EP06;
  K.EM07 = ONE / EP07;
  K.EM15 = ONE / EP15;
  K.EM20 = ONE / EP20;
  K.EM30 = ONE / EP30;
  K.ZEP1 = K.EM01;
  K.ZEP01 = K.EM02;
  K.ZEP03 = THREE / EP02;
  K.ZEP07 = K.ZEP1 - K.ZEP03;
  K.ZEP09 = K.ZEP1 - K.ZEP01;
  K.ZEP19 = K.ZEP1 + K.ZEP09;
  K.ZEP81 = ONE - K.ZEP19;
  K.ZEP4 = 4.0 / ten;
  K.TWOEM3 = TWO / EP03;
  K.FOUREM3 = 4.0 / EP03;
  K.FOUREM5 = 4.0 / EP05;
  K.SIXEM5 = K.EM04 - K.FOUREM5;
  K.NINEEM6 = K.EM05 - K.EM06;
  K.ZEP072169 = K.ZEP07 + K.TWOEM3 + K.EM04 + K.SIXEM5 + K.NINEEM6;
  K.ONEP4 = ONE + K.ZEP4;
  K.ONEP41 = K.ONEP4 + K.ZEP01;
  K.ONEP414 = K.ONEP41 + K.FOUREM3;
  K.ONE_OVER_9 = ONE / 9.0;
  K.ONE_OVER_12 = ONE / 12.0;
  K.FOUR_OVER_3 = 4.0 / THREE;
  K.FOUR_OVER_5 = 4.0 / 5.0;
  K.FIVE_OVER_6 = 5.0 / 6.0;
  K.THIRD = ONE / THREE;
  K.THREE_OVER_4 = THREE / 4.0;
  K.THREE_HALF = THREE / TWO;
  K.SR2D2 = sqrt(TWO) * HALF;               /* chvis3.F:118 */
  k_ready = 1;
}

void oracle_shell_constants(double *out) {
  init_constants();
  out[0] = K.ZEP81;       out[1] = K.ZEP072169;  out[2] = K.ONEP414;    out[3] = K.FIVE_OVER_6;
  out[4] = K.FOUR_OVER_3; out[5] = K.ONE_OVER_9; out[6] = K.ONE_OVER_12; out[7] = K.FOUR_OVER_5;
  out[8] = K.EM30;        out[9] = K.SR2D2;      out[10] = K.EM15;       out[11] = K.EM20;
  out[12] = K.EM07;       out[13] = K.EP20;      out[14] = K.THREE_OVER_4; out[15] = K.THIRD;
}

/* Through-thickness tables of COQINI / COQINI_WM (engine/source/elements/shell/
 * coqini.F:45-121) for NPT = 1..5: positions Z0, force weights WF, moment
 * weights WM.  The DATA statements are default-REAL literals, so the
 * double-precision build holds their single-precision roundings. */
static const float Z0T[5][5] = {
  {0.f, 0.f, 0.f, 0.f, 0.f},
  {-.5f, 0.5f, 0.f, 0.f, 0.f},
  {-.5f, 0.f, 0.5f, 0.f, 0.f},
  {-.5f, -.1666667f, 0.1666667f, 0.5f, 0.f},
  {-.5f, -.25f, 0.f, 0.25f, 0.5f}};
static const float WFT[5][5] = {
  {1.f, 0.f, 0.f, 0.f, 0.f},
  {0.5f, 0.5f, 0.f, 0.f, 0.f},
  {0.25f, 0.5f, 0.25f, 0.f, 0.f},
  {0.1666667f, 0.3333333f, 0.3333333f, 0.1666667f, 0.f},
  {0.125f, 0.25f, 0.25f, 0.25f, 0.125f}};
static const float WMT[5][5] = {
  {0.f, 0.f, 0.f, 0.f, 0.f},
  {-.0833333f, 0.0833333f, 0.f, 0.f, 0.f},
  {-.0833333f, 0.f, 0.0833333f, 0.f, 0.f},
  {-.0648148f, -.0555556f, 0.0555556f, 0.0648148f, 0.f},
  {-.0520833f, -.0625f, 0.f, 0.0625f, 0.0520833f}};

void oracle_shell_tables(int npt, double *z0, double *wf, double *wm) {
  for (int i = 0; i < npt; ++i) {
    z0[i] = (double)Z0T[npt - 1][i];
    wf[i] = (double)WFT[npt - 1][i];
    wm[i] = (double)WMT[npt - 1][i];
  }
}

static inline double dmin(double a, double b) { return (a < b) ? a : b; }
static inline double dmax(double a, double b) { return (a > b) ? a : b; }
static inline double dsign(double a, double b) { return (b >= 0.0) ? fabs(a) : -fabs(a); }

/* SIGEPS02C (engine/source/materials/mat/mat002/sigeps02c.F:90-226) with M2CPLR
 * (m2cplr.F:95-443) for one integration point: isotropic hardening
 * (FISOKIN = 0), Johnson-Cook (IFORM = 0), no temperature (TSTAR = 0).
 * off: element OFF (may be lowered by the ductile-rupture test); thkn: the
 * element thickness being updated point by point. */
static void sigeps02c(const orad_shmat_t *m, double gs, double dt1, double asrate,
                      double epsd_pg, double off_old, double *off, double *pla,
                      double *epsdp, double *epchk, double *sigy, double *thkn,
                      double thklyl, int npt, const double deps[5], const double sigo[5],
                      double sign[5]) {
  const double young = m->young, g = m->shear, nu = m->nu;
  const double a11 = young / (ONE - nu * nu);
  const double a12 = a11 * nu;
  const double ca0 = m->ca, cb0 = m->cb, cn = m->cn, epmx = m->epmx, ymax0 = m->ymax;
  const double cc = m->cc;
  const double epdr = dmax(m->epdr * dt1, K.EM20);
  double epsdot = 0.0;
  /* strain rate for the rate term (sigeps02c.F:131-163) */
  if (m->vp == 1) {
    epsdot = (*epsdp) * dt1;
  } else if (m->vp == 2) {
    *epsdp = asrate * epsd_pg + (ONE - asrate) * (*epsdp);
    epsdot = (*epsdp) * dt1;
  } else {
    const double dtinv = dt1 / dmax(dt1 * dt1, K.EM20);
    double ex = deps[0] * dtinv, ey = deps[1] * dtinv, exy = deps[2] * dtinv;
    double dav = (ex + ey) * K.THIRD;
    double d1 = ex - dav, d2 = ey - dav, d3 = -dav, d4 = HALF * exy;
    epsdot = HALF * (d1 * d1 + d2 * d2 + d3 * d3) + d4 * d4;
    epsdot = sqrt(THREE * epsdot) / K.THREE_HALF;
    if (m->israte > 0) epsdot = asrate * epsdot + (ONE - asrate) * (*epsdp);
    *epsdp = epsdot;
    epsdot = epsdot * dt1;
  }
  /* ---- M2CPLR ---- */
  double ezz = ZERO, etse = ONE, h = ZERO;
  double ca = ca0, cb = cb0, ymax = ymax0;
  double sxx = sigo[0], syy = sigo[1], sxy = sigo[2], syz = sigo[3], szx = sigo[4];
  sxx = sxx + a11 * deps[0] + a12 * deps[1];
  syy = syy + a12 * deps[0] + a11 * deps[1];
  sxy = sxy + g * deps[2];
  syz = syz + gs * deps[3];
  szx = szx + gs * deps[4];
  if (cc != ZERO) {
    double epsp = epsdot;
    if (m->israte == 0 && m->vp == 2)
      epsp = dmax(dmax(fabs(deps[0]), fabs(deps[1])), HALF * fabs(deps[2]));
    epsp = dmax(epsp, epdr);
    double logep = log(epsp / epdr);
    double q = (ONE + cc * logep);
    q = dmax(q, K.EM20);
    ca = ca * q;
    cb = cb * q;
    if (m->icc == 1) ymax = ymax * q;
  }
  double dpla = ZERO, yld;
  if (*pla == ZERO) {
    yld = ca;
  } else {
    double beta = cb * (ONE - ZERO);
    yld = ca + beta * exp(cn * log(*pla));
  }
  yld = dmin(yld, ymax);
  const double SMALL = K.EM07;
  if (m->ipla == 0) {
    double svm = sqrt(sxx * sxx + syy * syy - sxx * syy + THREE * sxy * sxy);
    double r = dmin(ONE, yld / (svm + K.EM15));
    if (r < ONE) {
      sxx = sxx * r;
      syy = syy * r;
      sxy = sxy * r;
      dpla = (*off) * dmax(ZERO, (svm - yld) / young);
      double s1 = HALF * (sxx + syy);
      ezz = dpla * s1 / yld;
      *pla = *pla + dpla;
      *epchk = dmax(*pla, *epchk);
      if (yld >= ymax) h = ZERO;
      else h = cn * cb * exp((cn - ONE) * log(*pla + SMALL));
      etse = h / (h + young);
    }
  } else if (m->ipla == 1) {
    double s1 = sxx + syy, s2 = sxx - syy, s3 = sxy;
    double A = FOURTH * s1 * s1;
    double B = K.THREE_OVER_4 * s2 * s2 + THREE * s3 * s3;
    double svm = sqrt(A + B);
    if (svm > yld && *off == ONE) {
      double nu1 = ONE / (ONE - nu), nu2 = ONE / (ONE + nu);
      if (yld >= ymax) h = ZERO;
      else h = cn * cb * exp((cn - ONE) * log(*pla + SMALL));
      double dpla_j = (svm - yld) / (THREE * g + h);
      etse = h / (h + young);
      double anu1 = A * nu1, bnu2 = THREE * B * nu2, h2 = TWO * h;
      double dpla_i = 0.0, dr = 0.0, p = 0.0, q = 0.0;
      for (int it = 0; it < 3; ++it) {
        dpla_i = dpla_j;
        double pla_i = *pla + dpla_i;
        dpla = dpla_j;
        double yld_i;
        if (pla_i == ZERO) yld_i = dmin(ymax, ca);
        else yld_i = dmin(ymax, ca + cb * exp(cn * log(pla_i)));
        dr = HALF * young * dpla_i / yld_i;
        p = ONE / (ONE + dr * nu1);
        q = ONE / (ONE + THREE * dr * nu2);
        double p2 = p * p, q2 = q * q;
        double f = A * p2 + B * q2 - yld_i * yld_i;
        double df = -(anu1 * p2 * p + bnu2 * q2 * q) * (young - dr * h2) / yld_i - h2 * yld_i;
        if (dpla_i > ZERO) dpla_j = dmax(ZERO, dpla_i - f / df);
        else dpla_j = ZERO;
      }
      *pla = *pla + dpla_i;
      *epchk = dmax(*pla, *epchk);
      double t1 = (sxx + syy) * p;
      double t2 = (sxx - syy) * q;
      sxx = HALF * (t1 + t2);
      syy = HALF * (t1 - t2);
      sxy = sxy * q;
      ezz = dr * t1 / young;
    }
  } else {
    double svm2 = sxx * sxx + syy * syy - sxx * syy + THREE * sxy * sxy;
    double svm = sqrt(svm2);
    double yld2 = yld * yld;
    if (svm2 > yld2 && *off == ONE) {
      if (yld >= ymax) h = ZERO;
      else h = cn * cb * exp((cn - ONE) * log(*pla + SMALL));
      etse = h / (h + young);
      double aa = (svm2 - yld2) / (5.0 * svm2 + THREE * (-sxx * syy + sxy * sxy));
      double s1 = (ONE - TWO * aa) * sxx + aa * syy;
      double s2 = aa * sxx + (ONE - TWO * aa) * syy;
      double s3 = (ONE - THREE * aa) * sxy;
      sxx = s1; syy = s2; sxy = s3;
      dpla = (*off) * (svm - yld) / (THREE * g + h);
      *pla = *pla + dpla;
      yld = yld + h * dpla;
      svm = sqrt(sxx * sxx + syy * syy - sxx * syy + THREE * sxy * sxy);
      double r = dmin(ONE, yld / dmax(K.EM20, svm));
      sxx = sxx * r; syy = syy * r; sxy = sxy * r;
      ezz = dpla * HALF * (sxx + syy) / yld;
    }
  }
  (void)etse;
  /* ---- back in SIGEPS02C (sigeps02c.F:176-213) ---- */
  if (m->vp == 1) {
    double epsdot_p = dpla / dmax(K.EM20, dt1);
    *epsdp = asrate * epsdot_p + (ONE - asrate) * (*epsdp);
  }
  *sigy = *sigy + yld / npt;
  if (*off == off_old && *off > ZERO) {
    if (*off == ONE && *epchk >= epmx) *off = K.FOUR_OVER_5;
    else if (*off < ONE) *off = (*off) * K.FOUR_OVER_5;
  }
  ezz = -(deps[0] + deps[1]) * nu - (ONE - TWO * nu) * ezz;
  ezz = ezz / (ONE - nu);
  *thkn = *thkn + ezz * thklyl * (*off);
  sign[0] = sxx; sign[1] = syy; sign[2] = sxy; sign[3] = syz; sign[4] = szx;
}

void oracle_shell_element(const double xg[3][4], const double vg[3][4], const double vrg[3][4],
                          double thke, int tri, orad_shstate_t *st, double dt1,
                          const orad_shmat_t *m, const orad_shprop_t *p, orad_shout_t *out) {
  init_constants();
  const double EM20 = K.EM20;
  /* ---- CCOOR3 (ccoor3.F:78-143) ---- */
  double vl[4][3], vrl[4][3];
  for (int k = 0; k < 4; ++k)
    for (int d = 0; d < 3; ++d) { vl[k][d] = vg[d][k]; vrl[k][d] = vrg[d][k]; }
  const double dt1c = dt1;
  double off = dmin(ONE, fabs(st->offg));
  if (st->offg < ZERO)
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) { vl[k][d] = ZERO; vrl[k][d] = ZERO; }
  const double thk0 = thke;
  const double x1 = xg[0][0], x2g = xg[0][1], x3g = xg[0][2], x4g = xg[0][3];
  const double y1 = xg[1][0], y2g = xg[1][1], y3g = xg[1][2], y4g = xg[1][3];
  const double z1 = xg[2][0], z2g = xg[2][1], z3g = xg[2][2], z4g = xg[2][3];
  /* ---- CNVEC3, ISHFRAM = 0 (cnvec3.F:79-128) ---- */
  double X21 = x2g - x1, X32 = x3g - x2g, X34 = x3g - x4g, X41 = x4g - x1;
  double Y21 = y2g - y1, Y32 = y3g - y2g, Y34 = y3g - y4g, Y41 = y4g - y1;
  double Z21 = z2g - z1, Z32 = z3g - z2g, Z34 = z3g - z4g, Z41 = z4g - z1;
  double E1X = (X21 + X34), E1Y = (Y21 + Y34), E1Z = (Z21 + Z34);
  double E2X = (X32 + X41), E2Y = (Y32 + Y41), E2Z = (Z32 + Z41);
  double E3X = E1Y * E2Z - E1Z * E2Y;
  double E3Y = E1Z * E2X - E1X * E2Z;
  double E3Z = E1X * E2Y - E1Y * E2X;
  double SUMA = E3X * E3X + E3Y * E3Y + E3Z * E3Z;
  SUMA = ONE / dmax(sqrt(SUMA), EM20);
  E3X = E3X * SUMA; E3Y = E3Y * SUMA; E3Z = E3Z * SUMA;
  double S1 = E1X * E1X + E1Y * E1Y + E1Z * E1Z;
  double S2 = E2X * E2X + E2Y * E2Y + E2Z * E2Z;
  SUMA = sqrt(S1 / S2);
  E1X = E1X + (E2Y * E3Z - E2Z * E3Y) * SUMA;
  E1Y = E1Y + (E2Z * E3X - E2X * E3Z) * SUMA;
  E1Z = E1Z + (E2X * E3Y - E2Y * E3X) * SUMA;
  SUMA = E1X * E1X + E1Y * E1Y + E1Z * E1Z;
  SUMA = ONE / dmax(sqrt(SUMA), EM20);
  E1X = E1X * SUMA; E1Y = E1Y * SUMA; E1Z = E1Z * SUMA;
  E2X = E3Y * E1Z - E3Z * E1Y;
  E2Y = E3Z * E1X - E3X * E1Z;
  E2Z = E3X * E1Y - E3Y * E1X;
  /* ---- CDERI3 (cderi3.F:89-188) ---- */
  double sti = ZERO, stir = ZERO;
  double X21GA = x2g - x1, Y21GA = y2g - y1, Z21GA = z2g - z1;
  double X31GA = x3g - x1, Y31GA = y3g - y1, Z31GA = z3g - z1;
  double X41GA = x4g - x1, Y41GA = y4g - y1, Z41GA = z4g - z1;
  double X2 = E1X * X21GA + E1Y * Y21GA + E1Z * Z21GA;
  double Y2 = E2X * X21GA + E2Y * Y21GA + E2Z * Z21GA;
  double Y3 = E2X * X31GA + E2Y * Y31GA + E2Z * Z31GA;
  double X3 = E1X * X31GA + E1Y * Y31GA + E1Z * Z31GA;
  double X4 = E1X * X41GA + E1Y * Y41GA + E1Z * Z41GA;
  double Y4 = E2X * X41GA + E2Y * Y41GA + E2Z * Z41GA;
  double Z2 = E3X * X21GA + E3Y * Y21GA + E3Z * Z21GA;
  (void)Z2;
  double PX1 = HALF * (Y2 - Y4);
  double PY1 = HALF * (X4 - X2);
  double PX2 = HALF * Y3;
  double PY2 = -HALF * X3;
  double AREA = dmax(TWO * (PY2 * PX1 - PY1 * PX2), EM20);
  double VHX = (-X2 + X3 - X4) / AREA;
  double VHY = (-Y2 + Y3 - Y4) / AREA;
  /* ---- CCOEF3, ITHK = 0 (ccoef3.F:76-160) ---- */
  const double ALPE = ONE;
  const double VOL00 = thk0 * AREA;
  const double VOL0 = VOL00;
  const double THK02 = thk0 * thk0;
  const double RHO = m->rho0, YM = m->young, NU = m->nu, G = m->shear, SSP = m->ssp;
  const double H1 = p->hm, H2 = p->hf, H3 = p->hr;
  const double SRH1 = ZERO, SRH2 = ZERO, SRH3 = ZERO;
  double SHF;
  if (p->npt == 1) {
    SHF = ZERO;
  } else {
    double FAC1TMP = 2.0 * (1.0 + NU) * THK02;
    int ISH = p->ish;
    double FSH = p->fsh;
    SHF = FSH * ((1.0 - ISH) + ISH * FAC1TMP / (FSH * AREA + FAC1TMP));
  }
  const double GS = G * SHF;
  /* ---- CDLEN3, IHBE = 1 (cdlen3.F:70-121) ---- */
  double AL1 = X2 * X2 + Y2 * Y2;
  double AL2 = (X3 - X2) * (X3 - X2) + (Y3 - Y2) * (Y3 - Y2);
  double AL6 = X3 * X3 + Y3 * Y3;
  double AL3 = (X4 - X3) * (X4 - X3) + (Y4 - Y3) * (Y4 - Y3);
  double AL4 = X4 * X4 + Y4 * Y4;
  double AL5 = (X4 - X2) * (X4 - X2) + (Y4 - Y2) * (Y4 - Y2);
  double ALMIN = dmin(dmin(AL1, AL2), AL4);
  double ALQUAD = dmin(dmin(AL3, AL5), AL6);
  if (AL3 != ZERO) ALMIN = dmin(ALMIN, ALQUAD);
  double DTDYN = AREA * AREA / dmax(dmax(AL5, AL6), EM20);
  double ALDT = dmax(DTDYN, ALMIN);
  {
    double DTHOUR = HALF * (ALMIN + ALDT) / dmax(H1, H2);
    if (DTHOUR < ALDT) ALDT = DTHOUR;
  }
  ALDT = sqrt(ALDT);
  /* ---- CDEFO3, IHBE <= 1 (cdefo3.F:79-127) ---- */
  double VX[4], VY[4], VZ[4];
  for (int k = 0; k < 4; ++k) {
    VX[k] = E1X * vl[k][0] + E1Y * vl[k][1] + E1Z * vl[k][2];
    VY[k] = E2X * vl[k][0] + E2Y * vl[k][1] + E2Z * vl[k][2];
    VZ[k] = E3X * vl[k][0] + E3Y * vl[k][1] + E3Z * vl[k][2];
  }
  double VZ13 = VZ[0] - VZ[2], VZ24 = VZ[1] - VZ[3];
  double EYZ = PY1 * VZ13 + PY2 * VZ24;
  double EXZ = PX1 * VZ13 + PX2 * VZ24;
  double EXX, EYY, EXY;
  {
    double DT1V4 = FOURTH * dt1c;
    double TMP2A = PY2 + PY1;
    double TMP3A = dsign(dmax(fabs(TMP2A), EM20), TMP2A);
    double TMP1A = DT1V4 * (VZ13 - VZ24) * (VZ13 - VZ24) / TMP3A;
    double VX13 = VX[0] - VX[2], VX24 = VX[1] - VX[3];
    VX13 = VX13 - TMP1A;
    VX24 = VX24 + TMP1A;
    EXX = PX1 * VX13 + PX2 * VX24;
    EXY = PY1 * VX13 + PY2 * VX24;
    double TMP1B = PX2 - PX1;
    double TMP3B = dsign(dmax(fabs(TMP1B), EM20), TMP1B);
    double TMP2B = DT1V4 * (VZ13 + VZ24) * (VZ13 + VZ24) / TMP3B;
    double VY13 = VY[0] - VY[2], VY24 = VY[1] - VY[3];
    VY13 = VY13 + TMP2B;
    VY24 = VY24 + TMP2B;
    EXY = EXY + PX1 * VY13 + PX2 * VY24;
    EYY = PY1 * VY13 + PY2 * VY24;
  }
  /* ---- CCURV3 (ccurv3.F:64-93) ---- */
  double RX[4], RY[4];
  for (int k = 0; k < 4; ++k) {
    RX[k] = E1X * vrl[k][0] + E1Y * vrl[k][1] + E1Z * vrl[k][2];
    RY[k] = E2X * vrl[k][0] + E2Y * vrl[k][1] + E2Z * vrl[k][2];
  }
  double KXX, KYY, KXY;
  {
    double RX13TA = RX[0] - RX[2];
    double RXAVTA = RX[0] + RX[1] + RX[2] + RX[3];
    double RX24TA = RX[1] - RX[3];
    KYY = -PY1 * RX13TA - PY2 * RX24TA;
    KXY = PX1 * RX13TA + PX2 * RX24TA;
    double RY13TA = RY[0] - RY[2];
    double RYAVTA = RY[0] + RY[1] + RY[2] + RY[3];
    double RY24TA = RY[1] - RY[3];
    KXX = PX1 * RY13TA + PX2 * RY24TA;
    KXY = PY1 * RY13TA + PY2 * RY24TA - KXY;
    EXZ = EXZ + RYAVTA * (0.25 * AREA);
    EYZ = EYZ - RXAVTA * (0.25 * AREA);
  }
  /* ---- CSTRA3 (cstra3.F:100-110) ---- */
  {
    double FAC1 = dt1c / AREA;
    EXX = EXX * FAC1; EYY = EYY * FAC1; EXY = EXY * FAC1;
    EYZ = EYZ * FAC1; EXZ = EXZ * FAC1;
    KXX = KXX * FAC1; KYY = KYY * FAC1; KXY = KXY * FAC1;
  }
  /* ---- element equivalent strain rate (cforc3.F:535-546) ---- */
  const double dtinv = dt1 / dmax(dt1 * dt1, EM20);
  double epsd_pg;
  {
    double eps_k2 = (KXX * KXX + KYY * KYY + KXX * KYY + FOURTH * (KXY * KXY)) * K.ONE_OVER_9 *
                    (st->thk * st->thk);
    double eps_m2 = K.FOUR_OVER_3 * (EXX * EXX + EYY * EYY + EXX * EYY + FOURTH * (EXY * EXY));
    epsd_pg = sqrt(eps_k2 + eps_m2) * dtinv;
    st->epsd = ONE * epsd_pg + (ONE - ONE) * st->epsd;
  }
  /* ---- CMAIN3 -> MULAWC, IGTYP = 1, law 2 (mulawc.F90:530-3046) ---- */
  double degmb = st->forc[0] * EXX + st->forc[1] * EYY + st->forc[2] * EXY +
                 st->forc[3] * EYZ + st->forc[4] * EXZ;
  double degfx = st->mom[0] * KXX + st->mom[1] * KYY + st->mom[2] * KXY;
  const double vol0 = AREA * thk0;
  double thkn = st->thk;
  double FOR[5] = {ZERO, ZERO, ZERO, ZERO, ZERO}, MOM[3] = {ZERO, ZERO, ZERO};
  double sigy = ZERO;
  const double off_old = off;
  double epchk = ZERO;
  double viscmx = ZERO;
  const double asrate = (m->israte > 0) ? dmin(ONE, m->asrate_pm9 * dt1) : ONE;
  const int npt = p->npt;
  double z0[5], wf[5], wm[5];
  oracle_shell_tables(npt, z0, wf, wm);
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
    double sign[5];
    sigeps02c(m, GS, dt1, asrate, epsd_pg, off_old, &off, &st->pla[it], &st->epsdp[it],
              &epchk, &sigy, &thkn, thklyl, npt, deps, st->sig[it], sign);
    viscmx = dmax(p->dm, viscmx);
    for (int c = 0; c < 5; ++c) st->sig[it][c] = sign[c] * ONE;
    for (int c = 0; c < 5; ++c) FOR[c] = FOR[c] + wf[it] * sign[c];
    for (int c = 0; c < 3; ++c) MOM[c] = MOM[c] + wmc * sign[c];
  }
  st->thk = dmax(thkn, K.EM30);
  {
    double fact = K.ONEP414 * p->dm;
    double visc = fact * SSP * sqrt(AREA) * dtinv * RHO;
    FOR[0] = FOR[0] + visc * (EXX + HALF * EYY);
    FOR[1] = FOR[1] + visc * (EYY + HALF * EXX);
    FOR[2] = FOR[2] + visc * EXY * K.THIRD;
  }
  for (int c = 0; c < 5; ++c) FOR[c] = FOR[c] * off * ONE;
  for (int c = 0; c < 3; ++c) MOM[c] = MOM[c] * off * ONE;
  degmb = degmb + FOR[0] * EXX + FOR[1] * EYY + FOR[2] * EXY + FOR[3] * EYZ + FOR[4] * EXZ;
  degfx = degfx + MOM[0] * KXX + MOM[1] * KYY + MOM[2] * KXY;
  {
    double vol2 = HALF * vol0;
    st->eint[0] = st->eint[0] + degmb * vol2;
    st->eint[1] = st->eint[1] + degfx * thk0 * vol2;
  }
  for (int c = 0; c < 5; ++c) st->forc[c] = FOR[c];
  for (int c = 0; c < 3; ++c) st->mom[c] = MOM[c];
  (void)sigy;
  /* cforc3.F:587-589 */
  viscmx = sqrt(ONE + viscmx * viscmx) - viscmx;
  /* ---- CHVIS3 (chvis3.F:120-330), ISMSTR = 2, IHBE = 1 ---- */
  double GAMA1, GAMA2, GAMA3, GAMA4;
  {
    double PX1V = PX1 * VHX, PX2V = PX2 * VHX, PY1V = PY1 * VHY, PY2V = PY2 * VHY;
    GAMA1 = off * (ONE - PX1V - PY1V);
    GAMA3 = off * (ONE + PX1V + PY1V);
    GAMA2 = off * (-ONE - PX2V - PY2V);
    GAMA4 = off * (-ONE + PX2V + PY2V);
  }
  const double HVISC = HALF, HELAS = HALF, HVLIN = ZERO;
  double A1, A2, A3, A4, A5, A6, A7, A8, HH3;
  double SHFPR3 = SHF / (THREE * (ONE + NU));
  {
    double HVISH1 = HVISC * H1, HVISH2 = HVISC * H2;
    double R0 = FOURTH * RHO;
    double R1 = R0 * HUNDRED;
    R0 = R0 * HVLIN;
    A1 = R1 * HVISH1;
    A2 = R0 * K.SR2D2 * SRH1;
    double SRSHFPR3 = sqrt(SHFPR3);
    A3 = R1 * HVISH2 * SRSHFPR3;
    A4 = R0 * K.SR2D2 * SRH2 * SRSHFPR3;
    HH3 = HELAS * H3;
    A5 = HH3 * R1 * K.ZEP072169;
    HH3 = K.SR2D2 * SRH3;
    A6 = HH3 * R0 * K.ZEP072169;
    R0 = FOURTH * YM * HELAS;
    A7 = H1 * R0;
    A8 = H2 * R0 * SHFPR3;
  }
  double T2A = THK02 * AREA;
  double TSA = sqrt(T2A);
  double H1Q = A1 * TSA, H1L = A2 * SSP * TSA;
  double H2Q = A3 * THK02, H2L = A4 * SSP * THK02;
  double H3Q = A5 * T2A, H3L = A6 * SSP * T2A;
  double TD = thk0 * dt1c;
  double HH1 = A7 * TD;
  double B1 = PX1 * PX1 + PY1 * PY1;
  double B2 = PX2 * PX2 + PY2 * PY2;
  double HH2 = A8 * THK02 * TD / (B1 + B2);
  if (tri) { H1Q = H1L = H2Q = H2L = H3Q = H3L = HH1 = HH2 = ZERO; }
  double HG1 = VX[0] * GAMA1 + VX[1] * GAMA2 + VX[2] * GAMA3 + VX[3] * GAMA4;
  double HG2 = VY[0] * GAMA1 + VY[1] * GAMA2 + VY[2] * GAMA3 + VY[3] * GAMA4;
  st->hour[0] = st->hour[0] + HG1 * HH1;
  st->hour[1] = st->hour[1] + HG2 * HH1;
  double HOUR1A = st->hour[0] + HG1 * (H1L + H1Q * fabs(HG1));
  double H11 = HOUR1A * GAMA1, H12 = HOUR1A * GAMA2, H13 = HOUR1A * GAMA3;
  double HOUR2A = st->hour[1] + HG2 * (H1L + H1Q * fabs(HG2));
  double H21 = HOUR2A * GAMA1, H22 = HOUR2A * GAMA2, H23 = HOUR2A * GAMA3;
  double EHOU = HOUR1A * HG1 + HOUR2A * HG2;
  HG1 = VZ[0] * GAMA1 + VZ[1] * GAMA2 + VZ[2] * GAMA3 + VZ[3] * GAMA4;
  st->hour[2] = st->hour[2] + HG1 * HH2;
  double HOUR3A = st->hour[2] + HG1 * (H2L + H2Q * fabs(HG1));
  double H31 = HOUR3A * GAMA1, H32 = HOUR3A * GAMA2, H33 = HOUR3A * GAMA3;
  EHOU = EHOU + HOUR3A * HG1;
  HG1 = RX[0] - RX[1] + RX[2] - RX[3];
  HG2 = RY[0] - RY[1] + RY[2] - RY[3];
  st->hour[3] = HG1 * (H3L + H3Q * fabs(HG1));
  st->hour[4] = HG2 * (H3L + H3Q * fabs(HG2));
  EHOU = EHOU + st->hour[3] * HG1 + st->hour[4] * HG2;
  EHOU = dt1c * EHOU * off;
  double B11 = st->hour[3] * off, B12 = -st->hour[3] * off;
  double B13 = st->hour[3] * off, B14 = -st->hour[3] * off;
  double B21 = st->hour[4] * off, B22 = -st->hour[4] * off;
  double B23 = st->hour[4] * off, B24 = -st->hour[4] * off;
  /* ---- CDT3, IDTMIN(3) = 2 (cdt3.F:110-229) ---- */
  ALDT = ALDT * viscmx / sqrt(ALPE);
  double DT = p->dtfac * ALDT / SSP;
  if (!(DT > p->dtmin || off < ONE || st->offg < ZERO)) off = ZERO;  /* element deletion */
  out->dt = DT;
  out->alive = (st->offg > ZERO && off != ZERO);
  {
    double DIVM = dmax(ALDT * ALDT, EM20);
    sti = HALF * VOL0 * YM / DIVM;
    sti = K.ZEP81 * sti * off;
    stir = ZERO;
  }
  /* ---- CFINT3, ISIGI = 0 (cfint3.F:165-253) ---- */
  double F[3][4], M[3][4];
  {
    double F1A = FOR[0] * thk0, F2A = FOR[1] * thk0, F3A = FOR[2] * thk0;
    double F4A = FOR[3] * thk0, F5A = FOR[4] * thk0;
    double M4 = F4A * AREA, M5 = F5A * AREA;
    double F12 = F1A * PX2 + F3A * PY2;
    double F22 = F2A * PY2 + F3A * PX2;
    double F32 = F5A * PX2 + F4A * PY2;
    double F11 = F1A * PX1 + F3A * PY1;
    double F21 = F2A * PY1 + F3A * PX1;
    double F31 = F5A * PX1 + F4A * PY1;
    double G11 = F11 + H11, G13 = H13 - F11, G21 = F21 + H21;
    double G23 = H23 - F21, G31 = F31 + H31, G33 = H33 - F31;
    double G12 = F12 + H12, G22 = F22 + H22, G32 = F32 + H32;
    F[0][0] = E1X * G11 + E2X * G21 + E3X * G31;
    F[0][1] = E1X * G12 + E2X * G22 + E3X * G32;
    F[0][2] = E1X * G13 + E2X * G23 + E3X * G33;
    F[1][0] = E1Y * G11 + E2Y * G21 + E3Y * G31;
    F[1][1] = E1Y * G12 + E2Y * G22 + E3Y * G32;
    F[1][2] = E1Y * G13 + E2Y * G23 + E3Y * G33;
    F[2][0] = E1Z * G11 + E2Z * G21 + E3Z * G31;
    F[2][1] = E1Z * G12 + E2Z * G22 + E3Z * G32;
    F[2][2] = E1Z * G13 + E2Z * G23 + E3Z * G33;
    for (int d = 0; d < 3; ++d) F[d][3] = -F[d][0] - F[d][1] - F[d][2];
    double M1A = MOM[0] * THK02, M2A = MOM[1] * THK02, M3A = MOM[2] * THK02;
    M4 = M4 * FOURTH;
    M5 = M5 * FOURTH;
    double M11 = -M2A * PY1 - M3A * PX1;
    double M21 = M1A * PX1 + M3A * PY1;
    double M12 = -M2A * PY2 - M3A * PX2;
    double M22 = M1A * PX2 + M3A * PY2;
    double Q11 = M11 - M4 + B11, Q13 = -M11 - M4 + B13;
    double Q12 = M12 - M4 + B12, Q14 = -M12 - M4 + B14;
    double Q21 = M21 + M5 + B21, Q23 = -M21 + M5 + B23;
    double Q22 = M22 + M5 + B22, Q24 = -M22 + M5 + B24;
    const double e1[3] = {E1X, E1Y, E1Z}, e2[3] = {E2X, E2Y, E2Z};
    for (int d = 0; d < 3; ++d) {
      M[d][0] = e1[d] * Q11 + e2[d] * Q21;
      M[d][1] = e1[d] * Q12 + e2[d] * Q22;
      M[d][2] = e1[d] * Q13 + e2[d] * Q23;
      M[d][3] = e1[d] * Q14 + e2[d] * Q24;
    }
  }
  /* ---- CUPDT3P (cupdt3.F:1000-1067) ---- */
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
  out->ehou = EHOU;
}

/* Generalised FILLCNE + W_PON ordering for nc corners per element. */
typedef struct { int32_t uid, idx; } uid_pair;
static int cmp_uid(const void *a, const void *b) {
  const uid_pair *p = a, *q = b;
  if (p->uid != q->uid) return (p->uid < q->uid) ? -1 : 1;
  return (p->idx < q->idx) ? -1 : (p->idx > q->idx);
}
int oracle_build_pon_nc(int numnod, int numel, int nc, const int32_t *conn,
                        const int32_t *uid, int32_t *adsky, int32_t *iads) {
  int32_t *cnt = calloc((size_t)numnod + 1, sizeof(int32_t));
  uid_pair *ord = malloc(sizeof(uid_pair) * (size_t)(numel > 0 ? numel : 1));
  if (!cnt || !ord) { free(cnt); free(ord); return -1; }
  for (int e = 0; e < numel; ++e)
    for (int k = 0; k < nc; ++k) {
      int n = conn[nc * e + k];
      if (n < 0 || n >= numnod) { free(cnt); free(ord); return -2; }
      cnt[n]++;
    }
  adsky[0] = 0;
  for (int n = 0; n < numnod; ++n) adsky[n + 1] = adsky[n] + cnt[n];
  for (int n = 0; n < numnod; ++n) cnt[n] = adsky[n];
  for (int e = 0; e < numel; ++e) { ord[e].uid = uid[e]; ord[e].idx = e; }
  qsort(ord, (size_t)numel, sizeof(uid_pair), cmp_uid);
  for (int j = 0; j < numel; ++j) {
    int e = ord[j].idx;
    for (int k = 0; k < nc; ++k) {
      int n = conn[nc * e + k];
      iads[nc * e + k] = cnt[n]++;
    }
  }
  free(cnt); free(ord);
  return 0;
}

/* ---- Whole shell cycle in RESOL order (resol.F:2586-9595) ---- */
void oracle_shell_cycle(orad_oshell_t *m) {
  init_constants();
  const int ne = m->numelc, nn = m->numnod;
  double dt2t = 1.0e300;
  double *ehou = malloc(sizeof(double) * (size_t)(ne > 0 ? ne : 1));
  m->dt1 = m->dt2;
  const double dt1 = m->dt1;
#ifdef _OPENMP
  int nth = m->nthreads > 0 ? m->nthreads : omp_get_max_threads();
#endif
#pragma omp parallel for num_threads(nth) schedule(static) reduction(min : dt2t)
  for (int e = 0; e < ne; ++e) {
    double xg[3][4], vg[3][4], vrg[3][4];
    for (int k = 0; k < 4; ++k) {
      int n = m->ixc[4 * e + k];
      for (int d = 0; d < 3; ++d) {
        xg[d][k] = m->x[3 * n + d];
        vg[d][k] = m->v[3 * n + d];
        vrg[d][k] = m->vr[3 * n + d];
      }
    }
    int tri = (m->ixc[4 * e + 2] == m->ixc[4 * e + 3]);
    orad_shout_t o;
    oracle_shell_element((const double(*)[4])xg, (const double(*)[4])vg, (const double(*)[4])vrg,
                         m->thke[e], tri, &m->st[e], dt1, &m->mat, &m->prop, &o);
    for (int k = 0; k < 4; ++k) {
      double *f = m->fsky + 8 * (size_t)m->iads[4 * e + k];
      f[0] = o.f[0][k]; f[1] = o.f[1][k]; f[2] = o.f[2][k];
      f[3] = o.m[0][k]; f[4] = o.m[1][k]; f[5] = o.m[2][k];
      f[6] = o.sti; f[7] = o.stir;
    }
    if (o.alive && o.dt < dt2t) dt2t = o.dt;
    ehou[e] = o.ehou;
  }
  for (int e = 0; e < ne; ++e) m->ehour += ehou[e];
  free(ehou);
  /* ASSPAR4 (engine/source/assembly/asspar4.F) ordered per-node gather, 6 DOF */
#pragma omp parallel for num_threads(nth) schedule(static)
  for (int n = 0; n < nn; ++n) {
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = m->adsky[n]; k < m->adsky[n + 1]; ++k) {
      const double *f = m->fsky + 8 * (size_t)k;
      for (int c = 0; c < 8; ++c) s[c] = s[c] + f[c];
    }
    for (int d = 0; d < 3; ++d) { m->a[3 * n + d] = s[d]; m->ar[3 * n + d] = s[3 + d]; }
    m->stifn[n] = s[6];
    m->stifr[n] = s[7];
  }
  double dt2 = 1.0e6;
  if (dt2t < dt2) dt2 = dt2t;
  /* resol.F:6080 writes 1.1*DT2OLD with a default-REAL literal */
  dt2 = dmin(dmin(dt2, (double)1.1f * m->dt2old), m->dtmx);
  m->dt2old = dt2;
  m->dt2 = dt2;
  m->dt12 = HALF * (dt1 + dt2);
  const double dt12 = m->dt12;
  /* ACCELE (accele.F:64-110) -> BCS10 -> FIXVEL -> VELOCITY -> DEPLA */
#pragma omp parallel for num_threads(nth) schedule(static)
  for (int n = 0; n < nn; ++n) {
    double a[3] = {m->a[3 * n], m->a[3 * n + 1], m->a[3 * n + 2]};
    double ar[3] = {m->ar[3 * n], m->ar[3 * n + 1], m->ar[3 * n + 2]};
    if (m->ms[n] > ZERO) {
      double RTMP = ONE / m->ms[n];
      a[0] = a[0] * RTMP; a[1] = a[1] * RTMP; a[2] = a[2] * RTMP;
    } else {
      a[0] = a[1] = a[2] = ZERO;
    }
    if (m->inr[n] > ZERO) {
      double RTMP = ONE / m->inr[n];
      ar[0] = ar[0] * RTMP; ar[1] = ar[1] * RTMP; ar[2] = ar[2] * RTMP;
    } else {
      ar[0] = ar[1] = ar[2] = ZERO;
    }
    double *v = m->v + 3 * n, *vr = m->vr + 3 * n;
    int c = m->icodt[n], r = m->icodr[n];
    if (c & 4) { v[0] = ZERO; a[0] = ZERO; }
    if (c & 2) { v[1] = ZERO; a[1] = ZERO; }
    if (c & 1) { v[2] = ZERO; a[2] = ZERO; }
    if (r & 4) { vr[0] = ZERO; ar[0] = ZERO; }
    if (r & 2) { vr[1] = ZERO; ar[1] = ZERO; }
    if (r & 1) { vr[2] = ZERO; ar[2] = ZERO; }
    if (m->fv_dir && m->fv_dir[n] > 0) {
      /* FIXVEL (constraints/general/impvel/fixvel.F:36): the imposed component's
       * acceleration is set so that V(t+dt12) equals the imposed value. */
      int d = m->fv_dir[n] - 1;
      a[d] = (m->fv_val[n] - v[d]) / dt12;
    }
    for (int d = 0; d < 3; ++d) {
      m->a[3 * n + d] = a[d];
      m->ar[3 * n + d] = ar[d];
      v[d] = v[d] + dt12 * a[d];
      vr[d] = vr[d] + dt12 * ar[d];
      double VDT = dt2 * v[d];
      m->d[3 * n + d] = m->d[3 * n + d] + VDT;
      m->x[3 * n + d] = m->x[3 * n + d] + VDT;
    }
  }
  m->tt = m->tt + dt2;
  m->ncycle += 1;
}

void oracle_shell_energies(const orad_oshell_t *m, double *eint, double *ekin) {
  double ei = 0.0, ek = 0.0;
  for (int e = 0; e < m->numelc; ++e) ei += m->st[e].eint[0] + m->st[e].eint[1];
  for (int n = 0; n < m->numnod; ++n) {
    const double *v = m->v + 3 * n, *w = m->vr + 3 * n;
    ek += HALF * m->ms[n] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) +
          HALF * m->inr[n] * (w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  }
  *eint = ei;
  *ekin = ek;
}

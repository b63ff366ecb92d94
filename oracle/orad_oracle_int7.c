/*
 * orad_oracle_int7.c -- TEST INFRASTRUCTURE (the oracle), not product code.
 *
 * CPU restatement of the reference's /INTER/TYPE7 node-to-segment penalty contact on
 * the path the GPU library implements: constant gap (IGAP=0), IGSTI<=1, INACTI=0,
 * no curvature (ICURV=0), no adaptive meshing (IADM=0), no heat exchange, no tied
 * contact, IVIS2=0 damping, Coulomb friction in the total (viscous) formulation
 * without filtering (IFQ=0, MFROT=0, isotropic), one domain.
 *
 *   candidates  I7TRI's box test (interfaces/intsort/i7tri.F:530-566): a secondary
 *               node that is not one of the segment's nodes and lies strictly inside
 *               the segment's bounding box widened by TZ; listed by (secondary
 *               node, segment).
 *   per pair    I7COR3 (int07/i7cor3.F:99-261), I7DST3 (int07/i7dst3.F:124-470)
 *               and its compression to the penetrating pairs, I7FOR3
 *               (int07/i7for3.F:230-1080 normal force, damping and contact time
 *               step, 1460-1466 normal impulse, 1866-1875 and 2444-2480 friction,
 *               4108-4121 corner forces) with FRICTIONPARTS_MODEL_ISOT's INTFRIC=0
 *               coefficients (frictionparts_model.F:394-399), I7ASS2
 *               (int07/i7ass3.F:1188-1285) for the skyline entries.
 *   batches     the penetrating pairs, in candidate order, in batches of NVSIZ=128:
 *               within a batch I7ASS2 writes main node 1 of every pair, then nodes 2,
 *               3, 4 (where H /= 0), then the secondary nodes; I7FOR3's contact time
 *               step and energies are per batch (i7mainf.F:529-737).
 *
 * Parity: bit for bit against the reference routines compiled from source (oracle/_ref,
 * oracle/ref/harness_int7.F) on the golden cases of tests/golden/int7_cases.npz.
 */
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>

#include "orad_oracle_int.h"

#define ONE 1.0
#define TWO 2.0
#define HALF 0.5
#define FOURTH 0.25
#define NVSIZ 128

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* One pair through I7COR3 and I7DST3: geometry, barycentric parameters of the
 * projections on the four triangles (X0, Xk, Xk+1), their distance vectors and squared
 * distances, STIF and GAPV.  Returns 1 when the pair penetrates (KEEP). */
typedef struct {
  int ix[4], nsvg, cand;
  double xi[3], vi[3], msi, x[4][3];
  double nx[4][3], lb[4], lc[4], p[4], stif, gapv;
} i7pair_t;

static int i7_dist(const orad_int7_t *it, const double *x, const double *v, const double *ms,
                   int ni, int l, i7pair_t *q) {
  const shconst_t *K = oracle_shk();
  /* I7COR3 (IGAP=0, ITY=7, INTTH=0, IGSTI<=1) */
  q->gapv = it->gap;
  const int ig = it->nsv[ni];
  q->nsvg = ig;
  for (int d = 0; d < 3; ++d) {
    q->xi[d] = x[3 * ig + d];
    q->vi[d] = v[3 * ig + d];
  }
  q->msi = ms[ig];
  for (int k = 0; k < 4; ++k) {
    q->ix[k] = it->irect[4 * l + k];
    for (int d = 0; d < 3; ++d) q->x[k][d] = x[3 * q->ix[k] + d];
  }
  q->stif = it->stfm[l] * fabs(it->stfns[ni]);
  /* I7DST3 */
  double x0[3];
  if (q->ix[2] != q->ix[3]) {
    for (int d = 0; d < 3; ++d)
      x0[d] = FOURTH * (q->x[0][d] + q->x[1][d] + q->x[2][d] + q->x[3][d]);
  } else {
    for (int d = 0; d < 3; ++d) x0[d] = q->x[2][d];
  }
  double x0k[4][3], xik[4][3], xi0v[3];
  for (int k = 0; k < 4; ++k)
    for (int d = 0; d < 3; ++d) x0k[k][d] = q->x[k][d] - x0[d];
  for (int d = 0; d < 3; ++d) xi0v[d] = x0[d] - q->xi[d];
  for (int k = 0; k < 4; ++k)
    for (int d = 0; d < 3; ++d) xik[k][d] = q->x[k][d] - q->xi[d];
  /* S_k = XI0V x XIk; S0 = X0k x X0(k+1) */
  double s[4][3];
  for (int k = 0; k < 4; ++k) {
    s[k][0] = xi0v[1] * xik[k][2] - xi0v[2] * xik[k][1];
    s[k][1] = xi0v[2] * xik[k][0] - xi0v[0] * xik[k][2];
    s[k][2] = xi0v[0] * xik[k][1] - xi0v[1] * xik[k][0];
  }
  double lb[4], lc[4];
  for (int k = 0; k < 4; ++k) {
    const int k1 = (k + 1) & 3;
    const double sx0 = x0k[k][1] * x0k[k1][2] - x0k[k][2] * x0k[k1][1];
    const double sy0 = x0k[k][2] * x0k[k1][0] - x0k[k][0] * x0k[k1][2];
    const double sz0 = x0k[k][0] * x0k[k1][1] - x0k[k][1] * x0k[k1][0];
    const double s2 = ONE / dmax(K->EM30, sx0 * sx0 + sy0 * sy0 + sz0 * sz0);
    lb[k] = -(sx0 * s[k1][0] + sy0 * s[k1][1] + sz0 * s[k1][2]) * s2;
    lc[k] = (sx0 * s[k][0] + sy0 * s[k][1] + sz0 * s[k][2]) * s2;
  }
  /* HLB/HLC and the clamped edge parameters AL (i7dst3.F:241-266) */
  double hlb[4], hlc[4], al[4];
  for (int k = 0; k < 4; ++k) {
    const double aaa = ONE / dmax(K->EM30, x0k[k][0] * x0k[k][0] + x0k[k][1] * x0k[k][1] +
                                               x0k[k][2] * x0k[k][2]);
    hlc[k] = lc[k] * fabs(lc[k]) * aaa;
    hlb[(k + 3) & 3] = lb[(k + 3) & 3] * fabs(lb[(k + 3) & 3]) * aaa;
    al[k] = -(xi0v[0] * x0k[k][0] + xi0v[1] * x0k[k][1] + xi0v[2] * x0k[k][2]) * aaa;
    al[k] = dmax(0.0, dmin(ONE, al[k]));
  }
  /* triangle k: outside its outer edge (LA < 0) or one of its inner edges (i7dst3.F:268-359) */
  for (int k = 0; k < 4; ++k) {
    const int k1 = (k + 1) & 3;
    const double ex = q->x[k1][0] - q->x[k][0];
    const double ey = q->x[k1][1] - q->x[k][1];
    const double ez = q->x[k1][2] - q->x[k][2];
    const double la = ONE - lb[k] - lc[k];
    const double aaa = ONE / dmax(K->EM20, ex * ex + ey * ey + ez * ez);
    const double hla = la * fabs(la) * aaa;
    if (la < 0.0 && hla <= hlb[k] && hla <= hlc[k]) {
      lb[k] = (xik[k1][0] * ex + xik[k1][1] * ey + xik[k1][2] * ez) * aaa;
      lb[k] = dmax(0.0, dmin(ONE, lb[k]));
      lc[k] = ONE - lb[k];
    } else if (lb[k] < 0.0 && hlb[k] <= hlc[k] && hlb[k] <= hla) {
      lb[k] = 0.0;
      lc[k] = al[k1];
    } else if (lc[k] < 0.0 && hlc[k] <= hla && hlc[k] <= hlb[k]) {
      lc[k] = 0.0;
      lb[k] = al[k];
    }
  }
  /* distances to the projections, penetration test (i7dst3.F:385-420) */
  const double gap2 = q->gapv * q->gapv;
  double pene2 = 0.0;
  for (int k = 0; k < 4; ++k) {
    const int k1 = (k + 1) & 3;
    for (int d = 0; d < 3; ++d)
      q->nx[k][d] = q->xi[d] - (x0[d] + lb[k] * x0k[k][d] + lc[k] * x0k[k1][d]);
    q->p[k] = q->nx[k][0] * q->nx[k][0] + q->nx[k][1] * q->nx[k][1] + q->nx[k][2] * q->nx[k][2];
    const double dk = dmax(0.0, gap2 - q->p[k]);
    pene2 = k == 0 ? dk : dmax(pene2, dk);
    q->lb[k] = lb[k];
    q->lc[k] = lc[k];
  }
  return pene2 != 0.0 && q->stif != 0.0;
}

/* I7FOR3 + I7ASS2 over one batch of penetrating pairs. */
static int i7_batch(orad_int7_t *it, const double *v, const double *ms, i7pair_t *q, int jlt,
                    double dt1, double *dt2t, double *fsky, int32_t *isky) {
  const shconst_t *K = oracle_shk();
  (void)ms;
  const double dt1inv = dt1 > 0.0 ? ONE / dt1 : 0.0;
  double econtt = 0.0, econvt = 0.0;
  double h[NVSIZ][4], n[NVSIZ][3], pene[NVSIZ], vn[NVSIZ], vx[NVSIZ][3], fni[NVSIZ];
  double stif[NVSIZ], vis2[NVSIZ], fxi[NVSIZ][3];
  for (int i = 0; i < jlt; ++i) {
    i7pair_t *p = &q[i];
    stif[i] = p->stif;
    double pp[4], a[4];
    for (int k = 0; k < 4; ++k) {
      const double dk = sqrt(p->p[k]);
      pp[k] = dmax(0.0, p->gapv - dk);
      a[k] = pp[k] / dmax(K->EM20, dk);
    }
    for (int d = 0; d < 3; ++d)
      n[i][d] = a[0] * p->nx[0][d] + a[1] * p->nx[1][d] + a[2] * p->nx[2][d] + a[3] * p->nx[3][d];
    if (p->ix[2] != p->ix[3]) {
      pene[i] = dmax(dmax(dmax(pp[0], pp[1]), pp[2]), pp[3]);
      double la[4];
      for (int k = 0; k < 4; ++k) la[k] = ONE - p->lb[k] - p->lc[k];
      const double h0 = FOURTH * (pp[0] * la[0] + pp[1] * la[1] + pp[2] * la[2] + pp[3] * la[3]);
      h[i][0] = h0 + pp[0] * p->lb[0] + pp[3] * p->lc[3];
      h[i][1] = h0 + pp[1] * p->lb[1] + pp[0] * p->lc[0];
      h[i][2] = h0 + pp[2] * p->lb[2] + pp[1] * p->lc[1];
      h[i][3] = h0 + pp[3] * p->lb[3] + pp[2] * p->lc[2];
      const double h00 = ONE / dmax(K->EM20, h[i][0] + h[i][1] + h[i][2] + h[i][3]);
      for (int k = 0; k < 4; ++k) h[i][k] = h[i][k] * h00;
    } else {
      pene[i] = pp[0];
      for (int d = 0; d < 3; ++d) n[i][d] = p->nx[0][d];
      h[i][0] = p->lb[0];
      h[i][1] = p->lc[0];
      h[i][2] = ONE - p->lb[0] - p->lc[0];
      h[i][3] = 0.0;
    }
  }
  for (int i = 0; i < jlt; ++i) {
    const double s2 = ONE / dmax(K->EM30, sqrt(n[i][0] * n[i][0] + n[i][1] * n[i][1] +
                                               n[i][2] * n[i][2]));
    for (int d = 0; d < 3; ++d) n[i][d] = n[i][d] * s2;
  }
  for (int i = 0; i < jlt; ++i) {
    const i7pair_t *p = &q[i];
    for (int d = 0; d < 3; ++d)
      vx[i][d] = p->vi[d] - h[i][0] * v[3 * p->ix[0] + d] - h[i][1] * v[3 * p->ix[1] + d] -
                 h[i][2] * v[3 * p->ix[2] + d] - h[i][3] * v[3 * p->ix[3] + d];
    vn[i] = n[i][0] * vx[i][0] + n[i][1] * vx[i][1] + n[i][2] * vx[i][2];
    /* hourglass correction of the shape functions */
    double h0 = -0.25 * (h[i][0] - h[i][1] + h[i][2] - h[i][3]);
    h0 = dmin(dmin(h0, h[i][1]), h[i][3]);
    h0 = dmax(dmax(h0, -h[i][0]), -h[i][2]);
    if (p->ix[2] == p->ix[3]) h0 = 0.0;
    h[i][0] = h[i][0] + h0;
    h[i][1] = h[i][1] - h0;
    h[i][2] = h[i][2] + h0;
    h[i][3] = h[i][3] - h0;
  }
  /* contact time step of the batch (i7for3.F:772-870; DTMIN1(10) = 0) */
  double dti = (double)1.E20f;
  for (int i = 0; i < jlt; ++i) {
    const double dist = q[i].gapv - pene[i];
    const double rdist = HALF * dist / dmax(K->EM30, -vn[i]);
    dti = dmin(rdist, dti);
  }
  if (dti <= 0.0) dti = (double)1.E20f;
  if (dti < *dt2t) *dt2t = dti;
  /* nonlinear penalty stiffness and normal force (i7for3.F:941-993) */
  for (int i = 0; i < jlt; ++i) {
    const double g = q[i].gapv;
    const double fac = g / dmax(K->EM10, g - pene[i]);
    const double facm1 = ONE / fac;
    if ((g - pene[i]) / g < K->EM10 && stif[i] > 0.0) {
      stif[i] = 0.0;
      it->stfns[q[i].cand] = -fabs(it->stfns[q[i].cand]);
    }
    econtt = econtt + HALF * stif[i] * (g * g) * (facm1 - ONE - log(dmax(DBL_MIN, facm1)));
    stif[i] = HALF * stif[i] * fac;
    fni[i] = -stif[i] * pene[i];
  }
  /* damping, IVIS2 = 0 (i7for3.F:1086-1121, 1205-1247, 1283-1288) */
  for (int i = 0; i < jlt; ++i) {
    const double g = q[i].gapv;
    const double msi = q[i].msi;
    if (it->visc != 0.0 || it->viscf != 0.0) {
      vis2[i] = TWO * stif[i] * msi;
      if (vn[i] < 0.0) vis2[i] = vis2[i] / dmax(K->EM10, (g - pene[i]) / g);
      const double visca = K->ZEP4;
      const double fac = stif[i] / dmax(K->EM30, stif[i]);
      const double vis = sqrt(vis2[i]);
      double ff = fac * (it->visc * vis + visca * visca * TWO * msi * dmax(0.0, -vn[i]) /
                                              dmax(g - pene[i], K->EM10));
      stif[i] = stif[i] * g / dmax(g - pene[i], K->EM10);
      stif[i] = stif[i] + ff * dt1inv;
      stif[i] = dmax(stif[i], fac * sqrt(it->viscf) * vis * dt1inv);
      ff = ff * vn[i];
      fni[i] = fni[i] + ff;
    } else {
      vis2[i] = 0.0;
      stif[i] = stif[i] * g / dmax(g - pene[i], K->EM10);
    }
  }
  for (int i = 0; i < jlt; ++i)
    for (int d = 0; d < 3; ++d) fxi[i][d] = n[i][d] * fni[i];
  /* Coulomb friction, total viscous formulation without filtering (i7for3.F:2444-2480) */
  for (int i = 0; i < jlt; ++i) {
    const double xmu = it->fric;
    double t[3];
    for (int d = 0; d < 3; ++d) t[d] = vx[i][d] - n[i][d] * vn[i];
    const double v2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    const double vis2f = it->viscf * vis2[i];
    const double fm2 = (xmu * fni[i]) * (xmu * fni[i]);
    const double f2 = vis2f * v2;
    const double a2 = dmin(f2, fm2) / dmax(K->EM30, f2);
    const double aa = sqrt(a2 * vis2f);
    for (int d = 0; d < 3; ++d) fxi[i][d] = fxi[i][d] + aa * t[d];
    econvt = econvt + aa * v2 * dt1;
  }
  it->econt = it->econt + econtt;
  it->econtv = it->econtv + econvt;
  /* I7ASS2: main nodes 1..4 of every pair (where H /= 0), then the secondary nodes */
  int nsky = 0;
  for (int k = 0; k < 4; ++k)
    for (int i = 0; i < jlt; ++i) {
      if (h[i][k] != 0.0) {
        double *f = fsky + 4 * (size_t)nsky;
        for (int d = 0; d < 3; ++d) f[d] = fxi[i][d] * h[i][k];
        f[3] = stif[i] * fabs(h[i][k]);
        isky[nsky++] = q[i].ix[k];
      }
    }
  for (int i = 0; i < jlt; ++i) {
    double *f = fsky + 4 * (size_t)nsky;
    for (int d = 0; d < 3; ++d) f[d] = -fxi[i][d];
    f[3] = stif[i];
    isky[nsky++] = q[i].nsvg;
  }
  return nsky;
}

int oracle_int7_forces(orad_int7_t *it, const double *x, const double *v, const double *ms,
                       int ncand, const int32_t *cand_n, const int32_t *cand_e, double dt1,
                       double *dt2t, double *fsky, int32_t *isky) {
  i7pair_t *q = malloc(sizeof(i7pair_t) * (size_t)(ncand > 0 ? ncand : 1));
  int jlt = 0, nsky = 0;
  for (int c = 0; c < ncand; ++c) {
    if (i7_dist(it, x, v, ms, cand_n[c], cand_e[c], &q[jlt])) {
      q[jlt].cand = cand_n[c];
      ++jlt;
    }
  }
  it->npen = jlt;
  for (int b = 0; b < jlt; b += NVSIZ) {
    const int n = jlt - b < NVSIZ ? jlt - b : NVSIZ;
    nsky += i7_batch(it, v, ms, q + b, n, dt1, dt2t, fsky + 4 * (size_t)nsky, isky + nsky);
  }
  it->nsky = nsky;
  free(q);
  return nsky;
}

int oracle_int7_candidates(const orad_int7_t *it, const double *x, double tz, int cap,
                           int32_t *cand_n, int32_t *cand_e) {
  int nc = 0;
  for (int i = 0; i < it->nsn; ++i) {
    if (it->stfns[i] == 0.0) continue;
    const int nn = it->nsv[i];
    for (int l = 0; l < it->nrtm; ++l) {
      if (it->stfm[l] == 0.0) continue;
      const int32_t *r = it->irect + 4 * l;
      if (nn == r[0] || nn == r[1] || nn == r[2] || nn == r[3]) continue;
      int inside = 1;
      for (int d = 0; d < 3 && inside; ++d) {
        double lo = x[3 * r[0] + d], hi = lo;
        for (int k = 1; k < 4; ++k) {
          lo = dmin(lo, x[3 * r[k] + d]);
          hi = dmax(hi, x[3 * r[k] + d]);
        }
        const double c = x[3 * nn + d];
        inside = c > lo - tz && c < hi + tz;
      }
      if (!inside) continue;
      if (nc < cap) {
        cand_n[nc] = i;
        cand_e[nc] = l;
      }
      ++nc;
    }
  }
  return nc;
}

/* One cycle's contact: candidates, forces, and the per-node order of the skyline entries
 * for ASSPAR4 (entries of a node in skyline order, asspar4.F:126-154).  Returns the entry
 * count; *ent, *first (numnod+1) are allocated here. */
int oracle_int7_cycle(orad_int7_t *it, int numnod, const double *x, const double *v,
                      const double *ms, double dt1, double *dt2t, double **fsky_out,
                      int32_t **order_out, int32_t **first_out) {
  const int cap = it->nsn * (it->nrtm > 0 ? it->nrtm : 1);
  int32_t *cn = malloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
  int32_t *ce = malloc(sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
  const int nc = oracle_int7_candidates(it, x, it->gap * 1.000001 + 1e-300, cap, cn, ce);
  double *fsky = malloc(sizeof(double) * 4 * (size_t)(5 * nc + 1));
  int32_t *isky = malloc(sizeof(int32_t) * (size_t)(5 * nc + 1));
  const int nsky = oracle_int7_forces(it, x, v, ms, nc, cn, ce, dt1, dt2t, fsky, isky);
  int32_t *first = calloc((size_t)numnod + 1, sizeof(int32_t));
  int32_t *order = malloc(sizeof(int32_t) * (size_t)(nsky + 1));
  for (int k = 0; k < nsky; ++k) first[isky[k] + 1]++;
  for (int n = 0; n < numnod; ++n) first[n + 1] += first[n];
  int32_t *pos = malloc(sizeof(int32_t) * (size_t)(numnod + 1));
  memcpy(pos, first, sizeof(int32_t) * (size_t)(numnod + 1));
  for (int k = 0; k < nsky; ++k) order[pos[isky[k]]++] = k;
  free(pos);
  free(cn);
  free(ce);
  free(isky);
  *fsky_out = fsky;
  *order_out = order;
  *first_out = first;
  return nsky;
}

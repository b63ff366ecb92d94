/*
 * orad_oracle_int.h -- internal to the oracle library (test infrastructure):
 * the shell path's constants and the law-2 integration point, shared by the
 * Belytschko-Tsay/QEPH restatement (orad_oracle_shell.c) and the QBAT one
 * (orad_oracle_qbat.c).
 */
#ifndef ORAD_ORACLE_INT_H
#define ORAD_ORACLE_INT_H
#include "orad_oracle.h"

/* common_source/modules/constant_mod.F, re-derived from their definitions
 * (each a double operation on exact or previously rounded values). */
typedef struct {
  double EM01, EM02, EM03, EM04, EM05, EM06, EM07, EM15, EM20, EM30, EP20;
  double ZEP1, ZEP01, ZEP03, ZEP07, ZEP09, ZEP19, ZEP81, ZEP4;
  double TWOEM3, FOUREM3, FOUREM5, SIXEM5, NINEEM6, ZEP072169;
  double ONEP4, ONEP41, ONEP414;
  double ONE_OVER_9, ONE_OVER_12, FOUR_OVER_3, FOUR_OVER_5, FIVE_OVER_6, THIRD;
  double THREE_OVER_4, THREE_HALF, SR2D2;
  /* QEPH (CZFORC3 chain) */
  double EM08, EM10, EM18, FIVE_OVER_4, ONE_OVER_16, ONE_OVER_64, ZEP85, ZEP999, FIVEP333;
  double THREEP464, ZEP9EP30;
  /* QBAT (CBAFORC3 chain) */
  double EM12, ZEP3;
} shconst_t;

/* The constants above, computed once. */
const shconst_t *oracle_shk(void);

/* SIGEPS02C with M2CPLR for one through-thickness point
 * (engine/source/materials/mat/mat002/sigeps02c.F:90-226, m2cplr.F:95-443). */
void oracle_sigeps02c_point(const orad_shmat_t *m, double gs, double dt1, double asrate,
                            double epsd_pg, double off_old, double *off, double *pla,
                            double *epsdp, double *epchk, double *sigy, double *thkn,
                            double thklyl, int npt, const double deps[5], const double sigo[5],
                            double sign[5], double *etse_out);
#endif

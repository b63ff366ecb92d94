/*
 * orad_oracle.h -- CPU restatement of the OpenRadioss explicit brick path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Nothing in openradioss_amd/ links or calls it.
 *
 * Semantics follow the reference's double-precision build (my_real = double,
 * IRESP = 0) for the default solid:  Isolid=1 (JHBE=1), Iframe automatic
 * (JCVT=1, co-rotational, starter/source/elements/solid/solide/sgrtails.F:857-859),
 * Ismstr=4, law 2 Johnson-Cook with viscous hourglass, PARITH/ON assembly.
 * Every function cites the reference routine it restates.
 */
#ifndef ORAD_ORACLE_H
#define ORAD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Law-2 parameters (starter/source/materials/mat/mat002/hm_read_mat02_jc.F90:146-192,319-339). */
typedef struct {
  double rho0, shear, bulk;       /* mat_param%rho, %shear, %bulk           */
  double ca, cb, cn;              /* uparam(1..3)  A, B, n                   */
  double epmx, sigmx;             /* uparam(4,5)   EPS_max, SIG_max          */
  double cc, epdr;                /* uparam(6,7)   C, EPS0                   */
  double fisokin;                 /* uparam(8)     (kinematic hardening, 0)  */
  double mexp;                    /* uparam(10)    thermal exponent m        */
  int    icc, vp, israte, pad;    /* iparam(2..4)                            */
} orad_law2_t;

/* Solid property + time-step control (hm_read_prop14.F:348-362, lectur.F:410-416). */
typedef struct {
  double qa, qb;                  /* GEO(14), GEO(15) bulk viscosity          */
  double hvis;                    /* GEO(13) hourglass coefficient h          */
  double dtfac;                   /* DTFAC1(1)                                */
} orad_sprop_t;

/* Element state of one brick (ELBUF GBUF for NPT=1, elbufdef_mod.F90:728-1069). */
typedef struct {
  double sig[6];                  /* co-rotational-frame stress xx yy zz xy yz zx */
  double eint;                    /* internal energy per unit initial volume      */
  double rho;                     /* current density                              */
  double vol0;                    /* GBUF%VOL (initial volume)                    */
  double pla;                     /* equivalent plastic strain                    */
  double qvis;                    /* bulk-viscosity pressure                      */
  double offg;                    /* GBUF%OFF                                     */
  double epsd;                    /* equivalent strain rate (output)              */
  double dmg;                     /* EPS_max reached flag                         */
} orad_bstate_t;

/* Per-element outputs of the element chain. */
typedef struct {
  double f[3][8];                 /* global corner forces F1j,F2j,F3j (j=node)    */
  double sti;                     /* STI after SCUMU3P's FOURTH scaling           */
  double dtx;                     /* DTFAC1 * DTX (element time-step candidate)   */
  double ehou;                    /* hourglass energy increment                   */
  double voln;                    /* current volume                               */
  int    mstop;                   /* 1 if a non-positive volume was found         */
  int    pad;
} orad_bout_t;

/* SFORC3 chain for one element (engine/source/elements/solid/solide/sforc3.F:125). */
void oracle_brick_element(const double xg[3][8], const double vg[3][8],
                          orad_bstate_t *st, double dt1,
                          const orad_law2_t *mat, const orad_sprop_t *prop,
                          orad_bout_t *out);

/* oracle_brick_element over n elements in turn, one thread (for timing the restatement
 * against the reference's own routines on the same group): xg, vg [n][3][8],
 * st [n] updated in place, out [n]. */
void oracle_brick_group(int n, const double *xg, const double *vg, orad_bstate_t *st,
                        double dt1, const orad_law2_t *mat, const orad_sprop_t *prop,
                        orad_bout_t *out);

/* Initial volume as the starter computes it (same determinant as SDERI3, global frame). */
double oracle_brick_volume(const double xg[3][8]);

/* ---- PARITH/ON element-corner CSR (FILLCNE + W_PON + GET_FSKY_ADDRESS) ---- */
/* ixs: 8*numels 0-based node ids, element-major ([e*8+k]); uid: user ids.
 * adsky: numnod+1 (0-based offsets); iads: 8*numels slot of (e,k).
 * Returns 0 on success. */
int oracle_build_pon(int numnod, int numels, const int32_t *ixs,
                     const int32_t *uid, int32_t *adsky, int32_t *iads);

/* ---- Nodal part of one explicit cycle, in RESOL order ----
 * time step (resol.F:2639-2640, 5857-5861, 6080, 6224-6226), ACCELE (accele.F:64-110),
 * BCS10/BCS1V (bcs10.F, bcs1.F), FIXVEL for a constant curve with its external work
 * (fixvel.F:333-374, 834-836), ECRIT's kinetic energies at V + DT1/2 A (ecrit.F:213-226,
 * 303-325), TT (resol.F:8256-8266), VELOCITY (velocity.F:56-83), DEPLA
 * (displacement.F:88-101).  a (ar) holds the assembled forces (moments) on entry and the
 * accelerations VELOCITY consumed on return. */
typedef struct {
  int numnod, rot;                /* rot != 0: rotational DOF (IRODDL = 1)          */
  double *x, *v, *d, *a;          /* 3*numnod each                                  */
  const double *ms;
  const int32_t *icodt;
  double *vr, *ar;                /* rot only                                       */
  const double *inr;
  const int32_t *icodr;
  const int32_t *fv_dir;          /* FIXVEL direction per node (1..3), 0 = none, or null */
  const double *fv_val;           /* imposed velocity per node                      */
  double *fv_mem;                 /* per node: FIXVEL's VEL(4,N) energy memory, or null */
  double tt, dt1, dt2, dt12, dt2old, dtmx;
  double wfext;                   /* external work of the imposed velocities (WFEXT) */
  int    ncycle, nthreads;
  double encin, enrot;            /* out: ECRIT's kinetic energies of this cycle     */
} orad_onodes_t;

void oracle_nodal_update(orad_onodes_t *m, double dt2t);

/* ---- /INTER/TYPE7 node-to-segment penalty contact (orad_oracle_int7.c) ---- */
typedef struct {
  int nsn, nrtm;
  const int32_t *nsv;             /* secondary nodes, 0-based                       */
  const int32_t *irect;           /* 4*nrtm segment nodes, 0-based; 3rd = 4th: triangle */
  const double *stfm;             /* segment stiffness STF                          */
  double *stfns;                  /* secondary node factor STFN (negated when a node
                                     is deactivated at full penetration)            */
  double gap, fric, visc, viscf;  /* GAP, FRIC, VISs (critical damping), VISf       */
  double econt, econtv, econtd;   /* cumulative elastic, friction, damping energies */
  int npen, nsky;                 /* last call: penetrating pairs, skyline entries  */
} orad_int7_t;

/* Candidates by I7TRI's box test (widened by tz), listed by (secondary, segment);
 * returns the count (entries beyond cap are counted, not written). */
int oracle_int7_candidates(const orad_int7_t *it, const double *x, double tz, int cap,
                           int32_t *cand_n, int32_t *cand_e);
/* I7COR3 + I7DST3 + I7FOR3 + I7ASS2 over a candidate list: skyline entries (fx, fy, fz,
 * stif per entry, node isky) in FSKYI order; lowers *dt2t; returns the entry count. */
int oracle_int7_forces(orad_int7_t *it, const double *x, const double *v, const double *ms,
                       int ncand, const int32_t *cand_n, const int32_t *cand_e, double dt1,
                       double *dt2t, double *fsky, int32_t *isky);
/* One cycle's contact for the whole-model cycles: entries and their per-node order. */
int oracle_int7_cycle(orad_int7_t *it, int numnod, const double *x, const double *v,
                      const double *ms, double dt1, double *dt2t, double **fsky,
                      int32_t **order, int32_t **first);

/* ---- Whole-model explicit cycle (RESOL order, resol.F:2586-9595) ---- */
typedef struct {
  int numnod, numels;
  double *x, *v, *d;              /* 3*numnod, node-major (Fortran X(3,NUMNOD))   */
  const double *ms;               /* numnod                                       */
  const int32_t *icodt;           /* numnod, BCS1V code (x=4,y=2,z=1)             */
  const int32_t *ixs;             /* 8*numels, 0-based                            */
  const int32_t *adsky, *iads;    /* PARITH/ON CSR                                */
  orad_bstate_t *st;              /* numels                                       */
  double *fsky;                   /* 4*8*numels: fx fy fz sti per slot            */
  double *a, *stifn;              /* 3*numnod, numnod (assembly outputs)          */
  orad_law2_t mat;
  orad_sprop_t prop;
  double tt, dt1, dt2, dt12, dt2old, dtmx;
  double ehour;                   /* cumulative hourglass energy                  */
  int    ncycle, mstop;
  int    nthreads;                /* OpenMP threads (0 = runtime default)         */
  int    pad;
  const int32_t *fv_dir;          /* FIXVEL direction per node (1..3), 0 = none, or null */
  const double *fv_val;           /* FIXVEL imposed velocity per node              */
  double *fv_mem;                 /* FIXVEL energy memory per node, or null        */
  double wfext;                   /* external work of the imposed velocities       */
  double encin, enrot;            /* ECRIT's kinetic energies of the last cycle    */
  orad_int7_t *contact;           /* /INTER/TYPE7, or null                          */
} orad_omodel_t;

/* One full cycle: element forces -> assembly -> dt -> accel -> BC -> vel -> disp. */
void oracle_cycle(orad_omodel_t *m);

/* Energies: internal sum(EINT*V0), kinetic 0.5*sum(m v^2). */
void oracle_energies(const orad_omodel_t *m, double *eint, double *ekin);

/* ======================= Belytschko-Tsay shells (CFORC3) ======================= */

/* Law 2 for shells: the PM slots and MAT_PARAM values the starter derives
 * (hm_read_mat02_jc.F90:234-339). */
typedef struct {
  double rho0, young, nu, shear;  /* PM(1), PM(20), PM(21), PM(22)            */
  double a11, a12, ssp;           /* PM(24), PM(25), PM(27) = sqrt(E/rho)     */
  double ca, cb, cn;              /* uparam(1..3)                              */
  double epmx, ymax;              /* uparam(4,5)  EPS_max, SIG_max             */
  double cc, epdr, mexp;          /* uparam(6,7,10)                            */
  double asrate_pm9;              /* PM(9): 2*pi*Fcut, EP20 when Fcut = 0      */
  int    icc, vp, israte, ipla;   /* iparam(2..4); IPARG(29) plasticity algo. */
} orad_shmat_t;

/* /PROP/SHELL (TYPE1) values used by CFORC3 (hm_read_prop01.F:155-244). */
typedef struct {
  double hm, hf, hr;              /* GEO(13..15) hourglass coefficients       */
  double dm;                      /* membrane damping GROUP_PARAM%VISC_DM     */
  double fsh;                     /* GEO(38) shear factor (5/6)               */
  double dtfac, dtmin;            /* DTFAC1(3), DTMIN1(3)                     */
  int    ish;                     /* GEO(37) ISHEAR                           */
  int    npt;                     /* through-thickness points, 2..5           */
  int    ishell;                  /* 1: Belytschko-Tsay (CFORC3), 24: QEPH (CZFORC3) */
  int    pad;
  double dn;                      /* QEPH: GEO(13) = Dn numerical damping (0.015) */
  double cvis;                    /* QEPH: GEO(17) = hourglass stiffness factor (1) */
} orad_shprop_t;

#define ORAD_SH_NPT_MAX 5
/* Element state (GBUF + LBUF of one BT shell, elbufdef_mod.F90). */
typedef struct {
  double forc[5], mom[3];         /* GBUF%FOR, GBUF%MOM (generalised stresses) */
  double thk;                     /* GBUF%THK                                   */
  double eint[2];                 /* GBUF%EINT membrane, bending (energy)       */
  double epsd;                    /* GBUF%EPSD element strain rate              */
  double hour[5];                 /* GBUF%HOURG                                 */
  double offg;                    /* GBUF%OFF                                   */
  double sig[ORAD_SH_NPT_MAX][5]; /* LBUF%SIG per point                         */
  double pla[ORAD_SH_NPT_MAX];    /* LBUF%PLA                                   */
  double epsdp[ORAD_SH_NPT_MAX];  /* LBUF%EPSD                                  */
  double hourx[7];                /* QEPH: GBUF%HOURG(6:12) (HOURG(1:5) in hour) */
} orad_shstate_t;

typedef struct {
  double f[3][4], m[3][4];        /* FSKY(1:6,corner) = -F, -M (CUPDT3P)        */
  double sti, stir;               /* FSKY(7), FSKY(8)                           */
  double dt;                      /* element time step (CDT3)                   */
  double ehou;                    /* hourglass energy increment                 */
  int    alive;                   /* dt enters DT2T (OFFG > 0 and OFF != 0)     */
  int    pad;
  double stic[4];                 /* FSKY(7) per corner (QEPH: STI*FACN)         */
} orad_shout_t;

/* CFORC3 chain for one element (engine/source/elements/shell/coque/cforc3.F:398-746).
 * tri: nodes 3 and 4 coincide (IXC(4)==IXC(5)). */
void oracle_shell_element(const double xg[3][4], const double vg[3][4], const double vrg[3][4],
                          double thke, int tri, orad_shstate_t *st, double dt1,
                          const orad_shmat_t *mat, const orad_shprop_t *prop,
                          orad_shout_t *out);

/* CZFORC3 chain (QEPH, Ishell=24) for one quadrilateral
 * (engine/source/elements/shell/coquez/czforc3.F:436-808). */
void oracle_qeph_element(const double xg[3][4], const double vg[3][4], const double vrg[3][4],
                         double thke, orad_shstate_t *st, double dt1,
                         const orad_shmat_t *mat, const orad_shprop_t *prop,
                         orad_shout_t *out);

/* QBAT (Ishell=12) element state: GBUF values of the BT layout (HOURG unused), then
 * per Gauss point g (the order of CBAFORC3's IS/IR loops) and layer p the LBUF
 * stress, plastic strain and strain rate, and the Gauss points' resultants
 * FORPG/MOMPG (elbufdef_mod.F90); 190 doubles. */
typedef struct {
  double forc[5], mom[3];         /* GBUF%FOR, GBUF%MOM: Gauss point averages   */
  double thk;                     /* GBUF%THK                                   */
  double eint[2];                 /* GBUF%EINT                                  */
  double epsd;                    /* GBUF%EPSD                                  */
  double unused[5];
  double offg;                    /* GBUF%OFF                                   */
  double sig[4][ORAD_SH_NPT_MAX][5];
  double pla[4][ORAD_SH_NPT_MAX];
  double epsdp[4][ORAD_SH_NPT_MAX];
  double forpg[4][5];             /* GBUF%FORPG                                 */
  double mompg[4][3];             /* GBUF%MOMPG                                 */
} orad_qbstate_t;

/* CBAFORC3 chain (QBAT, Ishell=12, 2 x 2 Gauss points) for one quadrilateral
 * (engine/source/elements/shell/coqueba/cbaforc3.F:440-1140).  out->ehou holds
 * CBAVISC's damping energy increment (PARTSAV(8)). */
void oracle_qbat_element(const double xg[3][4], const double vg[3][4], const double vrg[3][4],
                         double thke, orad_qbstate_t *st, double dt1,
                         const orad_shmat_t *mat, const orad_shprop_t *prop,
                         orad_shout_t *out);

/* COQINI tables for npt points (single-precision DATA values widened). */
void oracle_shell_tables(int npt, double *z0, double *wf, double *wm);
/* Derived constants used on the shell path, in a fixed order (for pinning). */
void oracle_shell_constants(double *out16);

/* PARITH/ON CSR for nc corners per element (FILLCNE + W_PON). */
int oracle_build_pon_nc(int numnod, int numel, int nc, const int32_t *conn,
                        const int32_t *uid, int32_t *adsky, int32_t *iads);

/* Shell model: 6 DOF nodes, FSKY slots of 8 values (F, M, STI, STIR). */
typedef struct {
  int numnod, numelc;
  double *x, *v, *vr, *d;         /* 3*numnod each                               */
  const double *ms, *inr;         /* nodal mass, rotational inertia             */
  const int32_t *icodt, *icodr;   /* BCS codes (x=4,y=2,z=1) on V and VR         */
  const int32_t *fv_dir;          /* FIXVEL direction per node (1..3), 0 = none  */
  const double *fv_val;           /* FIXVEL imposed velocity per node            */
  const int32_t *ixc;             /* 4*numelc, 0-based                           */
  const double *thke;             /* element thickness                           */
  const int32_t *adsky, *iads;    /* PARITH/ON CSR                               */
  orad_shstate_t *st;             /* Ishell 1 and 24                            */
  orad_qbstate_t *qst;            /* Ishell 12 (QBAT)                           */
  double *fsky;                   /* 8 per slot                                  */
  double *a, *ar, *stifn, *stifr;
  orad_shmat_t mat;
  orad_shprop_t prop;
  double tt, dt1, dt2, dt12, dt2old, dtmx;
  double ehour;
  int    ncycle, nthreads;
  double *fv_mem;                 /* FIXVEL energy memory per node, or null      */
  double wfext;                   /* external work of the imposed velocities     */
  double encin, enrot;            /* ECRIT's kinetic energies of the last cycle  */
  orad_int7_t *contact;           /* /INTER/TYPE7, or null                        */
} orad_oshell_t;

void oracle_shell_cycle(orad_oshell_t *m);
/* Energies: internal sum(EINT(1)+EINT(2)) (cbilan.F:182), kinetic with rotations. */
void oracle_shell_energies(const orad_oshell_t *m, double *eint, double *ekin);

#ifdef __cplusplus
}
#endif
#endif

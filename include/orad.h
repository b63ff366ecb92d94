/*
 * orad.h -- C ABI of the MI355X explicit brick element loop.
 *
 * The Fortran engine binds these through ISO_C_BINDING (see
 * openradioss_amd/fortran/orad_bind.f90); Python binds them through ctypes
 * (openradioss_amd/engine.py).  Every entry point returns an int status
 * (ORAD_OK = 0) in place of the reference's ANCMSG/ARRET error exits;
 * orad_last_error() gives the message.  Arrays are plain host pointers in the
 * reference's Fortran layouts; nothing here takes a device or torch type.
 *
 * Reference paths are relative to the OpenRadioss source root.
 */
#ifndef ORAD_H
#define ORAD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orad_ctx orad_ctx;

enum {
  ORAD_OK = 0,
  ORAD_EINVAL = 1,      /* bad argument (sizes, pointers, node ids)          */
  ORAD_EHIP = 2,        /* HIP runtime error                                 */
  ORAD_ESTATE = 3,      /* call out of order (mesh/material not set)         */
  ORAD_ENEGVOL = 4      /* non-positive element volume met (MSTOP=1)         */
};

enum { ORAD_FP64 = 0, ORAD_MIXED = 1 };  /* element arithmetic: FP64, or FP32 with FP64 nodes */

/* Cycle scalars as RESOL keeps them (resol.F:2639-2640, 6080, 6224-6226, 8256-8266). */
typedef struct {
  double tt, dt1, dt2, dt12, dt2old, dtmx;
  int ncycle, mstop;
} orad_time_t;

/* Create/destroy a context on one HIP device.  No reference counterpart: the
 * reference keeps these arrays in Fortran modules (RESOL locals, ELBUF_TAB). */
int orad_create(orad_ctx **ctx, int device, int precision);
int orad_destroy(orad_ctx *ctx);
const char *orad_last_error(const orad_ctx *ctx);

/* Nodes: X, V in X(3,NUMNOD) layout (FP64), nodal mass MS(NUMNOD), boundary
 * code ICODT(NUMNOD) with BCS1V's meaning (x=4, y=2, z=1; bcs1.F:46-76).
 * Replaces the NODES%X/V/MS/ICODT arrays read by RESOL (resol.F:4096-4131). */
int orad_set_nodes(orad_ctx *ctx, int numnod, const double *x, const double *v,
                   const double *ms, const int *icodt);

/* Bricks: IXS(NIXS=11,NUMELS) as the engine stores it (scoor3.F:106-121):
 * IXS(1)=material, IXS(2..9)=nodes (1-based), IXS(10)=property, IXS(11)=user id.
 * vol0 = initial volumes GBUF%VOL.  Builds the PARITH/ON CSR on the host with
 * FILLCNE/W_PON ordering (starter/source/spmd/domdec2.F:2100-2150,
 * starter/source/restart/ddsplit/w_pon.F:550-573) and initialises element state
 * (SIG=0, EINT=0, RHO=rho0, PLA=0, OFF=1) once a material is set. */
int orad_set_bricks(orad_ctx *ctx, int numels, const int *ixs, const double *vol0);

/* Law 2 (Johnson-Cook) parameters, layout of mat_param for M2LAW
 * (materials/mat/mat002/m2law.F:132-160):
 * p[0]=rho0 p[1]=G p[2]=K p[3]=A p[4]=B p[5]=n p[6]=EPS_max p[7]=SIG_max
 * p[8]=C p[9]=EPS0 p[10]=m;  ip[0]=ICC ip[1]=VP ip[2]=ISRATE. */
int orad_set_law2(orad_ctx *ctx, const double *p, const int *ip);

/* /PROP/SOLID viscosity and hourglass (GEO(14), GEO(15), GEO(13);
 * hm_read_prop14.F:348-362) and DTFAC1(1) (input/lectur.F:410-416). */
int orad_set_solid_prop(orad_ctx *ctx, double qa, double qb, double hvis, double dtfac);

/* PARITH/ON arrays as the starter writes them: ADSKY(NUMNOD+1) and IADS(8,NUMELS),
 * both 1-based (parith_on_mod.F90:38-73). */
int orad_get_pon(orad_ctx *ctx, int *adsky, int *iads);

/* Time state (RESOL scalars).  Defaults: TT=DT1=DT2=0, DT2OLD=DTINI/1.1 with
 * DTINI=1e6 (starter/source/starter/contrl.F:1108-1113), DTMX=1e20. */
int orad_set_time(orad_ctx *ctx, const orad_time_t *t);
int orad_get_time(orad_ctx *ctx, orad_time_t *t);

/* Drop-in for one element group, SFORC3 semantics (solide/sforc3.F:125-156,
 * called from forint.F:1122): elements NFT+1..NFT+NEL with the current X, V and
 * element state; writes their FSKY slots (SCUMU3P, scumu3p.F:259-308), updates
 * element state, and returns the group's DT2T candidate (min with *dt2t on entry). */
int orad_sforc3(orad_ctx *ctx, int nft, int nel, double dt1, double *dt2t);

/* ASSPAR4 (assembly/asspar4.F:157-175): ordered per-node sum of the FSKY slots
 * into A(3,NUMNOD) and STIFN(NUMNOD), returned to the host when pointers are
 * non-null. */
int orad_asspar4(orad_ctx *ctx, double *a, double *stifn);

/* Whole explicit cycles, device resident: forces (FORINT/SFORC3) -> ASSPAR4 ->
 * DT2 selection -> ACCELE -> BCS10 -> VELOCITY -> DEPLA (resol.F:2586-9595).
 * Replays a captured HIP graph; host traffic per cycle is nil, the time state is
 * read back once at the end.  Returns ORAD_ENEGVOL if an element volume went
 * non-positive (the reference sets MSTOP). */
int orad_run_cycles(orad_ctx *ctx, int ncycle, orad_time_t *t);

/* Same cycles launched eagerly with HIP events around each kernel on the
 * library stream; returns the mean kernel durations in milliseconds. */
int orad_profile_cycles(orad_ctx *ctx, int ncycle, double *elem_ms, double *node_ms);

/* Downloads (any pointer may be null). */
int orad_get_nodes(orad_ctx *ctx, double *x, double *v, double *d, double *a, double *stifn);
/* Element state, FP64 on output whatever the storage precision.
 * sig: SIG(NUMELS,6) component-major as in GBUF%SIG (elbufdef_mod.F90). */
int orad_get_brick_state(orad_ctx *ctx, double *sig, double *eint, double *rho,
                         double *pla, double *qvis, double *offg, double *epsd);
int orad_set_brick_state(orad_ctx *ctx, const double *sig, const double *eint,
                         const double *rho, const double *pla, const double *qvis,
                         const double *offg);
/* FSKY slots: fsky(4, 8*NUMELS) = (fx, fy, fz, stifn) in slot order. */
int orad_get_fsky(orad_ctx *ctx, double *fsky);

/* Energy balance terms (ECRIT, output/ecrit.F:500-508): internal sum(EINT*V0),
 * kinetic 0.5*sum(MS*V^2), cumulative hourglass energy (PARTSAV(8,:), shvis3.F:431-469).
 * Sums use a fixed reduction order, so they are run-to-run reproducible. */
int orad_energies(orad_ctx *ctx, double *eint, double *ekin, double *ehour);

#ifdef __cplusplus
}
#endif
#endif

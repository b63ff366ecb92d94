/*
 * orad.h -- C ABI of the MI355X explicit element loop (8-node bricks and
 * 4-node Belytschko-Tsay shells).
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

/* Launch configuration of the kernels (orad_set_launch, orad_autotune).
 * persistent: batched shells only -- a resident grid takes 256-element tiles from
 * a counter instead of one workgroup per tile; lds_kb: dynamic LDS reserved per
 * element workgroup, which caps resident workgroups per CU (-1 = the instance's
 * default; bricks and batched shells); xcd_remap: give each of the 8 XCDs a
 * contiguous run of workgroups; grid: the persistent grid size (output).
 * Bricks also choose: block and waves -- the element-kernel instance (threads per
 * workgroup, minimum waves per SIMD its registers are allocated for; mixed form
 * 256/2 (default), 256/3, 128/2, 64/2; FP64 form 64/1 (default), 128/1, 256/1,
 * 64/2; 0 = default); order_rows -- the library's element order, columns of that
 * many element rows walked layer by layer (1..128, default 10; 0 = the engine's order,
 * -1 = keep the current one).  Every family: node_tile -- PARITH/ON slots the nodal
 * kernel stages per LDS tile (1024, 2048, 4096 within 48 KB; 0 = default).
 * adj_vjp -- the element vector-Jacobian product of orad_adjoint_reverse: 0 = the
 * hand-written reverse pass where the element has one (bricks; Belytschko-Tsay and
 * QEPH shells with the law-2 radial return, IPLA = 0), forward-mode dual numbers
 * otherwise; 1 =
 * forward-mode dual numbers always (both give the same derivative to rounding).
 * orad_get_launch reports the values in use.  Nodes, slots and element state do
 * not depend on any of these; the element order changes only the rounding of the
 * cumulative hourglass energy (its per-workgroup partial sums). */
typedef struct {
  int persistent, lds_kb, xcd_remap, grid;
  int block, waves, order_rows, node_tile;
  int adj_vjp;
} orad_launch_t;

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
 * (SIG=0, EINT=0, RHO=rho0, PLA=0, OFF=1) once a material is set.  Slots are
 * addressed by 32-bit indices like the reference's IADS, so NUMELS <= 268435455
 * (ORAD_EINVAL beyond). */
int orad_set_bricks(orad_ctx *ctx, int numels, const int *ixs, const double *vol0);

/* Law 2 (Johnson-Cook) parameters, layout of mat_param for M2LAW
 * (materials/mat/mat002/m2law.F:132-160):
 * p[0]=rho0 p[1]=G p[2]=K p[3]=A p[4]=B p[5]=n p[6]=EPS_max p[7]=SIG_max
 * p[8]=C p[9]=EPS0 p[10]=m;  ip[0]=ICC ip[1]=VP ip[2]=ISRATE. */
int orad_set_law2(orad_ctx *ctx, const double *p, const int *ip);

/* /PROP/SOLID viscosity and hourglass (GEO(14), GEO(15), GEO(13);
 * hm_read_prop14.F:348-362) and DTFAC1(1) (input/lectur.F:410-416). */
int orad_set_solid_prop(orad_ctx *ctx, double qa, double qb, double hvis, double dtfac);

/* PARITH/ON arrays as the starter writes them: ADSKY(NUMNOD+1) and IADS(8,NUMELS)
 * for bricks or IADC(4,NUMELC) for shells, both 1-based (parith_on_mod.F90:38-73). */
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
/* The same group call without waiting (no reference counterpart: SFORC3 returns when
 * its group is done): records the element pass of elements nft+1..nft+nel and returns;
 * consecutive groups with the same DT1 run as one launch (each element is computed on
 * its own, so the results are those of separate calls), launched at the next call that
 * does not continue the run or at any other entry point.  orad_group_wait returns once
 * every group has run, lowers *dt2t to their minimum (the merge of the FORINT group
 * loop, resol.F:4134-4140) and returns ORAD_ENEGVOL if a brick saw a non-positive
 * volume.  orad_cforc3_async is the shell counterpart of orad_cforc3. */
int orad_sforc3_async(orad_ctx *ctx, int nft, int nel, double dt1);
int orad_cforc3_async(orad_ctx *ctx, int nft, int nel, double dt1);
int orad_group_wait(orad_ctx *ctx, double *dt2t);
/* The engine's X(3,NUMNOD), V(3,NUMNOD) and, for shells, VR(3,NUMNOD) (null for
 * bricks) after its own ACCELE/BCS10/VELOCITY/DEPLA (accele.F, bcs10.F, velocity.F,
 * displacement.F), handed to the device before the next group loop (no reference
 * counterpart: the engine's arrays are the kernels' in the reference).  Unlike
 * orad_set_nodes / orad_set_nodes_rot it keeps MS, IN, the boundary codes, D, A,
 * STIFN and the captured cycle graphs. */
int orad_update_xv(orad_ctx *ctx, const double *x, const double *v, const double *vr);

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
/* FSKY slots in slot order: bricks fsky(4, 8*NUMELS) = (fx, fy, fz, STI);
 * shells fsky(8, 4*NUMELC) = (fx, fy, fz, mx, my, mz, STI, STIR) as CUPDT3P
 * writes them (cupdt3.F:1060-1067). */
int orad_get_fsky(orad_ctx *ctx, double *fsky);

/* Energy balance terms (ECRIT, output/ecrit.F:500-508): internal energy (bricks
 * sum(EINT*V0); shells sum(EINT(1)+EINT(2)), cbilan.F:182), kinetic energy
 * 0.5*sum(MS*V^2) (+ 0.5*sum(IN*VR^2) with rotations), cumulative hourglass energy
 * (PARTSAV(8,:), shvis3.F:431-469, chvis3.F:366-393).  Sums use a fixed reduction
 * order, so they are run-to-run reproducible. */
int orad_energies(orad_ctx *ctx, double *eint, double *ekin, double *ehour);

/* Energy balance as ECRIT lists it (output/ecrit.F:192-226, 303-331, 480-508), for the
 * last cycle run (one model per context):
 *   encin, enrot  kinetic energies at the listing's velocity V + DT1/2 A (the nodal
 *                 kernel sums them per workgroup between FIXVEL and VELOCITY, where
 *                 SORTIE_MAIN calls ECRIT);
 *   enint         internal energy (as orad_energies); ehour the cumulative hourglass
 *                 energy, listed apart and not part of ENTOT, as in ECRIT;
 *   wfext         work of the imposed velocities since the context was created: per
 *                 imposed node DT1*DW of the cycle plus the DT2*DW kept from the one
 *                 before, DW = MS/4 (A DT12 + 2V)(A - AOLD)
 *                 (constraints/general/impvel/fixvel.F:333-374, 834-836);
 *   entot = encin + enint + enrot; entot0 = ENTOT - WFEXT of the first cycle run by
 *   orad_run_cycles; err = entot / (entot0 + wfext) - 1 (ERR, x100 in the listing).
 * Sums are in a fixed order (run-to-run reproducible); WFEXT is summed per node, then
 * over nodes, where the reference sums its entries each cycle. */
typedef struct {
  double encin, enrot, enint, ehour, wfext, entot, entot0, err, tt;
  int ncycle, pad;
} orad_balance_t;
int orad_energy_balance(orad_ctx *ctx, orad_balance_t *b);

/* orad_energies without any device allocation: the same kernels and results
 * (bitwise), with their partial sums in a fixed 1 MiB scratch array that the
 * library keeps on each GPU for the life of the process.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight monitoring loops of a serving process, harnesses that
 * forbid hipMalloc).  Limit: ceil(elements/256) + ceil(nodes/256) + 2 <= 131072
 * (about 33 million elements plus nodes), ORAD_EINVAL beyond.  Every context on
 * one GPU shares that GPU's array: calls that overlap on one GPU (from several host
 * threads) race; calls made one after another do not. */
int orad_energies_noalloc(orad_ctx *ctx, double *eint, double *ekin, double *ehour);

/* Imposed velocities, FIXVEL semantics for a constant velocity curve
 * (constraints/general/impvel/fixvel.F:362-375): after BCS10, the imposed
 * component's acceleration becomes (VEL - V)/DT12.  nodes 1-based, dir 1..3
 * (x, y, z).  Replaces the previous set; n = 0 clears it.  Both element families. */
int orad_set_fixvel(orad_ctx *ctx, int n, const int *nodes, const int *dir, const double *vel);

/* ------------------------------ shells ------------------------------ */

/* Rotational DOF of the nodes (call after orad_set_nodes): VR(3,NUMNOD),
 * nodal inertia IN(NUMNOD) and ICODR(NUMNOD) (same bit meaning as ICODT on
 * VR/AR).  Replaces NODES%VR, NODES%IN, NODES%ICODR read by RESOL
 * (resol.F:6639-6641 ACCELE, 7022 BCS10). */
int orad_set_nodes_rot(orad_ctx *ctx, const double *vr, const double *in, const int *icodr);

/* Shells: IXC(NIXC=7,NUMELC) as the engine stores it (ccoor3.F:69-78):
 * IXC(1)=material, IXC(2..5)=nodes (1-based, IXC(5)=IXC(4) for a triangle),
 * IXC(6)=property, IXC(7)=user id; thke = element thickness THKE.  Builds the
 * PARITH/ON CSR like orad_set_bricks and initialises the state (THK=THKE, OFF=1).
 * NUMELC <= 536870911 (32-bit slot indices, as IADC). */
int orad_set_shells(orad_ctx *ctx, int numelc, const int *ixc, const double *thke);

/* Law 2 for shells: the starter's derived values (hm_read_mat02_jc.F90:234-339)
 * p[0]=rho0 p[1]=E p[2]=nu p[3]=G p[4]=A11 p[5]=A12 p[6]=SSP p[7]=A p[8]=B
 * p[9]=n p[10]=EPS_max p[11]=SIG_max p[12]=C p[13]=EPS0 p[14]=m p[15]=PM(9);
 * ip[0]=ICC ip[1]=VP ip[2]=ISRATE ip[3]=IPLA (0 radial return, 1 Newton,
 * 2 approximate normal: m2cplr.F:174-377, IPARG(29) from cgrtails.F:823-829). */
int orad_set_shell_law2(orad_ctx *ctx, const double *p, const int *ip);

/* /PROP/SHELL TYPE1: g[0..2]=hm, hf, hr (GEO(13..15)), g[3]=DM membrane damping
 * (GROUP_PARAM%VISC_DM), g[4]=shear factor GEO(38), g[5]=DTFAC1(3),
 * g[6]=DTMIN1(3); ish=GEO(37); npt = 2..5 points through the thickness
 * (hm_read_prop01.F:155-244).  Ishell=1, Ismstr=2, Ithick=0 are implied. */
int orad_set_shell_prop(orad_ctx *ctx, const double *g, int ish, int npt);

/* Shell formulation, /PROP/SHELL Ishell (IPARG(23), hm_read_prop01.F:180-229):
 * 1 = Belytschko-Tsay (CFORC3, the default), 24 = QEPH (CZFORC3,
 * shell/coquez/czforc3.F:436-808) with numerical damping dn = Dn (GEO(13); 0
 * means the default 0.015) and CVIS = 1, 12 = QBAT (CBAFORC3,
 * shell/coqueba/cbaforc3.F:440-1140, 2 x 2 Gauss points) with dn = GEO(13) = Hm,
 * the membrane damping of CBAVISC (0 means EM3, cncoef3.F:421-431).  Must precede
 * orad_set_shells: QEPH elements keep 60 state values (HOURG(6:12) after the 53
 * of BT), QBAT elements 190 (columns 0..17 as BT, then per Gauss point g and layer
 * p: SIG at 18+5(5g+p), PLA at 118+5g+p, EPSD at 138+5g+p, FORPG at 158+5g,
 * MOMPG at 178+3g); both keep one STI per diagonal (FSKY(7) = STI*FACN, CUPDTN3P
 * cupdtn3.F:488-700).  Batched models and the adjoint take Ishell 1 or 24. */
int orad_set_shell_ishell(orad_ctx *ctx, int ishell, double dn);

/* Drop-in for one shell group, CFORC3 semantics (shell/coque/cforc3.F:398-746,
 * called from forintc.F:420), or CZFORC3 for QEPH: elements NFT+1..NFT+NEL;
 * writes their FSKY slots (CUPDT3P / CUPDTN3P), updates element state, lowers
 * *dt2t to the group minimum (CDT3 / CNDT3). */
int orad_cforc3(orad_ctx *ctx, int nft, int nel, double dt1, double *dt2t);

/* ASSPAR4 with rotations: A(3,NUMNOD), AR(3,NUMNOD), STIFN, STIFR. */
int orad_asspar4_rot(orad_ctx *ctx, double *a, double *ar, double *stifn, double *stifr);

/* Downloads of VR(3,NUMNOD) and AR(3,NUMNOD) (any pointer may be null). */
int orad_get_nodes_rot(orad_ctx *ctx, double *vr, double *ar);

/* Shell element state as rows of 53 FP64 values per element:
 * FOR(5) MOM(3) THK EINT(2) EPSD HOUR(5) OFF, then SIG(5) for points 1..5,
 * PLA(1..5), point EPSD(1..5) (GBUF/LBUF of elbufdef_mod.F90); QEPH rows have
 * 60 values, HOURG(6:12) appended (HOURG(1:5) in the HOUR slots). */
int orad_get_shell_state(orad_ctx *ctx, double *rows);
int orad_set_shell_state(orad_ctx *ctx, const double *rows);

/* Batched independent shell models in one context (no reference counterpart: a
 * reference run is one model; the nearest analogue is its dynamic schedule over
 * uneven element groups, forint.F:325, forintc.F:140-160).  Model m owns nodes
 * node_off[m]..node_off[m+1]-1 and shells elem_off[m]..elem_off[m+1]-1 (0-based,
 * offsets span the whole mesh; a shell may only use nodes of its own model) and
 * has its own law-2 record p[16*m..], ip[4*m..] (layout of orad_set_shell_law2;
 * null = the context's material for all) and its own time scalars and time step.
 * Every cycle runs all models in one element launch over 256-element tiles and one
 * nodal launch; results per model are bitwise those of the model run alone. */
int orad_set_shell_models(orad_ctx *ctx, int nmodel, const int *node_off, const int *elem_off,
                          const double *p, const int *ip);
int orad_get_model_time(orad_ctx *ctx, int model, orad_time_t *t);
/* Energies of one model; orad_energies sums all models. */
int orad_model_energies(orad_ctx *ctx, int model, double *eint, double *ekin, double *ehour);

/* Launch configuration (no reference counterpart; the reference fixes its group
 * size NVSIZ per architecture, archloops.inc:32-64).  Results do not depend on it. */
int orad_set_launch(orad_ctx *ctx, const orad_launch_t *l);
int orad_get_launch(orad_ctx *ctx, orad_launch_t *l);
/* Times `cycles` whole cycles under each candidate configuration from the current
 * state, restores the state after every trial, and keeps the fastest; *best_ms = its
 * time per cycle.  Bricks: element order (engine, 8, 10, 16, 32 rows), then element
 * instance (block, waves) with its LDS cap, XCD mapping and the nodal slot tile,
 * searched one axis at a time from the best so far; batched shells: LDS cap x XCD
 * mapping x persistent tiles, then the slot tile; one shell model: XCD mapping and
 * slot tile. */
int orad_autotune(orad_ctx *ctx, int cycles, orad_launch_t *best, double *best_ms);

/* /INTER/TYPE7 node-to-segment penalty contact, one interface per context (in place of
 * the interface part of I7MAINF, interfaces/int07/i7mainf.F:529-737, and ASSPAR4's
 * interface skyline, assembly/asspar4.F:126-154, 560-583).  The path: constant gap
 * (IGAP=0), IGSTI<=1 (STIF = STFM(segment) * |STFNS(node)|), INACTI=0, no curvature,
 * no heat exchange, no tied contact, IVIS2=0 critical damping, Coulomb friction in the
 * total viscous formulation without filtering (IFQ=0, MFROT=0, isotropic).
 *
 * orad_set_contact7: NSV(NSN) secondary nodes and IRECT(4,NRTM) main segments, 1-based
 * (IRECT(3)=IRECT(4): a triangle), their stiffness STFM(NRTM) and factor STFNS(NSN)
 * (as the starter writes INTBUF_TAB%STFM/STFNS), params = GAP, FRIC, VISs (critical
 * damping, the reference default 0.05), VISf (friction damping); cand_cap bounds the
 * penetrating pairs of one pass (<= 0: 8 per secondary node).  Every cycle of
 * orad_run_cycles then takes the candidates on the device (I7TRI's box test widened by
 * 1.5 GAP, searched again once a node has moved GAP / 4 since the last search, as the
 * reference re-sorts when its margin TZINF - GAP is used up), computes I7COR3/I7DST3/I7FOR3/I7ASS2 for the penetrating pairs in batches
 * of 128, lowers DT2T with the contact time step and adds each node's skyline entries
 * after its element slots, in skyline order.  The entries are the reference routines'
 * bit for bit on the same candidate batches; the batches are the penetrating pairs in
 * (secondary node, segment) order, where the reference batches the candidates its
 * bucket sort keeps.  A pass that outgrows cand_cap, or a secondary node with more
 * than 64 candidate segments or 16 penetrated ones, fails the run with ORAD_EINVAL.
 * With batched models (orad_set_shell_models first) the interface is one per model in
 * effect: NSV must increase, a pair's node and segment must lie in one model, a batch
 * never mixes models and each model's contact time step lowers its own DT2T, so every
 * model runs as it would alone (orad_contact7 is then refused: the contact runs in
 * orad_run_cycles).
 *
 * orad_contact7: one pass at the current X, V with the given DT1 (the group path:
 * call it where I7MAINF runs; the next orad_asspar4 adds the entries); lowers *dt2t.
 * orad_get_contact7: cumulative energies (ECONT elastic, ECONTV friction, ECONTD
 * damping = 0 on this path), counts[4] = candidates, penetrating pairs and skyline
 * entries of the last pass and the overflow flags; stfns (NSN) with the nodes that
 * reached full penetration negated, as I7FOR3 marks them (i7for3.F:945-961).
 * orad_get_contact7_sky: the last pass's skyline entries FSKYI(1:4) and their nodes
 * ISKY (1-based) in skyline order; *nsky their count (null outputs: count only). */
int orad_set_contact7(orad_ctx *ctx, int nsn, const int *nsv, int nrtm, const int *irect,
                      const double *stfm, const double *stfns, const double *params,
                      int cand_cap);
int orad_contact7(orad_ctx *ctx, double dt1, double *dt2t);
int orad_get_contact7(orad_ctx *ctx, double *energies, int *counts, double *stfns);
int orad_get_contact7_sky(orad_ctx *ctx, int cap, double *fsky, int *isky, int *nsky);

/* Restart files of the dynamic state (replace the element-loop part of the engine's
 * restart, output/restart/wrrestp.F:1142-1150 for X, D, V, VR and :1273 for the
 * element buffers, read back by output/restart/rdresb.F:1055-1059): orad_write_restart
 * writes the nodal X, V, D (and VR for shells), every element state column and the
 * per-element nodal stiffness, the time record of each model (TT, DT1, DT2, NCYCLE),
 * the energy balance's running sums and ENTOT0, in the layout the library holds them
 * in (one file, written to `path`.tmp and renamed).  orad_read_restart loads such a
 * file into a context built from the same model (same family, precision, Ishell,
 * sizes, models and PARITH/ON slot table: a fingerprint checks them) and the run
 * continues bit for bit as if it had not stopped.  Mesh, masses, materials,
 * properties, imposed velocities and launch settings are the model's and are not in
 * the file.  ORAD_EINVAL names the mismatch, a truncated or corrupt file (checksum)
 * or an I/O failure; nothing is changed on the device unless the whole file passed. */
int orad_write_restart(orad_ctx *ctx, const char *path);
int orad_read_restart(orad_ctx *ctx, const char *path);

/* Adjoint of a run of one shell or brick model (no reference counterpart: the
 * reference's only derivative is the implicit consistent tangent of law 2,
 * m2law.F:487-531, m2cplr.F:448-469).  orad_adjoint_forward runs ncycle cycles from
 * the current state as orad_run_cycles does, keeping a checkpoint every `every` cycles.
 * orad_adjoint_reverse takes the adjoints of the final state -- lam_x, lam_v,
 * lam_vr, lam_d as X(3,NUMNOD) arrays, lam_state as one row per element in the
 * engine's numbering: shells 33 for Ishell=1 (SIG(5) of points 1..5, PLA(1..5),
 * HOUR(1:3)) or 42 for QEPH (SIG, PLA, HOURG(1:12)), bricks 7 (SIG(6), PLA); null =
 * zero -- and, recomputing each segment from its checkpoint, returns the gradients
 * with respect to the law-2 A, B, n (g_jc[3]), the shell thicknesses THKE
 * (g_thk[NUMELC]; not touched for bricks), the nodal masses and rotational inertias
 * (g_ms, g_in [NUMNOD]; g_in zero for bricks), and the adjoints of the initial X, V,
 * VR (lam_vr0 zero for bricks); any output may be null.  The time-step sequence is
 * taken as fixed (exact when DTMX pins it).  The element derivatives are those of the
 * FP64 chain evaluated on the forward states, in FP64 for either form: bricks and
 * both shell formulations with IPLA = 0 by a hand-written reverse pass, IPLA 1/2 by
 * dual numbers (orad_launch_t.adj_vjp); the context is left at the end of the
 * forward run.  Law 2 without the strain-rate term only (C = 0; ORAD_EINVAL
 * otherwise): the rate filters (EPSD, EPSDP) and the shell thickness THK that the
 * rate factor reads are not part of the adjoint state. */
int orad_adjoint_forward(orad_ctx *ctx, int ncycle, int every);
int orad_adjoint_reverse(orad_ctx *ctx, const double *lam_x, const double *lam_v,
                         const double *lam_vr, const double *lam_d, const double *lam_state,
                         double *g_jc, double *g_thk, double *g_ms, double *g_in,
                         double *lam_x0, double *lam_v0, double *lam_vr0);

#ifdef __cplusplus
}
#endif
#endif

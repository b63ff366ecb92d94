// orad_common.hpp -- device data layout shared by the kernels and the C ABI.
//
// HBM layout (one context, one GPU):
//   nodes (FP64 always):  X4[n] = (x, y, z, MS), V4[n] = (vx, vy, vz, node word),
//                         D[3n] displacement, A[3n] / STIFN[n] assembled outputs.
//                         The node word holds the node's boundary codes (node_word).
//   bricks (T = double for the FP64 form, float for the mixed form), structure of
//   arrays indexed by the library's element order (a permutation of the engine's,
//   see orad_brick_order.hip):  conn[k*ne+e] (int32, k = corner), rank[e] (eight
//   8-bit corner ranks: corner k's slot is adsky[node] + byte k), sig[c*ne+e]
//   (c = 0..5), eint, rho, pla, qvis, offg, epsd; vol0[e] is FP64 in both forms
//   (the reference's single-precision build keeps VOL0DP too).
//   PARITH/ON slots: fsky[s] = (fx, fy, fz) in precision T, slots in node-major
//   CSR order (adsky[n]..adsky[n+1]-1); sti[e] = STI/4 once per element.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace orad {

struct Law2 {
  double rho0, shear, bulk, ca, cb, cn, epmx, sigmx, cc, epdr, mexp;
  int icc, vp, israte;
};

struct SProp {
  double qa, qb, hvis, dtfac;
};

// Cycle scalars, double-buffered by cycle parity so that one kernel can read the
// current cycle's values while the next cycle's are written.
struct Scal {
  double tt, dt1, dt2, dt12, dt2old, dtmx;
  int ncycle, mstop;
};

struct CycleCtl {
  Scal s[2];
  unsigned long long dt2t[2];   // bits of the positive element dt minimum
  int mstop;                    // sticky: non-positive volume seen
  int pad;
  double ehour;                 // cumulative hourglass energy
};

// One PARITH/ON slot: the corner force (FSKY(1:3,·)).  STI (FSKY(7,·)) is the same
// for the eight corners of a brick and is kept once per element.
template <typename T> struct Slot3 { T x, y, z; };

// Node word (V4.w): the node's codes as a small whole number held in a double.
__host__ __device__ __forceinline__ double node_word(int code) { return (double)code; }
__host__ __device__ __forceinline__ int word_code(double w) { return (int)w; }

template <typename T>
struct BrickArgs {
  int numels;
  const int32_t *conn;      // [8][ne], library element order
  const uint64_t *rank;     // [ne]: byte k = slot of corner k within its node's slots
  const int32_t *emap;      // engine element -> library element (group calls), or null
  const int32_t *adsky;     // [numnod+1] first slot of each node (ADSKY, 0-based)
  const double4 *X4;
  const double4 *V4;
  T *sig;                   // [6][ne]
  T *eint, *rho, *pla, *qvis, *offg, *epsd;
  const double *vol0;       // initial volumes, FP64 in both forms (VOL0DP)
  T *sti;                   // STI/4 per element (SCUMU3P's FSKY(7,·) value)
  Slot3<T> *fsky;
  double *ehou_part;        // per-block hourglass energy partials
  CycleCtl *ctl;
  Law2 mat;
  SProp prop;
};

// ---------------------------------------------------------------- shells --
// Law 2 as the shell chain reads it: the starter's PM values and MAT_PARAM
// (hm_read_mat02_jc.F90:234-339) and the plasticity algorithm IPLA (cgrtails.F:823-829).
struct ShMat {
  double rho0, young, nu, shear, a11, a12, ssp;
  double ca, cb, cn, epmx, ymax, cc, epdr, mexp, asrate_pm9;
  int icc, vp, israte, ipla;
};

// /PROP/SHELL values (hm_read_prop01.F:155-244) with the COQINI through-thickness
// tables for npt points (positions z0, force weights wf, moment weights wm).
struct ShProp {
  double hm, hf, hr, dm, fsh, dtfac, dtmin;
  int ish, npt;
  double z0[5], wf[5], wm[5];
  // formulation: 1 Belytschko-Tsay (CFORC3), 24 QEPH (CZFORC3) with its numerical
  // damping Dn (GEO(13)) and hourglass stiffness factor CVIS (GEO(17), always 1),
  // 12 QBAT (CBAFORC3) with GEO(13) = Hm as its membrane damping DN (EM3 when 0)
  int ishell;
  double dn, cvis;
};

// constant_mod values used on the shell path, derived on the host from their
// definitions (the same double operations as the reference's parameters).
struct ShConst {
  double zep81, zep072169, onep414, four_over_3, one_over_9, four_over_5, em30, sr2d2;
  double em15, em20, em07, three_over_4, third, three_half;
  // QEPH chain
  double em08, em10, em18, one_over_12, one_over_16, one_over_64, zep85, zep999, fivep333;
  double threep464, zep9ep30;
  // QBAT chain
  double em12, em03, zep3;
};

// One shell PARITH/ON slot: corner force and moment (FSKY(1:6,·)).  STI
// (FSKY(7,·)) is kept once per element (Belytschko-Tsay) or once per diagonal
// (QEPH: sti[e] for corners 1,3, sti[ne+e] for corners 2,4); STIR (FSKY(8,·)) is
// zero on this path.
template <typename T> struct Slot6 { T fx, fy, fz, mx, my, mz; };

// Shell element state columns (structure of arrays, column c at st[c*ne + e]).
enum {
  SC_FOR = 0, SC_MOM = 5, SC_THK = 8, SC_EINT = 9, SC_EPSD = 11, SC_HOUR = 12, SC_OFFG = 17,
  SC_SIG = 18,     // + 5*pt + component
  SC_PLA = 43,     // + pt
  SC_EPSDP = 48,   // + pt
  SC_NCOL = 53,
  // QEPH (Ishell=24) keeps HOURG(6:12) after the Belytschko-Tsay columns
  SC_HOURX = 53, SC_NCOL_QEPH = 60,
  // QBAT (Ishell=12): columns 0..17 as above (HOURG unused), then per Gauss point
  // g = 0..3 (CBAFORC3's IS/IR order) and layer p = 0..4
  QC_SIG = 18,     // + 5*(5*g + p) + component
  QC_PLA = 118,    // + 5*g + p
  QC_EPSDP = 138,  // + 5*g + p
  QC_FORPG = 158,  // + 5*g + component (GBUF%FORPG)
  QC_MOMPG = 178,  // + 3*g + component (GBUF%MOMPG)
  SC_NCOL_QBAT = 190
};

template <typename T>
struct ShellArgs {
  int numelc;
  const int32_t *conn;      // [4][ne]
  const int32_t *iadc;      // [4][ne]
  const double4 *X4;        // (x, y, z, MS)
  const double4 *V4;        // (vx, vy, vz, codes)
  const double4 *VR4;       // (vrx, vry, vrz, IN)
  const double *thke;       // element thickness THKE (FP64)
  T *st;                    // [SC_NCOL][ne]; [SC_NCOL_QEPH][ne] QEPH; [SC_NCOL_QBAT][ne] QBAT
  T *sti;                   // STI per element (FSKY(7,·)), per diagonal for QEPH and QBAT
  Slot6<T> *fsky;
  CycleCtl *ctl;
  ShMat mat;
  ShProp prop;
  ShConst k;
};

// Node record codes (V4.w): ICODT bits 0-2, ICODR bits 3-5, FIXVEL direction
// (1..3, 0 = none) bits 6-7.
__host__ __device__ __forceinline__ int code_icodt(int c) { return c & 7; }
__host__ __device__ __forceinline__ int code_icodr(int c) { return (c >> 3) & 7; }
__host__ __device__ __forceinline__ int code_fvdir(int c) { return (c >> 6) & 3; }

struct NodeArgs {
  int numnod;
  const int32_t *adsky;     // [numnod+1]
  double4 *X4;
  double4 *V4;
  double *D;
  double *A;
  double *STIFN;
  const void *fsky;         // Slot3<float> or Slot3<double> slots
  const void *sti;          // per-element STI (float or double), read by ASSPAR4 only
  const int32_t *slot_elem; // element of each slot, for STIFN in ASSPAR4
  const double *ehou_part;
  int n_ehou_part;
  CycleCtl *ctl;
  // rotational DOF (shells) and imposed velocities
  double4 *VR4;             // (vrx, vry, vrz, IN), null for bricks
  double *AR;               // angular accelerations (assembled moments in ASSPAR4)
  double *STIFR;
  const double *fv_val;     // FIXVEL value per node, null when none is imposed
  // energy balance (one model per context; null otherwise): per-workgroup kinetic
  // energies of the cycle at parity p, bal_part[p][2][nblocks] (translation, rotation),
  // and per imposed node the FIXVEL energy memory DT2*DW and the running external work
  double *bal_part;
  double *fv_mem;
  double *fv_work;
  // batched independent models: ctl[model], node -> model, first node of each model
  const int32_t *node_model;   // null: one model (ctl[0])
  const int32_t *model_node0;
  unsigned *tile_queue;        // persistent element-tile counter, reset here for the next cycle
  int nmodel;                  // batched: one extra workgroup per model sums its chunk energies
  const int32_t *model_chunks; // [nmodel+1]
  const double *chunk_ehou;
  // /INTER/TYPE7 skyline entries (null: none): node n's entries form the list
  // ci_head[n], ci_next[e], ... (-1 ends it), added after its element slots in
  // increasing entry order (ASSPAR4's skyline order)
  const int32_t *ci_head;
  const int32_t *ci_next;
  const double4 *ci_sky;
};

// Batched models (one launch over uneven models): elements of model m are
// [elem_off[m], elem_off[m+1]), cut into 256-element tiles from the model's first
// element (the last one ragged); tiles[t] = (model, first element, end element, t).  Each model has its own material, time scalars and
// time step; the hourglass energy is summed per chunk in a fixed order (four
// 64-lane wave sums, left to right) and, by the nodal kernel, per model over its
// chunks, so results do not depend on the workgroup size or the launch form.
struct BatchArgs {
  const int4 *tiles;
  int ntiles;
  const ShMat *mats;            // [nmodel]
  CycleCtl *ctl;                // [nmodel]
  unsigned *queue;              // next tile (persistent launch)
  double *chunk_ehou;           // [ntiles]
};

// Logical block index that gives each XCD a contiguous run of blocks (the
// hardware deals workgroups round-robin over the 8 XCDs, each with its own L2).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int NX = 8;
  int xcd = b % NX, local = b / NX, per = nb / NX, rem = nb % NX;
  return (xcd < rem) ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

}  // namespace orad

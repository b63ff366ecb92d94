// orad_api.hip -- the C ABI declared in include/orad.h.
//
// Host-side role of the reference's starter and RESOL driver for this path:
// owns device memory, builds the PARITH/ON element-corner CSR with the
// starter's ordering, and sequences the per-cycle kernels (captured once per
// cycle parity in a HIP graph and replayed).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/orad.h"
#include "orad_common.hpp"

#include "orad_ctx.hpp"
#include "orad_adjoint_args.hpp"

#include <mutex>

// Scratch of orad_energies_noalloc: the per-workgroup partial sums and the two
// results of the energy reduction, in a file-scope device array instead of a
// per-call allocation.  Each GPU the library runs on holds its own copy until the
// process ends; host code reaches it only through the per-device address taken
// with hipGetSymbolAddress at context creation (the array's name is not a device
// address on the host).
static constexpr size_t NOALLOC_SCRATCH_DOUBLES = 131072;   // 1 MiB
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];
static constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[MAX_DEVICES] = {};

// Address of this device's scratch array; called with the device current.
static hipError_t scratch_address(int device, double **out) {
  *out = nullptr;
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = (double *)p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

template <typename T>
static BrickArgs<T> brick_args(orad_ctx *c) {
  BrickArgs<T> a{};
  const size_t ne = (size_t)c->numels;
  T *base = (T *)c->state;
  a.numels = c->numels;
  a.conn = c->conn;
  a.rank = c->rank;
  a.emap = nullptr;
  a.adsky = c->adsky;
  a.X4 = c->X4;
  a.V4 = c->V4;
  a.sig = base;
  a.eint = base + 6 * ne;
  a.rho = base + 7 * ne;
  a.pla = base + 8 * ne;
  a.qvis = base + 9 * ne;
  a.offg = base + 10 * ne;
  a.epsd = base + 11 * ne;
  a.vol0 = c->vol0;
  a.sti = (T *)c->sti;
  a.fsky = (Slot3<T> *)c->fsky;
  a.ctl = c->ctl;
  a.mat = c->mat;
  a.prop = c->prop;
  return a;
}

static NodeArgs node_args(orad_ctx *c) {
  NodeArgs a{};
  a.numnod = c->numnod;
  a.adsky = c->adsky;
  a.X4 = c->X4;
  a.V4 = c->V4;
  a.D = c->D;
  a.A = c->A;
  a.STIFN = c->STIFN;
  a.fsky = c->fsky;
  a.sti = c->sti;
  a.slot_elem = c->slot_elem;
  a.ehou_part = c->ehou_part;
  a.n_ehou_part = c->n_ehou_part;
  a.ctl = c->ctl;
  a.VR4 = c->family == FAM_SHELL ? c->VR4 : nullptr;
  a.AR = c->AR;
  a.STIFR = c->STIFR;
  a.fv_val = c->fv_val;
  a.bal_part = c->bal_part;
  a.fv_mem = c->fv_mem;
  a.fv_work = c->fv_work;
  if (c->batched) {
    // the energy balance is kept for one model per context
    a.bal_part = nullptr;
    a.fv_mem = nullptr;
    a.fv_work = nullptr;
    a.node_model = c->node_model;
    a.model_node0 = c->model_node0;
    a.tile_queue = c->queue;
    a.ehou_part = nullptr;       // per-model totals from the chunk sums instead
    a.n_ehou_part = 0;
    a.nmodel = c->nmodel;
    a.model_chunks = c->model_tiles;
    a.chunk_ehou = c->chunk_ehou;
  }
  if (c->ci) {
    a.ci_head = c->ci->node_head;
    a.ci_next = c->ci->sky_next;
    a.ci_sky = c->ci->sky;
  }
  return a;
}

// The assembly-only launch (orad_asspar4): the interface entries of the last
// orad_contact7 call are added once, then dropped.
static NodeArgs assembly_args(orad_ctx *c) {
  NodeArgs a = node_args(c);
  if (!c->ci_fresh) a.ci_head = nullptr;
  c->ci_fresh = false;
  return a;
}

static BatchArgs batch_args(orad_ctx *c) {
  BatchArgs b{};
  b.tiles = c->tiles;
  b.ntiles = c->ntiles;
  b.mats = c->mats;
  b.ctl = c->ctl;
  b.queue = c->queue;
  b.chunk_ehou = c->chunk_ehou;
  return b;
}

template <typename T>
static ShellArgs<T> shell_args(orad_ctx *c) {
  ShellArgs<T> a{};
  a.numelc = c->numelc;
  a.conn = c->conn;
  a.iadc = c->iads;
  a.X4 = c->X4;
  a.V4 = c->V4;
  a.VR4 = c->VR4;
  a.thke = c->thke;
  a.st = (T *)c->state;
  a.sti = (T *)c->sti;
  a.fsky = (Slot6<T> *)c->fsky;
  a.ctl = c->ctl;
  a.mat = c->shmat;
  a.prop = c->shprop;
  a.k = c->shk;
  return a;
}

// [e0, e1): the library's element order, or with `group` one engine-numbered group
// (bricks map it through emap; shells keep the engine's order).
static hipError_t launch_elements(orad_ctx *c, int e0, int e1, const double *dt1p,
                                  unsigned long long *dt2t, double *ehou_part,
                                  bool group = false) {
  if (c->family == FAM_SHELL) {
    if (c->qbat()) {
      if (c->precision == ORAD_MIXED)
        return launch_qbat_mixed(shell_args<float>(c), e0, e1, dt1p, dt2t, ehou_part,
                                 c->xcd_remap, c->stream);
      return launch_qbat_fp64(shell_args<double>(c), e0, e1, dt1p, dt2t, ehou_part, c->xcd_remap,
                              c->stream);
    }
    if (c->qeph()) {
      if (c->precision == ORAD_MIXED)
        return launch_qeph_mixed(shell_args<float>(c), e0, e1, dt1p, dt2t, ehou_part,
                                 c->xcd_remap, c->stream);
      return launch_qeph_fp64(shell_args<double>(c), e0, e1, dt1p, dt2t, ehou_part, c->xcd_remap,
                              c->stream);
    }
    if (c->precision == ORAD_MIXED)
      return launch_shell_mixed(shell_args<float>(c), e0, e1, dt1p, dt2t, ehou_part, c->xcd_remap,
                                c->stream);
    return launch_shell_fp64(shell_args<double>(c), e0, e1, dt1p, dt2t, ehou_part, c->xcd_remap,
                             c->stream);
  }
  if (c->precision == ORAD_MIXED) {
    BrickArgs<float> a = brick_args<float>(c);
    if (group) a.emap = c->emap;
    return launch_brick_mixed(a, e0, e1, dt1p, dt2t, ehou_part, c->xcd_remap, c->elem_lds,
                              c->elem_block, c->elem_waves, c->stream);
  }
  BrickArgs<double> a = brick_args<double>(c);
  if (group) a.emap = c->emap;
  return launch_brick_fp64(a, e0, e1, dt1p, dt2t, ehou_part, c->xcd_remap, c->elem_lds,
                           c->elem_block, c->elem_waves, c->stream);
}

// One cycle's launches at parity p (no host synchronisation, graph-capturable).
static int nelem(const orad_ctx *c) { return c->family == FAM_SHELL ? c->numelc : c->numels; }

// The element launch of a whole cycle at parity p.
static hipError_t launch_cycle_elements(orad_ctx *c, int p) {
  if (c->batched) {
    const int grid = c->persist ? std::min(c->batch_grid, c->ntiles) : c->ntiles;
    if (c->qeph()) {
      if (c->precision == ORAD_MIXED)
        return launch_qeph_batch_mixed(shell_args<float>(c), batch_args(c), p, c->persist, grid,
                                       c->batch_lds, c->xcd_remap, c->stream);
      return launch_qeph_batch_fp64(shell_args<double>(c), batch_args(c), p, c->persist, grid,
                                    c->batch_lds, c->xcd_remap, c->stream);
    }
    if (c->precision == ORAD_MIXED)
      return launch_shell_batch_mixed(shell_args<float>(c), batch_args(c), p, c->persist, grid,
                                      c->batch_lds, c->xcd_remap, c->stream);
    return launch_shell_batch_fp64(shell_args<double>(c), batch_args(c), p, c->persist, grid,
                                   c->batch_lds, c->xcd_remap, c->stream);
  }
  return launch_elements(c, 0, nelem(c), &c->ctl->s[p].dt2, &c->ctl->dt2t[p], c->ehou_part);
}

// The contact pass of a cycle (I7MAINF's place, before ASSPAR4): DT1 of the cycle is
// the time record's DT2, the contact time step lowers the cycle's DT2T.
static hipError_t launch_cycle_contact(orad_ctx *c, int p) {
  if (!c->ci) return hipSuccess;
  return contact7_run(*c->ci, c->X4, c->V4, &c->ctl[0].s[p].dt2, &c->ctl[0].dt2t[p],
                      sizeof(CycleCtl), c->stream);
}

static hipError_t enqueue_cycle(orad_ctx *c, int p) {
  hipError_t e = launch_cycle_elements(c, p);
  if (e != hipSuccess) return e;
  e = launch_cycle_contact(c, p);
  if (e != hipSuccess) return e;
  return launch_node(node_args(c), c->precision == ORAD_MIXED, p, 1, c->cycle_write_a,
                     c->xcd_remap, c->stream, c->node_tile);
}

static int init_state(orad_ctx *c) {
  if (!(c->have_bricks && c->have_mat) || c->state_init) return ORAD_OK;
  const size_t ne = (size_t)c->numels;
  // SIG=0 EINT=0 RHO=rho0 PLA=0 QVIS=0 OFF=1 EPSD=0 (starter initial state).
  std::vector<char> all(NSTATE * ne * c->tsize(), 0);
  auto put = [&](int col, double v) {
    for (size_t e = 0; e < ne; ++e) {
      if (c->precision == ORAD_MIXED)
        ((float *)all.data())[col * ne + e] = (float)v;
      else
        ((double *)all.data())[col * ne + e] = v;
    }
  };
  put(C_RHO, c->mat.rho0);
  put(C_OFFG, 1.0);
  HIPCHK(c, hipMemcpy(c->state, all.data(), all.size(), hipMemcpyHostToDevice));
  c->state_init = true;
  return ORAD_OK;
}

static int get_fsky_shell(orad_ctx *c, double *fsky);
static void free_adjoint(orad_ctx *c);

static void free_contact(orad_ctx *c) {
  Contact7 *k = c->ci;
  if (!k) return;
  for (void *p : {(void *)k->nsv, (void *)k->irect, (void *)k->stfm, (void *)k->stfns,
                  (void *)k->bounds, (void *)k->bpart, (void *)k->grid, (void *)k->seg_lo,
                  (void *)k->seg_hi, (void *)k->seg_cell, (void *)k->cell_count,
                  (void *)k->cell_fill, (void *)k->cell_start, (void *)k->cell_lo,
                  (void *)k->cell_hi,
                  (void *)k->node_keep, (void *)k->node_nkeep,
                  (void *)k->node_cand, (void *)k->node_ncand, (void *)k->node_kstart,
                  (void *)k->counters, (void *)k->sky,
                  (void *)k->sky_node, (void *)k->sky_next, (void *)k->batch_cnt,
                  (void *)k->batch_e, (void *)k->energy, (void *)k->node_head,
                  (void *)k->scan_tmp, (void *)k->seg_model, (void *)k->nsv_off,
                  (void *)k->model_bstart, (void *)k->model_ext,
                  (void *)k->x_search, (void *)k->dpart})
    dfree(p);
  delete k;
  c->ci = nullptr;
  c->ci_fresh = false;
}

// Re-packs every node word (V4.w): ICODT kept, ICODR and the FIXVEL direction from
// h_code_hi.
static int pack_node_words(orad_ctx *c) {
  if (!c->have_nodes) return ORAD_OK;
  const size_t n = (size_t)c->numnod;
  std::vector<double4> hv(n);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(hv.data(), c->V4, n * sizeof(double4), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i) {
    const int code = word_code(hv[i].w);
    const int hi = c->h_code_hi.empty() ? 0 : c->h_code_hi[i];
    hv[i].w = node_word(code_icodt(code) | hi);
  }
  HIPCHK(c, hipMemcpy(c->V4, hv.data(), n * sizeof(double4), hipMemcpyHostToDevice));
  return ORAD_OK;
}

static int flush_groups(orad_ctx *c);

extern "C" {

int orad_create(orad_ctx **out, int device, int precision) {
  if (!out) return ORAD_EINVAL;
  *out = nullptr;
  if (precision != ORAD_FP64 && precision != ORAD_MIXED) return ORAD_EINVAL;
  orad_ctx *c = new (std::nothrow) orad_ctx();
  if (!c) return ORAD_EINVAL;
  c->device = device;
  c->precision = precision;
  if (const char *xr = std::getenv("ORAD_XCD_REMAP")) c->xcd_remap = std::atoi(xr) != 0;
  if (const char *el = std::getenv("ORAD_ELEM_LDS_KB")) c->elem_lds = 1024 * std::atoi(el);
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = scratch_address(device, &c->noalloc_scratch);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = dalloc(&c->ctl, sizeof(CycleCtl));
  if (e == hipSuccess) e = dalloc(&c->scratch_dt1, sizeof(double));
  if (e == hipSuccess) e = dalloc(&c->scratch_dt2t, sizeof(unsigned long long));
  for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipEventCreate(&c->ev[i]);
  if (e != hipSuccess) {
    orad_destroy(c);
    return ORAD_EHIP;
  }
  orad_time_t t{};
  t.dt2old = 1.0e6 / 1.1;  // DTINI/ONEP1
  t.dtmx = 1.0e20;
  CycleCtl h{};
  for (int p = 0; p < 2; ++p) {
    h.s[p].tt = t.tt; h.s[p].dt1 = t.dt1; h.s[p].dt2 = t.dt2; h.s[p].dt12 = t.dt12;
    h.s[p].dt2old = t.dt2old; h.s[p].dtmx = t.dtmx;
    h.dt2t[p] = 0x7FF0000000000000ull;
  }
  if (hipMemcpy(c->ctl, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) {
    orad_destroy(c);
    return ORAD_EHIP;
  }
  *out = c;
  return ORAD_OK;
}

int orad_destroy(orad_ctx *c) {
  if (!c) return ORAD_OK;
  c->pend = false;          // groups recorded and never waited for are dropped
  (void)hipSetDevice(c->device);
  drop_graphs(c);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_adjoint(c);
  for (void *p : {(void *)c->X4, (void *)c->V4, (void *)c->D, (void *)c->A, (void *)c->STIFN,
                  (void *)c->conn, (void *)c->iads, (void *)c->adsky, c->state, (void *)c->vol0,
                  c->fsky, c->sti, (void *)c->slot_elem,
                  (void *)c->ehou_part, (void *)c->ctl, (void *)c->scratch_dt1,
                  (void *)c->scratch_dt2t, (void *)c->VR4, (void *)c->AR, (void *)c->STIFR,
                  (void *)c->thke, (void *)c->fv_val, (void *)c->node_model,
                  (void *)c->model_node0, (void *)c->model_tiles, (void *)c->tiles,
                  (void *)c->mats, (void *)c->queue, (void *)c->chunk_ehou, (void *)c->rank,
                  (void *)c->emap, (void *)c->bal_part, (void *)c->fv_mem, (void *)c->fv_work,
                  (void *)c->xv_stage})
    dfree(p);
  free_contact(c);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return ORAD_OK;
}

const char *orad_last_error(const orad_ctx *c) { return c ? c->err.c_str() : "null context"; }

int orad_set_nodes(orad_ctx *c, int numnod, const double *x, const double *v, const double *ms,
                   const int *icodt) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (numnod <= 0 || !x || !v || !ms) return fail(c, ORAD_EINVAL, "set_nodes: bad arguments");
  if (c->have_nodes && numnod != c->numnod)
    return fail(c, ORAD_EINVAL, "set_nodes: node count differs from the mesh");
  const size_t n = (size_t)numnod;
  // validate and pack on the host before any device allocation
  std::vector<double4> hx(n), hv(n);
  for (size_t i = 0; i < n; ++i) {
    hx[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], ms[i]);
    const int code = icodt ? icodt[i] : 0;
    if (code < 0 || code > 7) return fail(c, ORAD_EINVAL, "set_nodes: ICODT outside 0..7");
    const int hi = c->h_code_hi.empty() ? 0 : c->h_code_hi[i];
    hv[i] = make_double4(v[3 * i], v[3 * i + 1], v[3 * i + 2], node_word(code | hi));
  }
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  if (!c->X4) HIPCHK(c, dalloc(&c->X4, n * sizeof(double4)));
  if (!c->V4) HIPCHK(c, dalloc(&c->V4, n * sizeof(double4)));
  if (!c->D) HIPCHK(c, dalloc(&c->D, 3 * n * sizeof(double)));
  if (!c->A) HIPCHK(c, dalloc(&c->A, 3 * n * sizeof(double)));
  if (!c->STIFN) HIPCHK(c, dalloc(&c->STIFN, n * sizeof(double)));
  if (!c->bal_part) {
    const size_t nbn = (n + 255) / 256;
    HIPCHK(c, dalloc(&c->bal_part, 4 * nbn * sizeof(double)));
    HIPCHK(c, hipMemset(c->bal_part, 0, 4 * nbn * sizeof(double)));
    HIPCHK(c, dalloc(&c->fv_mem, n * sizeof(double)));
    HIPCHK(c, hipMemset(c->fv_mem, 0, n * sizeof(double)));
    HIPCHK(c, dalloc(&c->fv_work, n * sizeof(double)));
    HIPCHK(c, hipMemset(c->fv_work, 0, n * sizeof(double)));
  }
  HIPCHK(c, hipMemcpy(c->X4, hx.data(), n * sizeof(double4), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->V4, hv.data(), n * sizeof(double4), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemset(c->D, 0, 3 * n * sizeof(double)));
  HIPCHK(c, hipMemset(c->A, 0, 3 * n * sizeof(double)));
  HIPCHK(c, hipMemset(c->STIFN, 0, n * sizeof(double)));
  c->numnod = numnod;
  c->have_nodes = true;
  return ORAD_OK;
}

int orad_set_bricks(orad_ctx *c, int numels, const int *ixs, const double *vol0) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!c->have_nodes) return fail(c, ORAD_ESTATE, "set_bricks: call orad_set_nodes first");
  if (c->family != FAM_NONE) return fail(c, ORAD_ESTATE, "set_bricks: mesh already set");
  if (numels <= 0 || !ixs || !vol0) return fail(c, ORAD_EINVAL, "set_bricks: bad arguments");
  // FSKY slots are addressed by 32-bit indices, as the reference's IADS (INTEGER)
  if (numels > INT_MAX / 8)
    return fail(c, ORAD_EINVAL, "set_bricks: more than 268435455 bricks (8 slots each, 32-bit IADS)");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t ne = (size_t)numels;
  const int nn = c->numnod;
  // Connectivity (IXS(2..9), 1-based) and user ids (IXS(11)).
  std::vector<int32_t> conn_em(8 * ne), uid(ne);
  for (size_t e = 0; e < ne; ++e) {
    for (int k = 0; k < 8; ++k) {
      const int nd = ixs[11 * e + 1 + k];
      if (nd < 1 || nd > nn) return fail(c, ORAD_EINVAL, "set_bricks: node id out of range");
      conn_em[8 * e + k] = nd - 1;
    }
    uid[e] = ixs[11 * e + 10];
  }
  // PARITH/ON CSR.  FILLCNE: elements in increasing user id (stable), each
  // appended to its corners' nodes; W_PON/GET_FSKY_ADDRESS: slot of (e,k) is
  // the next free position of node IXS(k+1,e) in that order.
  std::vector<int32_t> cnt(nn + 1, 0);
  for (size_t i = 0; i < 8 * ne; ++i) cnt[conn_em[i]]++;
  c->h_adsky.assign(nn + 1, 0);
  for (int n = 0; n < nn; ++n) c->h_adsky[n + 1] = c->h_adsky[n] + cnt[n];
  for (int n = 0; n < nn; ++n) cnt[n] = c->h_adsky[n];
  std::vector<int32_t> order(ne);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return uid[a] < uid[b]; });
  c->h_iads.assign(8 * ne, 0);
  for (size_t j = 0; j < ne; ++j) {
    const int32_t e = order[j];
    for (int k = 0; k < 8; ++k) c->h_iads[8 * (size_t)e + k] = cnt[conn_em[8 * (size_t)e + k]]++;
  }
  // Corner ranks: slot of (e,k) = ADSKY(node) + rank, rank < 256 (8 bits per corner).
  std::vector<uint64_t> rank(ne, 0);
  for (size_t e = 0; e < ne; ++e)
    for (int k = 0; k < 8; ++k) {
      const int32_t r = c->h_iads[8 * e + k] - c->h_adsky[conn_em[8 * e + k]];
      if (r < 0 || r > 255) {
        c->h_adsky.clear();
        c->h_iads.clear();
        return fail(c, ORAD_EINVAL, "set_bricks: a node is shared by more than 256 brick corners");
      }
      rank[e] |= (uint64_t)r << (8 * k);
    }
  // Device copies in the engine's order, corner-major for coalesced loads; the
  // library's column order is applied below.
  std::vector<int32_t> conn_cm(8 * ne);
  for (size_t e = 0; e < ne; ++e)
    for (int k = 0; k < 8; ++k) conn_cm[k * ne + e] = conn_em[8 * e + k];
  HIPCHK(c, dalloc(&c->conn, 8 * ne * sizeof(int32_t)));
  HIPCHK(c, dalloc(&c->rank, ne * sizeof(uint64_t)));
  HIPCHK(c, dalloc(&c->adsky, (nn + 1) * sizeof(int32_t)));
  HIPCHK(c, hipMemcpy(c->conn, conn_cm.data(), 8 * ne * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->rank, rank.data(), ne * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->adsky, c->h_adsky.data(), (nn + 1) * 4, hipMemcpyHostToDevice));
  const size_t ts = c->tsize();
  HIPCHK(c, dalloc(&c->state, NSTATE * ne * ts));
  HIPCHK(c, hipMemset(c->state, 0, NSTATE * ne * ts));
  HIPCHK(c, dalloc(&c->fsky, 8 * ne * 3 * ts));
  HIPCHK(c, hipMemset(c->fsky, 0, 8 * ne * 3 * ts));
  HIPCHK(c, dalloc(&c->sti, ne * ts));
  HIPCHK(c, hipMemset(c->sti, 0, ne * ts));
  {
    std::vector<int32_t> se(8 * ne);
    for (size_t e = 0; e < ne; ++e)
      for (int k = 0; k < 8; ++k) se[c->h_iads[8 * e + k]] = (int32_t)e;
    HIPCHK(c, dalloc(&c->slot_elem, 8 * ne * sizeof(int32_t)));
    HIPCHK(c, hipMemcpy(c->slot_elem, se.data(), 8 * ne * 4, hipMemcpyHostToDevice));
  }
  // hourglass-energy partials: one per element workgroup, room for the smallest (64)
  const int bb = c->precision == ORAD_MIXED ? brick_block_mixed() : brick_block_fp64();
  c->n_ehou_part = (int)((ne + bb - 1) / bb);
  HIPCHK(c, dalloc(&c->ehou_part, ((ne + 63) / 64) * sizeof(double)));
  HIPCHK(c, hipMemset(c->ehou_part, 0, ((ne + 63) / 64) * sizeof(double)));
  HIPCHK(c, dalloc(&c->vol0, ne * sizeof(double)));
  HIPCHK(c, hipMemcpy(c->vol0, vol0, ne * sizeof(double), hipMemcpyHostToDevice));
  c->numels = numels;
  c->have_bricks = true;
  c->family = FAM_BRICK;
  c->nc = 8;
  drop_graphs(c);
  int r = brick_bins(c, ne, conn_em, vol0);
  if (!r) r = brick_reorder(c, c->order_rows, {});
  if (!r) r = init_state(c);
  return r;
}

int orad_set_law2(orad_ctx *c, const double *p, const int *ip) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!p || !ip) return fail(c, ORAD_EINVAL, "set_law2: null parameters");
  if (!(p[0] > 0.0) || !(p[1] > 0.0) || !(p[2] > 0.0))
    return fail(c, ORAD_EINVAL, "set_law2: rho0, G and K must be positive");
  if (!(p[3] > 0.0)) return fail(c, ORAD_EINVAL, "set_law2: A must be positive (m2law input check)");
  Law2 m{};
  m.rho0 = p[0]; m.shear = p[1]; m.bulk = p[2]; m.ca = p[3]; m.cb = p[4]; m.cn = p[5];
  m.epmx = p[6]; m.sigmx = p[7]; m.cc = p[8]; m.epdr = p[9]; m.mexp = p[10];
  m.icc = ip[0]; m.vp = ip[1]; m.israte = ip[2];
  if (m.cc != 0.0 && !(m.epdr > 0.0)) return fail(c, ORAD_EINVAL, "set_law2: EPS0 must be > 0");
  c->mat = m;
  c->have_mat = true;
  drop_graphs(c);
  return init_state(c);
}

int orad_set_solid_prop(orad_ctx *c, double qa, double qb, double hvis, double dtfac) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (qa < 0.0 || qb < 0.0 || hvis < 0.0 || !(dtfac > 0.0))
    return fail(c, ORAD_EINVAL, "set_solid_prop: negative coefficient or DTFAC <= 0");
  c->prop = SProp{qa, qb, hvis, dtfac};
  c->have_prop = true;
  drop_graphs(c);
  return ORAD_OK;
}

int orad_get_pon(orad_ctx *c, int *adsky, int *iads) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (c->family == FAM_NONE) return fail(c, ORAD_ESTATE, "get_pon: no mesh");
  if (adsky)
    for (int n = 0; n <= c->numnod; ++n) adsky[n] = c->h_adsky[n] + 1;
  if (iads)
    for (size_t i = 0; i < c->h_iads.size(); ++i) iads[i] = c->h_iads[i] + 1;
  return ORAD_OK;
}

int orad_set_time(orad_ctx *c, const orad_time_t *t) {
  if (!c || !t) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->host_ncycle = t->ncycle;
  const int p = c->host_ncycle & 1;
  Scal s{t->tt, t->dt1, t->dt2, t->dt12, t->dt2old, t->dtmx, t->ncycle, t->mstop};
  unsigned long long inf = 0x7FF0000000000000ull;
  for (int m = 0; m < c->nmodel; ++m) {
    HIPCHK(c, hipMemcpy(&c->ctl[m].s[p], &s, sizeof(s), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(&c->ctl[m].dt2t[p], &inf, sizeof(inf), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(&c->ctl[m].mstop, &t->mstop, sizeof(int), hipMemcpyHostToDevice));
  }
  return ORAD_OK;
}

static int model_time(orad_ctx *c, int m, orad_time_t *t) {
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  Scal s;
  int ms = 0;
  HIPCHK(c, hipMemcpy(&s, &c->ctl[m].s[c->host_ncycle & 1], sizeof(s), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(&ms, &c->ctl[m].mstop, sizeof(int), hipMemcpyDeviceToHost));
  t->tt = s.tt; t->dt1 = s.dt1; t->dt2 = s.dt2; t->dt12 = s.dt12; t->dt2old = s.dt2old;
  t->dtmx = s.dtmx; t->ncycle = s.ncycle; t->mstop = ms;
  return ORAD_OK;
}

int orad_get_time(orad_ctx *c, orad_time_t *t) {
  if (!c || !t) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  return model_time(c, 0, t);
}

static int ready(orad_ctx *c) {
  if (c->family == FAM_SHELL) {
    if (!(c->have_nodes && c->have_rot && c->have_shells && c->have_shmat && c->have_shprop))
      return fail(c, ORAD_ESTATE,
                  "shell mesh, nodes with rotations, material and property must be set first");
    return ORAD_OK;
  }
  if (!(c->have_nodes && c->have_bricks && c->have_mat && c->have_prop))
    return fail(c, ORAD_ESTATE, "mesh, nodes, material and property must be set first");
  return ORAD_OK;
}

}  // extern "C"

// Launches the pending run of groups (bricks: SFORC3, shells: CFORC3 or CZFORC3) on the
// context's stream.  Its time-step minimum goes into the device accumulator
// scratch_dt2t, reset by the first launch after a wait; dt1 is uploaded when it changes.
static int group_launch(orad_ctx *c, int fam, int b, int e, double dt1) {
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->group_open) {
    c->group_bits = 0x7FF0000000000000ull;
    HIPCHK(c, hipMemcpyAsync(c->scratch_dt2t, &c->group_bits, sizeof(c->group_bits),
                             hipMemcpyHostToDevice, c->stream));
  }
  if (!c->group_open || dt1 != c->group_dt1[c->group_slot]) {
    // a ring of host values, so that an earlier upload's source outlives its copy
    c->group_slot = (c->group_slot + 1) % 8;
    c->group_dt1[c->group_slot] = dt1;
    HIPCHK(c, hipMemcpyAsync(c->scratch_dt1, &c->group_dt1[c->group_slot], sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
  }
  c->group_open = true;
  if (fam == FAM_BRICK) {
    HIPCHK(c, launch_elements(c, b, e, c->scratch_dt1, c->scratch_dt2t, nullptr, true));
    return ORAD_OK;
  }
  const ShMat keep = c->shmat;
  if (c->batched && e > b) {
    // a group of batched models lies within one model and takes that model's material
    const int m = (int)(std::upper_bound(c->h_elem_off.begin(), c->h_elem_off.end(), b) -
                        c->h_elem_off.begin()) - 1;
    if (e > c->h_elem_off[m + 1])
      return fail(c, ORAD_EINVAL, "cforc3: a group of batched models must lie in one model");
    c->shmat = c->h_mats[m];
  }
  hipError_t le = launch_elements(c, b, e, c->scratch_dt1, c->scratch_dt2t, nullptr);
  c->shmat = keep;
  HIPCHK(c, le);
  return ORAD_OK;
}

// Launches the groups recorded by the asynchronous group calls, if any.  Every entry
// point but the asynchronous group calls does this first, so that what it reads or
// enqueues follows the groups.
static int flush_groups(orad_ctx *c) {
  if (!c->pend) return ORAD_OK;
  c->pend = false;
  return group_launch(c, c->pend_fam, c->pend_b, c->pend_e, c->pend_dt1);
}

// One group call: validated, then recorded; consecutive groups of the engine's
// numbering with the same DT1 are run by one launch (a group's elements are computed
// independently, so the slots, state and time-step minimum are those of separate
// launches).
static int group_enqueue(orad_ctx *c, int fam, int nft, int nel, double dt1) {
  const char *who = fam == FAM_BRICK ? "sforc3" : "cforc3";
  if (c->family != fam)
    return fail(c, ORAD_ESTATE, std::string(who) + ": the mesh is not of " +
                                    (fam == FAM_BRICK ? "bricks" : "shells"));
  int r = ready(c);
  if (r) return r;
  const int ne = fam == FAM_BRICK ? c->numels : c->numelc;
  if (nft < 0 || nel < 0 || nel > ne - nft)
    return fail(c, ORAD_EINVAL, std::string(who) + ": group outside the element range");
  if (c->batched) {
    // batched models: one model's material per launch, so no merging across groups
    if ((r = flush_groups(c))) return r;
    return group_launch(c, fam, nft, nft + nel, dt1);
  }
  if (c->pend && c->pend_fam == fam && c->pend_e == nft && c->pend_dt1 == dt1) {
    c->pend_e = nft + nel;
    return ORAD_OK;
  }
  if ((r = flush_groups(c))) return r;
  c->pend = true;
  c->pend_fam = fam;
  c->pend_b = nft;
  c->pend_e = nft + nel;
  c->pend_dt1 = dt1;
  return ORAD_OK;
}

// Waits for the queued groups; lowers *dt2t to their minimum; MSTOP of bricks.
static int group_wait(orad_ctx *c, double *dt2t) {
  int r = flush_groups(c);
  if (r) return r;
  if (!c->group_open) return ORAD_OK;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(&c->group_bits, c->scratch_dt2t, sizeof(c->group_bits),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->group_open = false;
  double g;
  std::memcpy(&g, &c->group_bits, sizeof(g));
  if (dt2t && g < *dt2t) *dt2t = g;
  if (c->family == FAM_BRICK) {
    int ms = 0;
    HIPCHK(c, hipMemcpy(&ms, &c->ctl->mstop, sizeof(int), hipMemcpyDeviceToHost));
    if (ms) return fail(c, ORAD_ENEGVOL, "non-positive element volume (MSTOP)");
  }
  return ORAD_OK;
}

extern "C" {

int orad_sforc3(orad_ctx *c, int nft, int nel, double dt1, double *dt2t) {
  if (!c || !dt2t) return ORAD_EINVAL;
  int r = group_wait(c, nullptr);
  if (!r) r = group_enqueue(c, FAM_BRICK, nft, nel, dt1);
  if (!r) r = group_wait(c, dt2t);
  return r;
}

int orad_sforc3_async(orad_ctx *c, int nft, int nel, double dt1) {
  if (!c) return ORAD_EINVAL;
  return group_enqueue(c, FAM_BRICK, nft, nel, dt1);
}

int orad_cforc3_async(orad_ctx *c, int nft, int nel, double dt1) {
  if (!c) return ORAD_EINVAL;
  return group_enqueue(c, FAM_SHELL, nft, nel, dt1);
}

int orad_group_wait(orad_ctx *c, double *dt2t) {
  if (!c || !dt2t) return ORAD_EINVAL;
  return group_wait(c, dt2t);
}

int orad_update_xv(orad_ctx *c, const double *x, const double *v, const double *vr) {
  if (!c || !x || !v) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!c->have_nodes) return fail(c, ORAD_ESTATE, "update_xv: call orad_set_nodes first");
  if (vr && !c->have_rot)
    return fail(c, ORAD_ESTATE, "update_xv: rotational velocities need orad_set_nodes_rot");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t n = (size_t)c->numnod;
  if (!c->xv_stage) HIPCHK(c, dalloc(&c->xv_stage, 9 * n * sizeof(double)));
  HIPCHK(c, hipMemcpyAsync(c->xv_stage, x, 3 * n * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->xv_stage + 3 * n, v, 3 * n * 8, hipMemcpyHostToDevice, c->stream));
  if (vr)
    HIPCHK(c, hipMemcpyAsync(c->xv_stage + 6 * n, vr, 3 * n * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, launch_pack_xv(c->numnod, c->xv_stage, c->X4, c->V4, vr ? c->VR4 : nullptr,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ORAD_OK;
}

int orad_asspar4(orad_ctx *c, double *a, double *stifn) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  int r = ready(c);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, launch_node(assembly_args(c), c->precision == ORAD_MIXED, c->host_ncycle & 1, 0, 1,
                        c->xcd_remap, c->stream, c->node_tile));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (a) HIPCHK(c, hipMemcpy(a, c->A, 3 * (size_t)c->numnod * 8, hipMemcpyDeviceToHost));
  if (stifn) HIPCHK(c, hipMemcpy(stifn, c->STIFN, (size_t)c->numnod * 8, hipMemcpyDeviceToHost));
  return ORAD_OK;
}

static int energies_range(orad_ctx *c, size_t e0, size_t e1, size_t n0, size_t n1, int m0,
                          int m1, double *eint, double *ekin, double *ehour,
                          double *scr = nullptr);

// ECRIT's energy terms after the last cycle run (single-model contexts): the kinetic
// energies the nodal kernel summed per workgroup at the listing's velocity, the
// internal and hourglass energies, and the imposed-velocity work summed over nodes.
static int balance_now(orad_ctx *c, orad_balance_t *b) {
  const size_t nn = (size_t)c->numnod, nbn = (nn + 255) / 256;
  const int p = (c->host_ncycle - 1) & 1;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<double> part(2 * nbn);
  HIPCHK(c, hipMemcpy(part.data(), c->bal_part + (size_t)p * 2 * nbn, 2 * nbn * sizeof(double),
                      hipMemcpyDeviceToHost));
  double ek = 0.0, er = 0.0;
  for (size_t i = 0; i < nbn; ++i) ek += part[i];
  for (size_t i = 0; i < nbn; ++i) er += part[nbn + i];
  double *scr = nullptr;
  HIPCHK(c, dalloc(&scr, (nbn + 2) * sizeof(double)));
  hipError_t e = launch_sum(c->fv_work, (int)nn, scr, scr + nbn, c->stream);
  double wf[2] = {0.0, 0.0};
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(wf, scr + nbn, 2 * sizeof(double), hipMemcpyDeviceToHost);
  dfree(scr);
  if (e != hipSuccess) return fail(c, ORAD_EHIP, std::string("energy balance: ") + hipGetErrorString(e));
  double ei = 0.0, ekin_state = 0.0, eh = 0.0;
  int r = energies_range(c, 0, (size_t)nelem(c), 0, nn, 0, c->nmodel, &ei, &ekin_state, &eh);
  if (r) return r;
  orad_time_t t{};
  r = model_time(c, 0, &t);
  if (r) return r;
  b->encin = ek;
  b->enrot = er;
  b->enint = ei;
  b->ehour = eh;
  b->wfext = wf[0];
  b->entot = ek + ei + er;
  b->entot0 = c->entot0;
  const double den = c->entot0 + b->wfext;
  b->err = (c->have_entot0 && std::fabs(den) > 1.0e-20) ? b->entot / den - 1.0 : 0.0;
  b->tt = t.tt;
  b->ncycle = t.ncycle;
  return ORAD_OK;
}

static int build_graph(orad_ctx *c, int p) {
  hipGraph_t g = nullptr;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = enqueue_cycle(c, p);
  if (e == hipSuccess) e = enqueue_cycle(c, p ^ 1);
  hipError_t e2 = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(c, ORAD_EHIP, std::string("graph capture: ") +
                                  hipGetErrorString(e != hipSuccess ? e : e2));
  }
  e = hipGraphInstantiate(&c->graph[p], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(c, ORAD_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  return ORAD_OK;
}

// The contact pass's capacities are fixed at orad_set_contact7; a pass that outgrew
// them (sticky flag) has dropped candidates, so the run is reported as failed.
static int contact_overflow(orad_ctx *c) {
  if (!c->ci) return ORAD_OK;
  int ovf = 0;
  HIPCHK(c, hipMemcpy(&ovf, c->ci->counters + I7C_OVERFLOW, sizeof(int), hipMemcpyDeviceToHost));
  if (ovf & I7_OVF_NODE)
    return fail(c, ORAD_EINVAL, "contact: a secondary node has more than 64 candidate segments");
  if (ovf & I7_OVF_KEEP)
    return fail(c, ORAD_EINVAL, "contact: a secondary node penetrates more than 16 segments");
  if (ovf & I7_OVF_PAIRS)
    return fail(c, ORAD_EINVAL,
                "contact: penetrating-pair capacity exceeded (orad_set_contact7 cand_cap)");
  return ORAD_OK;
}

int orad_run_cycles(orad_ctx *c, int ncycle, orad_time_t *t) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  int r = ready(c);
  if (r) return r;
  if (ncycle < 0) return fail(c, ORAD_EINVAL, "run_cycles: negative cycle count");
  HIPCHK(c, hipSetDevice(c->device));
  int left = ncycle;
  if (c->host_ncycle == 0 && left > 0 && !c->batched) {
    // ECRIT's reference total at the first cycle: ENTOT0 = ENTOT - WFEXT (ecrit.F:500-503)
    HIPCHK(c, enqueue_cycle(c, 0));
    c->host_ncycle = 1;
    left -= 1;
    orad_balance_t b{};
    r = balance_now(c, &b);
    if (r) return r;
    c->entot0 = b.entot - b.wfext;
    c->have_entot0 = true;
  }
  while (left >= 2) {
    const int p = c->host_ncycle & 1;
    if (!c->graph[p]) {
      r = build_graph(c, p);
      if (r) return r;
    }
    HIPCHK(c, hipGraphLaunch(c->graph[p], c->stream));
    c->host_ncycle += 2;
    left -= 2;
  }
  if (left == 1) {
    HIPCHK(c, enqueue_cycle(c, c->host_ncycle & 1));
    c->host_ncycle += 1;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if ((r = contact_overflow(c))) return r;
  orad_time_t tt{};
  r = orad_get_time(c, &tt);
  if (r) return r;
  if (t) *t = tt;
  return tt.mstop ? fail(c, ORAD_ENEGVOL, "non-positive element volume (MSTOP)") : ORAD_OK;
}

int orad_profile_cycles(orad_ctx *c, int ncycle, double *elem_ms, double *node_ms) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  int r = ready(c);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  double se = 0.0, sn = 0.0;
  for (int i = 0; i < ncycle; ++i) {
    const int p = c->host_ncycle & 1;
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    HIPCHK(c, launch_cycle_elements(c, p));
    HIPCHK(c, launch_cycle_contact(c, p));
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, launch_node(node_args(c), c->precision == ORAD_MIXED, p, 1, 0, c->xcd_remap, c->stream,
                          c->node_tile));
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev[2]));
    float a = 0.f, b = 0.f;
    HIPCHK(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
    HIPCHK(c, hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
    se += a;
    sn += b;
    c->host_ncycle += 1;
  }
  if (elem_ms) *elem_ms = ncycle ? se / ncycle : 0.0;
  if (node_ms) *node_ms = ncycle ? sn / ncycle : 0.0;
  return ORAD_OK;
}

int orad_get_nodes(orad_ctx *c, double *x, double *v, double *d, double *a, double *stifn) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!c->have_nodes) return fail(c, ORAD_ESTATE, "get_nodes: no nodes");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t n = (size_t)c->numnod;
  if (x || v) {
    std::vector<double4> h(n);
    if (x) {
      HIPCHK(c, hipMemcpy(h.data(), c->X4, n * sizeof(double4), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) { x[3 * i] = h[i].x; x[3 * i + 1] = h[i].y; x[3 * i + 2] = h[i].z; }
    }
    if (v) {
      HIPCHK(c, hipMemcpy(h.data(), c->V4, n * sizeof(double4), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) { v[3 * i] = h[i].x; v[3 * i + 1] = h[i].y; v[3 * i + 2] = h[i].z; }
    }
  }
  if (d) HIPCHK(c, hipMemcpy(d, c->D, 3 * n * 8, hipMemcpyDeviceToHost));
  if (a) HIPCHK(c, hipMemcpy(a, c->A, 3 * n * 8, hipMemcpyDeviceToHost));
  if (stifn) HIPCHK(c, hipMemcpy(stifn, c->STIFN, n * 8, hipMemcpyDeviceToHost));
  return ORAD_OK;
}

int orad_get_brick_state(orad_ctx *c, double *sig, double *eint, double *rho, double *pla,
                         double *qvis, double *offg, double *epsd) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!c->have_bricks) return fail(c, ORAD_ESTATE, "get_brick_state: no mesh");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t ne = (size_t)c->numels, ts = c->tsize();
  std::vector<char> h(NSTATE * ne * ts);
  HIPCHK(c, hipMemcpy(h.data(), c->state, h.size(), hipMemcpyDeviceToHost));
  // e: engine element; its column entry is at its library position
  auto col = [&](int k, size_t e) -> double {
    const size_t i = c->h_iperm.empty() ? e : (size_t)c->h_iperm[e];
    return c->precision == ORAD_MIXED ? (double)((float *)h.data())[k * ne + i]
                                      : ((double *)h.data())[k * ne + i];
  };
  for (size_t e = 0; e < ne; ++e) {
    if (sig)
      for (int k = 0; k < 6; ++k) sig[k * ne + e] = col(k, e);
    if (eint) eint[e] = col(C_EINT, e);
    if (rho) rho[e] = col(C_RHO, e);
    if (pla) pla[e] = col(C_PLA, e);
    if (qvis) qvis[e] = col(C_QVIS, e);
    if (offg) offg[e] = col(C_OFFG, e);
    if (epsd) epsd[e] = col(C_EPSD, e);
  }
  return ORAD_OK;
}

int orad_set_brick_state(orad_ctx *c, const double *sig, const double *eint, const double *rho,
                         const double *pla, const double *qvis, const double *offg) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!c->have_bricks) return fail(c, ORAD_ESTATE, "set_brick_state: no mesh");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t ne = (size_t)c->numels, ts = c->tsize();
  std::vector<char> h(NSTATE * ne * ts);
  HIPCHK(c, hipMemcpy(h.data(), c->state, h.size(), hipMemcpyDeviceToHost));
  auto put = [&](int k, size_t e, double v) {
    const size_t i = c->h_iperm.empty() ? e : (size_t)c->h_iperm[e];
    if (c->precision == ORAD_MIXED) ((float *)h.data())[k * ne + i] = (float)v;
    else ((double *)h.data())[k * ne + i] = v;
  };
  for (size_t e = 0; e < ne; ++e) {
    if (sig)
      for (int k = 0; k < 6; ++k) put(k, e, sig[k * ne + e]);
    if (eint) put(C_EINT, e, eint[e]);
    if (rho) put(C_RHO, e, rho[e]);
    if (pla) put(C_PLA, e, pla[e]);
    if (qvis) put(C_QVIS, e, qvis[e]);
    if (offg) put(C_OFFG, e, offg[e]);
  }
  HIPCHK(c, hipMemcpy(c->state, h.data(), h.size(), hipMemcpyHostToDevice));
  c->state_init = true;
  return ORAD_OK;
}

int orad_get_fsky(orad_ctx *c, double *fsky) {
  if (!c || !fsky) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (c->family == FAM_SHELL) return get_fsky_shell(c, fsky);
  if (!c->have_bricks) return fail(c, ORAD_ESTATE, "get_fsky: no mesh");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t ne = (size_t)c->numels, ns = 8 * ne;
  std::vector<double> f(3 * ns), st(ne);
  if (c->precision == ORAD_MIXED) {
    std::vector<float> h(3 * ns), hs(ne);
    HIPCHK(c, hipMemcpy(h.data(), c->fsky, h.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(hs.data(), c->sti, ne * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < h.size(); ++i) f[i] = h[i];
    for (size_t i = 0; i < ne; ++i) st[i] = hs[i];
  } else {
    HIPCHK(c, hipMemcpy(f.data(), c->fsky, 3 * ns * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(st.data(), c->sti, ne * 8, hipMemcpyDeviceToHost));
  }
  for (size_t e = 0; e < ne; ++e) {
    const double sti = st[c->h_iperm.empty() ? e : (size_t)c->h_iperm[e]];
    for (int k = 0; k < 8; ++k) {
      const size_t sl = (size_t)c->h_iads[8 * e + k];
      fsky[4 * sl] = f[3 * sl];
      fsky[4 * sl + 1] = f[3 * sl + 1];
      fsky[4 * sl + 2] = f[3 * sl + 2];
      fsky[4 * sl + 3] = sti;
    }
  }
  return ORAD_OK;
}

// Internal energy of elements [e0, e1), kinetic energy of nodes [n0, n1) and the
// hourglass energy totals of models [m0, m1).
// `scr` (nullable): a buffer of at least energy_scratch_doubles() doubles to use
// instead of a per-call allocation.
static size_t energy_scratch_doubles(size_t me, size_t mn) {
  return (me + 255) / 256 + (mn + 255) / 256 + 2;
}

static int energies_range(orad_ctx *c, size_t e0, size_t e1, size_t n0, size_t n1, int m0,
                          int m1, double *eint, double *ekin, double *ehour, double *scr) {
  HIPCHK(c, hipSetDevice(c->device));
  const bool shell = c->family == FAM_SHELL;
  const size_t ne = (size_t)nelem(c), ts = c->tsize(), me = e1 - e0, mn = n1 - n0;
  const size_t nbe = (me + 255) / 256, nbk = (mn + 255) / 256;
  double *part = scr;
  if (!part) HIPCHK(c, dalloc(&part, energy_scratch_doubles(me, mn) * sizeof(double)));
  const char *st = (const char *)c->state;
  hipError_t e = launch_energy(c->precision == ORAD_MIXED, (int)me,
                               shell ? st + (SC_EINT * ne + e0) * ts : st + (C_EINT * ne + e0) * ts,
                               shell ? st + ((SC_EINT + 1) * ne + e0) * ts : nullptr,
                               shell ? nullptr : c->vol0 + e0, shell ? nullptr : c->emap,
                               (int)mn, c->X4 + n0, c->V4 + n0,
                               shell ? c->VR4 + n0 : nullptr, part, part + nbe + nbk, c->stream);
  double res[2] = {0, 0}, eh = 0;
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(res, part + nbe + nbk, 2 * sizeof(double), hipMemcpyDeviceToHost);
  for (int m = m0; m < m1 && e == hipSuccess; ++m) {
    double h = 0.0;
    e = hipMemcpy(&h, &c->ctl[m].ehour, sizeof(double), hipMemcpyDeviceToHost);
    eh += h;
  }
  if (!scr) dfree(part);
  if (e != hipSuccess) return fail(c, ORAD_EHIP, std::string("energies: ") + hipGetErrorString(e));
  if (eint) *eint = res[0];
  if (ekin) *ekin = res[1];
  if (ehour) *ehour = eh;
  return ORAD_OK;
}

int orad_energies(orad_ctx *c, double *eint, double *ekin, double *ehour) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  int r = ready(c);
  if (r) return r;
  return energies_range(c, 0, (size_t)nelem(c), 0, (size_t)c->numnod, 0, c->nmodel, eint, ekin,
                        ehour);
}

int orad_energy_balance(orad_ctx *c, orad_balance_t *b) {
  if (!c || !b) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  int r = ready(c);
  if (r) return r;
  if (c->batched) return fail(c, ORAD_ESTATE, "energy_balance: one model per context");
  if (c->host_ncycle < 1) return fail(c, ORAD_ESTATE, "energy_balance: run a cycle first");
  return balance_now(c, b);
}

int orad_energies_noalloc(orad_ctx *c, double *eint, double *ekin, double *ehour) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  int r = ready(c);
  if (r) return r;
  if (!c->noalloc_scratch) return fail(c, ORAD_ESTATE, "energies_noalloc: no device scratch");
  const size_t need = energy_scratch_doubles((size_t)nelem(c), (size_t)c->numnod);
  if (need > NOALLOC_SCRATCH_DOUBLES)
    return fail(c, ORAD_EINVAL,
                "energies_noalloc: mesh too large for the fixed 1 MiB device scratch "
                "(ceil(elements/256) + ceil(nodes/256) + 2 must not exceed 131072); "
                "use orad_energies");
  return energies_range(c, 0, (size_t)nelem(c), 0, (size_t)c->numnod, 0, c->nmodel, eint, ekin,
                        ehour, c->noalloc_scratch);
}

}  // extern "C"

// ================================================================ shells ===

// constant_mod values of the shell path, derived from their definitions
// (common_source/modules/constant_mod.F): each a double operation on exact or
// previously rounded values, as the compiler folds the reference's parameters.
static ShConst shell_constants() {
  volatile double ten = 10.0;
  double ep[31];
  ep[2] = 100.0;
  for (int i = 3; i <= 20; ++i) ep[i] = ep[i - 1] * ten;
  const double ep10 = ep[10], ep20 = ep[20], ep30 = ep20 * ep10;
  const double em01 = 1.0 / ten, em02 = 1.0 / ep[2], em04 = 1.0 / ep[4], em05 = 1.0 / ep[5];
  const double em06 = 1.0 / ep[6];
  const double zep1 = em01, zep01 = em02, zep03 = 3.0 / ep[2], zep07 = zep1 - zep03;
  const double zep09 = zep1 - zep01, zep19 = zep1 + zep09;
  const double twoem3 = 2.0 / ep[3], fourem3 = 4.0 / ep[3], fourem5 = 4.0 / ep[5];
  const double sixem5 = em04 - fourem5, nineem6 = em05 - em06;
  const double zep4 = 4.0 / ten, onep4 = 1.0 + zep4, onep41 = onep4 + zep01;
  ShConst k{};
  k.zep81 = 1.0 - zep19;
  k.zep072169 = zep07 + twoem3 + em04 + sixem5 + nineem6;
  k.onep414 = onep41 + fourem3;
  k.four_over_3 = 4.0 / 3.0;
  k.one_over_9 = 1.0 / 9.0;
  k.four_over_5 = 4.0 / 5.0;
  k.em30 = 1.0 / ep30;
  k.sr2d2 = std::sqrt(2.0) * 0.5;
  k.em15 = 1.0 / ep[15];
  k.em20 = 1.0 / ep20;
  k.em07 = 1.0 / ep[7];
  k.three_over_4 = 3.0 / 4.0;
  k.third = 1.0 / 3.0;
  k.three_half = 3.0 / 2.0;
  // QEPH chain (constant_mod.F: EM08 .. ZEP9EP30 from their definitions)
  k.em08 = 1.0 / ep[8];
  k.em10 = 1.0 / ep[10];
  k.em18 = 1.0 / ep[18];
  k.one_over_12 = 1.0 / 12.0;
  k.one_over_16 = 1.0 / 16.0;
  k.one_over_64 = 1.0 / 64.0;
  {
    const double em03 = 1.0 / ep[3];
    const double zep05 = 5.0 / ep[2], zep15 = zep1 + zep05;
    const double zep04 = 4.0 / ep[2], zep06 = zep1 - zep04;
    k.zep85 = 1.0 - zep15;
    k.zep999 = 1.0 - em03;
    k.fivep333 = 16.0 / 3.0;
    k.threep464 = 3.0 + zep4 + zep06 + fourem3;
    k.zep9ep30 = 9.0 * (ep20 * ep[9]);
    // QBAT chain
    k.em12 = 1.0 / ep[12];
    k.em03 = em03;
    k.zep3 = 3.0 / ten;
  }
  return k;
}

// COQINI tables (engine/source/elements/shell/coqini.F:45-121): default-REAL
// DATA literals, widened from single precision as the reference's double build holds them.
static void shell_tables(int npt, ShProp &p) {
  static const float z0[5][5] = {{0.f, 0.f, 0.f, 0.f, 0.f},
                                 {-.5f, 0.5f, 0.f, 0.f, 0.f},
                                 {-.5f, 0.f, 0.5f, 0.f, 0.f},
                                 {-.5f, -.1666667f, 0.1666667f, 0.5f, 0.f},
                                 {-.5f, -.25f, 0.f, 0.25f, 0.5f}};
  static const float wf[5][5] = {{1.f, 0.f, 0.f, 0.f, 0.f},
                                 {0.5f, 0.5f, 0.f, 0.f, 0.f},
                                 {0.25f, 0.5f, 0.25f, 0.f, 0.f},
                                 {0.1666667f, 0.3333333f, 0.3333333f, 0.1666667f, 0.f},
                                 {0.125f, 0.25f, 0.25f, 0.25f, 0.125f}};
  static const float wm[5][5] = {{0.f, 0.f, 0.f, 0.f, 0.f},
                                 {-.0833333f, 0.0833333f, 0.f, 0.f, 0.f},
                                 {-.0833333f, 0.f, 0.0833333f, 0.f, 0.f},
                                 {-.0648148f, -.0555556f, 0.0555556f, 0.0648148f, 0.f},
                                 {-.0520833f, -.0625f, 0.f, 0.0625f, 0.0520833f}};
  for (int i = 0; i < 5; ++i) {
    p.z0[i] = i < npt ? (double)z0[npt - 1][i] : 0.0;
    p.wf[i] = i < npt ? (double)wf[npt - 1][i] : 0.0;
    p.wm[i] = i < npt ? (double)wm[npt - 1][i] : 0.0;
  }
}

static int upload_codes(orad_ctx *c) {
  return pack_node_words(c);
}

static int init_shell_state(orad_ctx *c) {
  if (!c->have_shells || c->state_init) return ORAD_OK;
  const size_t ne = (size_t)c->numelc, ts = c->tsize();
  std::vector<char> all((size_t)c->shncol() * ne * ts, 0);
  std::vector<double> thk(ne);
  HIPCHK(c, hipMemcpy(thk.data(), c->thke, ne * 8, hipMemcpyDeviceToHost));
  for (size_t e = 0; e < ne; ++e) {
    if (c->precision == ORAD_MIXED) {
      ((float *)all.data())[SC_THK * ne + e] = (float)thk[e];
      ((float *)all.data())[SC_OFFG * ne + e] = 1.0f;
    } else {
      ((double *)all.data())[SC_THK * ne + e] = thk[e];
      ((double *)all.data())[SC_OFFG * ne + e] = 1.0;
    }
  }
  HIPCHK(c, hipMemcpy(c->state, all.data(), all.size(), hipMemcpyHostToDevice));
  c->state_init = true;
  return ORAD_OK;
}

extern "C" {

int orad_set_nodes_rot(orad_ctx *c, const double *vr, const double *in, const int *icodr) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!c->have_nodes) return fail(c, ORAD_ESTATE, "set_nodes_rot: call orad_set_nodes first");
  if (!in) return fail(c, ORAD_EINVAL, "set_nodes_rot: nodal inertia IN is required");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  const size_t n = (size_t)c->numnod;
  if (!c->have_rot) {
    HIPCHK(c, dalloc(&c->VR4, n * sizeof(double4)));
    HIPCHK(c, dalloc(&c->AR, 3 * n * sizeof(double)));
    HIPCHK(c, dalloc(&c->STIFR, n * sizeof(double)));
    HIPCHK(c, hipMemset(c->AR, 0, 3 * n * sizeof(double)));
    HIPCHK(c, hipMemset(c->STIFR, 0, n * sizeof(double)));
  }
  std::vector<double4> hw(n);
  if (c->h_code_hi.empty()) c->h_code_hi.assign(n, 0);
  for (size_t i = 0; i < n; ++i) {
    hw[i] = make_double4(vr ? vr[3 * i] : 0.0, vr ? vr[3 * i + 1] : 0.0, vr ? vr[3 * i + 2] : 0.0,
                         in[i]);
    const int r = icodr ? icodr[i] : 0;
    if (r < 0 || r > 7) return fail(c, ORAD_EINVAL, "set_nodes_rot: ICODR outside 0..7");
    c->h_code_hi[i] = (c->h_code_hi[i] & ~(7 << 3)) | (r << 3);
  }
  HIPCHK(c, hipMemcpy(c->VR4, hw.data(), n * sizeof(double4), hipMemcpyHostToDevice));
  c->have_rot = true;
  return upload_codes(c);
}

int orad_set_fixvel(orad_ctx *c, int n, const int *nodes, const int *dir, const double *vel) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!c->have_nodes) return fail(c, ORAD_ESTATE, "set_fixvel: call orad_set_nodes first");
  if (n < 0 || (n > 0 && (!nodes || !dir || !vel)))
    return fail(c, ORAD_EINVAL, "set_fixvel: bad arguments");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  const size_t nn = (size_t)c->numnod;
  if (c->h_code_hi.empty()) c->h_code_hi.assign(nn, 0);
  c->h_fv_val.assign(nn, 0.0);
  for (size_t i = 0; i < nn; ++i) c->h_code_hi[i] &= ~(3 << 6);
  for (int j = 0; j < n; ++j) {
    const int nd = nodes[j];
    if (nd < 1 || nd > c->numnod) return fail(c, ORAD_EINVAL, "set_fixvel: node id out of range");
    if (dir[j] < 1 || dir[j] > 3) return fail(c, ORAD_EINVAL, "set_fixvel: direction not 1..3");
    c->h_code_hi[nd - 1] = (c->h_code_hi[nd - 1] & ~(3 << 6)) | (dir[j] << 6);
    c->h_fv_val[nd - 1] = vel[j];
  }
  if (!c->fv_val) HIPCHK(c, dalloc(&c->fv_val, nn * sizeof(double)));
  HIPCHK(c, hipMemcpy(c->fv_val, c->h_fv_val.data(), nn * 8, hipMemcpyHostToDevice));
  return upload_codes(c);
}

int orad_set_shells(orad_ctx *c, int numelc, const int *ixc, const double *thke) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!c->have_nodes) return fail(c, ORAD_ESTATE, "set_shells: call orad_set_nodes first");
  if (c->family != FAM_NONE) return fail(c, ORAD_ESTATE, "set_shells: mesh already set");
  if (numelc <= 0 || !ixc || !thke) return fail(c, ORAD_EINVAL, "set_shells: bad arguments");
  if (numelc > INT_MAX / 4)
    return fail(c, ORAD_EINVAL, "set_shells: more than 536870911 shells (4 slots each, 32-bit IADC)");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t ne = (size_t)numelc;
  const int nn = c->numnod;
  std::vector<int32_t> conn_em(4 * ne), uid(ne);
  for (size_t e = 0; e < ne; ++e) {
    for (int k = 0; k < 4; ++k) {
      const int nd = ixc[7 * e + 1 + k];
      if (nd < 1 || nd > nn) return fail(c, ORAD_EINVAL, "set_shells: node id out of range");
      conn_em[4 * e + k] = nd - 1;
    }
    if (!(thke[e] > 0.0)) return fail(c, ORAD_EINVAL, "set_shells: thickness must be positive");
    uid[e] = ixc[7 * e + 6];
  }
  // PARITH/ON CSR with the starter's ordering (FILLCNE + W_PON), 4 corners.
  std::vector<int32_t> cnt(nn + 1, 0);
  for (size_t i = 0; i < 4 * ne; ++i) cnt[conn_em[i]]++;
  c->h_adsky.assign(nn + 1, 0);
  for (int n = 0; n < nn; ++n) c->h_adsky[n + 1] = c->h_adsky[n] + cnt[n];
  for (int n = 0; n < nn; ++n) cnt[n] = c->h_adsky[n];
  std::vector<int32_t> order(ne);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return uid[a] < uid[b]; });
  c->h_iads.assign(4 * ne, 0);
  for (size_t j = 0; j < ne; ++j) {
    const int32_t e = order[j];
    for (int k = 0; k < 4; ++k) c->h_iads[4 * (size_t)e + k] = cnt[conn_em[4 * (size_t)e + k]]++;
  }
  std::vector<int32_t> conn_cm(4 * ne), iads_cm(4 * ne), se(4 * ne);
  for (size_t e = 0; e < ne; ++e)
    for (int k = 0; k < 4; ++k) {
      conn_cm[k * ne + e] = conn_em[4 * e + k];
      iads_cm[k * ne + e] = c->h_iads[4 * e + k];
      // STI entry of the slot: the element's, or for QEPH its diagonal's
      se[c->h_iads[4 * e + k]] = (int32_t)(c->nsti() == 2 ? e + (k & 1) * ne : e);
    }
  const size_t ts = c->tsize();
  HIPCHK(c, dalloc(&c->conn, 4 * ne * 4));
  HIPCHK(c, dalloc(&c->iads, 4 * ne * 4));
  HIPCHK(c, dalloc(&c->adsky, (nn + 1) * 4));
  HIPCHK(c, dalloc(&c->slot_elem, 4 * ne * 4));
  HIPCHK(c, hipMemcpy(c->conn, conn_cm.data(), 4 * ne * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->iads, iads_cm.data(), 4 * ne * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->adsky, c->h_adsky.data(), (nn + 1) * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->slot_elem, se.data(), 4 * ne * 4, hipMemcpyHostToDevice));
  const size_t ncol = (size_t)c->shncol(), nsti = (size_t)c->nsti();
  HIPCHK(c, dalloc(&c->state, ncol * ne * ts));
  HIPCHK(c, hipMemset(c->state, 0, ncol * ne * ts));
  HIPCHK(c, dalloc(&c->fsky, 4 * ne * 6 * ts));
  HIPCHK(c, hipMemset(c->fsky, 0, 4 * ne * 6 * ts));
  HIPCHK(c, dalloc(&c->sti, nsti * ne * ts));
  HIPCHK(c, hipMemset(c->sti, 0, nsti * ne * ts));
  HIPCHK(c, dalloc(&c->thke, ne * 8));
  HIPCHK(c, hipMemcpy(c->thke, thke, ne * 8, hipMemcpyHostToDevice));
  const bool mixed = c->precision == ORAD_MIXED;
  const int bb = c->ishell == 24   ? (mixed ? qeph_block_mixed() : qeph_block_fp64())
                 : c->ishell == 12 ? (mixed ? qbat_block_mixed() : qbat_block_fp64())
                                   : (mixed ? shell_block_mixed() : shell_block_fp64());
  c->n_ehou_part = (int)((ne + bb - 1) / bb);
  HIPCHK(c, dalloc(&c->ehou_part, (size_t)c->n_ehou_part * sizeof(double)));
  HIPCHK(c, hipMemset(c->ehou_part, 0, (size_t)c->n_ehou_part * sizeof(double)));
  c->numelc = numelc;
  c->have_shells = true;
  c->family = FAM_SHELL;
  c->nc = 4;
  c->shk = shell_constants();
  drop_graphs(c);
  return init_shell_state(c);
}

}  // extern "C"

// One law-2 parameter record (the layout orad_set_shell_law2 documents); returns
// an empty string when valid, else the reason.
static std::string parse_shell_law2(const double *p, const int *ip, ShMat &m) {
  m = ShMat{};
  m.rho0 = p[0]; m.young = p[1]; m.nu = p[2]; m.shear = p[3]; m.a11 = p[4]; m.a12 = p[5];
  m.ssp = p[6]; m.ca = p[7]; m.cb = p[8]; m.cn = p[9]; m.epmx = p[10]; m.ymax = p[11];
  m.cc = p[12]; m.epdr = p[13]; m.mexp = p[14]; m.asrate_pm9 = p[15];
  m.icc = ip[0]; m.vp = ip[1]; m.israte = ip[2]; m.ipla = ip[3];
  if (!(m.rho0 > 0.0) || !(m.young > 0.0) || !(m.shear > 0.0) || !(m.ssp > 0.0))
    return "rho0, E, G and SSP must be positive";
  if (!(m.ca > 0.0)) return "A must be positive";
  if (m.vp < 1 || m.vp > 3 || m.ipla < 0 || m.ipla > 2) return "VP must be 1..3 and IPLA 0..2";
  if (m.cc != 0.0 && !(m.epdr > 0.0)) return "EPS0 must be > 0";
  return std::string();
}

extern "C" {

int orad_set_shell_law2(orad_ctx *c, const double *p, const int *ip) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!p || !ip) return fail(c, ORAD_EINVAL, "set_shell_law2: null parameters");
  ShMat m;
  const std::string why = parse_shell_law2(p, ip, m);
  if (!why.empty()) return fail(c, ORAD_EINVAL, "set_shell_law2: " + why);
  c->shmat = m;
  c->have_shmat = true;
  if (c->batched) {
    // one material for every model
    for (auto &x : c->h_mats) x = m;
    HIPCHK(c, hipMemcpy(c->mats, c->h_mats.data(), c->h_mats.size() * sizeof(ShMat),
                        hipMemcpyHostToDevice));
  }
  drop_graphs(c);
  return ORAD_OK;
}

int orad_set_shell_prop(orad_ctx *c, const double *g, int ish, int npt) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!g) return fail(c, ORAD_EINVAL, "set_shell_prop: null parameters");
  if (npt < 2 || npt > 5) return fail(c, ORAD_EINVAL, "set_shell_prop: NPT must be 2..5");
  if (ish < 0 || ish > 1) return fail(c, ORAD_EINVAL, "set_shell_prop: ISHEAR must be 0 or 1");
  ShProp p{};
  p.hm = g[0]; p.hf = g[1]; p.hr = g[2]; p.dm = g[3]; p.fsh = g[4]; p.dtfac = g[5]; p.dtmin = g[6];
  p.ish = ish;
  p.npt = npt;
  if (!(p.hm > 0.0 || p.hf > 0.0) || !(p.dtfac > 0.0) || p.dm < 0.0 || !(p.fsh > 0.0))
    return fail(c, ORAD_EINVAL, "set_shell_prop: hm/hf, DTFAC and the shear factor must be positive");
  shell_tables(npt, p);
  p.ishell = c->ishell;
  p.dn = c->qeph_dn;
  p.cvis = 1.0;
  c->shprop = p;
  c->have_shprop = true;
  drop_graphs(c);
  return ORAD_OK;
}

int orad_set_shell_ishell(orad_ctx *c, int ishell, double dn) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (c->have_shells)
    return fail(c, ORAD_ESTATE, "set_shell_ishell: choose the formulation before orad_set_shells");
  if (ishell != 1 && ishell != 24 && ishell != 12)
    return fail(c, ORAD_EINVAL, "set_shell_ishell: Ishell must be 1 (BT), 24 (QEPH) or 12 (QBAT)");
  if (!(dn >= 0.0)) return fail(c, ORAD_EINVAL, "set_shell_ishell: Dn must be >= 0");
  c->ishell = ishell;
  // QEPH: Dn = 0 on input means 0.015 (ZEP01 + FIVEEM3, hm_read_prop01.F:199,
  // cncoef3.F:259-264).  QBAT: GEO(13) as given; CNCOEF3 reads 0 as EM3
  // (cncoef3.F:421-431), which the element chain applies.
  c->qeph_dn = (ishell == 24 && dn == 0.0) ? 1.0 / 100.0 + 5.0 / 1000.0 : dn;
  c->shprop.ishell = ishell;
  c->shprop.dn = c->qeph_dn;
  c->shprop.cvis = 1.0;
  drop_graphs(c);
  return ORAD_OK;
}

int orad_cforc3(orad_ctx *c, int nft, int nel, double dt1, double *dt2t) {
  if (!c || !dt2t) return ORAD_EINVAL;
  int r = group_wait(c, nullptr);
  if (!r) r = group_enqueue(c, FAM_SHELL, nft, nel, dt1);
  if (!r) r = group_wait(c, dt2t);
  return r;
}

int orad_asspar4_rot(orad_ctx *c, double *a, double *ar, double *stifn, double *stifr) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (c->family != FAM_SHELL) return fail(c, ORAD_ESTATE, "asspar4_rot: the mesh is not of shells");
  int r = ready(c);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, launch_node(assembly_args(c), c->precision == ORAD_MIXED, c->host_ncycle & 1, 0, 1,
                        c->xcd_remap, c->stream, c->node_tile));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t n = (size_t)c->numnod;
  if (a) HIPCHK(c, hipMemcpy(a, c->A, 3 * n * 8, hipMemcpyDeviceToHost));
  if (ar) HIPCHK(c, hipMemcpy(ar, c->AR, 3 * n * 8, hipMemcpyDeviceToHost));
  if (stifn) HIPCHK(c, hipMemcpy(stifn, c->STIFN, n * 8, hipMemcpyDeviceToHost));
  if (stifr) HIPCHK(c, hipMemcpy(stifr, c->STIFR, n * 8, hipMemcpyDeviceToHost));
  return ORAD_OK;
}

int orad_get_nodes_rot(orad_ctx *c, double *vr, double *ar) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!c->have_rot) return fail(c, ORAD_ESTATE, "get_nodes_rot: no rotational DOF");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t n = (size_t)c->numnod;
  if (vr) {
    std::vector<double4> h(n);
    HIPCHK(c, hipMemcpy(h.data(), c->VR4, n * sizeof(double4), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) { vr[3 * i] = h[i].x; vr[3 * i + 1] = h[i].y; vr[3 * i + 2] = h[i].z; }
  }
  if (ar) HIPCHK(c, hipMemcpy(ar, c->AR, 3 * n * 8, hipMemcpyDeviceToHost));
  return ORAD_OK;
}

int orad_get_shell_state(orad_ctx *c, double *rows) {
  if (!c || !rows) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (c->family != FAM_SHELL) return fail(c, ORAD_ESTATE, "get_shell_state: no shell mesh");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t ne = (size_t)c->numelc, ts = c->tsize();
  const int nc = c->shncol();
  std::vector<char> h((size_t)nc * ne * ts);
  HIPCHK(c, hipMemcpy(h.data(), c->state, h.size(), hipMemcpyDeviceToHost));
  for (size_t e = 0; e < ne; ++e)
    for (int k = 0; k < nc; ++k)
      rows[e * nc + k] = c->precision == ORAD_MIXED ? (double)((float *)h.data())[k * ne + e]
                                                    : ((double *)h.data())[k * ne + e];
  return ORAD_OK;
}

int orad_set_shell_state(orad_ctx *c, const double *rows) {
  if (!c || !rows) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (c->family != FAM_SHELL) return fail(c, ORAD_ESTATE, "set_shell_state: no shell mesh");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t ne = (size_t)c->numelc, ts = c->tsize();
  const int nc = c->shncol();
  std::vector<char> h((size_t)nc * ne * ts);
  for (size_t e = 0; e < ne; ++e)
    for (int k = 0; k < nc; ++k) {
      if (c->precision == ORAD_MIXED) ((float *)h.data())[k * ne + e] = (float)rows[e * nc + k];
      else ((double *)h.data())[k * ne + e] = rows[e * nc + k];
    }
  HIPCHK(c, hipMemcpy(c->state, h.data(), h.size(), hipMemcpyHostToDevice));
  c->state_init = true;
  return ORAD_OK;
}

}  // extern "C"

// ======================================================= batched models ===

static int batch_occupancy(orad_ctx *c) {
  if (c->qeph())
    return c->precision == ORAD_MIXED ? qeph_batch_occupancy_mixed(c->batch_lds)
                                      : qeph_batch_occupancy_fp64(c->batch_lds);
  return c->precision == ORAD_MIXED ? shell_batch_occupancy_mixed(c->batch_lds)
                                    : shell_batch_occupancy_fp64(c->batch_lds);
}

static int update_batch_grid(orad_ctx *c) {
  if (!c->batched) return ORAD_OK;
  if (c->num_cu <= 0)
    HIPCHK(c, hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, c->device));
  const int occ = batch_occupancy(c);
  if (occ <= 0)
    return fail(c, ORAD_EINVAL, "launch: the batched kernel does not fit at this LDS reservation");
  c->batch_grid = occ * c->num_cu;
  return ORAD_OK;
}

extern "C" {

int orad_set_shell_models(orad_ctx *c, int nmodel, const int *node_off, const int *elem_off,
                          const double *p, const int *ip) {
  if (c && c->ci)
    return fail(c, ORAD_ESTATE, "set_shell_models: set the models before the contact interface");
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (c->family != FAM_SHELL || !c->have_shmat)
    return fail(c, ORAD_ESTATE, "set_shell_models: set the shells and a law-2 material first");
  if (c->batched) return fail(c, ORAD_ESTATE, "set_shell_models: models already set");
  if (c->qbat())
    return fail(c, ORAD_EINVAL, "set_shell_models: batched models take Ishell 1 or 24, not QBAT");
  if (nmodel < 1 || !node_off || !elem_off || (p == nullptr) != (ip == nullptr))
    return fail(c, ORAD_EINVAL, "set_shell_models: bad arguments");
  if (node_off[0] != 0 || elem_off[0] != 0 || node_off[nmodel] != c->numnod ||
      elem_off[nmodel] != c->numelc)
    return fail(c, ORAD_EINVAL, "set_shell_models: offsets must span all nodes and shells");
  for (int m = 0; m < nmodel; ++m)
    if (node_off[m + 1] <= node_off[m] || elem_off[m + 1] < elem_off[m])
      return fail(c, ORAD_EINVAL, "set_shell_models: offsets must increase (a model has nodes)");
  std::vector<ShMat> mats(nmodel, c->shmat);
  if (p)
    for (int m = 0; m < nmodel; ++m) {
      const std::string why = parse_shell_law2(p + 16 * (size_t)m, ip + 4 * (size_t)m, mats[m]);
      if (!why.empty())
        return fail(c, ORAD_EINVAL, "set_shell_models: model " + std::to_string(m) + ": " + why);
    }
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t ne = (size_t)c->numelc, nn = (size_t)c->numnod;
  // every shell's corners belong to its own model
  std::vector<int32_t> conn(4 * ne);
  HIPCHK(c, hipMemcpy(conn.data(), c->conn, 4 * ne * 4, hipMemcpyDeviceToHost));
  std::vector<int32_t> nmod(nn), node0(nmodel), mt(nmodel + 1, 0);
  std::vector<int4> tiles;
  for (int m = 0; m < nmodel; ++m) {
    for (int n = node_off[m]; n < node_off[m + 1]; ++n) nmod[n] = m;
    node0[m] = node_off[m];
  }
  for (int m = 0; m < nmodel; ++m) {
    for (int e = elem_off[m]; e < elem_off[m + 1]; ++e)
      for (int k = 0; k < 4; ++k)
        if (nmod[conn[(size_t)k * ne + e]] != m)
          return fail(c, ORAD_EINVAL, "set_shell_models: shell " + std::to_string(e + 1) +
                                          " uses a node of another model");
    mt[m] = (int)tiles.size();
    for (int e = elem_off[m]; e < elem_off[m + 1]; e += 256)
      tiles.push_back(make_int4(m, e, std::min(e + 256, elem_off[m + 1]), (int)tiles.size()));
  }
  mt[nmodel] = (int)tiles.size();
  // per-model time control, each starting from the context's current scalars
  CycleCtl h0{};
  HIPCHK(c, hipMemcpy(&h0, c->ctl, sizeof(CycleCtl), hipMemcpyDeviceToHost));
  std::vector<CycleCtl> hc(nmodel, h0);
  CycleCtl *ctl = nullptr;
  HIPCHK(c, dalloc(&ctl, nmodel * sizeof(CycleCtl)));
  HIPCHK(c, hipMemcpy(ctl, hc.data(), nmodel * sizeof(CycleCtl), hipMemcpyHostToDevice));
  dfree(c->ctl);
  c->ctl = ctl;
  HIPCHK(c, dalloc(&c->node_model, nn * 4));
  HIPCHK(c, dalloc(&c->model_node0, nmodel * 4));
  HIPCHK(c, dalloc(&c->model_tiles, (nmodel + 1) * 4));
  HIPCHK(c, dalloc(&c->tiles, tiles.size() * sizeof(int4)));
  HIPCHK(c, dalloc(&c->mats, nmodel * sizeof(ShMat)));
  HIPCHK(c, dalloc(&c->queue, sizeof(unsigned)));
  HIPCHK(c, dalloc(&c->chunk_ehou, tiles.size() * sizeof(double)));
  HIPCHK(c, hipMemcpy(c->node_model, nmod.data(), nn * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->model_node0, node0.data(), nmodel * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->model_tiles, mt.data(), (nmodel + 1) * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->tiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->mats, mats.data(), nmodel * sizeof(ShMat), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemset(c->queue, 0, sizeof(unsigned)));
  HIPCHK(c, hipMemset(c->chunk_ehou, 0, tiles.size() * sizeof(double)));
  c->nmodel = nmodel;
  c->h_node_off.assign(node_off, node_off + nmodel + 1);
  c->h_elem_off.assign(elem_off, elem_off + nmodel + 1);
  c->h_model_tiles = mt;
  c->h_mats = mats;
  c->ntiles = (int)tiles.size();
  c->batched = true;
  drop_graphs(c);
  return update_batch_grid(c);
}

int orad_get_model_time(orad_ctx *c, int model, orad_time_t *t) {
  if (!c || !t) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (model < 0 || model >= c->nmodel) return fail(c, ORAD_EINVAL, "get_model_time: no such model");
  return model_time(c, model, t);
}

int orad_model_energies(orad_ctx *c, int model, double *eint, double *ekin, double *ehour) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  int r = ready(c);
  if (r) return r;
  if (model < 0 || model >= c->nmodel) return fail(c, ORAD_EINVAL, "model_energies: no such model");
  if (!c->batched) return orad_energies(c, eint, ekin, ehour);
  return energies_range(c, c->h_elem_off[model], c->h_elem_off[model + 1], c->h_node_off[model],
                        c->h_node_off[model + 1], model, model + 1, eint, ekin, ehour);
}

int orad_get_launch(orad_ctx *c, orad_launch_t *l) {
  if (!c || !l) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  const bool mixed = c->precision == ORAD_MIXED;
  l->persistent = c->persist;
  l->lds_kb = c->batched ? c->batch_lds / 1024 : (c->elem_lds < 0 ? -1 : c->elem_lds / 1024);
  l->xcd_remap = c->xcd_remap;
  l->grid = c->batched && c->persist ? std::min(c->batch_grid, c->ntiles) : 0;
  l->block = l->waves = 0;
  l->order_rows = 0;
  if (c->family == FAM_BRICK || c->family == FAM_NONE) {
    l->block = c->elem_block > 0 ? c->elem_block : (mixed ? brick_block_mixed() : brick_block_fp64());
    l->waves = c->elem_waves > 0 ? c->elem_waves : (mixed ? brick_waves_mixed() : brick_waves_fp64());
    l->order_rows = c->order_rows;
  }
  l->node_tile = c->node_tile > 0 ? c->node_tile : node_tile_default(mixed);
  l->adj_vjp = c->adj_vjp;
  return ORAD_OK;
}

int orad_set_launch(orad_ctx *c, const orad_launch_t *l) {
  if (!c || !l) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (l->lds_kb > 160 || l->lds_kb < -1 || (l->persistent && !c->batched))
    return fail(c, ORAD_EINVAL, "set_launch: LDS in -1..160 KB; persistent needs batched models");
  const bool mixed = c->precision == ORAD_MIXED;
  const bool brick = c->family == FAM_BRICK || c->family == FAM_NONE;
  if (brick && (l->block || l->waves)) {
    const int b = l->block ? l->block : (mixed ? brick_block_mixed() : brick_block_fp64());
    const int w = l->waves ? l->waves : (mixed ? brick_waves_mixed() : brick_waves_fp64());
    if (!(mixed ? brick_shape_ok_mixed(b, w) : brick_shape_ok_fp64(b, w)))
      return fail(c, ORAD_EINVAL, "set_launch: no element-kernel instance of this block and waves");
  }
  if (!brick && (l->block || l->waves))
    return fail(c, ORAD_EINVAL, "set_launch: block and waves are chosen for bricks only");
  if (l->order_rows < -1 || l->order_rows > 128 || (!brick && l->order_rows > 0))
    return fail(c, ORAD_EINVAL, "set_launch: order_rows in -1..128, bricks only");
  if (l->node_tile && !node_tile_ok(mixed, c->family == FAM_SHELL, l->node_tile))
    return fail(c, ORAD_EINVAL, "set_launch: node_tile 1024, 2048 or 4096 slots within 48 KB");
  if (l->adj_vjp != 0 && l->adj_vjp != 1)
    return fail(c, ORAD_EINVAL, "set_launch: adj_vjp is 0 (reverse where available) or 1 (duals)");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int keep_p = c->persist, keep_b = c->batch_lds, keep_e = c->elem_lds;
  c->xcd_remap = l->xcd_remap != 0;
  c->node_tile = l->node_tile;
  c->adj_vjp = l->adj_vjp;
  if (brick) {
    c->elem_block = l->block;
    c->elem_waves = l->waves;
    if (c->have_bricks) {
      const int b = c->elem_block ? c->elem_block : (mixed ? brick_block_mixed() : brick_block_fp64());
      c->n_ehou_part = (c->numels + b - 1) / b;
    }
    if (l->order_rows >= 0) {
      if (c->have_bricks && l->order_rows != c->order_rows) {
        const int r = brick_reorder(c, l->order_rows, {});
        if (r) return r;
      }
      c->order_rows = l->order_rows;
    }
  }
  if (c->batched) {
    c->persist = l->persistent != 0;
    c->batch_lds = l->lds_kb > 0 ? 1024 * l->lds_kb : 0;
    const int r = update_batch_grid(c);
    if (r) {
      c->persist = keep_p;
      c->batch_lds = keep_b;
      (void)update_batch_grid(c);
      return r;
    }
  } else {
    c->elem_lds = l->lds_kb < 0 ? -1 : 1024 * l->lds_kb;
  }
  (void)keep_e;
  drop_graphs(c);
  return ORAD_OK;
}

}  // extern "C"

// Device state a cycle changes, saved and restored around autotuner trials.  The
// live buffers are looked up again at every restore: the brick order swaps them.
static std::vector<std::pair<void *, size_t>> cycle_buffers(orad_ctx *c) {
  const size_t nn = (size_t)c->numnod, ne = (size_t)nelem(c);
  const size_t ncol = c->family == FAM_SHELL ? (size_t)c->shncol() : NSTATE;
  const size_t nsti = c->family == FAM_SHELL ? (size_t)c->nsti() : 1;
  std::vector<std::pair<void *, size_t>> live = {
      {c->X4, nn * 32}, {c->V4, nn * 32}, {c->D, nn * 24}, {c->state, ncol * ne * c->tsize()},
      {c->sti, nsti * ne * c->tsize()}, {c->ctl, c->ctl_bytes()}};
  if (c->family == FAM_SHELL) live.push_back({c->VR4, nn * 32});
  live.push_back({c->bal_part, 4 * ((nn + 255) / 256) * sizeof(double)});
  live.push_back({c->fv_mem, nn * sizeof(double)});
  live.push_back({c->fv_work, nn * sizeof(double)});
  if (c->ci) {
    // contact: deactivated secondary nodes (negated STFN) and the cumulative energies
    live.push_back({c->ci->stfns, (size_t)c->ci->nsn * sizeof(double)});
    live.push_back({c->ci->energy, 3 * sizeof(double)});
  }
  return live;
}

// ---------------------------------------------------------------- contact --
// constant_mod.F's EM10, EM20, EM30 and ZEP4, each a double operation on exact or
// previously rounded values, as the reference defines them.
static void contact_constants(Contact7 &k) {
  volatile double ten = 10.0;
  double ep = 100.0;                      // EP02
  for (int i = 3; i <= 10; ++i) ep = ep * ten;
  const double ep10 = ep;
  for (int i = 11; i <= 20; ++i) ep = ep * ten;
  const double ep20 = ep, ep30 = ep20 * ep10;
  k.em10 = 1.0 / ep10;
  k.em20 = 1.0 / ep20;
  k.em30 = 1.0 / ep30;
  k.zep4 = 4.0 / ten;
}

extern "C" {

int orad_set_contact7(orad_ctx *c, int nsn, const int *nsv, int nrtm, const int *irect,
                      const double *stfm, const double *stfns, const double *params,
                      int cand_cap) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!c->have_nodes) return fail(c, ORAD_ESTATE, "set_contact7: nodes must be set first");
  if (nsn < 0 || nrtm < 0 || (nsn > 0 && (!nsv || !stfns)) ||
      (nrtm > 0 && (!irect || !stfm)) || !params)
    return fail(c, ORAD_EINVAL, "set_contact7: bad sizes or null arrays");
  const double gap = params[0], fric = params[1], visc = params[2], viscf = params[3];
  if (!(gap > 0.0) || !(fric >= 0.0) || !(visc >= 0.0) || !(viscf >= 0.0))
    return fail(c, ORAD_EINVAL, "set_contact7: GAP > 0, FRIC, VISs, VISf >= 0");
  const int nn = c->numnod;
  for (int i = 0; i < nsn; ++i)
    if (nsv[i] < 1 || nsv[i] > nn) return fail(c, ORAD_EINVAL, "set_contact7: NSV out of range");
  for (int i = 0; i < 4 * nrtm; ++i)
    if (irect[i] < 1 || irect[i] > nn)
      return fail(c, ORAD_EINVAL, "set_contact7: IRECT out of range");
  // batched models: a segment's nodes in one model, the secondary nodes in increasing
  // order so that each model's are contiguous
  const int nmodel = c->batched ? c->nmodel : 1;
  std::vector<int32_t> h_nsv_off(nmodel + 1, 0), h_seg_model;
  h_nsv_off[nmodel] = nsn;
  if (c->batched) {
    if (nmodel > 16384) return fail(c, ORAD_EINVAL, "set_contact7: at most 16384 models");
    auto model_of = [&](int node0) {
      return (int)(std::upper_bound(c->h_node_off.begin(), c->h_node_off.end(), node0) -
                   c->h_node_off.begin()) - 1;
    };
    for (int i = 1; i < nsn; ++i)
      if (nsv[i] <= nsv[i - 1])
        return fail(c, ORAD_EINVAL, "set_contact7: with batched models NSV must increase");
    for (int m = 0; m <= nmodel; ++m)
      h_nsv_off[m] = (int)(std::lower_bound(nsv, nsv + nsn, c->h_node_off[m] + 1) - nsv);
    h_seg_model.resize(nrtm);
    for (int l = 0; l < nrtm; ++l) {
      const int m = model_of(irect[4 * l] - 1);
      for (int k = 1; k < 4; ++k)
        if (model_of(irect[4 * l + k] - 1) != m)
          return fail(c, ORAD_EINVAL, "set_contact7: a segment's nodes lie in two models");
      h_seg_model[l] = m;
    }
  }
  const long long cap = cand_cap > 0 ? (long long)cand_cap
                                     : std::max<long long>(8LL * nsn, 1024LL);
  long long cells = 1024;   // a power of two, at least 2 per segment
  while (cells < 2LL * nrtm) cells *= 2;
  const long long lim = contact7_scan_max();
  if (cap > lim || cells > lim || nsn > lim || nn > lim ||
      (long long)nsn * I7_KEEP_CAP > (1LL << 31) - 1 ||
      ((cap + I7_NVSIZ - 1) / I7_NVSIZ + (c->batched ? c->nmodel : 1)) * I7_SLOTS >
          (1LL << 31) - 1)
    return fail(c, ORAD_EINVAL, "set_contact7: interface too large for one pass");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  free_contact(c);
  drop_graphs(c);
  Contact7 *k = new (std::nothrow) Contact7();
  if (!k) return fail(c, ORAD_EHIP, "set_contact7: out of host memory");
  c->ci = k;
  k->nsn = nsn;
  k->nrtm = nrtm;
  k->numnod = nn;
  k->gap = gap;
  k->fric = fric;
  k->visc = visc;
  k->viscf = viscf;
  k->tz = gap * 1.000001;
  // candidates searched with half a gap of margin and kept until a node has moved a
  // quarter gap (inter_check_sort.F:103-106 re-sorts when the margin TZINF - GAP is used up)
  k->margin = 0.5 * gap;
  k->tzs = k->tz + k->margin;
  contact_constants(*k);
  k->cell_cap = (int)cells;
  k->pair_cap = (int)cap;
  k->nmodel = nmodel;
  k->node_model = c->batched ? c->node_model : nullptr;
  k->nbatch_cap = (int)((cap + I7_NVSIZ - 1) / I7_NVSIZ) + nmodel;
  const size_t nslots = (size_t)k->nbatch_cap * I7_SLOTS;
  const size_t ntmp = contact7_scan_tmp_ints((int)std::max<long long>(cells, nsn));
  hipError_t e = hipSuccess;
  auto al = [&](auto **p, size_t bytes) {
    if (e == hipSuccess) e = dalloc(p, bytes);
  };
  al(&k->nsv, (size_t)nsn * 4);
  al(&k->irect, (size_t)nrtm * sizeof(int4));
  al(&k->stfm, (size_t)nrtm * 8);
  al(&k->stfns, (size_t)nsn * 8);
  al(&k->bounds, 8 * 8);
  k->nbound_blocks = std::max(1, (nrtm + 255) / 256);
  al(&k->bpart, (size_t)7 * k->nbound_blocks * 8);
  al(&k->grid, 4 * 8);
  al(&k->seg_lo, (size_t)nrtm * sizeof(float4));
  al(&k->seg_hi, (size_t)nrtm * sizeof(float4));
  al(&k->seg_cell, (size_t)nrtm * 4);
  al(&k->cell_count, (size_t)cells * 4);
  al(&k->cell_fill, (size_t)cells * 4);
  al(&k->cell_start, (size_t)(cells + 1) * 4);
  al(&k->cell_lo, (size_t)nrtm * sizeof(float4));
  al(&k->cell_hi, (size_t)nrtm * sizeof(float4));
  al(&k->node_keep, (size_t)nsn * I7_KEEP_CAP * 4);
  al(&k->node_cand, (size_t)nsn * I7_NODE_CAP * 4);
  al(&k->node_ncand, (size_t)nsn * 4);
  al(&k->node_nkeep, (size_t)nsn * 4);
  al(&k->node_kstart, (size_t)(nsn + 1) * 4);
  al(&k->counters, I7C_N * 4);
  al(&k->sky, nslots * sizeof(double4));
  al(&k->sky_node, nslots * 4);
  al(&k->sky_next, nslots * 4);
  al(&k->batch_cnt, (size_t)k->nbatch_cap * 4);
  al(&k->batch_e, (size_t)k->nbatch_cap * 16);
  al(&k->energy, 3 * 8);
  al(&k->node_head, (size_t)nn * 4);
  al(&k->scan_tmp, ntmp * 4);
  al(&k->nsv_off, (size_t)(nmodel + 1) * 4);
  al(&k->model_bstart, (size_t)(nmodel + 1) * 4);
  al(&k->model_ext, (size_t)nmodel * 8);
  k->ndisp_blocks = std::max(1, (nn + 255) / 256);
  al(&k->x_search, (size_t)nn * sizeof(double4));
  al(&k->dpart, (size_t)k->ndisp_blocks * 8);
  if (c->batched) al(&k->seg_model, (size_t)nrtm * 4);
  if (e != hipSuccess) {
    free_contact(c);
    return fail(c, ORAD_EHIP, std::string("set_contact7: ") + hipGetErrorString(e));
  }
  std::vector<int32_t> hn(nsn);
  for (int i = 0; i < nsn; ++i) hn[i] = nsv[i] - 1;
  std::vector<int4> hr(nrtm);
  for (int l = 0; l < nrtm; ++l)
    hr[l] = make_int4(irect[4 * l] - 1, irect[4 * l + 1] - 1, irect[4 * l + 2] - 1,
                      irect[4 * l + 3] - 1);
  if (nsn) {
    HIPCHK(c, hipMemcpy(k->nsv, hn.data(), (size_t)nsn * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(k->stfns, stfns, (size_t)nsn * 8, hipMemcpyHostToDevice));
  }
  if (nrtm) {
    HIPCHK(c, hipMemcpy(k->irect, hr.data(), (size_t)nrtm * sizeof(int4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(k->stfm, stfm, (size_t)nrtm * 8, hipMemcpyHostToDevice));
  }
  HIPCHK(c, hipMemcpy(k->nsv_off, h_nsv_off.data(), (size_t)(nmodel + 1) * 4,
                      hipMemcpyHostToDevice));
  HIPCHK(c, hipMemset(k->model_bstart, 0, (size_t)(nmodel + 1) * 4));
  if (c->batched && nrtm)
    HIPCHK(c, hipMemcpy(k->seg_model, h_seg_model.data(), (size_t)nrtm * 4,
                        hipMemcpyHostToDevice));
  HIPCHK(c, hipMemset(k->counters, 0, I7C_N * 4));
  HIPCHK(c, hipMemset(k->energy, 0, 3 * 8));
  HIPCHK(c, hipMemset(k->node_head, 0xFF, (size_t)nn * 4));      // empty lists (-1)
  HIPCHK(c, hipMemset(k->x_search, 0xFF, (size_t)nn * sizeof(double4)));   // NaN: no search
  HIPCHK(c, hipMemset(k->node_ncand, 0, (size_t)nsn * 4));
  HIPCHK(c, hipDeviceSynchronize());
  return ORAD_OK;
}

int orad_contact7(orad_ctx *c, double dt1, double *dt2t) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!c->ci) return fail(c, ORAD_ESTATE, "contact7: no contact interface (orad_set_contact7)");
  if (c->batched)
    return fail(c, ORAD_EINVAL, "contact7: batched models run their contact in orad_run_cycles");
  int r = ready(c);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  const unsigned long long inf = 0x7FF0000000000000ull;
  HIPCHK(c, hipMemcpyAsync(c->scratch_dt1, &dt1, 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->scratch_dt2t, &inf, 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, contact7_run(*c->ci, c->X4, c->V4, c->scratch_dt1, c->scratch_dt2t, 0, c->stream));
  unsigned long long bits = inf;
  HIPCHK(c, hipMemcpyAsync(&bits, c->scratch_dt2t, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if ((r = contact_overflow(c))) return r;
  c->ci_fresh = true;
  double d;
  std::memcpy(&d, &bits, 8);
  if (dt2t && d < *dt2t) *dt2t = d;
  return ORAD_OK;
}

int orad_get_contact7(orad_ctx *c, double *energies, int *counts, double *stfns) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!c->ci) return fail(c, ORAD_ESTATE, "get_contact7: no contact interface");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (energies) HIPCHK(c, hipMemcpy(energies, c->ci->energy, 3 * 8, hipMemcpyDeviceToHost));
  if (counts) {
    int k[I7C_N];
    HIPCHK(c, hipMemcpy(k, c->ci->counters, sizeof k, hipMemcpyDeviceToHost));
    counts[0] = k[I7C_NCAND];
    counts[1] = k[I7C_NPEN];
    counts[2] = k[I7C_NSKY];
    counts[3] = k[I7C_OVERFLOW];
  }
  if (stfns && c->ci->nsn)
    HIPCHK(c, hipMemcpy(stfns, c->ci->stfns, (size_t)c->ci->nsn * 8, hipMemcpyDeviceToHost));
  return ORAD_OK;
}

int orad_get_contact7_sky(orad_ctx *c, int cap, double *fsky, int *isky, int *nsky) {
  if (!c || !nsky) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!c->ci) return fail(c, ORAD_ESTATE, "get_contact7_sky: no contact interface");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const Contact7 &k = *c->ci;
  int cnt[I7C_N];
  HIPCHK(c, hipMemcpy(cnt, k.counters, sizeof cnt, hipMemcpyDeviceToHost));
  const int nb = cnt[I7C_NBATCH];
  std::vector<int32_t> bc(nb > 0 ? nb : 1);
  if (nb) HIPCHK(c, hipMemcpy(bc.data(), k.batch_cnt, (size_t)nb * 4, hipMemcpyDeviceToHost));
  int n = 0;
  for (int b = 0; b < nb; ++b) n += bc[b];
  *nsky = n;
  if (!fsky && !isky) return ORAD_OK;
  if (n > cap) return fail(c, ORAD_EINVAL, "get_contact7_sky: output capacity too small");
  std::vector<double4> sk((size_t)nb * I7_SLOTS + 1);
  std::vector<int32_t> nd((size_t)nb * I7_SLOTS + 1);
  if (nb) {
    HIPCHK(c, hipMemcpy(sk.data(), k.sky, (size_t)nb * I7_SLOTS * sizeof(double4),
                        hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(nd.data(), k.sky_node, (size_t)nb * I7_SLOTS * 4, hipMemcpyDeviceToHost));
  }
  int o = 0;
  for (int b = 0; b < nb; ++b)
    for (int j = 0; j < bc[b]; ++j, ++o) {
      const double4 &v = sk[(size_t)b * I7_SLOTS + j];
      if (fsky) {
        fsky[4 * o] = v.x; fsky[4 * o + 1] = v.y; fsky[4 * o + 2] = v.z; fsky[4 * o + 3] = v.w;
      }
      if (isky) isky[o] = nd[(size_t)b * I7_SLOTS + j] + 1;
    }
  return ORAD_OK;
}

}  // extern "C"

// ------------------------------------------------------------ restart files --
// The dynamic state of a context (the cycle buffers above, the cycle counter and the
// energy balance's ENTOT0) in a file, so a run can stop and resume bit for bit in a
// context built from the same model.  Layout (little-endian): 8-byte magic
// "ORADRST1"; uint32 version; uint32 number of sections; a header of int64 values
// (family, precision, Ishell, numnod, elements, models, state columns, value size,
// host cycle count, ENTOT0 known) and ENTOT0 as a double; the model fingerprint
// (uint64); then per section its byte count (uint64) and bytes, in cycle_buffers
// order; last a 64-bit checksum of every section's bytes.

static const char RESTART_MAGIC[8] = {'O', 'R', 'A', 'D', 'R', 'S', 'T', '1'};
static constexpr uint32_t RESTART_VERSION = 1;

// 64-bit FNV-1a over bytes.
static uint64_t fnv1a(uint64_t h, const void *p, size_t n) {
  const unsigned char *b = (const unsigned char *)p;
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}

// Fingerprint of what the state's layout depends on: the sizes, the PARITH/ON slot
// table (ADSKY, IADS/IADC) and, for bricks, the library's element order.
static uint64_t model_fingerprint(const orad_ctx *c) {
  uint64_t h = 1469598103934665603ull;
  const int64_t v[8] = {c->family, c->precision, c->ishell, c->numnod, nelem(c), c->nmodel,
                        c->family == FAM_SHELL ? c->shncol() : NSTATE, (int64_t)c->tsize()};
  h = fnv1a(h, v, sizeof v);
  h = fnv1a(h, c->h_adsky.data(), c->h_adsky.size() * sizeof(int32_t));
  h = fnv1a(h, c->h_iads.data(), c->h_iads.size() * sizeof(int32_t));
  h = fnv1a(h, c->h_perm.data(), c->h_perm.size() * sizeof(int32_t));
  const int64_t ci[2] = {c->ci ? c->ci->nsn : -1, c->ci ? c->ci->nrtm : -1};
  h = fnv1a(h, ci, sizeof ci);
  return h;
}

static void restart_header(const orad_ctx *c, int64_t hdr[10]) {
  hdr[0] = c->family;
  hdr[1] = c->precision;
  hdr[2] = c->ishell;
  hdr[3] = c->numnod;
  hdr[4] = nelem(c);
  hdr[5] = c->nmodel;
  hdr[6] = c->family == FAM_SHELL ? c->shncol() : NSTATE;
  hdr[7] = (int64_t)c->tsize();
  hdr[8] = c->host_ncycle;
  hdr[9] = c->have_entot0 ? 1 : 0;
}

extern "C" {

int orad_write_restart(orad_ctx *c, const char *path) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!path) return fail(c, ORAD_EINVAL, "write_restart: null path");
  int r = ready(c);
  if (r) return r;
  if (c->family == FAM_BRICK ? (r = init_state(c)) : (r = init_shell_state(c))) return r;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const auto live = cycle_buffers(c);
  std::vector<std::vector<char>> host(live.size());
  uint64_t sum = 1469598103934665603ull;
  for (size_t i = 0; i < live.size(); ++i) {
    host[i].resize(live[i].second);
    HIPCHK(c, hipMemcpy(host[i].data(), live[i].first, live[i].second, hipMemcpyDeviceToHost));
    sum = fnv1a(sum, host[i].data(), host[i].size());
  }
  int64_t hdr[10];
  restart_header(c, hdr);
  const double entot0 = c->entot0;
  const uint64_t fp = model_fingerprint(c);
  const uint32_t ver = RESTART_VERSION, nsec = (uint32_t)live.size();
  const std::string tmp = std::string(path) + ".tmp";
  FILE *fpo = std::fopen(tmp.c_str(), "wb");
  if (!fpo) return fail(c, ORAD_EINVAL, std::string("write_restart: cannot open ") + tmp);
  bool ok = std::fwrite(RESTART_MAGIC, 1, 8, fpo) == 8 && std::fwrite(&ver, 4, 1, fpo) == 1 &&
            std::fwrite(&nsec, 4, 1, fpo) == 1 && std::fwrite(hdr, 8, 10, fpo) == 10 &&
            std::fwrite(&entot0, 8, 1, fpo) == 1 && std::fwrite(&fp, 8, 1, fpo) == 1;
  for (size_t i = 0; ok && i < host.size(); ++i) {
    const uint64_t n = host[i].size();
    ok = std::fwrite(&n, 8, 1, fpo) == 1 && std::fwrite(host[i].data(), 1, n, fpo) == n;
  }
  ok = ok && std::fwrite(&sum, 8, 1, fpo) == 1;
  ok = (std::fclose(fpo) == 0) && ok;
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return fail(c, ORAD_EINVAL, std::string("write_restart: cannot write ") + path);
  }
  return ORAD_OK;
}

int orad_read_restart(orad_ctx *c, const char *path) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!path) return fail(c, ORAD_EINVAL, "read_restart: null path");
  int r = ready(c);
  if (r) return r;
  if (c->family == FAM_BRICK ? (r = init_state(c)) : (r = init_shell_state(c))) return r;
  FILE *fpi = std::fopen(path, "rb");
  if (!fpi) return fail(c, ORAD_EINVAL, std::string("read_restart: cannot open ") + path);
  auto bad = [&](const std::string &why) {
    std::fclose(fpi);
    return fail(c, ORAD_EINVAL, "read_restart: " + why);
  };
  char magic[8];
  uint32_t ver = 0, nsec = 0;
  int64_t hdr[10], want[10];
  double entot0 = 0.0;
  uint64_t fp = 0;
  if (std::fread(magic, 1, 8, fpi) != 8 || std::memcmp(magic, RESTART_MAGIC, 8) != 0)
    return bad("not a restart file of this library");
  if (std::fread(&ver, 4, 1, fpi) != 1 || ver != RESTART_VERSION)
    return bad("unsupported restart version");
  if (std::fread(&nsec, 4, 1, fpi) != 1 || std::fread(hdr, 8, 10, fpi) != 10 ||
      std::fread(&entot0, 8, 1, fpi) != 1 || std::fread(&fp, 8, 1, fpi) != 1)
    return bad("truncated header");
  restart_header(c, want);
  for (int i = 0; i < 8; ++i)
    if (hdr[i] != want[i])
      return bad("the file was written by a different model (family, precision, Ishell, "
                 "sizes or state layout)");
  if (fp != model_fingerprint(c))
    return bad("the file was written by a model with another mesh or element order");
  const auto live = cycle_buffers(c);
  if (nsec != live.size()) return bad("section count does not match this model");
  std::vector<std::vector<char>> host(live.size());
  uint64_t sum = 1469598103934665603ull;
  for (size_t i = 0; i < live.size(); ++i) {
    uint64_t n = 0;
    if (std::fread(&n, 8, 1, fpi) != 1 || n != live[i].second)
      return bad("section size does not match this model");
    host[i].resize(n);
    if (std::fread(host[i].data(), 1, n, fpi) != n) return bad("truncated section");
    sum = fnv1a(sum, host[i].data(), n);
  }
  uint64_t want_sum = 0;
  if (std::fread(&want_sum, 8, 1, fpi) != 1 || want_sum != sum)
    return bad("checksum mismatch (corrupt or truncated file)");
  std::fclose(fpi);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < live.size(); ++i)
    HIPCHK(c, hipMemcpy(live[i].first, host[i].data(), host[i].size(), hipMemcpyHostToDevice));
  c->host_ncycle = (int)hdr[8];
  c->have_entot0 = hdr[9] != 0;
  c->entot0 = entot0;
  drop_graphs(c);
  return ORAD_OK;
}

}  // extern "C"

struct Snapshot {
  std::vector<void *> copies;   // in cycle_buffers order
  std::vector<size_t> sizes;
  int host_ncycle = 0;
  hipError_t take(orad_ctx *c) {
    host_ncycle = c->host_ncycle;
    for (auto &b : cycle_buffers(c)) {
      void *p = nullptr;
      hipError_t e = dalloc(&p, b.second);
      if (e == hipSuccess)
        e = hipMemcpyAsync(p, b.first, b.second, hipMemcpyDeviceToDevice, c->stream);
      if (e != hipSuccess) {
        dfree(p);
        return e;
      }
      copies.push_back(p);
      sizes.push_back(b.second);
    }
    return hipStreamSynchronize(c->stream);
  }
  // copies on the context's stream, ordered with its kernels
  hipError_t restore(orad_ctx *c) {
    const auto live = cycle_buffers(c);
    for (size_t i = 0; i < copies.size(); ++i) {
      hipError_t e = hipMemcpyAsync(live[i].first, copies[i], sizes[i], hipMemcpyDeviceToDevice,
                                    c->stream);
      if (e != hipSuccess) return e;
    }
    c->host_ncycle = host_ncycle;
    return hipStreamSynchronize(c->stream);
  }
  void *state_copy() const { return copies.size() > 3 ? copies[3] : nullptr; }
  void *sti_copy() const { return copies.size() > 4 ? copies[4] : nullptr; }
  ~Snapshot() {
    for (void *p : copies) dfree(p);
  }
};

// Time per cycle of `cycles` graph-replayed cycles from the snapshot under the current
// configuration; +inf when a cycle reports a non-positive volume.
static int time_trial(orad_ctx *c, Snapshot &snap, int cycles, double *per) {
  *per = 1e300;
  HIPCHK(c, snap.restore(c));
  int r = orad_run_cycles(c, 2, nullptr);                     // captures the graphs
  if (r == ORAD_OK) {
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    r = orad_run_cycles(c, cycles, nullptr);
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev[1]));
  }
  if (r == ORAD_ENEGVOL) return ORAD_OK;
  if (r != ORAD_OK) return r;
  float ms = 0.f;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  *per = ms / cycles;
  return ORAD_OK;
}

extern "C" {

int orad_autotune(orad_ctx *c, int cycles, orad_launch_t *best, double *best_ms) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  int r = ready(c);
  if (r) return r;
  if (cycles < 2) return fail(c, ORAD_EINVAL, "autotune: at least 2 timed cycles");
  cycles += cycles & 1;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  orad_launch_t keep{};
  orad_get_launch(c, &keep);
  Snapshot snap;
  HIPCHK(c, snap.take(c));
  double tbest = 1e300;
  orad_launch_t lbest = keep;
  lbest.order_rows = -1;
  // one candidate: `l` applied, timed, kept when faster
  auto trial = [&](orad_launch_t l) -> int {
    if (orad_set_launch(c, &l) != ORAD_OK) return ORAD_OK;        // does not fit: skip
    double per = 1e300;
    const int rr = time_trial(c, snap, cycles, &per);
    if (rr) return rr;
    if (per < tbest) {
      tbest = per;
      lbest = l;
    }
    return ORAD_OK;
  };
  const bool mixed = c->precision == ORAD_MIXED;
  if (c->family == FAM_BRICK) {
    // 1. element order: the state in the snapshot is permuted along with the live one
    int best_rows = c->order_rows;
    for (int rows : {0, 8, 10, 16, 32}) {
      r = brick_reorder(c, rows, {{snap.state_copy(), snap.sti_copy()}});
      if (r) return r;
      orad_launch_t l = lbest;
      l.order_rows = -1;
      const double before = tbest;
      r = trial(l);
      if (r) return r;
      if (tbest < before) best_rows = rows;
    }
    r = brick_reorder(c, best_rows, {{snap.state_copy(), snap.sti_copy()}});
    if (r) return r;
    // 2. element-kernel instance, 3. LDS cap, 4. XCD mapping, 5. nodal slot tile
    const int shapes[][2] = {{256, 2}, {256, 3}, {128, 2}, {64, 2}, {64, 1}, {128, 1}, {256, 1}};
    orad_launch_t base = lbest;
    for (const auto &sh : shapes) {
      if (!(mixed ? brick_shape_ok_mixed(sh[0], sh[1]) : brick_shape_ok_fp64(sh[0], sh[1])))
        continue;
      orad_launch_t l = base;
      l.block = sh[0]; l.waves = sh[1]; l.lds_kb = -1;
      if ((r = trial(l))) return r;
    }
    base = lbest;
    for (int lds : {0, 40, 54, 80}) {
      orad_launch_t l = base;
      l.lds_kb = lds;
      if ((r = trial(l))) return r;
    }
    base = lbest;
    {
      orad_launch_t l = base;
      l.xcd_remap = !base.xcd_remap;
      if ((r = trial(l))) return r;
    }
  } else if (c->batched) {
    for (int lds : {0, 40, 80}) {
      for (int xcd : {0, 1}) {
        orad_launch_t l = keep;
        l.persistent = 0; l.lds_kb = lds; l.xcd_remap = xcd;
        if ((r = trial(l))) return r;
      }
      orad_launch_t l = keep;
      l.persistent = 1; l.lds_kb = lds; l.xcd_remap = 0;
      if ((r = trial(l))) return r;
    }
  } else {
    for (int xcd : {0, 1}) {
      orad_launch_t l = keep;
      l.xcd_remap = xcd;
      if ((r = trial(l))) return r;
    }
  }
  {
    const orad_launch_t base = lbest;
    for (int tile : {1024, 2048, 4096}) {
      if (!node_tile_ok(mixed, c->family == FAM_SHELL, tile) || tile == base.node_tile) continue;
      orad_launch_t l = base;
      l.node_tile = tile;
      if ((r = trial(l))) return r;
    }
  }
  HIPCHK(c, snap.restore(c));
  r = orad_set_launch(c, tbest < 1e300 ? &lbest : &keep);
  if (r) return r;
  if (best) orad_get_launch(c, best);
  if (best_ms) *best_ms = tbest;
  c->err.clear();
  return ORAD_OK;
}

}  // extern "C"

// ============================================================== adjoint ===

// Checkpoints of an adjoint forward run: checkpoint s holds the state before cycle
// s * every; fin the state after the last cycle.
struct AdjStore {
  int ncycle = 0, every = 0;
  std::vector<Snapshot *> ck;
  Snapshot *fin = nullptr;
  ~AdjStore() {
    for (Snapshot *p : ck) delete p;
    delete fin;
  }
};

static void free_adjoint(orad_ctx *c) {
  delete c->adj;
  c->adj = nullptr;
}

// Device buffers of one reverse sweep (freed on scope exit).
struct DevBufs {
  std::vector<void *> ptrs;
  template <typename P>
  hipError_t get(P **p, size_t bytes, bool zero = true) {
    hipError_t e = dalloc(p, bytes);
    if (e != hipSuccess) return e;
    ptrs.push_back((void *)*p);
    return zero ? hipMemset(*p, 0, bytes) : hipSuccess;
  }
  ~DevBufs() {
    for (void *p : ptrs) dfree(p);
  }
};

// The reverse sweep's stream (default priority, like the context's) and the events
// that order it against the recompute on the context's stream.  Both streams are
// drained before the buffers of the sweep are released.
struct AdjStreams {
  hipStream_t ctx_stream, s = nullptr;
  hipEvent_t stashed[2] = {nullptr, nullptr}, read_done[2] = {nullptr, nullptr};
  explicit AdjStreams(hipStream_t cs) : ctx_stream(cs) {}
  hipError_t init() {
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
      e = hipEventCreateWithFlags(&stashed[i], hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&read_done[i], hipEventDisableTiming);
    }
    return e;
  }
  ~AdjStreams() {
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
    (void)hipStreamSynchronize(ctx_stream);
    for (int i = 0; i < 2; ++i) {
      if (stashed[i]) (void)hipEventDestroy(stashed[i]);
      if (read_done[i]) (void)hipEventDestroy(read_done[i]);
    }
  }
};

extern "C" {

int orad_adjoint_forward(orad_ctx *c, int ncycle, int every) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if ((c->family != FAM_SHELL && c->family != FAM_BRICK) || c->batched)
    return fail(c, ORAD_ESTATE, "adjoint_forward: one shell or brick model only");
  int r = ready(c);
  if (r) return r;
  if (ncycle < 1 || every < 1) return fail(c, ORAD_EINVAL, "adjoint_forward: bad cycle counts");
  // The adjoint state carries SIG, PLA and (shells) the hourglass forces.  The
  // strain-rate filters (EPSD, EPSDP) and the shell thickness THK are also carried from
  // cycle to cycle, but they reach the forces only through the Johnson-Cook rate
  // factor, i.e. when C != 0.
  if ((c->family == FAM_SHELL ? c->shmat.cc : c->mat.cc) != 0.0)
    return fail(c, ORAD_EINVAL, "adjoint_forward: rate-dependent law 2 (C != 0) not differentiated");
  if (c->qbat())
    return fail(c, ORAD_EINVAL, "adjoint_forward: QBAT shells (Ishell 12) are not differentiated");
  HIPCHK(c, hipSetDevice(c->device));
  free_adjoint(c);
  AdjStore *a = new (std::nothrow) AdjStore();
  if (!a) return fail(c, ORAD_EINVAL, "adjoint_forward: out of host memory");
  c->adj = a;
  a->ncycle = ncycle;
  a->every = every;
  for (int k = 0; k < ncycle; ++k) {
    if (k % every == 0) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      Snapshot *sp = new (std::nothrow) Snapshot();
      if (!sp) return fail(c, ORAD_EINVAL, "adjoint_forward: out of host memory");
      a->ck.push_back(sp);
      HIPCHK(c, sp->take(c));
    }
    HIPCHK(c, enqueue_cycle(c, c->host_ncycle & 1));
    c->host_ncycle += 1;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  a->fin = new (std::nothrow) Snapshot();
  if (!a->fin) return fail(c, ORAD_EINVAL, "adjoint_forward: out of host memory");
  HIPCHK(c, a->fin->take(c));
  return ORAD_OK;
}

int orad_adjoint_reverse(orad_ctx *c, const double *lam_x, const double *lam_v,
                         const double *lam_vr, const double *lam_d, const double *lam_state,
                         double *g_jc, double *g_thk, double *g_ms, double *g_in,
                         double *lam_x0, double *lam_v0, double *lam_vr0) {
  if (!c) return ORAD_EINVAL;
  if (int fr = flush_groups(c)) return fr;
  if (!c->adj || !c->adj->fin) return fail(c, ORAD_ESTATE, "adjoint_reverse: no forward run");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  AdjStore *A = c->adj;
  const bool brick = c->family == FAM_BRICK;
  const size_t nn = (size_t)c->numnod, ts = c->tsize();
  const size_t ne = brick ? (size_t)c->numels : (size_t)c->numelc;
  const size_t stb = (size_t)(brick ? NSTATE : c->shncol()) * ne * ts;
  const int nst = brick ? ADJB_NST : c->qeph() ? AdjShape<true>::NST : AdjShape<false>::NST;
  const int nco = brick ? 8 : 4, per = brick ? 6 : 9;   // corners, inputs per corner
  const int L = A->every;
  DevBufs B;
  double *lx, *lv, *ld, *lvr, *lamS, *lamR, *gms, *gin, *lst[2], *lcorner, *gthk, *gjc;
  int32_t *scorner, *selem;
  HIPCHK(c, B.get(&lx, 3 * nn * 8));
  HIPCHK(c, B.get(&lv, 3 * nn * 8));
  HIPCHK(c, B.get(&ld, 3 * nn * 8));
  HIPCHK(c, B.get(&lvr, 3 * nn * 8));
  HIPCHK(c, B.get(&lamS, 3 * nn * 8));
  HIPCHK(c, B.get(&lamR, 3 * nn * 8));
  HIPCHK(c, B.get(&gms, nn * 8));
  HIPCHK(c, B.get(&gin, nn * 8));
  HIPCHK(c, B.get(&lst[0], (size_t)nst * ne * 8));
  HIPCHK(c, B.get(&lst[1], (size_t)nst * ne * 8));
  HIPCHK(c, B.get(&lcorner, (size_t)nco * per * ne * 8));
  HIPCHK(c, B.get(&gthk, ne * 8));
  HIPCHK(c, B.get(&gjc, 3 * ne * 8));
  HIPCHK(c, B.get(&scorner, (size_t)nco * ne * 4, false));
  HIPCHK(c, B.get(&selem, (size_t)nco * ne * 4, false));
  // per-cycle stash slots, strides rounded to 256 bytes so every slot is 16-byte aligned
  const size_t sa_stride = (3 * nn + 31) / 32 * 32;        // doubles
  const size_t st_stride = (stb + 255) / 256 * 256;        // bytes
  struct Stash {
    double4 *sx, *sv, *svr;
    double *sa, *sar;
    char *sst;
    Scal *sscal;
  };
  auto get_stash = [&](Stash &t) {
    hipError_t e = B.get(&t.sx, (size_t)L * nn * 32, false);
    if (e == hipSuccess) e = B.get(&t.sv, (size_t)L * nn * 32, false);
    if (e == hipSuccess) e = B.get(&t.svr, (size_t)L * nn * 32 * (brick ? 0 : 1) + 256, false);
    if (e == hipSuccess) e = B.get(&t.sa, (size_t)L * sa_stride * 8, false);
    if (e == hipSuccess) e = B.get(&t.sar, (size_t)L * sa_stride * 8 * (brick ? 0 : 1) + 256, false);
    if (e == hipSuccess) e = B.get(&t.sst, (size_t)L * st_stride, false);
    if (e == hipSuccess) e = B.get(&t.sscal, (size_t)L * sizeof(Scal), false);
    return e;
  };
  // Two stash sets when memory allows: the recompute of segment s-1 (on the context's
  // stream) then runs while the reverse sweep of segment s (on its own stream) reads
  // the other set; with one set the two alternate.
  Stash SS[2];
  HIPCHK(c, get_stash(SS[0]));
  int nbuf = 1;
  {
    const size_t mark = B.ptrs.size();
    if (get_stash(SS[1]) == hipSuccess) {
      nbuf = 2;
    } else {
      (void)hipGetLastError();   // the failed allocation is not an error of the run
      while (B.ptrs.size() > mark) { dfree(B.ptrs.back()); B.ptrs.pop_back(); }
    }
  }
  AdjStreams R(c->stream);
  HIPCHK(c, R.init());
  hipStream_t as = R.s;
  // device element index of an engine element (bricks are stored in the library's order)
  auto dev_elem = [&](size_t e) { return brick ? (size_t)c->h_iperm[e] : e; };
  {
    // element and corner of every slot, for the ordered gather of corner adjoints
    std::vector<int32_t> sc((size_t)nco * ne), se((size_t)nco * ne);
    for (size_t e = 0; e < ne; ++e)
      for (int k = 0; k < nco; ++k) {
        sc[c->h_iads[(size_t)nco * e + k]] = k;
        se[c->h_iads[(size_t)nco * e + k]] = (int32_t)dev_elem(e);
      }
    HIPCHK(c, hipMemcpy(scorner, sc.data(), (size_t)nco * ne * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(selem, se.data(), (size_t)nco * ne * 4, hipMemcpyHostToDevice));
  }
  if (lam_x) HIPCHK(c, hipMemcpy(lx, lam_x, 3 * nn * 8, hipMemcpyHostToDevice));
  if (lam_v) HIPCHK(c, hipMemcpy(lv, lam_v, 3 * nn * 8, hipMemcpyHostToDevice));
  if (lam_vr && !brick) HIPCHK(c, hipMemcpy(lvr, lam_vr, 3 * nn * 8, hipMemcpyHostToDevice));
  if (lam_d) HIPCHK(c, hipMemcpy(ld, lam_d, 3 * nn * 8, hipMemcpyHostToDevice));
  if (lam_state) {
    // rows of nst (SIG, PLA, then HOUR(1:3) or HOURG(1:12)) -> columns
    std::vector<double> col((size_t)nst * ne);
    for (size_t e = 0; e < ne; ++e)
      for (int j = 0; j < nst; ++j) col[j * ne + dev_elem(e)] = lam_state[(size_t)nst * e + j];
    HIPCHK(c, hipMemcpy(lst[0], col.data(), (size_t)nst * ne * 8, hipMemcpyHostToDevice));
  }
  int cur = 0;
  AdjArgs<double> ag{};
  ag.numelc = c->numelc;
  ag.conn = c->conn;
  ag.thke = c->thke;
  ag.lamS = lamS;
  ag.lamR = lamR;
  ag.lcorner = lcorner;
  ag.g_thk = gthk;
  ag.g_jc = gjc;
  ag.mat = c->shmat;
  ag.prop = c->shprop;
  ag.k = c->shk;
  AdjBrickArgs<double> ab{};
  ab.numels = c->numels;
  ab.conn = c->conn;
  ab.vol0 = c->vol0;
  ab.lamS = lamS;
  ab.lcorner = lcorner;
  ab.g_jc = gjc;
  ab.mat = c->mat;
  ab.prop = c->prop;
  // the hand-written reverse pass covers bricks and both shell formulations with
  // radial return
  const bool rev = c->adj_vjp == 0 && (brick || c->shmat.ipla == 0);
  const int nseg = (int)A->ck.size();
  std::vector<Scal> hs[2] = {std::vector<Scal>(L), std::vector<Scal>(L)};
  bool read_pending[2] = {false, false};   // a reverse sweep enqueued on `as` reads the set
  auto seg_len = [&](int sgi) { return std::min(A->every, A->ncycle - sgi * A->every); };
  auto buf_of = [&](int sgi) { return nbuf == 2 ? (sgi & 1) : 0; };
  // recompute segment sgi into its stash set, on the context's stream: before each
  // cycle one launch stashes its inputs (X, V, VR, element state) and the previous
  // cycle's accelerations and scalars; ends with the scalars copied to the host
  auto recompute = [&](int sgi) -> int {
    const int b = buf_of(sgi), len = seg_len(sgi);
    const Stash &t = SS[b];
    if (read_pending[b]) HIPCHK(c, hipStreamWaitEvent(c->stream, R.read_done[b], 0));
    HIPCHK(c, A->ck[sgi]->restore(c));
    drop_graphs(c);
    c->cycle_write_a = 1;
    hipError_t e = hipSuccess;
    auto stash = [&](int j, bool inputs) {
      const void *src[7];
      void *dst[7];
      size_t bytes[7];
      int n = 0;
      auto add = [&](const void *a, void *b2, size_t nb) {
        src[n] = a; dst[n] = b2; bytes[n] = nb; ++n;
      };
      if (inputs) {
        add(c->X4, t.sx + (size_t)j * nn, nn * 32);
        add(c->V4, t.sv + (size_t)j * nn, nn * 32);
        if (!brick) add(c->VR4, t.svr + (size_t)j * nn, nn * 32);
        add(c->state, t.sst + (size_t)j * st_stride, stb);
      }
      if (j > 0 || !inputs) {
        const int jp = inputs ? j - 1 : j;
        add(c->A, t.sa + (size_t)jp * sa_stride, 3 * nn * 8);
        if (!brick) add(c->AR, t.sar + (size_t)jp * sa_stride, 3 * nn * 8);
        // the scalars the previous cycle wrote sit in the current parity's slot
        add(&c->ctl->s[c->host_ncycle & 1], t.sscal + jp, sizeof(Scal));
      }
      return launch_adj_stash(n, src, dst, bytes, c->stream);
    };
    for (int j = 0; j < len && e == hipSuccess; ++j) {
      e = stash(j, true);
      if (e == hipSuccess) e = enqueue_cycle(c, c->host_ncycle & 1);
      c->host_ncycle += 1;
    }
    if (e == hipSuccess) e = stash(len - 1, false);
    c->cycle_write_a = 0;
    HIPCHK(c, e);
    HIPCHK(c, hipMemcpyAsync(hs[b].data(), t.sscal, len * sizeof(Scal), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipEventRecord(R.stashed[b], c->stream));
    return ORAD_OK;
  };
  if (nseg > 0) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (int r = recompute(nseg - 1)) return r;
  }
  for (int sgi = nseg - 1; sgi >= 0; --sgi) {
    const int b = buf_of(sgi), len = seg_len(sgi);
    const Stash &t = SS[b];
    HIPCHK(c, hipEventSynchronize(R.stashed[b]));   // the segment's scalars are on the host
    const std::vector<Scal> &h = hs[b];
    // reverse through the segment on `as`: the nodal reverse of the segment's last
    // cycle, then per cycle the element VJP and the gather, fused with the nodal
    // reverse of the cycle before it inside the segment
    auto slotX = [&](int j) { return t.sx + (size_t)j * nn; };
    auto slotV = [&](int j) { return t.sv + (size_t)j * nn; };
    // no rotational DOF for bricks: the nodal reverse skips them on null VR4 / AR
    auto slotW = [&](int j) { return brick ? (const double4 *)nullptr : t.svr + (size_t)j * nn; };
    auto slotA = [&](int j) { return t.sa + (size_t)j * sa_stride; };
    auto slotAR = [&](int j) { return brick ? (const double *)nullptr : t.sar + (size_t)j * sa_stride; };
    HIPCHK(c, launch_adj_node(c->numnod, slotX(len - 1), slotV(len - 1), slotW(len - 1),
                              slotA(len - 1), slotAR(len - 1), h[len - 1].dt2, h[len - 1].dt12,
                              lx, lv, ld, lvr, lamS, lamR, gms, gin, as));
    for (int j = len - 1; j >= 0; --j) {
      const double *stj = reinterpret_cast<const double *>(t.sst + (size_t)j * st_stride);
      if (brick) {
        ab.X4 = slotX(j);
        ab.V4 = slotV(j);
        ab.st = stj;
        ab.lst_in = lst[cur];
        ab.lst_out = lst[cur ^ 1];
        ab.dt1 = h[j].dt1;
        HIPCHK(c, launch_adj_brick_elem(c->precision == ORAD_MIXED, rev, ab, as));
      } else {
        ag.X4 = slotX(j);
        ag.V4 = slotV(j);
        ag.VR4 = slotW(j);
        ag.st = stj;
        ag.lst_in = lst[cur];
        ag.lst_out = lst[cur ^ 1];
        ag.dt1 = h[j].dt1;
        HIPCHK(c, launch_adj_elem(c->precision == ORAD_MIXED, c->qeph(), rev, ag, as));
      }
      if (j > 0) {
        HIPCHK(c, launch_adj_gather_node(c->numnod, (int)ne, per, c->adsky, selem, scorner, lcorner,
                                         slotX(j - 1), slotV(j - 1), slotW(j - 1), slotA(j - 1),
                                         slotAR(j - 1), h[j - 1].dt2, h[j - 1].dt12, lx, lv, ld,
                                         lvr, lamS, lamR, gms, gin, as));
      } else {
        HIPCHK(c, launch_adj_gather(c->numnod, (int)ne, per, c->adsky, selem, scorner, lcorner,
                                    lx, lv, lvr, as));
      }
      cur ^= 1;
    }
    HIPCHK(c, hipEventRecord(R.read_done[b], as));
    read_pending[b] = true;
    // the next segment's recompute overlaps this sweep (or, with one stash set, waits
    // for it)
    if (sgi > 0)
      if (int r = recompute(sgi - 1)) return r;
  }
  HIPCHK(c, hipStreamSynchronize(as));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (g_jc) {
    std::vector<double> h(3 * ne);
    HIPCHK(c, hipMemcpy(h.data(), gjc, 3 * ne * 8, hipMemcpyDeviceToHost));
    // summed in the engine's element order, whatever the device order
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
      for (size_t e2 = 0; e2 < ne; ++e2) acc += h[j * ne + dev_elem(e2)];
      g_jc[j] = acc;
    }
  }
  if (g_thk && !brick) HIPCHK(c, hipMemcpy(g_thk, gthk, ne * 8, hipMemcpyDeviceToHost));
  if (g_ms) HIPCHK(c, hipMemcpy(g_ms, gms, nn * 8, hipMemcpyDeviceToHost));
  if (g_in) {
    if (brick) std::fill(g_in, g_in + nn, 0.0);
    else HIPCHK(c, hipMemcpy(g_in, gin, nn * 8, hipMemcpyDeviceToHost));
  }
  if (lam_x0) HIPCHK(c, hipMemcpy(lam_x0, lx, 3 * nn * 8, hipMemcpyDeviceToHost));
  if (lam_v0) HIPCHK(c, hipMemcpy(lam_v0, lv, 3 * nn * 8, hipMemcpyDeviceToHost));
  if (lam_vr0) {
    if (brick) std::fill(lam_vr0, lam_vr0 + 3 * nn, 0.0);
    else HIPCHK(c, hipMemcpy(lam_vr0, lvr, 3 * nn * 8, hipMemcpyDeviceToHost));
  }
  // leave the context at the end of the forward run
  HIPCHK(c, A->fin->restore(c));
  drop_graphs(c);
  return ORAD_OK;
}

}  // extern "C"

static int get_fsky_shell(orad_ctx *c, double *fsky) {
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t ne = (size_t)c->numelc, ns = 4 * ne, nst = (size_t)c->nsti() * ne;
  std::vector<double> f(6 * ns), st(nst);
  if (c->precision == ORAD_MIXED) {
    std::vector<float> h(6 * ns), hs(nst);
    HIPCHK(c, hipMemcpy(h.data(), c->fsky, h.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(hs.data(), c->sti, nst * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < h.size(); ++i) f[i] = h[i];
    for (size_t i = 0; i < nst; ++i) st[i] = hs[i];
  } else {
    HIPCHK(c, hipMemcpy(f.data(), c->fsky, 6 * ns * 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(st.data(), c->sti, nst * 8, hipMemcpyDeviceToHost));
  }
  const bool two = c->nsti() == 2;   // STI per diagonal (QEPH, QBAT)
  for (size_t e = 0; e < ne; ++e)
    for (int k = 0; k < 4; ++k) {
      const size_t sl = (size_t)c->h_iads[4 * e + k];
      for (int j = 0; j < 6; ++j) fsky[8 * sl + j] = f[6 * sl + j];
      fsky[8 * sl + 6] = two ? st[e + (k & 1) * ne] : st[e];
      fsky[8 * sl + 7] = 0.0;
    }
  return ORAD_OK;
}

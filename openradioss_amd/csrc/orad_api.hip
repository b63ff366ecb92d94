// orad_api.hip -- the C ABI declared in include/orad.h.
//
// Host-side role of the reference's starter and RESOL driver for this path:
// owns device memory, builds the PARITH/ON element-corner CSR with the
// starter's ordering, and sequences the per-cycle kernels (captured once per
// cycle parity in a HIP graph and replayed).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/orad.h"
#include "orad_common.hpp"

namespace orad {
hipError_t launch_brick_fp64(const BrickArgs<double> &, int, int, const double *,
                             unsigned long long *, double *, int, int, hipStream_t);
hipError_t launch_brick_mixed(const BrickArgs<float> &, int, int, const double *,
                              unsigned long long *, double *, int, int, hipStream_t);
int brick_block_fp64();
int brick_block_mixed();
hipError_t launch_node(const NodeArgs &, bool, int, int, int, int, hipStream_t);
hipError_t launch_energy(bool, int, const void *, const double *, int, const double4 *,
                         const double4 *, double *, double *, hipStream_t);
}  // namespace orad

using namespace orad;

struct orad_ctx {
  int device = 0;
  int precision = ORAD_FP64;
  hipStream_t stream = nullptr;
  std::string err;

  int numnod = 0, numels = 0;
  bool have_nodes = false, have_bricks = false, have_mat = false, have_prop = false;
  bool state_init = false;
  Law2 mat{};
  SProp prop{};

  std::vector<int32_t> h_adsky, h_iads;  // 0-based, element-major iads [e*8+k]

  // device buffers
  double4 *X4 = nullptr, *V4 = nullptr;
  double *D = nullptr, *A = nullptr, *STIFN = nullptr;
  int32_t *conn = nullptr, *iads = nullptr, *adsky = nullptr;
  void *state = nullptr;  // SoA element state of the storage type (NSTATE columns)
  double *vol0 = nullptr; // initial volumes, FP64
  void *fsky = nullptr;
  void *sti = nullptr;              // per-element STI/4
  int32_t *slot_elem = nullptr;     // element owning each slot
  double *ehou_part = nullptr;
  int n_ehou_part = 0;
  CycleCtl *ctl = nullptr;
  double *scratch_dt1 = nullptr;
  unsigned long long *scratch_dt2t = nullptr;

  int host_ncycle = 0;  // cycle counter mirrored on the host (its parity selects buffers)
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // two cycles starting at parity 0 / 1
  int xcd_remap = 1;
  int elem_lds = -1;        // dynamic LDS bytes per element workgroup; < 0: the form's default
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};

  size_t tsize() const { return precision == ORAD_MIXED ? sizeof(float) : sizeof(double); }
};

static int fail(orad_ctx *c, int code, const std::string &m) {
  if (c) c->err = m;
  return code;
}

#define HIPCHK(c, call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail((c), ORAD_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename P>
static hipError_t dalloc(P **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  return hipMalloc((void **)p, bytes);
}

static void dfree(void *p) {
  if (p) (void)hipFree(p);
}

static void drop_graphs(orad_ctx *c) {
  for (auto &g : c->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

static constexpr int NSTATE = 12;  // sig(6) eint rho pla qvis offg epsd
enum { C_EINT = 6, C_RHO = 7, C_PLA = 8, C_QVIS = 9, C_OFFG = 10, C_EPSD = 11 };

template <typename T>
static BrickArgs<T> brick_args(orad_ctx *c) {
  BrickArgs<T> a{};
  const size_t ne = (size_t)c->numels;
  T *base = (T *)c->state;
  a.numels = c->numels;
  a.conn = c->conn;
  a.iads = c->iads;
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
  return a;
}

static hipError_t launch_elements(orad_ctx *c, int e0, int e1, const double *dt1p,
                                  unsigned long long *dt2t, double *ehou_part) {
  if (c->precision == ORAD_MIXED)
    return launch_brick_mixed(brick_args<float>(c), e0, e1, dt1p, dt2t, ehou_part, c->xcd_remap,
                              c->elem_lds, c->stream);
  return launch_brick_fp64(brick_args<double>(c), e0, e1, dt1p, dt2t, ehou_part, c->xcd_remap,
                           c->elem_lds, c->stream);
}

// One cycle's launches at parity p (no host synchronisation, graph-capturable).
static hipError_t enqueue_cycle(orad_ctx *c, int p) {
  hipError_t e = launch_elements(c, 0, c->numels, &c->ctl->s[p].dt2, &c->ctl->dt2t[p],
                                 c->ehou_part);
  if (e != hipSuccess) return e;
  return launch_node(node_args(c), c->precision == ORAD_MIXED, p, 1, 0, c->xcd_remap, c->stream);
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
  (void)hipSetDevice(c->device);
  drop_graphs(c);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void *p : {(void *)c->X4, (void *)c->V4, (void *)c->D, (void *)c->A, (void *)c->STIFN,
                  (void *)c->conn, (void *)c->iads, (void *)c->adsky, c->state, (void *)c->vol0,
                  c->fsky, c->sti, (void *)c->slot_elem,
                  (void *)c->ehou_part, (void *)c->ctl, (void *)c->scratch_dt1,
                  (void *)c->scratch_dt2t})
    dfree(p);
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
  if (numnod <= 0 || !x || !v || !ms) return fail(c, ORAD_EINVAL, "set_nodes: bad arguments");
  if (c->have_nodes && numnod != c->numnod)
    return fail(c, ORAD_EINVAL, "set_nodes: node count differs from the mesh");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  const size_t n = (size_t)numnod;
  if (!c->have_nodes) {
    HIPCHK(c, dalloc(&c->X4, n * sizeof(double4)));
    HIPCHK(c, dalloc(&c->V4, n * sizeof(double4)));
    HIPCHK(c, dalloc(&c->D, 3 * n * sizeof(double)));
    HIPCHK(c, dalloc(&c->A, 3 * n * sizeof(double)));
    HIPCHK(c, dalloc(&c->STIFN, n * sizeof(double)));
  }
  std::vector<double4> hx(n), hv(n);
  for (size_t i = 0; i < n; ++i) {
    hx[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], ms[i]);
    const int code = icodt ? icodt[i] : 0;
    if (code < 0 || code > 7) return fail(c, ORAD_EINVAL, "set_nodes: ICODT outside 0..7");
    hv[i] = make_double4(v[3 * i], v[3 * i + 1], v[3 * i + 2], (double)code);
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
  if (!c->have_nodes) return fail(c, ORAD_ESTATE, "set_bricks: call orad_set_nodes first");
  if (c->have_bricks) return fail(c, ORAD_ESTATE, "set_bricks: mesh already set");
  if (numels <= 0 || !ixs || !vol0) return fail(c, ORAD_EINVAL, "set_bricks: bad arguments");
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
  // Device copies, corner-major for coalesced loads.
  std::vector<int32_t> conn_cm(8 * ne), iads_cm(8 * ne);
  for (size_t e = 0; e < ne; ++e)
    for (int k = 0; k < 8; ++k) {
      conn_cm[k * ne + e] = conn_em[8 * e + k];
      iads_cm[k * ne + e] = c->h_iads[8 * e + k];
    }
  HIPCHK(c, dalloc(&c->conn, 8 * ne * sizeof(int32_t)));
  HIPCHK(c, dalloc(&c->iads, 8 * ne * sizeof(int32_t)));
  HIPCHK(c, dalloc(&c->adsky, (nn + 1) * sizeof(int32_t)));
  HIPCHK(c, hipMemcpy(c->conn, conn_cm.data(), 8 * ne * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->iads, iads_cm.data(), 8 * ne * 4, hipMemcpyHostToDevice));
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
  const int bb = c->precision == ORAD_MIXED ? brick_block_mixed() : brick_block_fp64();
  c->n_ehou_part = (int)((ne + bb - 1) / bb);
  HIPCHK(c, dalloc(&c->ehou_part, (size_t)c->n_ehou_part * sizeof(double)));
  HIPCHK(c, hipMemset(c->ehou_part, 0, (size_t)c->n_ehou_part * sizeof(double)));
  HIPCHK(c, dalloc(&c->vol0, ne * sizeof(double)));
  HIPCHK(c, hipMemcpy(c->vol0, vol0, ne * sizeof(double), hipMemcpyHostToDevice));
  c->numels = numels;
  c->have_bricks = true;
  drop_graphs(c);
  return init_state(c);
}

int orad_set_law2(orad_ctx *c, const double *p, const int *ip) {
  if (!c) return ORAD_EINVAL;
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
  if (qa < 0.0 || qb < 0.0 || hvis < 0.0 || !(dtfac > 0.0))
    return fail(c, ORAD_EINVAL, "set_solid_prop: negative coefficient or DTFAC <= 0");
  c->prop = SProp{qa, qb, hvis, dtfac};
  c->have_prop = true;
  drop_graphs(c);
  return ORAD_OK;
}

int orad_get_pon(orad_ctx *c, int *adsky, int *iads) {
  if (!c) return ORAD_EINVAL;
  if (!c->have_bricks) return fail(c, ORAD_ESTATE, "get_pon: no mesh");
  if (adsky)
    for (int n = 0; n <= c->numnod; ++n) adsky[n] = c->h_adsky[n] + 1;
  if (iads)
    for (size_t i = 0; i < c->h_iads.size(); ++i) iads[i] = c->h_iads[i] + 1;
  return ORAD_OK;
}

int orad_set_time(orad_ctx *c, const orad_time_t *t) {
  if (!c || !t) return ORAD_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->host_ncycle = t->ncycle;
  const int p = c->host_ncycle & 1;
  Scal s{t->tt, t->dt1, t->dt2, t->dt12, t->dt2old, t->dtmx, t->ncycle, t->mstop};
  unsigned long long inf = 0x7FF0000000000000ull;
  HIPCHK(c, hipMemcpy(&c->ctl->s[p], &s, sizeof(s), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(&c->ctl->dt2t[p], &inf, sizeof(inf), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(&c->ctl->mstop, &t->mstop, sizeof(int), hipMemcpyHostToDevice));
  return ORAD_OK;
}

int orad_get_time(orad_ctx *c, orad_time_t *t) {
  if (!c || !t) return ORAD_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  Scal s;
  int ms = 0;
  HIPCHK(c, hipMemcpy(&s, &c->ctl->s[c->host_ncycle & 1], sizeof(s), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(&ms, &c->ctl->mstop, sizeof(int), hipMemcpyDeviceToHost));
  t->tt = s.tt; t->dt1 = s.dt1; t->dt2 = s.dt2; t->dt12 = s.dt12; t->dt2old = s.dt2old;
  t->dtmx = s.dtmx; t->ncycle = s.ncycle; t->mstop = ms;
  return ORAD_OK;
}

static int ready(orad_ctx *c) {
  if (!(c->have_nodes && c->have_bricks && c->have_mat && c->have_prop))
    return fail(c, ORAD_ESTATE, "mesh, nodes, material and property must be set first");
  return ORAD_OK;
}

int orad_sforc3(orad_ctx *c, int nft, int nel, double dt1, double *dt2t) {
  if (!c || !dt2t) return ORAD_EINVAL;
  int r = ready(c);
  if (r) return r;
  if (nft < 0 || nel < 0 || nft + nel > c->numels)
    return fail(c, ORAD_EINVAL, "sforc3: group outside the element range");
  HIPCHK(c, hipSetDevice(c->device));
  unsigned long long bits = 0x7FF0000000000000ull;
  HIPCHK(c, hipMemcpyAsync(c->scratch_dt1, &dt1, sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->scratch_dt2t, &bits, sizeof(bits), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, launch_elements(c, nft, nft + nel, c->scratch_dt1, c->scratch_dt2t, nullptr));
  HIPCHK(c, hipMemcpyAsync(&bits, c->scratch_dt2t, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double g;
  std::memcpy(&g, &bits, sizeof(g));
  if (g < *dt2t) *dt2t = g;
  int ms = 0;
  HIPCHK(c, hipMemcpy(&ms, &c->ctl->mstop, sizeof(int), hipMemcpyDeviceToHost));
  return ms ? fail(c, ORAD_ENEGVOL, "non-positive element volume (MSTOP)") : ORAD_OK;
}

int orad_asspar4(orad_ctx *c, double *a, double *stifn) {
  if (!c) return ORAD_EINVAL;
  int r = ready(c);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, launch_node(node_args(c), c->precision == ORAD_MIXED, c->host_ncycle & 1, 0, 1,
                        c->xcd_remap, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (a) HIPCHK(c, hipMemcpy(a, c->A, 3 * (size_t)c->numnod * 8, hipMemcpyDeviceToHost));
  if (stifn) HIPCHK(c, hipMemcpy(stifn, c->STIFN, (size_t)c->numnod * 8, hipMemcpyDeviceToHost));
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

int orad_run_cycles(orad_ctx *c, int ncycle, orad_time_t *t) {
  if (!c) return ORAD_EINVAL;
  int r = ready(c);
  if (r) return r;
  if (ncycle < 0) return fail(c, ORAD_EINVAL, "run_cycles: negative cycle count");
  HIPCHK(c, hipSetDevice(c->device));
  int left = ncycle;
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
  orad_time_t tt{};
  r = orad_get_time(c, &tt);
  if (r) return r;
  if (t) *t = tt;
  return tt.mstop ? fail(c, ORAD_ENEGVOL, "non-positive element volume (MSTOP)") : ORAD_OK;
}

int orad_profile_cycles(orad_ctx *c, int ncycle, double *elem_ms, double *node_ms) {
  if (!c) return ORAD_EINVAL;
  int r = ready(c);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  double se = 0.0, sn = 0.0;
  for (int i = 0; i < ncycle; ++i) {
    const int p = c->host_ncycle & 1;
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    HIPCHK(c, launch_elements(c, 0, c->numels, &c->ctl->s[p].dt2, &c->ctl->dt2t[p], c->ehou_part));
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, launch_node(node_args(c), c->precision == ORAD_MIXED, p, 1, 0, c->xcd_remap, c->stream));
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
  if (!c->have_bricks) return fail(c, ORAD_ESTATE, "get_brick_state: no mesh");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t ne = (size_t)c->numels, ts = c->tsize();
  std::vector<char> h(NSTATE * ne * ts);
  HIPCHK(c, hipMemcpy(h.data(), c->state, h.size(), hipMemcpyDeviceToHost));
  auto col = [&](int k, size_t e) -> double {
    return c->precision == ORAD_MIXED ? (double)((float *)h.data())[k * ne + e]
                                      : ((double *)h.data())[k * ne + e];
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
  if (!c->have_bricks) return fail(c, ORAD_ESTATE, "set_brick_state: no mesh");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t ne = (size_t)c->numels, ts = c->tsize();
  std::vector<char> h(NSTATE * ne * ts);
  HIPCHK(c, hipMemcpy(h.data(), c->state, h.size(), hipMemcpyDeviceToHost));
  auto put = [&](int k, size_t e, double v) {
    if (c->precision == ORAD_MIXED) ((float *)h.data())[k * ne + e] = (float)v;
    else ((double *)h.data())[k * ne + e] = v;
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
  for (size_t e = 0; e < ne; ++e)
    for (int k = 0; k < 8; ++k) {
      const size_t sl = (size_t)c->h_iads[8 * e + k];
      fsky[4 * sl] = f[3 * sl];
      fsky[4 * sl + 1] = f[3 * sl + 1];
      fsky[4 * sl + 2] = f[3 * sl + 2];
      fsky[4 * sl + 3] = st[e];
    }
  return ORAD_OK;
}

int orad_energies(orad_ctx *c, double *eint, double *ekin, double *ehour) {
  if (!c) return ORAD_EINVAL;
  int r = ready(c);
  if (r) return r;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t ne = (size_t)c->numels, ts = c->tsize();
  const size_t nbe = (ne + 255) / 256, nbk = ((size_t)c->numnod + 255) / 256;
  double *part = nullptr;
  HIPCHK(c, dalloc(&part, (nbe + nbk + 2) * sizeof(double)));
  hipError_t e = launch_energy(c->precision == ORAD_MIXED, c->numels,
                               (char *)c->state + C_EINT * ne * ts, c->vol0, c->numnod, c->X4,
                               c->V4, part, part + nbe + nbk, c->stream);
  double res[2] = {0, 0}, eh = 0;
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(res, part + nbe + nbk, 2 * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&eh, &c->ctl->ehour, sizeof(double), hipMemcpyDeviceToHost);
  dfree(part);
  if (e != hipSuccess) return fail(c, ORAD_EHIP, std::string("energies: ") + hipGetErrorString(e));
  if (eint) *eint = res[0];
  if (ekin) *ekin = res[1];
  if (ehour) *ehour = eh;
  return ORAD_OK;
}

}  // extern "C"

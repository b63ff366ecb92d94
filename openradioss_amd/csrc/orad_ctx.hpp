// orad_ctx.hpp -- the context behind the C ABI (internal to liborad).
//
// One context models one mesh of one element family (8-node bricks or 4-node
// shells) on one GPU: device arrays, the PARITH/ON CSR built on the host with
// the starter's ordering, and the captured cycle graphs.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/orad.h"
#include "orad_common.hpp"
#include "orad_contact7.hpp"

namespace orad {
hipError_t launch_brick_fp64(const BrickArgs<double> &, int, int, const double *,
                             unsigned long long *, double *, int, int, int, int, hipStream_t);
hipError_t launch_brick_mixed(const BrickArgs<float> &, int, int, const double *,
                              unsigned long long *, double *, int, int, int, int, hipStream_t);
int brick_block_fp64();
int brick_block_mixed();
int brick_waves_fp64();
int brick_waves_mixed();
bool brick_shape_ok_fp64(int block, int waves);
bool brick_shape_ok_mixed(int block, int waves);
hipError_t launch_node(const NodeArgs &, bool, int, int, int, int, hipStream_t, int tile);
int node_tile_default(bool slot_fp32);
bool node_tile_ok(bool slot_fp32, bool rot, int tile);
hipError_t launch_energy(bool, int, const void *, const void *, const double *,
                         const int32_t *, int, const double4 *, const double4 *, const double4 *,
                         double *, double *, hipStream_t);
hipError_t launch_sum(const double *x, int n, double *part, double *res, hipStream_t s);
hipError_t launch_shell_fp64(const ShellArgs<double> &, int, int, const double *,
                             unsigned long long *, double *, int, hipStream_t);
hipError_t launch_shell_mixed(const ShellArgs<float> &, int, int, const double *,
                              unsigned long long *, double *, int, hipStream_t);
int shell_block_fp64();
int shell_block_mixed();
hipError_t launch_shell_batch_fp64(const ShellArgs<double> &, const BatchArgs &, int, int, int,
                                   int, int, hipStream_t);
hipError_t launch_shell_batch_mixed(const ShellArgs<float> &, const BatchArgs &, int, int, int,
                                    int, int, hipStream_t);
hipError_t launch_qeph_fp64(const ShellArgs<double> &, int, int, const double *,
                            unsigned long long *, double *, int, hipStream_t);
hipError_t launch_qeph_mixed(const ShellArgs<float> &, int, int, const double *,
                             unsigned long long *, double *, int, hipStream_t);
int qeph_block_fp64();
int qeph_block_mixed();
hipError_t launch_qeph_batch_fp64(const ShellArgs<double> &, const BatchArgs &, int, int, int,
                                  int, int, hipStream_t);
hipError_t launch_qeph_batch_mixed(const ShellArgs<float> &, const BatchArgs &, int, int, int,
                                   int, int, hipStream_t);
int qeph_batch_occupancy_fp64(int);
hipError_t launch_qbat_fp64(const ShellArgs<double> &, int, int, const double *,
                            unsigned long long *, double *, int, hipStream_t);
hipError_t launch_qbat_mixed(const ShellArgs<float> &, int, int, const double *,
                             unsigned long long *, double *, int, hipStream_t);
int qbat_block_fp64();
int qbat_block_mixed();
int qeph_batch_occupancy_mixed(int);
int shell_batch_occupancy_fp64(int);
int shell_batch_occupancy_mixed(int);
template <typename S> struct AdjArgs;
template <typename S> struct AdjBrickArgs;
// fp32 state, QEPH, hand-written reverse (Belytschko-Tsay, IPLA = 0) instead of duals
hipError_t launch_adj_elem(bool, bool, bool, const AdjArgs<double> &, hipStream_t);
// bricks: fp32 state, hand-written reverse (else dual numbers)
hipError_t launch_adj_brick_elem(bool, bool, const AdjBrickArgs<double> &, hipStream_t);
hipError_t launch_adj_node(int, const double4 *, const double4 *, const double4 *,
                           const double *, const double *, double, double, const double *,
                           double *, const double *, double *, double *, double *, double *,
                           double *, hipStream_t);
hipError_t launch_adj_gather(int, int, int, const int32_t *, const int32_t *, const int32_t *,
                             const double *, double *, double *, double *, hipStream_t);
hipError_t launch_adj_gather_node(int numnod, int numelc, int per, const int32_t *adsky,
                                  const int32_t *slot_elem, const int32_t *slot_corner,
                                  const double *lcorner, const double4 *X4, const double4 *V4,
                                  const double4 *VR4, const double *A, const double *AR,
                                  double dt2, double dt12, double *lx, double *lv,
                                  const double *ld, double *lvr, double *lamS, double *lamR,
                                  double *g_ms, double *g_in, hipStream_t s);
// up to 7 device-to-device copies (sizes multiples of 4 bytes) in one launch
hipError_t launch_pack_xv(int numnod, const double *xv, double4 *X4, double4 *V4, double4 *VR4,
                          hipStream_t s);
hipError_t launch_adj_stash(int n, const void *const *src, void *const *dst,
                            const size_t *bytes, hipStream_t s);
}  // namespace orad

struct AdjStore;   // checkpoints of an adjoint forward run (orad_api.hip)
struct orad_ctx;
// orad_brick_order.hip
int brick_bins(orad_ctx *c, size_t ne, const std::vector<int32_t> &conn_em, const double *vol0);
int brick_reorder(orad_ctx *c, int rows, const std::vector<std::pair<void *, void *>> &extra);

using namespace orad;

enum { FAM_NONE = 0, FAM_BRICK = 1, FAM_SHELL = 2 };

struct orad_ctx {
  int device = 0;
  int family = FAM_NONE;
  int precision = ORAD_FP64;
  hipStream_t stream = nullptr;
  std::string err;

  int numnod = 0, numels = 0;
  bool have_nodes = false, have_bricks = false, have_mat = false, have_prop = false;
  bool state_init = false;
  Law2 mat{};
  SProp prop{};

  std::vector<int32_t> h_adsky, h_iads;  // 0-based, element-major iads [e*nc+k]
  int nc = 8;                             // corners per element of the family

  // shells
  int numelc = 0;
  bool have_shells = false, have_shmat = false, have_shprop = false, have_rot = false;
  ShMat shmat{};
  ShProp shprop{};
  ShConst shk{};
  int ishell = 1;           // shell formulation (orad_set_shell_ishell): 1 BT, 24 QEPH, 12 QBAT
  double qeph_dn = 0.015;   // GEO(13): QEPH numerical damping Dn, QBAT membrane damping Hm
  double4 *VR4 = nullptr;
  double *AR = nullptr, *STIFR = nullptr, *thke = nullptr;
  // imposed velocities (FIXVEL) and rotational codes, folded into V4.w on upload
  std::vector<int> h_code_hi;   // per node: ICODR << 3 | FIXVEL direction << 6
  std::vector<double> h_fv_val;
  double *fv_val = nullptr;
  // energy balance (ECRIT): per-workgroup kinetic energies of the last two cycles, FIXVEL
  // energy memory and running external work per node, ENTOT0 of the first cycle
  double *bal_part = nullptr, *fv_mem = nullptr, *fv_work = nullptr;
  double entot0 = 0.0;
  bool have_entot0 = false;

  // bricks: the library's element order (orad_brick_order.hip)
  int order_rows = 10;                // column height in element rows; 0 = engine order
  std::vector<int32_t> h_perm;        // library element -> engine element
  std::vector<int32_t> h_iperm;       // engine element -> library element
  std::vector<unsigned long long> h_bin;  // centroid bins, engine order
  uint64_t *rank = nullptr;           // eight 8-bit corner ranks per brick, library order
  int32_t *emap = nullptr;            // engine element -> library element (device)

  // device buffers
  double4 *X4 = nullptr, *V4 = nullptr;
  double *D = nullptr, *A = nullptr, *STIFN = nullptr;
  int32_t *conn = nullptr, *iads = nullptr, *adsky = nullptr;   // iads: shells' IADC
  void *state = nullptr;  // SoA element state of the storage type (NSTATE columns)
  double *vol0 = nullptr; // initial volumes, FP64
  void *fsky = nullptr;
  void *sti = nullptr;              // per-element STI (QEPH: per diagonal, 2*numelc)
  int32_t *slot_elem = nullptr;     // STI entry of each slot (its element; QEPH: + diagonal)
  double *ehou_part = nullptr;
  int n_ehou_part = 0;
  CycleCtl *ctl = nullptr;
  double *scratch_dt1 = nullptr;
  unsigned long long *scratch_dt2t = nullptr;
  // asynchronous group calls (orad_sforc3_async / orad_cforc3_async, orad_group_wait)
  bool group_open = false;
  unsigned long long group_bits = 0;
  double group_dt1[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int group_slot = 0;
  bool pend = false;                  // groups recorded, not yet launched
  int pend_fam = 0, pend_b = 0, pend_e = 0;
  double pend_dt1 = 0.0;
  double *xv_stage = nullptr;         // orad_update_xv staging: X, V, VR (3 * numnod each)
  double *noalloc_scratch = nullptr;   // this device's file-scope scratch array (not owned)

  // batched independent shell models (orad_set_shell_models): per-model material,
  // time scalars and time step; element tiles of 256 elements within one model
  int nmodel = 1;
  bool batched = false;
  std::vector<int32_t> h_node_off, h_elem_off, h_model_tiles;
  std::vector<ShMat> h_mats;
  int32_t *node_model = nullptr, *model_node0 = nullptr, *model_tiles = nullptr;
  int4 *tiles = nullptr;
  int ntiles = 0;
  ShMat *mats = nullptr;
  unsigned *queue = nullptr;
  double *chunk_ehou = nullptr;
  // launch configuration of the element kernels (orad_set_launch / orad_autotune)
  int persist = 0;          // batched shells: persistent grid taking tiles from a counter
  int batch_lds = 0;        // batched shells: dynamic LDS bytes per workgroup (residency cap)
  int batch_grid = 0;       // persistent grid size (resident capacity at batch_lds)
  int num_cu = 0;

  // /INTER/TYPE7 contact (orad_set_contact7), or null
  Contact7 *ci = nullptr;
  bool ci_fresh = false;    // orad_contact7 ran since the last orad_asspar4

  AdjStore *adj = nullptr;
  int cycle_write_a = 0;    // cycles also store accelerations A / AR (adjoint recomputation)

  int host_ncycle = 0;  // cycle counter mirrored on the host (its parity selects buffers)
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // two cycles starting at parity 0 / 1
  int xcd_remap = 1;
  int elem_lds = -1;        // dynamic LDS bytes per element workgroup; < 0: the form's default
  int elem_block = 0;       // bricks: element workgroup size (0: the form's default)
  int elem_waves = 0;       // bricks: waves per SIMD of the element-kernel instance (0: default)
  int node_tile = 0;        // nodal kernel: slots per LDS tile (0: default)
  int adj_vjp = 0;          // shell adjoint: 0 reverse pass where available, 1 duals
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};

  size_t ctl_bytes() const { return (size_t)nmodel * sizeof(CycleCtl); }
  bool qeph() const { return family == FAM_SHELL && ishell == 24; }
  bool qbat() const { return family == FAM_SHELL && ishell == 12; }
  int shncol() const {                                                  // shell state columns
    return ishell == 24 ? SC_NCOL_QEPH : ishell == 12 ? SC_NCOL_QBAT : SC_NCOL;
  }
  int nsti() const { return (ishell == 24 || ishell == 12) ? 2 : 1; }    // STI entries per shell
  size_t tsize() const { return precision == ORAD_MIXED ? sizeof(float) : sizeof(double); }
};

static inline int fail(orad_ctx *c, int code, const std::string &m) {
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
static inline hipError_t dalloc(P **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  return hipMalloc((void **)p, bytes);
}

static inline void dfree(void *p) {
  if (p) (void)hipFree(p);
}

static inline void drop_graphs(orad_ctx *c) {
  for (auto &g : c->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

static constexpr int NSTATE = 12;  // sig(6) eint rho pla qvis offg epsd
enum { C_EINT = 6, C_RHO = 7, C_PLA = 8, C_QVIS = 9, C_OFFG = 10, C_EPSD = 11 };


// orad_contact7.hpp -- /INTER/TYPE7 node-to-segment penalty contact on the device
// (internal to liborad).
//
// Path: constant gap, IGSTI<=1, INACTI=0, no curvature, no heat exchange, no tied
// contact, IVIS2=0 damping, Coulomb friction in the total viscous formulation without
// filtering (IFQ=0, MFROT=0, isotropic).  Every cycle, on the context's stream and in
// its captured graph:
//   check       the largest coordinate change of any node since the last search; a
//               new search when twice it reaches the margin (so every pair within TZ
//               now was within TZ + margin then), otherwise the last candidates stand;
//   search      segments binned on a grid by their box centre (per-model cell size,
//               cells hashed with the model); each secondary node tests the segments of
//               the 2 x 2 x 2 cells around it with I7TRI's box test widened by TZS;
//               a second kernel filters each node's candidates with I7COR3 + I7DST3
//               and keeps the penetrated ones in increasing segment order, so the
//               penetrating pairs come out in (secondary node, segment) order without
//               a global sort;
//   forces      one 128-lane workgroup per batch of NVSIZ = 128 penetrating pairs:
//               I7FOR3 and I7ASS2's skyline layout (main node 1 of every pair, then
//               nodes 2, 3, 4 where H /= 0, then the secondary nodes) in the batch's
//               640-entry region; the batch's contact time step lowers DT2T;
//   assembly    each entry is pushed on its node's list; the nodal kernel adds a node's
//               entries after its element slots in increasing entry order, i.e.
//               ASSPAR4's skyline order.
// Every order is fixed (atomics only count, place and link; lists are read sorted), so
// a run is bitwise reproducible and equal to the FP64 oracle's.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace orad {

constexpr int I7_NVSIZ = 128;             // pairs per batch (the reference's NVSIZ)
constexpr int I7_SLOTS = 5 * I7_NVSIZ;    // skyline entries one batch can write
constexpr int I7_NODE_CAP = 64;           // candidate segments of one secondary node
constexpr int I7_KEEP_CAP = 16;           // penetrated segments of one secondary node

// counters[] of a pass
enum {
  I7C_NCAND = 0, I7C_NPEN = 1, I7C_OVERFLOW = 2, I7C_NSKY = 3, I7C_DONE = 4, I7C_NBATCH = 5,
  I7C_SEARCH = 6,   // 1: this pass searches
  I7C_NCELL = 7,    // cells to scan this pass (cell_cap or 0)
  I7C_N = 8
};
// overflow bits (sticky until the interface is set again)
enum { I7_OVF_NODE = 1, I7_OVF_PAIRS = 2, I7_OVF_KEEP = 4 };

struct Contact7 {
  int nsn = 0, nrtm = 0, numnod = 0;
  int32_t *nsv = nullptr;       // [nsn] secondary nodes, 0-based
  int4 *irect = nullptr;        // [nrtm] segment nodes, 0-based (3rd = 4th: triangle)
  double *stfm = nullptr;       // [nrtm] segment stiffness STF
  double *stfns = nullptr;      // [nsn] secondary factor STFN, negated on deactivation
  double gap = 0.0, fric = 0.0, visc = 0.0, viscf = 0.0, tz = 0.0;
  double em10 = 0.0, em20 = 0.0, em30 = 0.0, zep4 = 0.0;   // constant_mod.F values
  // search margin: the candidates are searched with the reach TZS = TZ + margin and kept
  // until a node has moved by margin / 2 (largest coordinate change) since that search
  double margin = 0.0, tzs = 0.0;
  double4 *x_search = nullptr;           // [numnod] positions at the last search (NaN: none)
  unsigned long long *dpart = nullptr;   // [ndisp_blocks] the workgroups' largest move
  int ndisp_blocks = 0;
  // search grid over the segments' box centres, cells hashed with the model into
  // cell_cap buckets (a power of two)
  int cell_cap = 0;
  unsigned long long *bounds = nullptr;  // [7] ordered bits: min xyz, max xyz, max extent
  unsigned long long *bpart = nullptr;   // [7 * nbound_blocks] the workgroups' bounds
  int nbound_blocks = 0;
  double *grid = nullptr;                // [4] origin xyz, cell size
  float4 *seg_lo = nullptr, *seg_hi = nullptr;   // [nrtm] segment boxes of the pass,
                                                  // rounded outward to FP32
  int32_t *seg_cell = nullptr;           // [nrtm]
  int32_t *cell_count = nullptr, *cell_fill = nullptr;   // [cell_cap]
  int32_t *cell_start = nullptr;         // [cell_cap + 1]
  // the cells' segment boxes in cell order (a cell's in any order); cell_lo[j].w holds
  // the segment's index as int bits.  Segments with STF = 0 are left out.
  float4 *cell_lo = nullptr, *cell_hi = nullptr;   // [nrtm]
  // penetrating pairs: per secondary node its penetrated segments in increasing order;
  // pair r of the pass is node i's (r - node_kstart[i])-th, node_kstart[i] <= r
  int pair_cap = 0, nbatch_cap = 0;
  // models (batched independent models; one otherwise): a pair's two ends belong to one
  // model, the secondary nodes of model m are nsv[nsv_off[m] .. nsv_off[m+1]), and a
  // batch never mixes models, so each model's pairs batch as they would alone
  int nmodel = 1;
  const int32_t *node_model = nullptr;   // the context's node -> model (null: one model)
  int32_t *seg_model = nullptr;          // [nrtm] (null: one model)
  int32_t *nsv_off = nullptr;            // [nmodel + 1]
  int32_t *model_bstart = nullptr;       // [nmodel + 1] first batch of each model
  unsigned long long *model_ext = nullptr;   // [nmodel] ordered bits: largest segment edge
  int32_t *node_cand = nullptr;          // [nsn * I7_NODE_CAP] candidates, found order
  int32_t *node_ncand = nullptr;         // [nsn]
  int32_t *node_keep = nullptr;          // [nsn * I7_KEEP_CAP]
  int32_t *node_nkeep = nullptr;         // [nsn]
  int32_t *node_kstart = nullptr;        // [nsn + 1]
  int *counters = nullptr;               // [I7C_N]
  // skyline: batch b owns entries [b * I7_SLOTS, b * I7_SLOTS + batch_cnt[b])
  double4 *sky = nullptr;                // (fx, fy, fz, stif)
  int32_t *sky_node = nullptr;
  int32_t *sky_next = nullptr;           // per entry: the next entry of its node's list
  int32_t *batch_cnt = nullptr;          // [nbatch_cap]
  double *batch_e = nullptr;             // [2 * nbatch_cap] elastic, friction energy
  double *energy = nullptr;              // [3] cumulative ECONT, ECONTV, ECONTD
  int32_t *node_head = nullptr;          // [numnod] first entry of each node's list, -1
  int32_t *scan_tmp = nullptr;           // block sums of the scans
};

// One contact pass: search, filter, forces, lists.  DT1 is read on the device; the
// batches' contact time steps lower DT2T (bits of a positive double, atomicMin).
// Model m's DT1 is at (char *)dt1 + m * stride and its DT2T accumulator at
// (char *)dt2t + m * stride (the models' time records; stride 0 with one model).
hipError_t contact7_run(const Contact7 &ci, const double4 *X4, const double4 *V4,
                        const double *dt1, unsigned long long *dt2t, size_t stride,
                        hipStream_t s);
// Scan workspace (ints) for n elements, and the longest scan one pass can do.
size_t contact7_scan_tmp_ints(int n);
int contact7_scan_max();

}  // namespace orad

// orad_contact7.hip -- /INTER/TYPE7 node-to-segment penalty contact (see
// orad_contact7.hpp for the pass).  FP64 throughout, built with -ffp-contract=off: the
// per-pair arithmetic keeps the reference routines' operation order (I7COR3
// i7cor3.F:99-261, I7DST3 i7dst3.F:124-470, I7FOR3 i7for3.F:230-1121, 1460-1466,
// 1866-1875, 2444-2480, 4108-4121, I7ASS2 i7ass3.F:1188-1285) so the skyline entries
// equal the FP64 oracle's bit for bit.
#include <hip/hip_runtime.h>

#include "orad_contact7.hpp"

namespace orad {

// ------------------------------------------------------------------ scans --
// Exclusive prefix sums of int32 with the length read on the device (or given), so
// the launches are fixed for graph capture: out[i] = sum in[0..i), out[n] = total.
constexpr int SCAN_T = 1024, SCAN_E = 4, SCAN_B = SCAN_T * SCAN_E;
constexpr int SCAN_TOP_E = 16;                 // block sums per thread of the top scan
constexpr int SCAN_MAX = SCAN_B * SCAN_T * SCAN_TOP_E;

size_t contact7_scan_tmp_ints(int n) { return (size_t)(n + SCAN_B - 1) / SCAN_B + 1; }
int contact7_scan_max() { return SCAN_MAX; }

// Exclusive scan of one value per thread over a 1024-thread workgroup.
__device__ __forceinline__ int block_excl(int v, int *s_w, int *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int i = 0; i < SCAN_T / 64; ++i) {
      const int t = s_w[i];
      s_w[i] = acc;
      acc += t;
    }
    s_w[SCAN_T / 64] = acc;
  }
  __syncthreads();
  *total = s_w[SCAN_T / 64];
  return x - v + s_w[w];
}

__global__ __launch_bounds__(SCAN_T) void scan_blocks(const int32_t *in, int32_t *out,
                                                      const int *n_dev, int n_static,
                                                      int32_t *bsum) {
  __shared__ int s_w[SCAN_T / 64 + 1];
  const int n = n_dev ? *n_dev : n_static;
  const int base = (int)blockIdx.x * SCAN_B;
  if (base >= n) return;
  const int i0 = base + (int)threadIdx.x * SCAN_E;
  int v[SCAN_E], sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_E; ++k) {
    v[k] = (i0 + k < n) ? in[i0 + k] : 0;
    sum += v[k];
  }
  int total;
  int pre = block_excl(sum, s_w, &total);
#pragma unroll
  for (int k = 0; k < SCAN_E; ++k) {
    if (i0 + k < n) out[i0 + k] = pre;
    pre += v[k];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_T) void scan_top(int32_t *bsum, const int *n_dev,
                                                   int n_static, int32_t *out) {
  __shared__ int s_w[SCAN_T / 64 + 1];
  const int n = n_dev ? *n_dev : n_static;
  const int nb = (n + SCAN_B - 1) / SCAN_B;
  const int i0 = (int)threadIdx.x * SCAN_TOP_E;
  int v[SCAN_TOP_E], sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_TOP_E; ++k) {
    v[k] = (i0 + k < nb) ? bsum[i0 + k] : 0;
    sum += v[k];
  }
  int total;
  int pre = block_excl(sum, s_w, &total);
#pragma unroll
  for (int k = 0; k < SCAN_TOP_E; ++k) {
    if (i0 + k < nb) bsum[i0 + k] = pre;
    pre += v[k];
  }
  if (threadIdx.x == 0) out[n] = total;
}

__global__ __launch_bounds__(256) void scan_add(int32_t *out, const int *n_dev, int n_static,
                                                const int32_t *bsum) {
  const int n = n_dev ? *n_dev : n_static;
  const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (i >= n) return;
  out[i] += bsum[i / SCAN_B];
}

static hipError_t scan_excl(const int32_t *in, int32_t *out, const int *n_dev, int n_static,
                            int max_n, int32_t *tmp, hipStream_t s) {
  if (max_n <= 0) max_n = 1;
  const int nb = (max_n + SCAN_B - 1) / SCAN_B;
  hipLaunchKernelGGL(scan_blocks, dim3(nb), dim3(SCAN_T), 0, s, in, out, n_dev, n_static, tmp);
  hipLaunchKernelGGL(scan_top, dim3(1), dim3(SCAN_T), 0, s, tmp, n_dev, n_static, out);
  hipLaunchKernelGGL(scan_add, dim3((max_n + 255) / 256), dim3(256), 0, s, out, n_dev,
                     n_static, tmp);
  return hipGetLastError();
}

// ---------------------------------------------------------------- helpers --
// Order-preserving 64-bit keys of doubles (for atomicMin/atomicMax of a bound).
__device__ __forceinline__ unsigned long long okey(double d) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(d);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double odec(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  return __longlong_as_double((long long)b);
}
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

__device__ __forceinline__ void seg_box(const Contact7 &c, const double4 *X4, int l,
                                        double lo[3], double hi[3]) {
  const int4 r = c.irect[l];
  const double4 p0 = X4[r.x], p1 = X4[r.y], p2 = X4[r.z], p3 = X4[r.w];
  lo[0] = dmin(dmin(dmin(p0.x, p1.x), p2.x), p3.x);
  lo[1] = dmin(dmin(dmin(p0.y, p1.y), p2.y), p3.y);
  lo[2] = dmin(dmin(dmin(p0.z, p1.z), p2.z), p3.z);
  hi[0] = dmax(dmax(dmax(p0.x, p1.x), p2.x), p3.x);
  hi[1] = dmax(dmax(dmax(p0.y, p1.y), p2.y), p3.y);
  hi[2] = dmax(dmax(dmax(p0.z, p1.z), p2.z), p3.z);
}


// ------------------------------------------------------------ search grid --
// Every pass: list heads, counters, and each workgroup's largest coordinate change of a
// node since the last search.
__global__ __launch_bounds__(256) void i7_reset(Contact7 c, const double4 *X4) {
  __shared__ unsigned long long s_d[4];
  const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
  unsigned long long d = 0ull;
  if (i < c.numnod) {
    c.node_head[i] = -1;
    const double4 x = X4[i], o = c.x_search[i];
    // NaN (no search yet, or non-finite coordinates) orders above every number
    const double m = fmax(fmax(fabs(x.x - o.x), fabs(x.y - o.y)), fabs(x.z - o.z));
    d = okey(m == m ? m : __longlong_as_double(0x7FF8000000000000ll));
  }
  if (i < I7C_N && i != I7C_OVERFLOW && i != I7C_NCAND && i != I7C_SEARCH && i != I7C_NCELL)
    c.counters[i] = 0;   // the overflow flag is sticky; NCAND is the last search's
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long b = __shfl_xor(d, o, 64);
    d = b > d ? b : d;
  }
  if ((threadIdx.x & 63) == 0) s_d[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0 && (int)blockIdx.x < c.ndisp_blocks) {
    unsigned long long v = s_d[0];
    for (int w = 1; w < 4; ++w) v = s_d[w] > v ? s_d[w] : v;
    c.dpart[blockIdx.x] = v;
  }
}

// One workgroup: does this pass search?
__global__ __launch_bounds__(1024) void i7_decide(Contact7 c) {
  __shared__ unsigned long long s_d[16];
  unsigned long long d = 0ull;
  for (int b = threadIdx.x; b < c.ndisp_blocks; b += 1024) d = c.dpart[b] > d ? c.dpart[b] : d;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long b = __shfl_xor(d, o, 64);
    d = b > d ? b : d;
  }
  if ((threadIdx.x & 63) == 0) s_d[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 16; ++w) d = s_d[w] > d ? s_d[w] : d;
    const double m = odec(d);
    // with every node within margin / 2 of its place at the last search, a node and a
    // segment box have closed in by less than the margin on every axis
    const int search = !(2.0 * m * (1.0 + 1.0e-6) < c.margin);
    c.counters[I7C_SEARCH] = search;
    c.counters[I7C_NCELL] = search ? c.cell_cap : 0;
  }
}

// A searching pass: cell counts, candidate rows, bounds, and the positions it searches.
__global__ __launch_bounds__(256) void i7_search_reset(Contact7 c, const double4 *X4) {
  if (!c.counters[I7C_SEARCH]) return;
  const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (i < c.cell_cap) {
    c.cell_count[i] = 0;
    c.cell_fill[i] = 0;
  }
  if (i < c.numnod) c.x_search[i] = X4[i];
  if (i < c.nsn) c.node_ncand[i] = 0;
  if (i < c.nmodel) c.model_ext[i] = 0ull;
  if (i < 3) {
    c.bounds[i] = ~0ull;
    c.bounds[3 + i] = 0ull;
  }
  if (i == 3) c.bounds[6] = 0ull;
  if (i == 4) c.counters[I7C_NCAND] = 0;
}

// Cell size (per model): the model's largest segment box edge plus 2 TZS, so the centre
// of every segment whose box widened by TZS holds a node lies within half a cell of the
// node on each axis: in the node's cell or the neighbour on the side of the node's half,
// 2 x 2 x 2 cells.  Cells are
// hashed with the model into cell_cap buckets: models that overlap in space (batched
// independent models) share no bucket but by collision, and the grid needs no extent.
struct I7Grid {
  double o[3], s, inv;
};

__device__ __forceinline__ I7Grid i7_grid(const Contact7 &c) {
  I7Grid g;
  for (int d = 0; d < 3; ++d) g.o[d] = c.nrtm > 0 ? odec(c.bounds[d]) : 0.0;
  // padded by 1e-6 so a candidate's centre lies strictly within half a cell of the
  // node on every axis, with room for the rounding of the cell coordinates
  double s = ((c.nrtm > 0 ? odec(c.bounds[6]) : 0.0) + 2.0 * c.tzs) * (1.0 + 1.0e-6);
  if (!(s > 0.0)) s = 1.0;
  if (!(s < 1.0e300) || !(g.o[0] == g.o[0] && g.o[1] == g.o[1] && g.o[2] == g.o[2])) {
    // non-finite coordinates: one cell per model
    s = 1.0e300;
    g.o[0] = g.o[1] = g.o[2] = 0.0;
  }
  g.s = s;
  g.inv = 1.0 / s;
  return g;
}

// Model m's cell size: its largest segment edge plus 2 TZS, padded as above.
__device__ __forceinline__ double model_cell(const Contact7 &c, int m) {
  const unsigned long long e = c.seg_model ? c.model_ext[m] : c.bounds[6];
  double s = (odec(e) + 2.0 * c.tzs) * (1.0 + 1.0e-6);
  if (!(s > 0.0)) s = 1.0;
  if (!(s < 1.0e300) || !(c.grid[3] < 1.0e300)) s = 1.0e300;
  return s;
}

__device__ __forceinline__ long long cell_coord(double x, double o, double inv) {
  double f = floor((x - o) * inv);
  if (!(f > -1.0e15)) f = -1.0e15;      // far below, or not a number
  if (f > 1.0e15) f = 1.0e15;
  return (long long)f;
}

// Bucket of cell (kx, ky, kz) of a model: blocks of 4 x 4 x 4 cells are hashed with the
// model, the cells of a block are consecutive buckets, so neighbouring cells (and the
// segments binned in them) lie close in memory.  cap: a power of two, >= 64.
__device__ __forceinline__ int cell_hash(long long kx, long long ky, long long kz, int model,
                                         int cap) {
  unsigned long long h = (unsigned long long)(kx >> 2) * 0x9E3779B97F4A7C15ull;
  h ^= (unsigned long long)(ky >> 2) * 0xC2B2AE3D27D4EB4Full;
  h ^= (unsigned long long)(kz >> 2) * 0x165667B19E3779F9ull;
  h ^= (unsigned long long)(unsigned)model * 0x27D4EB2F165667C5ull;
  h ^= h >> 29;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 32;
  const int in_block = (int)(((kz & 3) << 4) | ((ky & 3) << 2) | (kx & 3));
  return (int)((h & (unsigned long long)((cap >> 6) - 1)) << 6) | in_block;
}

// Segment boxes of the pass, and the bounds of their centres and their largest edge:
// reduced over each wave by shuffles, then over the workgroup, then one atomic per bound
// and workgroup.
__global__ __launch_bounds__(256) void i7_seg_bounds(Contact7 c, const double4 *X4) {
  if (!c.counters[I7C_SEARCH]) return;
  __shared__ unsigned long long s_b[7][4];
  const int l = (int)blockIdx.x * 256 + (int)threadIdx.x;
  unsigned long long k[7] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull, 0ull};
  const int model = (l < c.nrtm && c.seg_model) ? c.seg_model[l] : 0;
  if (l < c.nrtm) {
    double lo[3], hi[3];
    seg_box(c, X4, l, lo, hi);
    // rounded outward: the FP32 box holds the FP64 one, so its test keeps every pair
    // the exact test keeps; centres and edges below are the FP32 box's
    const float4 fl = make_float4(__double2float_rd(lo[0]), __double2float_rd(lo[1]),
                                  __double2float_rd(lo[2]), 0.0f);
    const float4 fh = make_float4(__double2float_ru(hi[0]), __double2float_ru(hi[1]),
                                  __double2float_ru(hi[2]), 0.0f);
    c.seg_lo[l] = fl;
    c.seg_hi[l] = fh;
    lo[0] = fl.x; lo[1] = fl.y; lo[2] = fl.z;
    hi[0] = fh.x; hi[1] = fh.y; hi[2] = fh.z;
    double ext = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double m = 0.5 * (lo[d] + hi[d]);
      k[d] = okey(m);
      k[3 + d] = k[d];
      ext = dmax(ext, hi[d] - lo[d]);
    }
    k[6] = okey(ext);
  }
  // batched models: the model's largest edge, one atomic per wave when the wave holds
  // one model (one model: the largest edge of all, reduced below)
  if (c.seg_model) {
    const int m0 = __shfl(model, 0, 64);
    const bool one = __ballot(l < c.nrtm && model != m0) == 0ull;
    if (one) {
      unsigned long long e = k[6];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long b = __shfl_xor(e, o, 64);
        e = b > e ? b : e;
      }
      if ((threadIdx.x & 63) == 0 && e) atomicMax(&c.model_ext[m0], e);
    } else if (l < c.nrtm) {
      atomicMax(&c.model_ext[model], k[6]);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int j = 0; j < 7; ++j) {
      const unsigned long long b = __shfl_xor(k[j], o, 64);
      k[j] = j < 3 ? (b < k[j] ? b : k[j]) : (b > k[j] ? b : k[j]);
    }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int j = 0; j < 7; ++j) s_b[j][w] = k[j];
  __syncthreads();
  if (threadIdx.x < 7) {
    const int j = threadIdx.x;
    unsigned long long v = s_b[j][0];
    for (int q = 1; q < 4; ++q) v = j < 3 ? (s_b[j][q] < v ? s_b[j][q] : v)
                                          : (s_b[j][q] > v ? s_b[j][q] : v);
    c.bpart[(size_t)j * c.nbound_blocks + blockIdx.x] = v;
  }
}

// The workgroups' bounds reduced by one workgroup, then the grid of the pass.
__global__ __launch_bounds__(1024) void i7_bounds_grid(Contact7 c) {
  if (!c.counters[I7C_SEARCH]) return;
  __shared__ unsigned long long s_b[7][16];
  unsigned long long k[7] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull, 0ull};
  for (int b = threadIdx.x; b < c.nbound_blocks; b += 1024)
    for (int j = 0; j < 7; ++j) {
      const unsigned long long v = c.bpart[(size_t)j * c.nbound_blocks + b];
      k[j] = j < 3 ? (v < k[j] ? v : k[j]) : (v > k[j] ? v : k[j]);
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int j = 0; j < 7; ++j) {
      const unsigned long long b = __shfl_xor(k[j], o, 64);
      k[j] = j < 3 ? (b < k[j] ? b : k[j]) : (b > k[j] ? b : k[j]);
    }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int j = 0; j < 7; ++j) s_b[j][w] = k[j];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 0; j < 7; ++j) {
      unsigned long long v = s_b[j][0];
      for (int q = 1; q < 16; ++q) v = j < 3 ? (s_b[j][q] < v ? s_b[j][q] : v)
                                             : (s_b[j][q] > v ? s_b[j][q] : v);
      c.bounds[j] = v;
    }
    const I7Grid g = i7_grid(c);
    for (int d = 0; d < 3; ++d) c.grid[d] = g.o[d];
    c.grid[3] = g.s;
  }
}

__global__ __launch_bounds__(256) void i7_seg_count(Contact7 c) {
  if (!c.counters[I7C_SEARCH]) return;
  const int l = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (l >= c.nrtm) return;
  if (c.stfm[l] == 0.0) {   // never a candidate
    c.seg_cell[l] = -1;
    return;
  }
  const int model = c.seg_model ? c.seg_model[l] : 0;
  const double inv = 1.0 / model_cell(c, model);
  const float4 lo = c.seg_lo[l], hi = c.seg_hi[l];
  const long long kx = cell_coord(0.5 * ((double)lo.x + (double)hi.x), c.grid[0], inv);
  const long long ky = cell_coord(0.5 * ((double)lo.y + (double)hi.y), c.grid[1], inv);
  const long long kz = cell_coord(0.5 * ((double)lo.z + (double)hi.z), c.grid[2], inv);
  const int cell = cell_hash(kx, ky, kz, model, c.cell_cap);
  c.seg_cell[l] = cell;
  atomicAdd(&c.cell_count[cell], 1);
}

__global__ __launch_bounds__(256) void i7_seg_place(Contact7 c) {
  if (!c.counters[I7C_SEARCH]) return;
  const int l = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (l >= c.nrtm) return;
  const int cell = c.seg_cell[l];
  if (cell < 0) return;
  const int j = c.cell_start[cell] + atomicAdd(&c.cell_fill[cell], 1);
  float4 lo = c.seg_lo[l];
  lo.w = __int_as_float(l);
  c.cell_lo[j] = lo;
  c.cell_hi[j] = c.seg_hi[l];
}

// ------------------------------------------------------------- per pair --
struct I7Pair {
  int ix[4], nsvg;
  double xi[3], vi[3], msi;
  double x[4][3];
  double nx[4][3], lb[4], lc[4], p[4], stif, gapv;
};

// I7COR3 + I7DST3 for candidate (ni, l); true when the pair penetrates (KEEP).
__device__ __forceinline__ bool i7_dist(const Contact7 &c, const double4 *X4, const double4 *V4,
                                     int ni, int l, I7Pair &q) {
  q.gapv = c.gap;
  const int ig = c.nsv[ni];
  q.nsvg = ig;
  const double4 xi = X4[ig], vi = V4[ig];
  q.xi[0] = xi.x; q.xi[1] = xi.y; q.xi[2] = xi.z;
  q.vi[0] = vi.x; q.vi[1] = vi.y; q.vi[2] = vi.z;
  q.msi = xi.w;
  const int4 r = c.irect[l];
  q.ix[0] = r.x; q.ix[1] = r.y; q.ix[2] = r.z; q.ix[3] = r.w;
  for (int k = 0; k < 4; ++k) {
    const double4 p = X4[q.ix[k]];
    q.x[k][0] = p.x; q.x[k][1] = p.y; q.x[k][2] = p.z;
  }
  q.stif = c.stfm[l] * fabs(c.stfns[ni]);
  double x0[3];
  if (q.ix[2] != q.ix[3]) {
    for (int d = 0; d < 3; ++d) x0[d] = 0.25 * (q.x[0][d] + q.x[1][d] + q.x[2][d] + q.x[3][d]);
  } else {
    for (int d = 0; d < 3; ++d) x0[d] = q.x[2][d];
  }
  double x0k[4][3], xik[4][3], xi0v[3];
  for (int k = 0; k < 4; ++k)
    for (int d = 0; d < 3; ++d) x0k[k][d] = q.x[k][d] - x0[d];
  for (int d = 0; d < 3; ++d) xi0v[d] = x0[d] - q.xi[d];
  for (int k = 0; k < 4; ++k)
    for (int d = 0; d < 3; ++d) xik[k][d] = q.x[k][d] - q.xi[d];
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
    const double s2 = 1.0 / dmax(c.em30, sx0 * sx0 + sy0 * sy0 + sz0 * sz0);
    lb[k] = -(sx0 * s[k1][0] + sy0 * s[k1][1] + sz0 * s[k1][2]) * s2;
    lc[k] = (sx0 * s[k][0] + sy0 * s[k][1] + sz0 * s[k][2]) * s2;
  }
  double hlb[4], hlc[4], al[4];
  for (int k = 0; k < 4; ++k) {
    const int km = (k + 3) & 3;
    const double aaa = 1.0 / dmax(c.em30, x0k[k][0] * x0k[k][0] + x0k[k][1] * x0k[k][1] +
                                              x0k[k][2] * x0k[k][2]);
    hlc[k] = lc[k] * fabs(lc[k]) * aaa;
    hlb[km] = lb[km] * fabs(lb[km]) * aaa;
    double a = -(xi0v[0] * x0k[k][0] + xi0v[1] * x0k[k][1] + xi0v[2] * x0k[k][2]) * aaa;
    al[k] = dmax(0.0, dmin(1.0, a));
  }
  for (int k = 0; k < 4; ++k) {
    const int k1 = (k + 1) & 3;
    const double ex = q.x[k1][0] - q.x[k][0];
    const double ey = q.x[k1][1] - q.x[k][1];
    const double ez = q.x[k1][2] - q.x[k][2];
    const double la = 1.0 - lb[k] - lc[k];
    const double aaa = 1.0 / dmax(c.em20, ex * ex + ey * ey + ez * ez);
    const double hla = la * fabs(la) * aaa;
    if (la < 0.0 && hla <= hlb[k] && hla <= hlc[k]) {
      double t = (xik[k1][0] * ex + xik[k1][1] * ey + xik[k1][2] * ez) * aaa;
      lb[k] = dmax(0.0, dmin(1.0, t));
      lc[k] = 1.0 - lb[k];
    } else if (lb[k] < 0.0 && hlb[k] <= hlc[k] && hlb[k] <= hla) {
      lb[k] = 0.0;
      lc[k] = al[k1];
    } else if (lc[k] < 0.0 && hlc[k] <= hla && hlc[k] <= hlb[k]) {
      lc[k] = 0.0;
      lb[k] = al[k];
    }
  }
  const double gap2 = q.gapv * q.gapv;
  double pene2 = 0.0;
  for (int k = 0; k < 4; ++k) {
    const int k1 = (k + 1) & 3;
    for (int d = 0; d < 3; ++d)
      q.nx[k][d] = q.xi[d] - (x0[d] + lb[k] * x0k[k][d] + lc[k] * x0k[k1][d]);
    q.p[k] = q.nx[k][0] * q.nx[k][0] + q.nx[k][1] * q.nx[k][1] + q.nx[k][2] * q.nx[k][2];
    const double dk = dmax(0.0, gap2 - q.p[k]);
    pene2 = k == 0 ? dk : dmax(pene2, dk);
    q.lb[k] = lb[k];
    q.lc[k] = lc[k];
  }
  return pene2 != 0.0 && q.stif != 0.0;
}

// Each secondary node's candidates (I7TRI's box test, i7tri.F:530-566, widened by
// TZS = TZ + margin): the segments of the 2 x 2 x 2 cells around it whose box widened by
// TZS holds it, but its own segments and other models'.  Eight lanes per node, one cell
// each; a lane appends to the node's row (node_ncand cleared by i7_search_reset), so a row
// is in no particular order -- the filter keeps the penetrated ones in segment order.
__global__ __launch_bounds__(256) void i7_node_search(Contact7 c, const double4 *X4) {
  if (!c.counters[I7C_SEARCH]) return;
  const long t = (long)blockIdx.x * 256 + (long)threadIdx.x;
  const int i = (int)(t >> 3), q = (int)(t & 7);
  if (i >= c.nsn || c.stfns[i] == 0.0) return;
  const int nn = c.nsv[i];
  const double4 x = X4[nn];
  const int model = c.node_model ? c.node_model[nn] : 0;
  const double inv = 1.0 / model_cell(c, model);
  const long long kx = cell_coord(x.x, c.grid[0], inv);
  const long long ky = cell_coord(x.y, c.grid[1], inv);
  const long long kz = cell_coord(x.z, c.grid[2], inv);
  // the neighbour on the side of the node's half of its cell, per axis
  const long long nx = ((x.x - c.grid[0]) * inv - (double)kx < 0.5) ? kx - 1 : kx + 1;
  const long long ny = ((x.y - c.grid[1]) * inv - (double)ky < 0.5) ? ky - 1 : ky + 1;
  const long long nz = ((x.z - c.grid[2]) * inv - (double)kz < 0.5) ? kz - 1 : kz + 1;
  const int cell = cell_hash((q & 1) ? nx : kx, (q & 2) ? ny : ky, (q & 4) ? nz : kz, model,
                             c.cell_cap);
  // The eight cells lie in eight different buckets (on each axis the two coordinates
  // differ by one, so the cells' places in their 4 x 4 x 4 blocks differ), so every
  // segment is met once.
  int32_t *cand = c.node_cand + (size_t)i * I7_NODE_CAP;
  int found = 0;
  for (int j = c.cell_start[cell]; j < c.cell_start[cell + 1]; ++j) {
    const float4 lo = c.cell_lo[j], hi = c.cell_hi[j];
    if (!(x.x > (double)lo.x - c.tzs && x.x < (double)hi.x + c.tzs &&
          x.y > (double)lo.y - c.tzs && x.y < (double)hi.y + c.tzs &&
          x.z > (double)lo.z - c.tzs && x.z < (double)hi.z + c.tzs))
      continue;
    const int l = __float_as_int(lo.w);
    const int4 r = c.irect[l];
    if (nn == r.x || nn == r.y || nn == r.z || nn == r.w) continue;
    if (c.node_model && model != c.seg_model[l]) continue;
    const int k = atomicAdd(&c.node_ncand[i], 1);
    if (k < I7_NODE_CAP) cand[k] = l;
    else atomicOr(&c.counters[I7C_OVERFLOW], I7_OVF_NODE);
    ++found;
  }
  if (found) atomicAdd(&c.counters[I7C_NCAND], found);
}

// I7COR3 + I7DST3 over the candidates: one lane per node reads its count, the wave's
// candidates are dealt out over its 64 lanes (so a wave with one busy node and one with
// many cost alike), the penetrated ones marked in a 64-bit mask per node in LDS, then
// each node's lane lists them in increasing segment order.
__global__ __launch_bounds__(256) void i7_node_filter(Contact7 c, const double4 *X4,
                                                     const double4 *V4) {
  __shared__ unsigned long long s_mask[256];
  __shared__ int s_start[256];
  const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
  const int lane = threadIdx.x & 63, w0 = threadIdx.x & ~63;
  const bool live = i < c.nsn;
  const int n = live ? min(c.node_ncand[i], I7_NODE_CAP) : 0;
  int x = n;   // inclusive scan of the counts over the wave
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  const int total = __shfl(x, 63, 64);
  s_start[threadIdx.x] = x - n;
  s_mask[threadIdx.x] = 0ull;
  __syncthreads();
  for (int k = lane; k < total; k += 64) {
    // the lane whose range holds k: the last one starting at or before k
    int lo = 0, hi = 63;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_start[w0 + mid] <= k) lo = mid;
      else hi = mid - 1;
    }
    const int j = k - s_start[w0 + lo];
    const int ni = (int)blockIdx.x * 256 + w0 + lo;
    I7Pair p;
    if (i7_dist(c, X4, V4, ni, c.node_cand[(size_t)ni * I7_NODE_CAP + j], p))
      atomicOr(&s_mask[w0 + lo], 1ull << j);
  }
  __syncthreads();
  if (!live) return;
  unsigned long long mask = s_mask[threadIdx.x];
  const int32_t *cand = c.node_cand + (size_t)i * I7_NODE_CAP;
  int32_t *keep = c.node_keep + (size_t)i * I7_KEEP_CAP;
  int nkeep = 0;
  while (mask) {
    const int j = __ffsll((long long)mask) - 1;
    mask &= mask - 1;
    const int l = cand[j];
    if (nkeep < I7_KEEP_CAP) {
      int p = nkeep;
      for (; p > 0 && keep[p - 1] > l; --p) keep[p] = keep[p - 1];
      keep[p] = l;
    }
    ++nkeep;
  }
  if (nkeep > I7_KEEP_CAP) {
    atomicOr(&c.counters[I7C_OVERFLOW], I7_OVF_KEEP);
    nkeep = I7_KEEP_CAP;
  }
  c.node_nkeep[i] = nkeep;
}

// The penetrating pairs before secondary node k, capped at the pair capacity.
__device__ __forceinline__ int i7_pairs_before(const Contact7 &c, int k) {
  const int v = c.node_kstart[k];
  return v > c.pair_cap ? c.pair_cap : v;
}

// Batches per model (NVSIZ pairs each, a model's last one ragged) and their prefix: one
// workgroup.
__global__ __launch_bounds__(1024) void i7_model_batches(Contact7 c) {
  __shared__ int s_w[17];
  constexpr int PER = 16;
  const int m0 = (int)threadIdx.x * PER;
  int nb[PER], sum = 0;
  for (int k = 0; k < PER; ++k) {
    const int m = m0 + k;
    nb[k] = 0;
    if (m < c.nmodel) {
      const int n = i7_pairs_before(c, c.nsv_off[m + 1]) - i7_pairs_before(c, c.nsv_off[m]);
      nb[k] = (n + I7_NVSIZ - 1) / I7_NVSIZ;
    }
    sum += nb[k];
  }
  // exclusive scan of the threads' sums
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = sum;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int i = 0; i < 16; ++i) {
      const int t = s_w[i];
      s_w[i] = acc;
      acc += t;
    }
    s_w[16] = acc;
    const int tot = c.node_kstart[c.nsn];
    if (tot > c.pair_cap) atomicOr(&c.counters[I7C_OVERFLOW], I7_OVF_PAIRS);
    c.counters[I7C_NPEN] = tot > c.pair_cap ? c.pair_cap : tot;
    c.counters[I7C_NBATCH] = acc;
    c.model_bstart[c.nmodel] = acc;
  }
  __syncthreads();
  int pre = x - sum + s_w[w];
  for (int k = 0; k < PER; ++k) {
    const int m = m0 + k;
    if (m < c.nmodel) c.model_bstart[m] = pre;
    pre += nb[k];
  }
}

// I7FOR3 + I7ASS2 for one batch of NVSIZ penetrating pairs, one lane per pair.
__global__ __launch_bounds__(I7_NVSIZ) void i7_batch(Contact7 c, const double4 *X4,
                                                     const double4 *V4, const double *dt1b,
                                                     unsigned long long *dt2tb, size_t stride) {
  __shared__ double s_rd[I7_NVSIZ], s_ec[I7_NVSIZ], s_ev[I7_NVSIZ];
  __shared__ int s_cnt[5][2];
  const int nbt = c.model_bstart[c.nmodel];
  // a bounded grid strides over the batches (most of a large interface's capacity is
  // unused); the workgroup's shared arrays are reused batch after batch
  for (int b = (int)blockIdx.x; b < nbt; b += (int)gridDim.x) {
    // the batch's model: the last m with model_bstart[m] <= b
    int mlo = 0, mhi = c.nmodel - 1;
    while (mlo < mhi) {
      const int mid = (mlo + mhi + 1) >> 1;
      if (c.model_bstart[mid] <= b) mlo = mid;
      else mhi = mid - 1;
    }
    const int model = mlo;
    const int p0 = i7_pairs_before(c, c.nsv_off[model]) + (b - c.model_bstart[model]) * I7_NVSIZ;
    const int p1 = i7_pairs_before(c, c.nsv_off[model + 1]);
    const int t = (int)threadIdx.x;
    const int r = p0 + t;
    const bool live = r < p1;
    const int nlive = p1 - p0 < I7_NVSIZ ? p1 - p0 : I7_NVSIZ;
    const double dt1 = *(const double *)((const char *)dt1b + (size_t)model * stride);
    unsigned long long *dt2t =
        (unsigned long long *)((char *)dt2tb + (size_t)model * stride);
    const double dt1inv = dt1 > 0.0 ? 1.0 / dt1 : 0.0;
    double h[4] = {0.0, 0.0, 0.0, 0.0}, fxi[3] = {0.0, 0.0, 0.0}, stif = 0.0;
    double rdist = 1.0e300, econtt = 0.0, econvt = 0.0;
    int ix[4] = {0, 0, 0, 0}, nsvg = 0;
    if (live) {
      // the secondary node of pair r: the last i with node_kstart[i] <= r
      int lo = 0, hi = c.nsn - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (c.node_kstart[mid] <= r) lo = mid;
        else hi = mid - 1;
      }
      const int ni = lo;
      const int l = c.node_keep[(size_t)ni * I7_KEEP_CAP + (r - c.node_kstart[ni])];
      I7Pair q;
      i7_dist(c, X4, V4, ni, l, q);
      for (int j = 0; j < 4; ++j) ix[j] = q.ix[j];
      nsvg = q.nsvg;
      stif = q.stif;
      const double g = q.gapv;
      double pp[4], a[4], n[3], pene;
      for (int j = 0; j < 4; ++j) {
        const double dk = sqrt(q.p[j]);
        pp[j] = dmax(0.0, g - dk);
        a[j] = pp[j] / dmax(c.em20, dk);
      }
      for (int d = 0; d < 3; ++d)
        n[d] = a[0] * q.nx[0][d] + a[1] * q.nx[1][d] + a[2] * q.nx[2][d] + a[3] * q.nx[3][d];
      if (q.ix[2] != q.ix[3]) {
        pene = dmax(dmax(dmax(pp[0], pp[1]), pp[2]), pp[3]);
        double la[4];
        for (int j = 0; j < 4; ++j) la[j] = 1.0 - q.lb[j] - q.lc[j];
        const double h0 = 0.25 * (pp[0] * la[0] + pp[1] * la[1] + pp[2] * la[2] + pp[3] * la[3]);
        h[0] = h0 + pp[0] * q.lb[0] + pp[3] * q.lc[3];
        h[1] = h0 + pp[1] * q.lb[1] + pp[0] * q.lc[0];
        h[2] = h0 + pp[2] * q.lb[2] + pp[1] * q.lc[1];
        h[3] = h0 + pp[3] * q.lb[3] + pp[2] * q.lc[2];
        const double h00 = 1.0 / dmax(c.em20, h[0] + h[1] + h[2] + h[3]);
        for (int j = 0; j < 4; ++j) h[j] = h[j] * h00;
      } else {
        pene = pp[0];
        for (int d = 0; d < 3; ++d) n[d] = q.nx[0][d];
        h[0] = q.lb[0];
        h[1] = q.lc[0];
        h[2] = 1.0 - q.lb[0] - q.lc[0];
        h[3] = 0.0;
      }
      const double s2 = 1.0 / dmax(c.em30, sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]));
      for (int d = 0; d < 3; ++d) n[d] = n[d] * s2;
      double vx[3];
      for (int d = 0; d < 3; ++d) {
        const double v0 = d == 0 ? V4[ix[0]].x : (d == 1 ? V4[ix[0]].y : V4[ix[0]].z);
        const double v1 = d == 0 ? V4[ix[1]].x : (d == 1 ? V4[ix[1]].y : V4[ix[1]].z);
        const double v2 = d == 0 ? V4[ix[2]].x : (d == 1 ? V4[ix[2]].y : V4[ix[2]].z);
        const double v3 = d == 0 ? V4[ix[3]].x : (d == 1 ? V4[ix[3]].y : V4[ix[3]].z);
        vx[d] = q.vi[d] - h[0] * v0 - h[1] * v1 - h[2] * v2 - h[3] * v3;
      }
      const double vn = n[0] * vx[0] + n[1] * vx[1] + n[2] * vx[2];
      double hc = -0.25 * (h[0] - h[1] + h[2] - h[3]);
      hc = dmin(dmin(hc, h[1]), h[3]);
      hc = dmax(dmax(hc, -h[0]), -h[2]);
      if (q.ix[2] == q.ix[3]) hc = 0.0;
      h[0] = h[0] + hc;
      h[1] = h[1] - hc;
      h[2] = h[2] + hc;
      h[3] = h[3] - hc;
      // contact time step of the pair
      rdist = 0.5 * (g - pene) / dmax(c.em30, -vn);
      // nonlinear penalty
      const double fac = g / dmax(c.em10, g - pene);
      const double facm1 = 1.0 / fac;
      if ((g - pene) / g < c.em10 && stif > 0.0) {
        stif = 0.0;
        c.stfns[ni] = -fabs(c.stfns[ni]);
      }
      econtt = 0.5 * stif * (g * g) * (facm1 - 1.0 - log(dmax(2.2250738585072014e-308, facm1)));
      stif = 0.5 * stif * fac;
      double fni = -stif * pene;
      double vis2 = 0.0;
      if (c.visc != 0.0 || c.viscf != 0.0) {
        vis2 = 2.0 * stif * q.msi;
        if (vn < 0.0) vis2 = vis2 / dmax(c.em10, (g - pene) / g);
        const double fac2 = stif / dmax(c.em30, stif);
        const double vis = sqrt(vis2);
        double ff = fac2 * (c.visc * vis + c.zep4 * c.zep4 * 2.0 * q.msi * dmax(0.0, -vn) /
                                                dmax(g - pene, c.em10));
        stif = stif * g / dmax(g - pene, c.em10);
        stif = stif + ff * dt1inv;
        stif = dmax(stif, fac2 * sqrt(c.viscf) * vis * dt1inv);
        ff = ff * vn;
        fni = fni + ff;
      } else {
        stif = stif * g / dmax(g - pene, c.em10);
      }
      for (int d = 0; d < 3; ++d) fxi[d] = n[d] * fni;
      // Coulomb friction, total viscous formulation
      double tv[3];
      for (int d = 0; d < 3; ++d) tv[d] = vx[d] - n[d] * vn;
      const double v2 = tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2];
      const double vis2f = c.viscf * vis2;
      const double fm2 = (c.fric * fni) * (c.fric * fni);
      const double f2 = vis2f * v2;
      const double a2 = dmin(f2, fm2) / dmax(c.em30, f2);
      const double aa = sqrt(a2 * vis2f);
      for (int d = 0; d < 3; ++d) fxi[d] = fxi[d] + aa * tv[d];
      econvt = aa * v2 * dt1;
    }
    s_rd[t] = rdist;
    s_ec[t] = econtt;
    s_ev[t] = econvt;
    // role counts per wave (skyline layout of I7ASS2)
    const int lane = t & 63, w = t >> 6;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    unsigned long long m[4];
    for (int j = 0; j < 4; ++j) {
      m[j] = __ballot(live && h[j] != 0.0);
      if (lane == 0) s_cnt[j][w] = __popcll(m[j]);
    }
    __syncthreads();
    if (t == 0) {
      // batch contact time step (i7for3.F:772-870) and energies, in pair order
      double dti = (double)1.E20f, ec = 0.0, ev = 0.0;
      for (int i = 0; i < nlive; ++i) dti = dmin(s_rd[i], dti);
      if (dti <= 0.0) dti = (double)1.E20f;
      atomicMin(dt2t, (unsigned long long)__double_as_longlong(dti));
      for (int i = 0; i < nlive; ++i) ec = ec + s_ec[i];
      for (int i = 0; i < nlive; ++i) ev = ev + s_ev[i];
      c.batch_e[2 * b] = ec;
      c.batch_e[2 * b + 1] = ev;
      int tot = 0;
      for (int j = 0; j < 4; ++j) tot += s_cnt[j][0] + s_cnt[j][1];
      c.batch_cnt[b] = tot + nlive;
      // the last batch to finish adds the batches' energies in batch order
      // (ECONT = ECONT + ECONTT per I7FOR3 call)
      __threadfence();
      const int nb = nbt;
      if (atomicAdd(&c.counters[I7C_DONE], 1) == nb - 1) {
        __threadfence();
        double et = c.energy[0], vt = c.energy[1];
        int nsky = 0;
        for (int q = 0; q < nb; ++q) {
          et = et + __hip_atomic_load(&c.batch_e[2 * q], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
          vt = vt + __hip_atomic_load(&c.batch_e[2 * q + 1], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
          nsky += __hip_atomic_load(&c.batch_cnt[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        c.energy[0] = et;
        c.energy[1] = vt;
        c.counters[I7C_NSKY] = nsky;
      }
    }
    if (live) {
      const int e0 = b * I7_SLOTS;
      int base = 0;
      for (int j = 0; j < 4; ++j) {
        if (h[j] != 0.0) {
          const int e = e0 + base + (w ? s_cnt[j][0] : 0) + __popcll(m[j] & lt);
          c.sky[e] = make_double4(fxi[0] * h[j], fxi[1] * h[j], fxi[2] * h[j], stif * fabs(h[j]));
          c.sky_node[e] = ix[j];
          c.sky_next[e] = atomicExch(&c.node_head[ix[j]], e);
        }
        base += s_cnt[j][0] + s_cnt[j][1];
      }
      const int e = e0 + base + t;
      c.sky[e] = make_double4(-fxi[0], -fxi[1], -fxi[2], stif);
      c.sky_node[e] = nsvg;
      c.sky_next[e] = atomicExch(&c.node_head[nsvg], e);
    }
    __syncthreads();
  }
}

hipError_t contact7_run(const Contact7 &c, const double4 *X4, const double4 *V4,
                        const double *dt1, unsigned long long *dt2t, size_t stride,
                        hipStream_t s) {
  auto g = [](long n, int blk) {
    const long b = (n + blk - 1) / blk;
    return dim3((unsigned)(b > 0 ? b : 1));
  };
  long nsr = c.cell_cap;
  if (c.numnod > nsr) nsr = c.numnod;
  if (c.nmodel > nsr) nsr = c.nmodel;
  if (c.nsn > nsr) nsr = c.nsn;
  if (nsr < 8) nsr = 8;
  hipLaunchKernelGGL(i7_reset, g(c.numnod > 8 ? c.numnod : 8, 256), dim3(256), 0, s, c, X4);
  hipLaunchKernelGGL(i7_decide, dim3(1), dim3(1024), 0, s, c);
  hipLaunchKernelGGL(i7_search_reset, g(nsr, 256), dim3(256), 0, s, c, X4);
  hipLaunchKernelGGL(i7_seg_bounds, dim3(c.nbound_blocks), dim3(256), 0, s, c, X4);
  hipLaunchKernelGGL(i7_bounds_grid, dim3(1), dim3(1024), 0, s, c);
  hipLaunchKernelGGL(i7_seg_count, g(c.nrtm, 256), dim3(256), 0, s, c);
  hipError_t e = scan_excl(c.cell_count, c.cell_start, c.counters + I7C_NCELL, 0, c.cell_cap,
                           c.scan_tmp, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(i7_seg_place, g(c.nrtm, 256), dim3(256), 0, s, c);
  hipLaunchKernelGGL(i7_node_search, g(8L * c.nsn, 256), dim3(256), 0, s, c, X4);
  hipLaunchKernelGGL(i7_node_filter, g(c.nsn, 256), dim3(256), 0, s, c, X4, V4);
  e = scan_excl(c.node_nkeep, c.node_kstart, nullptr, c.nsn, c.nsn, c.scan_tmp, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(i7_model_batches, dim3(1), dim3(1024), 0, s, c);
  hipLaunchKernelGGL(i7_batch,
                     dim3(c.nbatch_cap < 1 ? 1 : c.nbatch_cap > 2048 ? 2048 : c.nbatch_cap),
                     dim3(I7_NVSIZ), 0, s,
                     c, X4, V4, dt1, dt2t, stride);
  return hipGetLastError();
}

}  // namespace orad

// orad_brick_order.hip -- the library's own order of the brick elements.
//
// The element kernel streams the element arrays in their storage order and gathers
// the corner nodes.  A node is shared by the elements of two consecutive element
// layers; in the engine's numbering of a block (x fastest, then y, then z) those two
// layers are a whole layer of elements apart, more than an XCD's L2 holds, so the
// corner records and the partly written slot lines of a node layer leave the cache
// between their two uses.  The library therefore stores bricks in column order:
// element centroids are binned on the mean element size, the mesh is cut into
// columns of `rows` element rows (y), and each column is walked layer by layer (z),
// row by row, along x.  The two users of a node layer are then one column layer
// (rows x nx elements) apart.  Any permutation gives bitwise the same slots, nodes
// and state (each slot is written by exactly one element; ASSPAR4 sums in slot
// order); only the per-workgroup hourglass-energy partials are summed in a different
// order.  rows = 0 keeps the engine's order.
//
// Arrays kept in library order: conn, rank, the element state, vol0, sti.
// h_perm[i] = engine element of library element i; emap (device) = the inverse,
// used by the per-group drop-in (engine-numbered groups) and the energy sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <tuple>
#include <vector>

#include "orad_ctx.hpp"

namespace orad {

template <typename T>
__global__ __launch_bounds__(256) void gather_columns(const T *__restrict__ in,
                                                      T *__restrict__ out,
                                                      const int32_t *__restrict__ map, int ne,
                                                      int ncol) {
  const int i = blockIdx.x * 256 + (int)threadIdx.x;
  if (i >= ne) return;
  const int j = map[i];
  for (int c = 0; c < ncol; ++c) out[(size_t)c * ne + i] = in[(size_t)c * ne + j];
}

// out[c*ne + i] = in[c*ne + map[i]] for ncol columns of `bytes`-byte values.
static hipError_t gather(const void *in, void *out, const int32_t *map, int ne, int ncol,
                         size_t bytes, hipStream_t s) {
  const int nb = (ne + 255) / 256;
  if (nb == 0) return hipSuccess;
  if (bytes == 8)
    hipLaunchKernelGGL(gather_columns<unsigned long long>, dim3(nb), dim3(256), 0, s,
                       (const unsigned long long *)in, (unsigned long long *)out, map, ne, ncol);
  else
    hipLaunchKernelGGL(gather_columns<unsigned>, dim3(nb), dim3(256), 0, s,
                       (const unsigned *)in, (unsigned *)out, map, ne, ncol);
  return hipGetLastError();
}

}  // namespace orad

// Centroid bins of the bricks on the mean element size, engine order, packed
// (z << 42 | y << 21 | x).  Called once the connectivity and X are known.
int brick_bins(orad_ctx *c, size_t ne, const std::vector<int32_t> &conn_em, const double *vol0) {
  const size_t nn = (size_t)c->numnod;
  std::vector<double4> X(nn);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(X.data(), c->X4, nn * sizeof(double4), hipMemcpyDeviceToHost));
  double vs = 0.0;
  for (size_t e = 0; e < ne; ++e) vs += std::fabs(vol0[e]);
  double h = std::cbrt(vs / (double)std::max<size_t>(ne, 1));
  if (!(h > 0.0) || !std::isfinite(h)) h = 1.0;
  std::vector<double> cx(ne), cy(ne), cz(ne);
  double lo[3] = {1e300, 1e300, 1e300};
  for (size_t e = 0; e < ne; ++e) {
    double s[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 8; ++k) {
      const double4 &p = X[conn_em[8 * e + k]];
      s[0] += p.x; s[1] += p.y; s[2] += p.z;
    }
    cx[e] = 0.125 * s[0]; cy[e] = 0.125 * s[1]; cz[e] = 0.125 * s[2];
    lo[0] = std::min(lo[0], cx[e]); lo[1] = std::min(lo[1], cy[e]); lo[2] = std::min(lo[2], cz[e]);
  }
  const double lim = (double)((1 << 21) - 1);
  auto bin = [&](double v, double l) {
    const double b = std::floor((v - l) / h + 0.5);
    return (unsigned long long)std::min(std::max(b, 0.0), lim);
  };
  c->h_bin.resize(ne);
  for (size_t e = 0; e < ne; ++e)
    c->h_bin[e] = bin(cz[e], lo[2]) << 42 | bin(cy[e], lo[1]) << 21 | bin(cx[e], lo[0]);
  return ORAD_OK;
}

// Engine element of each library element for column height `rows` (0: identity).
static std::vector<int32_t> column_perm(const orad_ctx *c, int rows) {
  const size_t ne = c->h_bin.size();
  std::vector<int32_t> p(ne);
  std::iota(p.begin(), p.end(), 0);
  if (rows <= 0) return p;
  const unsigned long long m21 = (1ull << 21) - 1;
  std::vector<unsigned long long> key(ne);
  for (size_t e = 0; e < ne; ++e) {
    const unsigned long long b = c->h_bin[e];
    const unsigned long long bz = b >> 42, by = (b >> 21) & m21, bx = b & m21;
    // column, then layer, then row within the column, then x
    key[e] = ((by / (unsigned)rows) << 42) | (bz << 21) | ((by % (unsigned)rows) << 14) |
             std::min<unsigned long long>(bx, (1ull << 14) - 1);
  }
  std::stable_sort(p.begin(), p.end(), [&](int32_t a, int32_t b) {
    if (key[a] != key[b]) return key[a] < key[b];
    return c->h_bin[a] < c->h_bin[b];
  });
  return p;
}

// Puts the bricks in the column order of height `rows`, permuting every device array
// kept in library order and, when given, the state and STI copies of a snapshot
// (`extra`: pairs of (state copy, sti copy)).
int brick_reorder(orad_ctx *c, int rows, const std::vector<std::pair<void *, void *>> &extra) {
  const size_t ne = (size_t)c->numels;
  std::vector<int32_t> np = column_perm(c, rows);
  std::vector<int32_t> old_iperm(ne);
  if (c->h_perm.empty())
    std::iota(old_iperm.begin(), old_iperm.end(), 0);
  else
    for (size_t i = 0; i < ne; ++i) old_iperm[c->h_perm[i]] = (int32_t)i;
  std::vector<int32_t> map(ne);
  bool same = true;
  for (size_t i = 0; i < ne; ++i) {
    map[i] = old_iperm[np[i]];
    same = same && map[i] == (int32_t)i;
  }
  c->order_rows = rows;
  HIPCHK(c, hipSetDevice(c->device));
  if (same) {
    // already in this order: only make sure the inverse map exists
    if (!c->emap) {
      HIPCHK(c, dalloc(&c->emap, ne * 4));
      HIPCHK(c, hipMemcpy(c->emap, old_iperm.data(), ne * 4, hipMemcpyHostToDevice));
    }
    c->h_perm = np;
    c->h_iperm = old_iperm;
    return ORAD_OK;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  int32_t *dmap = nullptr;
  HIPCHK(c, dalloc(&dmap, ne * 4));
  hipError_t e = hipMemcpy(dmap, map.data(), ne * 4, hipMemcpyHostToDevice);
  const size_t ts = c->tsize();
  // (array, columns, bytes per value): each gathered into a fresh buffer, then swapped
  std::vector<std::tuple<void **, int, size_t>> arrays = {
      {(void **)&c->conn, 8, 4}, {(void **)&c->rank, 1, 8}, {&c->state, NSTATE, ts},
      {(void **)&c->vol0, 1, 8}, {&c->sti, 1, ts}};
  for (auto &a : arrays) {
    if (e != hipSuccess) break;
    void *out = nullptr;
    e = dalloc(&out, (size_t)std::get<1>(a) * ne * std::get<2>(a));
    if (e == hipSuccess)
      e = orad::gather(*std::get<0>(a), out, dmap, (int)ne, std::get<1>(a), std::get<2>(a),
                       c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) {
      dfree(*std::get<0>(a));
      *std::get<0>(a) = out;
    } else {
      dfree(out);
    }
  }
  for (const auto &x : extra) {
    for (int which = 0; which < 2 && e == hipSuccess; ++which) {
      void *buf = which == 0 ? x.first : x.second;
      const int ncol = which == 0 ? NSTATE : 1;
      if (!buf) continue;
      void *tmp = nullptr;
      e = dalloc(&tmp, (size_t)ncol * ne * ts);
      if (e == hipSuccess) e = orad::gather(buf, tmp, dmap, (int)ne, ncol, ts, c->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(buf, tmp, (size_t)ncol * ne * ts, hipMemcpyDeviceToDevice, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      dfree(tmp);
    }
  }
  dfree(dmap);
  if (e != hipSuccess) return fail(c, ORAD_EHIP, std::string("brick reorder: ") + hipGetErrorString(e));
  // new permutation, its inverse on the device, and the slot -> element map
  c->h_perm = np;
  std::vector<int32_t> iperm(ne);
  for (size_t i = 0; i < ne; ++i) iperm[np[i]] = (int32_t)i;
  std::vector<int32_t> se(8 * ne);
  for (size_t en = 0; en < ne; ++en)
    for (int k = 0; k < 8; ++k) se[c->h_iads[8 * en + k]] = iperm[en];
  if (!c->emap) HIPCHK(c, dalloc(&c->emap, ne * 4));
  HIPCHK(c, hipMemcpy(c->emap, iperm.data(), ne * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->slot_elem, se.data(), 8 * ne * 4, hipMemcpyHostToDevice));
  c->h_iperm = std::move(iperm);
  return ORAD_OK;
}

// orad_brick_io.hpp -- memory accesses of the element-force kernel for one element:
// corner connectivity and slot addresses, corner node records, the element state
// columns, and the eight PARITH/ON slot stores.
//
// Element state is read once and written once per cycle: those accesses are
// non-temporal, which keeps them from displacing the corner-node records that
// neighbouring elements reuse.
#pragma once
#include "orad_common.hpp"

namespace orad {

struct D3 { double x, y, z; };

template <typename T> __device__ __forceinline__ T ldnt(const T *p) { return __builtin_nontemporal_load(p); }
template <typename T> __device__ __forceinline__ void stnt(T *p, T v) { __builtin_nontemporal_store(v, p); }

// Scalar state fields after the six stress columns.
enum { F_EINT, F_RHO, F_PLA, F_QVIS, F_OFFG, F_EPSD, NFIELD };
template <typename T>
__device__ __forceinline__ T *field_ptr(const BrickArgs<T> &a, int f) {
  return f == F_EINT ? a.eint : f == F_RHO ? a.rho : f == F_PLA ? a.pla
       : f == F_QVIS ? a.qvis : f == F_OFFG ? a.offg : a.epsd;
}

template <typename T>
struct BrickIO {
  const BrickArgs<T> &a;
  int e, ne;

  __device__ __forceinline__ BrickIO(const BrickArgs<T> &args, int el) : a(args), e(el), ne(args.numels) {}
  __device__ __forceinline__ int conn(int k) const { return ldnt(&a.conn[(size_t)k * ne + e]); }
  // the eight corner ranks, byte k for corner k
  __device__ __forceinline__ uint64_t ranks() const { return ldnt(&a.rank[e]); }
  __device__ __forceinline__ D3 x(int n) const { return rec(&a.X4[n]); }
  __device__ __forceinline__ D3 v(int n) const { return rec(&a.V4[n]); }
  // first PARITH/ON slot of node n (ADSKY; a small array that stays in cache)
  __device__ __forceinline__ int slot0(int n) const { return a.adsky[n]; }
  __device__ __forceinline__ T sig(int c) const { return ldnt(&a.sig[(size_t)c * ne + e]); }
  __device__ __forceinline__ void set_sig(int c, T s) const { stnt(&a.sig[(size_t)c * ne + e], s); }
  __device__ __forceinline__ T get(int f) const { return ldnt(&field_ptr(a, f)[e]); }
  __device__ __forceinline__ void put(int f, T s) const { stnt(&field_ptr(a, f)[e], s); }
  __device__ __forceinline__ double vol0() const { return ldnt(&a.vol0[e]); }
  __device__ __forceinline__ void set_sti(T s) const { stnt(&a.sti[e], s); }
  // slot store at a slot address formed beforehand (the node's first slot plus the
  // corner's rank), so no address load follows the chain
  __device__ __forceinline__ void slot_at(int sl, T fx, T fy, T fz) const {
    Slot3<T> f;
    f.x = fx; f.y = fy; f.z = fz;
    a.fsky[sl] = f;
  }

 private:
  __device__ __forceinline__ static D3 rec(const double4 *r) {
    const double *p = reinterpret_cast<const double *>(r);
    return D3{p[0], p[1], p[2]};
  }
};

}  // namespace orad

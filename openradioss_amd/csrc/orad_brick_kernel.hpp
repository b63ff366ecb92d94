// orad_brick_kernel.hpp -- element-force kernel: one thread per brick.
//
// Per element it gathers the eight corners' X and V (FP64, 32-byte records),
// runs the SFORC3 chain (orad_brick_chain.hpp), writes the element state back
// (structure of arrays, coalesced, non-temporal), STI once per element, and the
// eight PARITH/ON corner force slots at IADS (SCUMU3P).  The group's stable time step is a wavefront shuffle min,
// then one 64-bit atomicMin per workgroup on the bits of a positive double
// (order independent, so deterministic).  Hourglass energy is summed per
// workgroup in a fixed tree order into ehou_part[block].
#pragma once
#include "orad_brick_chain.hpp"

namespace orad {

constexpr int BRICK_BLOCK = 256;

// Element state is read once and written once per cycle: non-temporal accesses
// keep it from displacing the corner-node records that neighbouring elements reuse.
template <typename T> __device__ __forceinline__ T ldnt(const T *p) { return __builtin_nontemporal_load(p); }
template <typename T> __device__ __forceinline__ void stnt(T *p, T v) { __builtin_nontemporal_store(v, p); }

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_xor(v, o, 64);
    v = (w < v) ? w : v;
  }
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// SDERI3's determinant (sderi3.F:104-155) of FP64 corner coordinates: the mixed
// form's VOLDP.  Rotation-invariant, so it is taken in the global frame.
__device__ __forceinline__ double jacobian_det(const double x[8], const double y[8],
                                               const double z[8]) {
  double X17 = x[6] - x[0], X28 = x[7] - x[1], X35 = x[4] - x[2], X46 = x[5] - x[3];
  double Y17 = y[6] - y[0], Y28 = y[7] - y[1], Y35 = y[4] - y[2], Y46 = y[5] - y[3];
  double Z17 = z[6] - z[0], Z28 = z[7] - z[1], Z35 = z[4] - z[2], Z46 = z[5] - z[3];
  double J1 = X17 + X28 - X35 - X46, J2 = Y17 + Y28 - Y35 - Y46, J3 = Z17 + Z28 - Z35 - Z46;
  double a1 = X17 + X46, a2 = X28 + X35, b1 = Y17 + Y46, b2 = Y28 + Y35;
  double c1 = Z17 + Z46, c2 = Z28 + Z35;
  double J4 = a1 + a2, J5 = b1 + b2, J6 = c1 + c2, J7 = a1 - a2, J8 = b1 - b2, J9 = c1 - c2;
  return 0.015625 * (J1 * (J5 * J9 - J6 * J8) + J2 * (J6 * J7 - J4 * J9) + J3 * (J4 * J8 - J5 * J7));
}

// REL: coordinates/velocities relative to corner 1 (mixed form).
template <typename T, bool REL, int MINW>
__global__ __launch_bounds__(BRICK_BLOCK) __attribute__((amdgpu_waves_per_eu(MINW)))
void brick_forces_kernel(
    BrickArgs<T> a, int e_begin, int e_end, const double *dt1p,
    unsigned long long *dt2t_out, double *ehou_part, int xcd_remap) {
  __shared__ double s_dt[BRICK_BLOCK / 64];
  __shared__ double s_eh[BRICK_BLOCK / 64];
  const int nb = gridDim.x;
  const int b = xcd_remap ? xcd_block(blockIdx.x, nb) : (int)blockIdx.x;
  const int e = e_begin + b * BRICK_BLOCK + (int)threadIdx.x;
  const T dt1 = T(*dt1p);
  double my_dt = 1.0e300, my_eh = 0.0;
  if (e < e_end) {
    const int ne = a.numels;
    int nd[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) nd[k] = ldnt(&a.conn[k * ne + e]);
    T xd[8], yd[8], zd[8], vx[8], vy[8], vz[8];
    double voldp = 0.0;
    if (REL) {
      const double4 x0 = a.X4[nd[0]], v0 = a.V4[nd[0]];
      double rx[8], ry[8], rz[8];
      rx[0] = 0.0; ry[0] = 0.0; rz[0] = 0.0;
      vx[0] = T(0); vy[0] = T(0); vz[0] = T(0);
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        const double4 xk = a.X4[nd[k]], vk = a.V4[nd[k]];
        rx[k] = xk.x - x0.x; ry[k] = xk.y - x0.y; rz[k] = xk.z - x0.z;
        vx[k] = T(vk.x - v0.x); vy[k] = T(vk.y - v0.y); vz[k] = T(vk.z - v0.z);
      }
      voldp = jacobian_det(rx, ry, rz);
#pragma unroll
      for (int k = 0; k < 8; ++k) { xd[k] = T(rx[k]); yd[k] = T(ry[k]); zd[k] = T(rz[k]); }
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const double4 xk = a.X4[nd[k]], vk = a.V4[nd[k]];
        xd[k] = T(xk.x); yd[k] = T(xk.y); zd[k] = T(xk.z);
        vx[k] = T(vk.x); vy[k] = T(vk.y); vz[k] = T(vk.z);
      }
    }
    T sig[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) sig[c] = ldnt(&a.sig[c * ne + e]);
    T eint = ldnt(&a.eint[e]), rho = ldnt(&a.rho[e]), pla = ldnt(&a.pla[e]);
    T qvis = ldnt(&a.qvis[e]), offg = ldnt(&a.offg[e]), epsd = ldnt(&a.epsd[e]);
    const double vol0 = ldnt(&a.vol0[e]);
    const T offg_in = offg;
    BrickOut<T> o;
    brick_chain<T>(xd, yd, zd, vx, vy, vz, sig, eint, rho, vol0, voldp, pla, qvis, offg, epsd,
                   dt1, a.mat, a.prop, o);
#pragma unroll
    for (int c = 0; c < 6; ++c) stnt(&a.sig[c * ne + e], sig[c]);
    stnt(&a.eint[e], eint); stnt(&a.rho[e], rho); stnt(&a.pla[e], pla);
    stnt(&a.qvis[e], qvis); stnt(&a.epsd[e], epsd); stnt(&a.sti[e], o.sti);
    if (offg != offg_in) a.offg[e] = offg;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int slot = ldnt(&a.iads[k * ne + e]);
      Slot3<T> f;
      f.x = o.f[0][k]; f.y = o.f[1][k]; f.z = o.f[2][k];
      a.fsky[slot] = f;
    }
    if (o.mstop) atomicOr(&a.ctl->mstop, 1);
    my_dt = double(o.dtx);
    my_eh = double(o.ehou);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double wd = wave_min(my_dt);
  double we = wave_sum(my_eh);
  if (lane == 0) { s_dt[wv] = wd; s_eh[wv] = we; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double d = s_dt[0], h = s_eh[0];
#pragma unroll
    for (int w = 1; w < BRICK_BLOCK / 64; ++w) { d = (s_dt[w] < d) ? s_dt[w] : d; h += s_eh[w]; }
    if (d < 1.0e300) atomicMin(dt2t_out, (unsigned long long)__double_as_longlong(d));
    if (ehou_part) ehou_part[b] = h;
  }
}

}  // namespace orad

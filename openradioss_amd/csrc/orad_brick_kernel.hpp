// orad_brick_kernel.hpp -- element-force kernel: one thread per brick.
//
// Per element it gathers the eight corners' X and V (FP64, 32-byte records),
// runs the SFORC3 chain (orad_brick_chain.hpp), writes the element state back
// (structure of arrays, coalesced, non-temporal), STI once per element, and the
// eight PARITH/ON corner force slots (SCUMU3P): corner k of the element writes slot
// ADSKY(node) + rank_k (an 8-bit rank per corner in place of a 32-bit IADS).  The group's stable time step is a wavefront shuffle min,
// then one 64-bit atomicMin per workgroup on the bits of a positive double
// (order independent, so deterministic).  Hourglass energy is summed per
// workgroup in a fixed tree order into ehou_part[block].
#pragma once
#include <type_traits>

#include "orad_brick_chain.hpp"
#include "orad_brick_io.hpp"

namespace orad {

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

// REL: coordinates and velocities relative to corner 1 (mixed form).
// MINW: minimum waves per SIMD asked of the register allocator; BLOCK: threads per group.
template <typename T, bool REL, int MINW, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(MINW)))
void brick_forces_kernel(
    BrickArgs<T> a, int e_begin, int e_end, const double *dt1p,
    unsigned long long *dt2t_out, double *ehou_part, int xcd_remap) {
  __shared__ double s_dt[BLOCK / 64];
  __shared__ double s_eh[BLOCK / 64];
  const int nb = gridDim.x;
  const int b = xcd_remap ? xcd_block(blockIdx.x, nb) : (int)blockIdx.x;
  // t runs over [e_begin, e_end) in the library's element order, or in the engine's
  // when emap is set (one engine group, SFORC3's NFT+1..NFT+NEL)
  const int t = e_begin + b * BLOCK + (int)threadIdx.x;
  const T dt1 = T(*dt1p);
  double my_dt = 1.0e300, my_eh = 0.0;
  if (t < e_end) {
    const int e = a.emap ? a.emap[t] : t;
    const BrickIO<T> io(a, e);
    int nd[8], sl[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) nd[k] = io.conn(k);
    // slot of corner k: the node's first slot plus the corner's rank, formed with
    // the corner gathers so no address load follows the chain
    const uint64_t rk = io.ranks();
#pragma unroll
    for (int k = 0; k < 8; ++k) sl[k] = io.slot0(nd[k]) + (int)((rk >> (8 * k)) & 0xFF);
    T xd[8], yd[8], zd[8], vx[8], vy[8], vz[8];
    double voldp = 0.0;
    if (REL) {
      // coordinates first: FP64 differences, the FP64 determinant, then FP32 copies
      const D3 x0 = io.x(nd[0]);
      double rx[8], ry[8], rz[8];
      rx[0] = 0.0; ry[0] = 0.0; rz[0] = 0.0;
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        const D3 xk = io.x(nd[k]);
        rx[k] = xk.x - x0.x; ry[k] = xk.y - x0.y; rz[k] = xk.z - x0.z;
      }
      voldp = jacobian_det(rx, ry, rz);
#pragma unroll
      for (int k = 0; k < 8; ++k) { xd[k] = T(rx[k]); yd[k] = T(ry[k]); zd[k] = T(rz[k]); }
      const D3 v0 = io.v(nd[0]);
      vx[0] = T(0); vy[0] = T(0); vz[0] = T(0);
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        const D3 vk = io.v(nd[k]);
        vx[k] = T(vk.x - v0.x); vy[k] = T(vk.y - v0.y); vz[k] = T(vk.z - v0.z);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const D3 xk = io.x(nd[k]), vk = io.v(nd[k]);
        xd[k] = T(xk.x); yd[k] = T(xk.y); zd[k] = T(xk.z);
        vx[k] = T(vk.x); vy[k] = T(vk.y); vz[k] = T(vk.z);
      }
    }
    T sig[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) sig[c] = io.sig(c);
    T eint = io.get(F_EINT), rho = io.get(F_RHO), pla = io.get(F_PLA);
    T qvis = io.get(F_QVIS), offg = io.get(F_OFFG), epsd = io.get(F_EPSD);
    const double vol0 = io.vol0();
    const T offg_in = offg;
    BrickOut<T> o;
    const T jc[3] = {T(a.mat.ca), T(a.mat.cb), T(a.mat.cn)};
    brick_chain<T, T *>(xd, yd, zd, vx, vy, vz, sig, eint, rho, vol0, voldp, pla, qvis, offg, epsd,
                        dt1, a.mat, jc, a.prop, o,
                        [&](int k, T fx, T fy, T fz) { io.slot_at(sl[k], fx, fy, fz); });
#pragma unroll
    for (int c = 0; c < 6; ++c) io.set_sig(c, sig[c]);
    io.put(F_EINT, eint); io.put(F_RHO, rho); io.put(F_PLA, pla);
    io.put(F_QVIS, qvis); io.put(F_EPSD, epsd); io.set_sti(o.sti);
    if (offg != offg_in) io.put(F_OFFG, offg);
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
    for (int w = 1; w < BLOCK / 64; ++w) { d = (s_dt[w] < d) ? s_dt[w] : d; h += s_eh[w]; }
    if (d < 1.0e300) atomicMin(dt2t_out, (unsigned long long)__double_as_longlong(d));
    if (ehou_part) ehou_part[b] = h;
  }
}

}  // namespace orad

namespace orad {

// One instance of the element kernel: W = minimum waves per SIMD asked of the
// register allocator, B = threads per workgroup; `lds` = dynamic LDS per workgroup,
// which caps the workgroups resident on a CU (160 KB of LDS per CU).
template <typename T, bool REL, int W, int B>
static hipError_t launch_brick_instance(const BrickArgs<T> &a, int e_begin, int e_end,
                                        const double *dt1p, unsigned long long *dt2t,
                                        double *ehou_part, int xcd_remap, unsigned lds,
                                        hipStream_t s) {
  const int n = e_end - e_begin;
  if (n <= 0) return hipSuccess;
  const int nb = (n + B - 1) / B;
  hipLaunchKernelGGL((brick_forces_kernel<T, REL, W, B>), dim3(nb), dim3(B), lds, s, a,
                     e_begin, e_end, dt1p, dt2t, ehou_part, xcd_remap);
  return hipGetLastError();
}

// Dynamic LDS that holds a CU to `waves` resident waves per SIMD for workgroups of
// `block` threads: (waves * 256 / block) workgroups fit 160 KB, one more does not.
inline unsigned lds_for_waves(int block, int waves) {
  const int wgs = waves * 256 / block;
  return 1024u * (unsigned)(160 / (wgs + 1) + 1);
}

}  // namespace orad

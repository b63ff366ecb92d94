// orad_shell_kernel.hpp -- shell element-force kernel: one thread per shell.
//
// Per element it gathers the four corners' X, V, VR (FP64 32-byte records),
// runs the CFORC3 chain (orad_shell_chain.hpp) on the element's state columns
// (structure of arrays, coalesced, non-temporal), writes STI once per element and
// the four PARITH/ON corner slots (force and moment) at IADC (CUPDT3P).  The
// group's stable time step is a wavefront shuffle min plus one 64-bit atomicMin
// per workgroup on the bits of a positive double (order independent); hourglass
// energy is summed per workgroup in a fixed order into ehou_part[block].
#pragma once
#include "orad_brick_io.hpp"
#include "orad_shell_chain.hpp"

namespace orad {

__device__ __forceinline__ double shell_wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_xor(v, o, 64);
    v = (w < v) ? w : v;
  }
  return v;
}
__device__ __forceinline__ double shell_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename T>
struct ShellIO {
  T *st;
  int e, ne;
  __device__ __forceinline__ T get(int c) const { return ldnt(&st[(size_t)c * ne + e]); }
  __device__ __forceinline__ void put(int c, T v) const { stnt(&st[(size_t)c * ne + e], v); }
};

// Gather of one element's corners in the kinematics precision TG (REL: relative
// to corner 1, formed in FP64).
template <typename T, typename TG, bool REL>
__device__ __forceinline__ void shell_gather(const ShellArgs<T> &a, const int nd[4], TG xg[3][4],
                                             TG vg[3][4], TG vrg[3][4]) {
  const double4 x0 = a.X4[nd[0]], v0 = a.V4[nd[0]];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double4 xc = a.X4[nd[c]], vc = a.V4[nd[c]], wc = a.VR4[nd[c]];
    if (REL) {
      xg[0][c] = TG(xc.x - x0.x); xg[1][c] = TG(xc.y - x0.y); xg[2][c] = TG(xc.z - x0.z);
      vg[0][c] = TG(vc.x - v0.x); vg[1][c] = TG(vc.y - v0.y); vg[2][c] = TG(vc.z - v0.z);
    } else {
      xg[0][c] = TG(xc.x); xg[1][c] = TG(xc.y); xg[2][c] = TG(xc.z);
      vg[0][c] = TG(vc.x); vg[1][c] = TG(vc.y); vg[2][c] = TG(vc.z);
    }
    vrg[0][c] = TG(wc.x); vrg[1][c] = TG(wc.y); vrg[2][c] = TG(wc.z);
  }
}

// REL: coordinates and translational velocities relative to corner 1 (mixed form).
// TG: precision of the frame and kinematics stage (shell_kinematics).
template <typename T, typename TG, bool REL, int BLOCK>
__global__ __launch_bounds__(BLOCK) void shell_forces_kernel(ShellArgs<T> a, int e_begin, int e_end,
                                                             const double *dt1p,
                                                             unsigned long long *dt2t_out,
                                                             double *ehou_part, int xcd_remap) {
  __shared__ double s_dt[BLOCK / 64];
  __shared__ double s_eh[BLOCK / 64];
  const int nb = gridDim.x;
  const int b = xcd_remap ? xcd_block(blockIdx.x, nb) : (int)blockIdx.x;
  const int e = e_begin + b * BLOCK + (int)threadIdx.x;
  const double dt1 = *dt1p;
  double my_dt = 1.0e300, my_eh = 0.0;
  if (e < e_end) {
    const int ne = a.numelc;
    // slot addresses loaded with the connectivity, so their latency overlaps the
    // corner gathers instead of following the chain
    int nd[4], sl[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) nd[c] = ldnt(&a.conn[(size_t)c * ne + e]);
#pragma unroll
    for (int c = 0; c < 4; ++c) sl[c] = ldnt(&a.iadc[(size_t)c * ne + e]);
    TG xg[3][4], vg[3][4], vrg[3][4];
    shell_gather<T, TG, REL>(a, nd, xg, vg, vrg);
    const ShellIO<T> io{a.st, e, ne};
    const T thk0 = T(ldnt(&a.thke[e]));
    ShellOut<T> o;
    Slot6<T> *fsky = a.fsky;
    const T jc[3] = {T(a.mat.ca), T(a.mat.cb), T(a.mat.cn)};
    shell_chain<T, TG>(xg, vg, vrg, thk0, nd[2] == nd[3], io, dt1, a.mat, jc, a.prop, a.k, o,
                   [&](int c, T fx, T fy, T fz, T mx, T my, T mz) {
                     Slot6<T> f;
                     f.fx = fx; f.fy = fy; f.fz = fz; f.mx = mx; f.my = my; f.mz = mz;
                     fsky[sl[c]] = f;
                   });
    stnt(&a.sti[e], o.sti);
    if (o.alive) my_dt = double(o.dt);
    my_eh = double(o.ehou);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double wd = shell_wave_min(my_dt);
  const double we = shell_wave_sum(my_eh);
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

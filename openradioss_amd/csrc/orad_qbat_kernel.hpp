// orad_qbat_kernel.hpp -- QBAT shell element-force kernel: one thread per shell.
//
// Per element it gathers the four corners' X, V, VR (FP64 32-byte records), runs
// the CBAFORC3 chain (orad_qbat_chain.hpp) on the element's state columns
// (structure of arrays, coalesced, non-temporal: 190 columns, the per Gauss point
// and layer stresses among them), writes the two diagonal STI values and the four
// PARITH/ON corner slots (force and moment) at IADC (CUPDTN3P).  The group's
// stable time step and the damping energy are reduced as in the other shell
// kernels: a wavefront shuffle minimum and one 64-bit atomicMin per workgroup, and
// a fixed-order per-workgroup energy sum.
#pragma once
#include "orad_qbat_chain.hpp"
#include "orad_shell_kernel.hpp"

namespace orad {

template <typename T, typename TG, bool REL, int BLOCK>
__global__ __launch_bounds__(BLOCK) void qbat_forces_kernel(ShellArgs<T> a, int e_begin, int e_end,
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
    int nd[4], sl[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) nd[c] = ldnt(&a.conn[(size_t)c * ne + e]);
#pragma unroll
    for (int c = 0; c < 4; ++c) sl[c] = ldnt(&a.iadc[(size_t)c * ne + e]);
    TG xg[3][4], vg[3][4], vrg[3][4];
    shell_gather<T, TG, REL>(a, nd, xg, vg, vrg);
    const ShellIO<T> io{a.st, e, ne};
    const T thk0 = T(ldnt(&a.thke[e]));
    QbatOut<T> o;
    Slot6<T> *fsky = a.fsky;
    const T jc[3] = {T(a.mat.ca), T(a.mat.cb), T(a.mat.cn)};
    qbat_chain<T, TG>(xg, vg, vrg, thk0, io, dt1, a.mat, jc, a.prop, a.k, o,
                      [&](int c, T fx, T fy, T fz, T mx, T my, T mz) {
                        Slot6<T> f;
                        f.fx = fx; f.fy = fy; f.fz = fz; f.mx = mx; f.my = my; f.mz = mz;
                        fsky[sl[c]] = f;
                      });
    stnt(&a.sti[e], o.sti13);
    stnt(&a.sti[(size_t)ne + e], o.sti24);
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

// orad_batch_kernel.hpp -- element-force kernel over batched independent shell
// models (BatchArgs), for either shell formulation: QEPH = false runs the CFORC3
// chain (Belytschko-Tsay), QEPH = true the CZFORC3 chain (Ishell=24).
#pragma once
#include "orad_qeph_chain.hpp"
#include "orad_shell_kernel.hpp"

namespace orad {

// One 256-element chunk per tile (BatchArgs); a tile loop over
// several chunks made the compiler keep the chain's invariants live across
// iterations (256 registers and one wave per SIMD instead of 225 and two).
// PERSIST: a grid sized
// to the resident capacity takes tiles from an atomic counter until none is left
// (the reference's dynamic group schedule, forint.F:325); otherwise one workgroup
// per tile.  Per tile: each element's CFORC3 chain with its model's material and
// DT1, slots written at IADC as in the one-model kernel, the model's DT2T by an
// atomic minimum, and the chunk's hourglass energy as four 64-lane wave sums added
// left to right (the nodal kernel adds each model's chunks to its total).  No
// workgroup waits for or reads another's results within the launch, so no
// device-scope fence is needed (a fence per tile, which writes back the XCD's L2,
// made the kernel three times slower).
template <typename T, typename TG, bool REL, int BLOCK, bool PERSIST, bool QEPH>
__global__ __launch_bounds__(BLOCK) void shell_batch_kernel(ShellArgs<T> a, BatchArgs bt,
                                                            int parity, int xcd_remap) {
  static_assert(BLOCK == 64 || BLOCK == 128 || BLOCK == 256, "BLOCK divides a 256-element chunk");
  constexpr int NW = BLOCK / 64, SUB = 256 / BLOCK;
  __shared__ double s_dt[NW], s_eh[NW];
  __shared__ int s_tile;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ne = a.numelc;
  int pass = 0;
  for (;;) {
    int t;
    if (PERSIST) {
      if (threadIdx.x == 0) s_tile = (int)atomicAdd(bt.queue, 1u);
      __syncthreads();
      t = s_tile;
      __syncthreads();
    } else {
      if (pass > 0) break;
      t = xcd_remap ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    }
    ++pass;
    if (t >= bt.ntiles) break;
    t = __builtin_amdgcn_readfirstlane(t);
    const int4 td = bt.tiles[t];
    const int m = td.x;
    const ShMat &mat = bt.mats[m];
    CycleCtl *ctl = bt.ctl + m;
    const double dt1 = ctl->s[parity].dt2;
    double my_dt = 1.0e300, ch_eh = 0.0;
    const int chunk = td.w;
    {
      const int c0 = td.y;
#pragma unroll 1
      for (int sub = 0; sub < SUB; ++sub) {
        const int e = c0 + sub * BLOCK + (int)threadIdx.x;
        double my_eh = 0.0;
        if (e < td.z) {
          int nd[4], sl[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) nd[c] = ldnt(&a.conn[(size_t)c * ne + e]);
#pragma unroll
          for (int c = 0; c < 4; ++c) sl[c] = ldnt(&a.iadc[(size_t)c * ne + e]);
          TG xg[3][4], vg[3][4], vrg[3][4];
          shell_gather<T, TG, REL>(a, nd, xg, vg, vrg);
          const ShellIO<T> io{a.st, e, ne};
          const T thk0 = T(ldnt(&a.thke[e]));
          Slot6<T> *fsky = a.fsky;
          const T jc[3] = {T(mat.ca), T(mat.cb), T(mat.cn)};
          auto emit = [&](int c, T fx, T fy, T fz, T mx, T my, T mz) {
            Slot6<T> f;
            f.fx = fx; f.fy = fy; f.fz = fz; f.mx = mx; f.my = my; f.mz = mz;
            fsky[sl[c]] = f;
          };
          if constexpr (QEPH) {
            QephOut<T> o;
            qeph_chain<T, TG>(xg, vg, vrg, thk0, io, dt1, mat, jc, a.prop, a.k, o, emit);
            stnt(&a.sti[e], o.sti13);
            stnt(&a.sti[(size_t)ne + e], o.sti24);
            if (o.alive) my_dt = (double(o.dt) < my_dt) ? double(o.dt) : my_dt;
            my_eh = double(o.ehou);
          } else {
            ShellOut<T> o;
            shell_chain<T, TG>(xg, vg, vrg, thk0, nd[2] == nd[3], io, dt1, mat, jc, a.prop,
                               a.k, o, emit);
            stnt(&a.sti[e], o.sti);
            if (o.alive) my_dt = (double(o.dt) < my_dt) ? double(o.dt) : my_dt;
            my_eh = double(o.ehou);
          }
        }
        const double we = shell_wave_sum(my_eh);   // identical in every lane
        if (NW == 1) {
          ch_eh = (sub == 0) ? we : ch_eh + we;
        } else {
          if (lane == 0) s_eh[wv] = we;
          __syncthreads();
          if (threadIdx.x == 0) {
            double h = (sub == 0) ? s_eh[0] : ch_eh + s_eh[0];
#pragma unroll
            for (int w = 1; w < NW; ++w) h += s_eh[w];
            ch_eh = h;
          }
          __syncthreads();
        }
      }
      if (threadIdx.x == 0) bt.chunk_ehou[chunk] = ch_eh;
    }
    const double wd = shell_wave_min(my_dt);
    if (lane == 0) s_dt[wv] = wd;
    __syncthreads();
    if (threadIdx.x == 0) {
      double d = s_dt[0];
#pragma unroll
      for (int w = 1; w < NW; ++w) d = (s_dt[w] < d) ? s_dt[w] : d;
      if (d < 1.0e300) atomicMin(&ctl->dt2t[parity], (unsigned long long)__double_as_longlong(d));
    }
    __syncthreads();
  }
}

}  // namespace orad

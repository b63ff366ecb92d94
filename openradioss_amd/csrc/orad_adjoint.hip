// orad_adjoint.hip -- reverse (adjoint) sweep of the explicit shell cycle, FP64,
// built without FMA contraction.
//
// One forward cycle k maps (nodes x_k, v_k, vr_k, d_k; element state s_k) to
// cycle k+1: the element kernel turns (x_k, v_k, vr_k, s_k, thickness, A, B, n)
// into PARITH/ON slots and s_{k+1}; the nodal kernel sums the slots (S, R per
// node), divides by MS / IN, applies BCS10 / FIXVEL and integrates
// v_{k+1} = v_k + DT12 a, x_{k+1} = x_k + DT2 v_{k+1}, d_{k+1} = d_k + DT2 v_{k+1}.
// The time-step sequence is taken as given (exact when DTMX pins it).
//
// Reverse of cycle k, three launches:
//   adj_node_kernel    adjoints of v_{k+1}, x, d -> adjoint of S and R per node,
//                      of v_k through the update, and the MS / IN sensitivities;
//   adj_elem_rev_kernel  Belytschko-Tsay or QEPH with radial return (IPLA = 0): the
//                      hand-written reverse pass of the CFORC3 chain
//                      (orad_shell_adjoint.hpp) or of the CZFORC3 chain
//                      (orad_qeph_adjoint.hpp), one thread per element;
//   adj_elem_kernel    otherwise, vector-Jacobian product of the element chain: the CFORC3
//                      chain (orad_shell_chain.hpp), or for QEPH the CZFORC3 chain
//                      (orad_qeph_chain.hpp), instantiated on Dual<K>, seeds K of the
//                      element's inputs per thread (4 corners x (x, v, vr), 25
//                      stresses, 5 plastic strains, the incremental hourglass forces
//                      -- 3 for BT, 12 for QEPH --, thickness, A, B, n: 73 or 82;
//                      K = 1 for the corner inputs, 2 for the others) and
//                      contracts the tangents of its outputs (slots, new state) with
//                      their adjoints;
//   adj_gather_kernel  per node, the corner contributions in slot order (the
//                      PARITH/ON CSR), so the sums are deterministic; within a
//                      checkpoint segment fused with the next (earlier) cycle's
//                      adj_node_kernel (adj_gather_node_kernel).
// The recompute of a segment stashes each cycle's inputs and the previous cycle's
// accelerations and scalars with one copy launch per cycle (adj_stash_kernel).
#include "orad_dual.hpp"
#include "orad_qeph_chain.hpp"
#include "orad_shell_adjoint.hpp"
#include "orad_qeph_adjoint.hpp"
#include "orad_brick_adjoint.hpp"
#include "orad_brick_io.hpp"
#include "orad_adjoint_args.hpp"

#include <algorithm>
#include <type_traits>

namespace orad {

// Workgroup size of the adjoint kernels; 64, 128 and 256 measured within 2% of each
// other on the 128^2 plate's reverse sweep (633 / 635 / 646 ms).
#ifndef ORAD_ADJ_BLOCK
#define ORAD_ADJ_BLOCK 64
#endif
constexpr int ADJ_BLOCK = ORAD_ADJ_BLOCK;
// Tangents per thread, for the chunks of corner inputs and of the other inputs.
// Measured on MI355X (128^2 plate, 2000 cycles, reverse sweep BT / QEPH): 1 and 2
// give 632 / 1164 ms, 1 and 3 663 / 1177, 1 and 4 712 / 1325, 2 and 2 701 / 1350,
// 4 and 4 1071 / 2372 ms -- wider duals spill (up to 3 KB per lane at 4), which costs
// more than the extra primal evaluations of narrow ones.
#ifndef ORAD_ADJ_K_KIN
#define ORAD_ADJ_K_KIN 1
#endif
#ifndef ORAD_ADJ_K_ST
#define ORAD_ADJ_K_ST 2
#endif
constexpr int ADJ_K_KIN = ORAD_ADJ_K_KIN, ADJ_K_ST = ORAD_ADJ_K_ST;

template <typename S, int K, bool QEPH>
struct AdjIO {
  const S *st;
  int e, ne, i0;
  const double *lst_in;
  double *acc;
  __device__ __forceinline__ Dual<K> get(int c) const {
    Dual<K> r{double(st[(size_t)c * ne + e])};
    const int j = adj_state_index<QEPH>(c);
    if (j >= 0) {
      const int q = 36 + j - i0;
#pragma unroll
      for (int t = 0; t < K; ++t) r.d[t] = (q == t) ? 1.0 : 0.0;
    }
    return r;
  }
  __device__ __forceinline__ void put(int c, const Dual<K> &v) const {
    const int j = adj_state_index<QEPH>(c);
    if (j < 0) return;
    const double l = lst_in[(size_t)j * ne + e];
#pragma unroll
    for (int t = 0; t < K; ++t) acc[t] += l * v.d[t];
  }
};

// KIN: the chunks seed corner inputs (i < 36), so the kinematics carry tangents; the
// chunks of state, thickness and material inputs run the kinematics in plain FP64
// (TG = double), whose values are the same and whose tangents would all be zero.
template <typename S, int K, bool QEPH, bool KIN>
__global__ __launch_bounds__(ADJ_BLOCK) void adj_elem_kernel(AdjArgs<S> a, int base,
                                                             int nchunk) {
  typedef Dual<K> D;
  typedef typename std::conditional<KIN, D, double>::type G;
  typedef AdjShape<QEPH> SH;
  const int ne = a.numelc;
  const long idx = (long)blockIdx.x * ADJ_BLOCK + threadIdx.x;
  if (idx >= (long)ne * nchunk) return;
  const int cl = (int)(idx / ne), e = (int)(idx - (long)cl * ne);
  const int i0 = base + cl * K;     // first input of this thread's chunk
  int nd[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) nd[c] = a.conn[(size_t)c * ne + e];
  G xg[3][4], vg[3][4], vrg[3][4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double4 x = a.X4[nd[c]], v = a.V4[nd[c]], w = a.VR4[nd[c]];
    const double xv[3] = {x.x, x.y, x.z}, vv[3] = {v.x, v.y, v.z}, wv[3] = {w.x, w.y, w.z};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      xg[i][c] = G(xv[i]);
      vg[i][c] = G(vv[i]);
      vrg[i][c] = G(wv[i]);
      if constexpr (KIN) {
#pragma unroll
        for (int t = 0; t < K; ++t) {
          xg[i][c].d[t] = (9 * c + i - i0 == t) ? 1.0 : 0.0;
          vg[i][c].d[t] = (9 * c + 3 + i - i0 == t) ? 1.0 : 0.0;
          vrg[i][c].d[t] = (9 * c + 6 + i - i0 == t) ? 1.0 : 0.0;
        }
      }
    }
  }
  D thk0(a.thke[e]);
  D jc[3] = {D(a.mat.ca), D(a.mat.cb), D(a.mat.cn)};
#pragma unroll
  for (int t = 0; t < K; ++t) {
    thk0.d[t] = (SH::I_THK - i0 == t) ? 1.0 : 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) jc[j].d[t] = (SH::I_JC + j - i0 == t) ? 1.0 : 0.0;
  }
  double acc[K];
#pragma unroll
  for (int t = 0; t < K; ++t) acc[t] = 0.0;
  const AdjIO<S, K, QEPH> io{a.st, e, ne, i0, a.lst_in, acc};
  const double *lamS = a.lamS, *lamR = a.lamR;
  auto emit = [&](int c, const D &fx, const D &fy, const D &fz, const D &mx, const D &my,
                  const D &mz) {
    const size_t n3 = 3 * (size_t)nd[c];
    const double l0 = lamS[n3], l1 = lamS[n3 + 1], l2 = lamS[n3 + 2];
    const double r0 = lamR[n3], r1 = lamR[n3 + 1], r2 = lamR[n3 + 2];
#pragma unroll
    for (int t = 0; t < K; ++t)
      acc[t] += l0 * fx.d[t] + l1 * fy.d[t] + l2 * fz.d[t] + r0 * mx.d[t] + r1 * my.d[t] +
                r2 * mz.d[t];
  };
  if constexpr (QEPH) {
    QephOut<D> o;
    qeph_chain<D, G>(xg, vg, vrg, thk0, io, a.dt1, a.mat, jc, a.prop, a.k, o, emit);
  } else {
    ShellOut<D> o;
    shell_chain<D, G>(xg, vg, vrg, thk0, nd[2] == nd[3], io, a.dt1, a.mat, jc, a.prop, a.k, o,
                      emit);
  }
#pragma unroll
  for (int t = 0; t < K; ++t) {
    const int i = i0 + t;
    if (i < 36) a.lcorner[(size_t)i * ne + e] = acc[t];
    else if (i < SH::I_THK) a.lst_out[(size_t)(i - 36) * ne + e] = acc[t];
    else if (i == SH::I_THK) a.g_thk[e] += acc[t];
    else if (i < SH::NIN) a.g_jc[(size_t)(i - SH::I_JC) * ne + e] += acc[t];
  }
}

// One Belytschko-Tsay or QEPH element per thread through the hand-written reverse
// pass: the same outputs as the dual chunks (lcorner and lst_out set, g_thk and g_jc
// added to).
template <typename S, bool QEPH>
__global__ __launch_bounds__(ADJ_BLOCK) void adj_elem_rev_kernel(AdjArgs<S> a) {
  const int ne = a.numelc;
  const int e = blockIdx.x * ADJ_BLOCK + threadIdx.x;
  if (e >= ne) return;
  int nd[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) nd[c] = a.conn[(size_t)c * ne + e];
  double xg[3][4], vg[3][4], vrg[3][4];
  double lx[3][4], lv[3][4], lvr[3][4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double4 x = a.X4[nd[c]], v = a.V4[nd[c]], w = a.VR4[nd[c]];
    xg[0][c] = x.x; xg[1][c] = x.y; xg[2][c] = x.z;
    vg[0][c] = v.x; vg[1][c] = v.y; vg[2][c] = v.z;
    vrg[0][c] = w.x; vrg[1][c] = w.y; vrg[2][c] = w.z;
#pragma unroll
    for (int i = 0; i < 3; ++i) lx[i][c] = lv[i][c] = lvr[i][c] = 0.0;
  }
  const double jc[3] = {a.mat.ca, a.mat.cb, a.mat.cn};
  double gthk = 0.0, gjc[3] = {0.0, 0.0, 0.0};
  const S *st = a.st;
  const double *lin = a.lst_in, *lamS = a.lamS, *lamR = a.lamR;
  double *lout = a.lst_out;
  auto stv = [&](int col) { return double(st[(size_t)col * ne + e]); };
  auto lpost = [&](int j) { return lin[(size_t)j * ne + e]; };
  auto lslot = [&](int c, int i) {
    return i < 3 ? lamS[3 * (size_t)nd[c] + i] : lamR[3 * (size_t)nd[c] + i - 3];
  };
  auto lpre = [&](int j, double v) { lout[(size_t)j * ne + e] = v; };
  if constexpr (QEPH)
    qeph_vjp(xg, vg, vrg, a.thke[e], stv, a.dt1, a.mat, jc, a.prop, a.k, lpost, lslot, lpre, lx, lv,
             lvr, gthk, gjc);
  else
    shell_vjp_bt(xg, vg, vrg, a.thke[e], nd[2] == nd[3], stv, a.dt1, a.mat, jc, a.prop, a.k,
                 lpost, lslot, lpre, lx, lv, lvr, gthk, gjc);
  for (int it = a.prop.npt; it < 5; ++it) {     // points not in use carry no adjoint
#pragma unroll
    for (int c = 0; c < 5; ++c) lout[(size_t)(5 * it + c) * ne + e] = 0.0;
    lout[(size_t)(25 + it) * ne + e] = 0.0;
  }
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      a.lcorner[(size_t)(9 * c + i) * ne + e] = lx[i][c];
      a.lcorner[(size_t)(9 * c + 3 + i) * ne + e] = lv[i][c];
      a.lcorner[(size_t)(9 * c + 6 + i) * ne + e] = lvr[i][c];
    }
  a.g_thk[e] += gthk;
#pragma unroll
  for (int j = 0; j < 3; ++j) a.g_jc[(size_t)j * ne + e] += gjc[j];
}

// Bricks, state columns: SIG(6), then EINT, RHO, PLA, QVIS, OFF, EPSD.
constexpr int BC_EINT = 6, BC_RHO = 7, BC_PLA = 8, BC_QVIS = 9, BC_OFFG = 10, BC_EPSD = 11;

// One brick per thread through the hand-written reverse pass (orad_brick_adjoint.hpp).
template <typename S>
__global__ __launch_bounds__(ADJ_BLOCK) void adj_brick_rev_kernel(AdjBrickArgs<S> a) {
  const int ne = a.numels;
  const int e = blockIdx.x * ADJ_BLOCK + threadIdx.x;
  if (e >= ne) return;
  int nd[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) nd[k] = a.conn[(size_t)k * ne + e];
  const double4 *X4 = a.X4, *V4 = a.V4;
  const S *st = a.st;
  auto col = [&](int c) { return double(st[(size_t)c * ne + e]); };
  double sig[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) sig[i] = col(i);
  const double jc[3] = {a.mat.ca, a.mat.cb, a.mat.cn};
  double gjc[3] = {0.0, 0.0, 0.0};
  const double *lin = a.lst_in, *lamS = a.lamS;
  double *lout = a.lst_out;
  brick_vjp([&](int k, double *xk, double *vk) {
              const double4 p = X4[nd[k]], q = V4[nd[k]];
              xk[0] = p.x; xk[1] = p.y; xk[2] = p.z;
              vk[0] = q.x; vk[1] = q.y; vk[2] = q.z;
            },
            sig, col(BC_PLA), col(BC_RHO), col(BC_OFFG), a.vol0[e], a.dt1, a.mat, jc,
            a.prop, [&](int j) { return lin[(size_t)j * ne + e]; },
            [&](int k, int i) { return lamS[3 * (size_t)nd[k] + i]; },
            [&](int j, double val) { lout[(size_t)j * ne + e] = val; },
            [&](int k, const double *lxk, const double *lvk) {
#pragma unroll
              for (int d = 0; d < 3; ++d) {
                a.lcorner[(size_t)(6 * k + d) * ne + e] = lxk[d];
                a.lcorner[(size_t)(6 * k + 3 + d) * ne + e] = lvk[d];
              }
            },
            gjc);
#pragma unroll
  for (int j = 0; j < 3; ++j) a.g_jc[(size_t)j * ne + e] += gjc[j];
}

// The same VJP by forward-mode dual numbers, K inputs per thread: the brick chain
// (FP64 form) instantiated on Dual<K>.
template <typename S, int K>
__global__ __launch_bounds__(ADJ_BLOCK) void adj_brick_dual_kernel(AdjBrickArgs<S> a,
                                                                   int nchunk) {
  typedef Dual<K> D;
  const int ne = a.numels;
  const long idx = (long)blockIdx.x * ADJ_BLOCK + threadIdx.x;
  if (idx >= (long)ne * nchunk) return;
  const int cl = (int)(idx / ne), e = (int)(idx - (long)cl * ne);
  const int i0 = cl * K;
  auto seed = [&](D &u, int i) {
#pragma unroll
    for (int t = 0; t < K; ++t) u.d[t] = (i - i0 == t) ? 1.0 : 0.0;
  };
  int nd[8];
  D xd[8], yd[8], zd[8], vx[8], vy[8], vz[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    nd[k] = a.conn[(size_t)k * ne + e];
    const double4 p = a.X4[nd[k]], q = a.V4[nd[k]];
    xd[k] = D(p.x); yd[k] = D(p.y); zd[k] = D(p.z);
    vx[k] = D(q.x); vy[k] = D(q.y); vz[k] = D(q.z);
    seed(xd[k], 6 * k); seed(yd[k], 6 * k + 1); seed(zd[k], 6 * k + 2);
    seed(vx[k], 6 * k + 3); seed(vy[k], 6 * k + 4); seed(vz[k], 6 * k + 5);
  }
  const S *st = a.st;
  auto col = [&](int c) { return D(double(st[(size_t)c * ne + e])); };
  D sig[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) { sig[i] = col(i); seed(sig[i], 48 + i); }
  D pla = col(BC_PLA);
  seed(pla, 54);
  D jc[3] = {D(a.mat.ca), D(a.mat.cb), D(a.mat.cn)};
#pragma unroll
  for (int j = 0; j < 3; ++j) seed(jc[j], ADJB_I_JC + j);
  D eint = col(BC_EINT), rho = col(BC_RHO), qvis = col(BC_QVIS), offg = col(BC_OFFG);
  D epsd = col(BC_EPSD);
  double acc[K];
#pragma unroll
  for (int t = 0; t < K; ++t) acc[t] = 0.0;
  const double *lamS = a.lamS;
  BrickOut<D> o;
  brick_chain<D, D *>(xd, yd, zd, vx, vy, vz, sig, eint, rho, a.vol0[e], 0.0, pla, qvis, offg,
                      epsd, D(a.dt1), a.mat, jc, a.prop, o,
                      [&](int k, const D &fx, const D &fy, const D &fz) {
                        const size_t n3 = 3 * (size_t)nd[k];
#pragma unroll
                        for (int t = 0; t < K; ++t)
                          acc[t] += lamS[n3] * fx.d[t] + lamS[n3 + 1] * fy.d[t] +
                                    lamS[n3 + 2] * fz.d[t];
                      });
#pragma unroll
  for (int t = 0; t < K; ++t) {
    double s2 = a.lst_in[(size_t)6 * ne + e] * pla.d[t];
#pragma unroll
    for (int i = 0; i < 6; ++i) s2 += a.lst_in[(size_t)i * ne + e] * sig[i].d[t];
    acc[t] += s2;
    const int i = i0 + t;
    if (i < 48) a.lcorner[(size_t)i * ne + e] = acc[t];
    else if (i < ADJB_I_JC) a.lst_out[(size_t)(i - 48) * ne + e] = acc[t];
    else if (i < ADJB_NIN) a.g_jc[(size_t)(i - ADJB_I_JC) * ne + e] += acc[t];
  }
}

template <typename S>
static void launch_adj_brick_form(bool rev, const AdjBrickArgs<S> &a, hipStream_t s) {
  if (rev) {
    const int nb = (a.numels + ADJ_BLOCK - 1) / ADJ_BLOCK;
    hipLaunchKernelGGL(adj_brick_rev_kernel<S>, dim3(nb), dim3(ADJ_BLOCK), 0, s, a);
  } else {
    const int nchunk = (ADJB_NIN + 1) / 2;
    const long n = (long)a.numels * nchunk;
    const int nb = (int)((n + ADJ_BLOCK - 1) / ADJ_BLOCK);
    hipLaunchKernelGGL((adj_brick_dual_kernel<S, 2>), dim3(nb), dim3(ADJ_BLOCK), 0, s, a,
                       nchunk);
  }
}

// Reverse of the nodal update of one cycle for node n.  In: lx, lv, ld, lvr = adjoints
// of x, v, d, vr after the cycle; A, AR = the cycle's accelerations (after BCS10).
// Out: lv, lvr = adjoints of v_k, vr_k through the update (element terms are added
// by the gather), lamS / lamR = adjoints of the slot sums, MS / IN sensitivities
// accumulated into g_ms / g_in.  Without rotational DOF (bricks) VR4 is null and
// lvr, lamR, g_in are not touched.
struct AdjNodeArgs {
  const double4 *X4, *V4, *VR4;
  const double *A, *AR;
  double dt2, dt12;
};

__device__ __forceinline__ void adj_node_one(int n, const AdjNodeArgs &q, const double *lx,
                                             double *lv, const double *ld, double *lvr,
                                             double *lamS, double *lamR, double *g_ms,
                                             double *g_in) {
  const int code = word_code(q.V4[n].w);
  const int ct = code_icodt(code), cr = code_icodr(code), fd = code_fvdir(code);
  const bool rot = q.VR4 != nullptr;
  const double ms = q.X4[n].w, in = rot ? q.VR4[n].w : 0.0;
  const double dt2 = q.dt2, dt12 = q.dt12;
  double gm = 0.0, gi = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const size_t j = 3 * (size_t)n + i;
    const double u = lv[j] + dt2 * (lx[j] + ld[j]);
    const bool fixed = (ct & (4 >> i)) != 0 || fd == i + 1;
    double ls = 0.0, lvn = 0.0;
    if (!fixed) {
      lvn = u;
      if (ms > 0.0) {
        ls = dt12 * u / ms;
        gm = gm - dt12 * u * q.A[j] / ms;
      }
    }
    lv[j] = lvn;
    lamS[j] = ls;
    if (!rot) continue;
    const double ur = lvr[j];
    const bool rfixed = (cr & (4 >> i)) != 0;
    double lr = 0.0, lvrn = 0.0;
    if (!rfixed) {
      lvrn = ur;
      if (in > 0.0) {
        lr = dt12 * ur / in;
        gi = gi - dt12 * ur * q.AR[j] / in;
      }
    }
    lvr[j] = lvrn;
    lamR[j] = lr;
  }
  g_ms[n] += gm;
  if (rot) g_in[n] += gi;
}

__global__ __launch_bounds__(256) void adj_node_kernel(int numnod, AdjNodeArgs q,
                                                       const double *lx, double *lv,
                                                       const double *ld, double *lvr,
                                                       double *lamS, double *lamR, double *g_ms,
                                                       double *g_in) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= numnod) return;
  adj_node_one(n, q, lx, lv, ld, lvr, lamS, lamR, g_ms, g_in);
}

// Element corner contributions to the adjoints of node n, in slot order.  `per`
// inputs per corner: 9 for shells (x, v, vr), 6 for bricks (x, v; lvr untouched).
__device__ __forceinline__ void adj_gather_one(int n, int numelc, int per, const int32_t *adsky,
                                               const int32_t *slot_elem,
                                               const int32_t *slot_corner,
                                               const double *lcorner, double *lx, double *lv,
                                               double *lvr) {
  double ax[3] = {0, 0, 0}, av[3] = {0, 0, 0}, ar[3] = {0, 0, 0};
  const bool rot = per == 9;
  for (int s = adsky[n]; s < adsky[n + 1]; ++s) {
    const int e = slot_elem[s], c = slot_corner[s];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      ax[i] += lcorner[(size_t)(per * c + i) * numelc + e];
      av[i] += lcorner[(size_t)(per * c + 3 + i) * numelc + e];
      if (rot) ar[i] += lcorner[(size_t)(per * c + 6 + i) * numelc + e];
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    lx[3 * (size_t)n + i] += ax[i];
    lv[3 * (size_t)n + i] += av[i];
    if (rot) lvr[3 * (size_t)n + i] += ar[i];
  }
}

__global__ __launch_bounds__(256) void adj_gather_kernel(int numnod, int numelc, int per,
                                                         const int32_t *adsky,
                                                         const int32_t *slot_elem,
                                                         const int32_t *slot_corner,
                                                         const double *lcorner, double *lx,
                                                         double *lv, double *lvr) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= numnod) return;
  adj_gather_one(n, numelc, per, adsky, slot_elem, slot_corner, lcorner, lx, lv, lvr);
}

// The gather of cycle k followed, node by node, by the nodal reverse of cycle k-1
// (both touch only node n's entries, so one thread does both in order).
__global__ __launch_bounds__(256) void adj_gather_node_kernel(
    int numnod, int numelc, int per, const int32_t *adsky, const int32_t *slot_elem,
    const int32_t *slot_corner, const double *lcorner, AdjNodeArgs q, double *lx, double *lv,
    const double *ld, double *lvr, double *lamS, double *lamR, double *g_ms, double *g_in) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= numnod) return;
  adj_gather_one(n, numelc, per, adsky, slot_elem, slot_corner, lcorner, lx, lv, lvr);
  adj_node_one(n, q, lx, lv, ld, lvr, lamS, lamR, g_ms, g_in);
}

// Copies of up to 7 device ranges in one launch (blockIdx.y = range): the inputs of
// a recomputed cycle and the accelerations and scalars of the one before it.  A range
// whose ends are 16-byte aligned moves in 16-byte words, others in 4-byte words.
struct StashJob {
  const void *src[7];
  void *dst[7];
  unsigned long long bytes[7];
};

__global__ __launch_bounds__(256) void adj_stash_kernel(StashJob j) {
  const int r = blockIdx.y;
  const unsigned long long nb = j.bytes[r];
  const char *src = (const char *)j.src[r];
  char *dst = (char *)j.dst[r];
  const size_t stride = (size_t)gridDim.x * 256;
  const bool wide = (((uintptr_t)src | (uintptr_t)dst | nb) & 15) == 0;
  if (wide) {
    const uint4 *s4 = (const uint4 *)src;
    uint4 *d4 = (uint4 *)dst;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < nb / 16; i += stride) d4[i] = s4[i];
  } else {
    const uint32_t *s1 = (const uint32_t *)src;
    uint32_t *d1 = (uint32_t *)dst;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < nb / 4; i += stride) d1[i] = s1[i];
  }
}

static_assert(36 % ADJ_K_KIN == 0, "corner inputs must fill whole chunks");

// One launch for the chunks of corner inputs 0..35, one for the inputs from 36 on
// (state, thickness, material).
template <typename S, bool QEPH>
static void launch_adj_elem_chunks(const AdjArgs<S> &a, hipStream_t s) {
  const int nkin = 36 / ADJ_K_KIN;
  const int nst = (AdjShape<QEPH>::NIN - 36 + ADJ_K_ST - 1) / ADJ_K_ST;
  const long ne = a.numelc;
  const int nb0 = (int)((ne * nkin + ADJ_BLOCK - 1) / ADJ_BLOCK);
  const int nb1 = (int)((ne * nst + ADJ_BLOCK - 1) / ADJ_BLOCK);
  hipLaunchKernelGGL((adj_elem_kernel<S, ADJ_K_KIN, QEPH, true>), dim3(nb0), dim3(ADJ_BLOCK), 0,
                     s, a, 0, nkin);
  hipLaunchKernelGGL((adj_elem_kernel<S, ADJ_K_ST, QEPH, false>), dim3(nb1), dim3(ADJ_BLOCK), 0,
                     s, a, 36, nst);
}

template <typename S>
static void launch_adj_elem_form(bool qeph, bool rev, const AdjArgs<S> &a, hipStream_t s) {
  const int nb = (a.numelc + ADJ_BLOCK - 1) / ADJ_BLOCK;
  if (rev && qeph) {
    hipLaunchKernelGGL((adj_elem_rev_kernel<S, true>), dim3(nb), dim3(ADJ_BLOCK), 0, s, a);
  } else if (rev) {
    hipLaunchKernelGGL((adj_elem_rev_kernel<S, false>), dim3(nb), dim3(ADJ_BLOCK), 0, s, a);
  } else if (qeph) {
    launch_adj_elem_chunks<S, true>(a, s);
  } else {
    launch_adj_elem_chunks<S, false>(a, s);
  }
}

hipError_t launch_adj_elem(bool fp32_state, bool qeph, bool rev, const AdjArgs<double> &a64,
                           hipStream_t s) {
  if (a64.numelc <= 0) return hipSuccess;
  if (fp32_state) {
    AdjArgs<float> a;
    a.numelc = a64.numelc; a.conn = a64.conn; a.X4 = a64.X4; a.V4 = a64.V4; a.VR4 = a64.VR4;
    a.st = reinterpret_cast<const float *>(a64.st);
    a.thke = a64.thke; a.lamS = a64.lamS; a.lamR = a64.lamR; a.lst_in = a64.lst_in;
    a.lst_out = a64.lst_out; a.lcorner = a64.lcorner; a.g_thk = a64.g_thk; a.g_jc = a64.g_jc;
    a.mat = a64.mat; a.prop = a64.prop; a.k = a64.k; a.dt1 = a64.dt1;
    launch_adj_elem_form(qeph, rev, a, s);
  } else {
    launch_adj_elem_form(qeph, rev, a64, s);
  }
  return hipGetLastError();
}

hipError_t launch_adj_brick_elem(bool fp32_state, bool rev, const AdjBrickArgs<double> &a64,
                                 hipStream_t s) {
  if (a64.numels <= 0) return hipSuccess;
  if (fp32_state) {
    AdjBrickArgs<float> a;
    a.numels = a64.numels; a.conn = a64.conn; a.X4 = a64.X4; a.V4 = a64.V4;
    a.st = reinterpret_cast<const float *>(a64.st);
    a.vol0 = a64.vol0; a.lamS = a64.lamS; a.lst_in = a64.lst_in; a.lst_out = a64.lst_out;
    a.lcorner = a64.lcorner; a.g_jc = a64.g_jc; a.mat = a64.mat; a.prop = a64.prop;
    a.dt1 = a64.dt1;
    launch_adj_brick_form(rev, a, s);
  } else {
    launch_adj_brick_form(rev, a64, s);
  }
  return hipGetLastError();
}

hipError_t launch_adj_node(int numnod, const double4 *X4, const double4 *V4, const double4 *VR4,
                           const double *A, const double *AR, double dt2, double dt12,
                           const double *lx, double *lv, const double *ld, double *lvr,
                           double *lamS, double *lamR, double *g_ms, double *g_in,
                           hipStream_t s) {
  const int nb = (numnod + 255) / 256;
  if (nb <= 0) return hipSuccess;
  const AdjNodeArgs q{X4, V4, VR4, A, AR, dt2, dt12};
  hipLaunchKernelGGL(adj_node_kernel, dim3(nb), dim3(256), 0, s, numnod, q, lx, lv, ld, lvr,
                     lamS, lamR, g_ms, g_in);
  return hipGetLastError();
}

hipError_t launch_adj_gather(int numnod, int numelc, int per, const int32_t *adsky,
                             const int32_t *slot_elem, const int32_t *slot_corner,
                             const double *lcorner, double *lx, double *lv, double *lvr,
                             hipStream_t s) {
  const int nb = (numnod + 255) / 256;
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(adj_gather_kernel, dim3(nb), dim3(256), 0, s, numnod, numelc, per, adsky,
                     slot_elem, slot_corner, lcorner, lx, lv, lvr);
  return hipGetLastError();
}

hipError_t launch_adj_gather_node(int numnod, int numelc, int per, const int32_t *adsky,
                                  const int32_t *slot_elem, const int32_t *slot_corner,
                                  const double *lcorner, const double4 *X4, const double4 *V4,
                                  const double4 *VR4, const double *A, const double *AR,
                                  double dt2, double dt12, double *lx, double *lv,
                                  const double *ld, double *lvr, double *lamS, double *lamR,
                                  double *g_ms, double *g_in, hipStream_t s) {
  const int nb = (numnod + 255) / 256;
  if (nb <= 0) return hipSuccess;
  const AdjNodeArgs q{X4, V4, VR4, A, AR, dt2, dt12};
  hipLaunchKernelGGL(adj_gather_node_kernel, dim3(nb), dim3(256), 0, s, numnod, numelc, per, adsky,
                     slot_elem, slot_corner, lcorner, q, lx, lv, ld, lvr, lamS, lamR, g_ms, g_in);
  return hipGetLastError();
}

hipError_t launch_adj_stash(int n, const void *const *src, void *const *dst,
                            const size_t *bytes, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (n > 7) return hipErrorInvalidValue;
  StashJob j{};
  size_t most = 0;
  for (int r = 0; r < n; ++r) {
    if (bytes[r] % 4) return hipErrorInvalidValue;
    j.src[r] = src[r];
    j.dst[r] = dst[r];
    j.bytes[r] = bytes[r];
    most = std::max(most, bytes[r]);
  }
  const int gx = (int)std::min<size_t>(1024, std::max<size_t>(1, (most / 16 + 255) / 256));
  hipLaunchKernelGGL(adj_stash_kernel, dim3(gx, n), dim3(256), 0, s, j);
  return hipGetLastError();
}

}  // namespace orad

// orad_nodal.hip -- node-side kernels, FP64, built without FMA contraction.
//
// node_kernel fuses, per node n and in the reference's order:
//   ASSPAR4  (assembly/asspar4.F:157-175)  a = sum of FSKY slots adsky[n]..adsky[n+1]-1
//            in slot order (PARITH/ON: bitwise independent of launch shape); STIFN
//            only when ASSPAR4 is called on its own (mode 0);
//   DT2      (resol.F:2640,5857-5861,6080,6224-6226) from the element minimum;
//   ACCELE   (assembly/accele.F:64-82)       a = a / MS, 0 where MS <= 0;
//   BCS10    (constraints/general/bcs/bcs1.F) zero fixed components of a and v;
//   FIXVEL   (constraints/general/impvel/fixvel.F:362-375) a = (v_imp - v)/DT12
//            on the imposed component;
//   VELOCITY (assembly/velocity.F:56-64,77-83) v = v + DT12 a (and vr = vr + DT12 ar);
//   DEPLA    (assembly/displacement.F:88-101) d = d + DT2 v, x = x + DT2 v.
// Energy balance (one model per context): between FIXVEL and VELOCITY, where RESOL's
// SORTIE_MAIN calls ECRIT, each workgroup sums the kinetic energies at the listing's
// velocity V + DT1/2 A (output/ecrit.F:213-226, 303-325) in a fixed order into
// bal_part[parity]; each imposed node adds its FIXVEL external work -- the energy
// memory DT2*DW of the last cycle plus DT1*DW of this one, DW = MS/4 (A DT12 + 2V)
// (A - AOLD) (constraints/general/impvel/fixvel.F:333-374, 834-836) -- to its own
// running total fv_work[n].
// ROT = true adds the rotational DOF of shells: the slots carry moments, ACCELE
// divides by the nodal inertia IN (accele.F:95-109) and BCS10 applies ICODR.
// Workgroup 0 also publishes the next cycle's time scalars (double-buffered by
// cycle parity), resets the next cycle's dt minimum and adds this cycle's
// hourglass-energy partials to the running total in a fixed order.  With batched
// models (node_model set) every node reads its own model's scalars, the thread of
// each model's first node publishes that model's next scalars, and one extra
// workgroup per model adds that model's per-chunk hourglass energies to its total.
#include "orad_common.hpp"

namespace orad {

constexpr int NODE_BLOCK = 256;

// Slots staged per LDS tile: a launch-time choice among instances whose tile holds at
// most 48 KB (default 2048 FP32 or 1024 FP64 slots).
template <typename R> struct SlotTile;
template <> struct SlotTile<float> { static constexpr int TILE = 2048; };
template <> struct SlotTile<double> { static constexpr int TILE = 1024; };
constexpr int NODE_TILE_BYTES = 49152;

// LDS index swizzle: permutes each aligned run of 32 words, so the per-node
// strided reads below (about 8 words apart) spread over the banks.
__device__ __forceinline__ int lds_sw(int i) { return i ^ ((i >> 5) & 31); }

template <typename R, bool ROT> struct SlotOf { typedef Slot3<R> type; };
template <typename R> struct SlotOf<R, true> { typedef Slot6<R> type; };

template <typename R, bool ROT, int TILE>
__global__ __launch_bounds__(NODE_BLOCK) void node_kernel(NodeArgs a, int parity, int mode,
                                                          int write_a, int xcd_remap) {
  typedef typename SlotOf<R, ROT>::type S;
  constexpr int NC = ROT ? 6 : 3;
  __shared__ R t[NC][TILE];
  const int nb = gridDim.x;
  const int b = xcd_remap ? xcd_block(blockIdx.x, nb) : (int)blockIdx.x;
  const int nbn = (a.numnod + NODE_BLOCK - 1) / NODE_BLOCK;
  if (b >= nbn) {
    // batched models: workgroup nbn + m adds model m's chunk hourglass energies to
    // its total, lanes striding the chunks then a butterfly sum (fixed order)
    const int m = b - nbn;
    if (mode == 1 && m < a.nmodel && threadIdx.x < 64) {
      const int k0 = a.model_chunks[m], k1 = a.model_chunks[m + 1];
      double h = 0.0;
      for (int k = k0 + (int)threadIdx.x; k < k1; k += 64) h += a.chunk_ehou[k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
      if (threadIdx.x == 0) a.ctl[m].ehour += h;
    }
    return;
  }
  const int n0 = b * NODE_BLOCK;
  const int n = n0 + (int)threadIdx.x;
  const S *fsky = reinterpret_cast<const S *>(a.fsky);
  const bool live = n < a.numnod;
  // batched models: each node steps with its own model's time scalars
  const int m = a.node_model ? a.node_model[live ? n : a.numnod - 1] : 0;
  CycleCtl *ctl = a.ctl + m;
  const Scal sc = ctl->s[parity];

  double dt1 = 0.0, dt2 = 0.0, dt12 = 0.0;
  if (mode == 1) {
    dt1 = sc.dt2;
    const double dt2t = __longlong_as_double((long long)ctl->dt2t[parity]);
    dt2 = 1.0e6;                                  // DT2 = EP06
    if (dt2t < dt2) dt2 = dt2t;
    const double lim = (double)1.1f * sc.dt2old;  // 1.1 is a default-real literal
    dt2 = (lim < dt2) ? lim : dt2;
    dt2 = (sc.dtmx < dt2) ? sc.dtmx : dt2;
    dt12 = 0.5 * (dt1 + dt2);
    const bool publisher = a.node_model ? (live && a.model_node0[m] == n)
                                        : (b == 0 && threadIdx.x == 0);
    if (publisher) {
      Scal nx = sc;
      nx.dt1 = dt1;
      nx.dt2 = dt2;
      nx.dt12 = dt12;
      nx.dt2old = dt2;
      nx.tt = sc.tt + dt2;
      nx.ncycle = sc.ncycle + 1;
      nx.mstop = ctl->mstop;
      ctl->s[parity ^ 1] = nx;
      ctl->dt2t[parity ^ 1] = 0x7FF0000000000000ull;
    }
    if (a.tile_queue && b == 0 && threadIdx.x == 0) *a.tile_queue = 0u;
    if (b == 0 && a.ehou_part) {
      __shared__ double s_part[NODE_BLOCK];
      double acc = 0.0;
      for (int i = threadIdx.x; i < a.n_ehou_part; i += NODE_BLOCK) acc += a.ehou_part[i];
      s_part[threadIdx.x] = acc;
      __syncthreads();
      for (int w = NODE_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s_part[threadIdx.x] += s_part[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0) ctl->ehour += s_part[0];
    }
  }

  // ASSPAR4: the workgroup's nodes own the contiguous slot range [s0, s1); it is
  // staged through LDS in tiles with coalesced loads, and each thread folds its
  // own node's slots left to right, in slot order, across tile boundaries.
  const int n1 = (n0 + NODE_BLOCK < a.numnod) ? n0 + NODE_BLOCK : a.numnod;
  const int s0 = a.adsky[n0], s1 = a.adsky[n1];
  const int k0 = live ? a.adsky[n] : 0, k1 = live ? a.adsky[n + 1] : 0;
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.0;
  for (int ts = s0; ts < s1; ts += TILE) {
    const int te = (ts + TILE < s1) ? ts + TILE : s1;
    __syncthreads();
    for (int i = threadIdx.x; i < te - ts; i += NODE_BLOCK) {
      const int j = lds_sw(i);
      const R *f = reinterpret_cast<const R *>(&fsky[ts + i]);
#pragma unroll
      for (int c = 0; c < NC; ++c) t[c][j] = f[c];
    }
    __syncthreads();
    const int lo = (k0 > ts) ? k0 : ts, hi = (k1 < te) ? k1 : te;
    for (int k = lo; k < hi; ++k) {
      const int j = lds_sw(k - ts);
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] = acc[c] + double(t[c][j]);
    }
  }
  double ax = acc[0], ay = acc[1], az = acc[2];
  // /INTER/TYPE7: the node's skyline entries in increasing entry order
  // (asspar4.F:560-583), selected one by one from its list
  const int head = (a.ci_head && live) ? a.ci_head[n] : -1;
  for (int last = -1;;) {
    int e = 0x7FFFFFFF;
    for (int q = head; q >= 0; q = a.ci_next[q])
      if (q > last && q < e) e = q;
    if (e == 0x7FFFFFFF) break;
    last = e;
    const double4 f = a.ci_sky[e];
    ax = ax + f.x; ay = ay + f.y; az = az + f.z;
  }
  if (mode == 0) {
    if (!live) return;
    // STIFN (FSKY(7,·) summed in slot order, then the interface entries); not consumed
    // by the explicit cycle.
    double sf = 0.0;
    const R *sti = reinterpret_cast<const R *>(a.sti);
    for (int k = k0; k < k1; ++k) sf = sf + double(sti[a.slot_elem[k]]);
    for (int last = -1;;) {
      int e = 0x7FFFFFFF;
      for (int q = head; q >= 0; q = a.ci_next[q])
        if (q > last && q < e) e = q;
      if (e == 0x7FFFFFFF) break;
      last = e;
      sf = sf + a.ci_sky[e].w;
    }
    a.A[3 * n] = ax; a.A[3 * n + 1] = ay; a.A[3 * n + 2] = az;
    a.STIFN[n] = sf;
    if (ROT) {
      a.AR[3 * n] = acc[NC - 3]; a.AR[3 * n + 1] = acc[NC - 2]; a.AR[3 * n + 2] = acc[NC - 1];
      a.STIFR[n] = 0.0;   // FSKY(8,·) = STIR is zero on this path (cdt3.F:219-227)
    }
    return;
  }
  double ek = 0.0, er = 0.0;
  if (live) {
  double4 X = a.X4[n];
  double4 V = a.V4[n];
  const double ms = X.w;
  if (ms > 0.0) {
    const double RTMP = 1.0 / ms;
    ax = ax * RTMP; ay = ay * RTMP; az = az * RTMP;
  } else {
    ax = 0.0; ay = 0.0; az = 0.0;
  }
  const int code = word_code(V.w);
  double rx = 0.0, ry = 0.0, rz = 0.0;
  double4 W = make_double4(0.0, 0.0, 0.0, 0.0);
  if (ROT) {
    W = a.VR4[n];
    rx = acc[NC - 3]; ry = acc[NC - 2]; rz = acc[NC - 1];
    const double in = W.w;
    if (in > 0.0) {
      const double RTMP = 1.0 / in;
      rx = rx * RTMP; ry = ry * RTMP; rz = rz * RTMP;
    } else {
      rx = 0.0; ry = 0.0; rz = 0.0;
    }
  }
  const int ct = code_icodt(code);
  if (ct & 4) { V.x = 0.0; ax = 0.0; }
  if (ct & 2) { V.y = 0.0; ay = 0.0; }
  if (ct & 1) { V.z = 0.0; az = 0.0; }
  if (ROT) {
    const int cr = code_icodr(code);
    if (cr & 4) { W.x = 0.0; rx = 0.0; }
    if (cr & 2) { W.y = 0.0; ry = 0.0; }
    if (cr & 1) { W.z = 0.0; rz = 0.0; }
  }
  const int fd = code_fvdir(code);
  if (fd) {
    double &aj = fd == 1 ? ax : (fd == 2 ? ay : az);
    const double vj = fd == 1 ? V.x : (fd == 2 ? V.y : V.z);
    const double aold = aj;
    aj = (a.fv_val[n] - vj) / dt12;
    if (a.fv_mem) {
      const double dw = 0.25 * ms * (aj * dt12 + 2.0 * vj) * (aj - aold);
      a.fv_work[n] = a.fv_work[n] + (a.fv_mem[n] + dt1 * dw);
      a.fv_mem[n] = dt2 * dw;
    }
  }
  if (a.bal_part) {
    const double dt05 = 0.5 * dt1;
    const double ux = V.x + dt05 * ax, uy = V.y + dt05 * ay, uz = V.z + dt05 * az;
    ek = (ux * ux + uy * uy + uz * uz) * 0.5 * ms;
    if (ROT) {
      const double wx = W.x + dt05 * rx, wy = W.y + dt05 * ry, wz = W.z + dt05 * rz;
      er = (wx * wx + wy * wy + wz * wz) * 0.5 * W.w;
    }
  }
  if (write_a) {
    a.A[3 * n] = ax; a.A[3 * n + 1] = ay; a.A[3 * n + 2] = az;
    if (ROT) { a.AR[3 * n] = rx; a.AR[3 * n + 1] = ry; a.AR[3 * n + 2] = rz; }
  }
  V.x = V.x + dt12 * ax;
  V.y = V.y + dt12 * ay;
  V.z = V.z + dt12 * az;
  if (ROT) {
    W.x = W.x + dt12 * rx;
    W.y = W.y + dt12 * ry;
    W.z = W.z + dt12 * rz;
    a.VR4[n] = W;
  }
  const double vdx = dt2 * V.x, vdy = dt2 * V.y, vdz = dt2 * V.z;
  double *D = a.D + 3 * n;
  D[0] = D[0] + vdx; D[1] = D[1] + vdy; D[2] = D[2] + vdz;
  X.x = X.x + vdx; X.y = X.y + vdy; X.z = X.z + vdz;
  a.X4[n] = X;
  a.V4[n] = V;
  }  // live
  if (a.bal_part) {
    // workgroup sums in a fixed order: 64-lane butterflies, then the four waves
    __shared__ double s_bal[2][NODE_BLOCK / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ek += __shfl_xor(ek, o, 64);
      er += __shfl_xor(er, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      s_bal[0][threadIdx.x >> 6] = ek;
      s_bal[1][threadIdx.x >> 6] = er;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double se = s_bal[0][0], sr = s_bal[1][0];
#pragma unroll
      for (int w = 1; w < NODE_BLOCK / 64; ++w) { se += s_bal[0][w]; sr += s_bal[1][w]; }
      double *bp = a.bal_part + (size_t)parity * 2 * nbn;
      bp[b] = se;
      bp[nbn + b] = sr;
    }
  }
}

// Energy partial sums: block b sums its range in a fixed tree order.
// Bricks: eint per unit initial volume times vol0.  Shells (vol0 null): membrane
// plus bending energy eint[e] + eint2[e]; kinetic energy adds 0.5*IN*|vr|^2.
template <typename T>
__global__ __launch_bounds__(NODE_BLOCK) void energy_partial_kernel(
    int numels, const T *eint, const T *eint2, const double *vol0, const int32_t *emap,
    int numnod, const double4 *X4, const double4 *V4, const double4 *VR4, double *part_e,
    double *part_k, int nbe) {
  __shared__ double s[NODE_BLOCK];
  const int b = blockIdx.x;
  double v = 0.0;
  double *out;
  if (b < nbe) {
    const int e = b * NODE_BLOCK + threadIdx.x;
    if (e < numels) {
      // engine order whatever the storage order (emap: engine -> library element)
      const int s = emap ? emap[e] : e;
      v = vol0 ? double(eint[s]) * double(vol0[s]) : double(eint[s]) + double(eint2[s]);
    }
    out = part_e + b;
  } else {
    const int n = (b - nbe) * NODE_BLOCK + threadIdx.x;
    if (n < numnod) {
      const double4 V = V4[n];
      v = 0.5 * X4[n].w * (V.x * V.x + V.y * V.y + V.z * V.z);
      if (VR4) {
        const double4 W = VR4[n];
        v += 0.5 * W.w * (W.x * W.x + W.y * W.y + W.z * W.z);
      }
    }
    out = part_k + (b - nbe);
  }
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = NODE_BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

__global__ __launch_bounds__(NODE_BLOCK) void energy_final_kernel(const double *part_e, int ne,
                                                                  const double *part_k, int nk,
                                                                  double *res) {
  __shared__ double s[2][NODE_BLOCK];
  double ae = 0.0, ak = 0.0;
  for (int i = threadIdx.x; i < ne; i += NODE_BLOCK) ae += part_e[i];
  for (int i = threadIdx.x; i < nk; i += NODE_BLOCK) ak += part_k[i];
  s[0][threadIdx.x] = ae;
  s[1][threadIdx.x] = ak;
  __syncthreads();
  for (int w = NODE_BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s[0][threadIdx.x] += s[0][threadIdx.x + w];
      s[1][threadIdx.x] += s[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { res[0] = s[0][0]; res[1] = s[1][0]; }
}

template <typename R, bool ROT, int TILE>
static hipError_t node_instance(const NodeArgs &a, int nb, int parity, int mode, int write_a,
                                int xcd_remap, hipStream_t s) {
  static_assert((ROT ? 6 : 3) * TILE * sizeof(R) <= NODE_TILE_BYTES, "tile over 48 KB");
  hipLaunchKernelGGL((node_kernel<R, ROT, TILE>), dim3(nb), dim3(NODE_BLOCK), 0, s, a, parity,
                     mode, write_a, xcd_remap);
  return hipGetLastError();
}

int node_tile_default(bool slot_fp32) {
  return slot_fp32 ? SlotTile<float>::TILE : SlotTile<double>::TILE;
}

// Tiles built: 1024, 2048, 4096 slots where the tile fits 48 KB of LDS.
bool node_tile_ok(bool slot_fp32, bool rot, int tile) {
  if (tile != 1024 && tile != 2048 && tile != 4096) return false;
  return (rot ? 6 : 3) * tile * (slot_fp32 ? 4 : 8) <= NODE_TILE_BYTES;
}

hipError_t launch_node(const NodeArgs &a, bool slot_fp32, int parity, int mode, int write_a,
                       int xcd_remap, hipStream_t s, int tile) {
  const int nb = (a.numnod + NODE_BLOCK - 1) / NODE_BLOCK + (mode == 1 ? a.nmodel : 0);
  if (nb <= 0) return hipSuccess;
  const bool rot = a.VR4 != nullptr;
  if (tile <= 0) tile = node_tile_default(slot_fp32);
  if (!node_tile_ok(slot_fp32, rot, tile)) return hipErrorInvalidConfiguration;
#define ORAD_NODE(R, ROT, T)                                                              \
  if ((sizeof(R) == 4) == slot_fp32 && rot == ROT && tile == T)                           \
    return node_instance<R, ROT, T>(a, nb, parity, mode, write_a, xcd_remap, s);
  ORAD_NODE(float, false, 1024)
  ORAD_NODE(float, false, 2048)
  ORAD_NODE(float, false, 4096)
  ORAD_NODE(float, true, 1024)
  ORAD_NODE(float, true, 2048)
  ORAD_NODE(double, false, 1024)
  ORAD_NODE(double, false, 2048)
  ORAD_NODE(double, true, 1024)
#undef ORAD_NODE
  return hipErrorInvalidConfiguration;
}

template <typename T>
static hipError_t launch_energy_t(int numels, const T *eint, const T *eint2, const double *vol0,
                                  const int32_t *emap, int numnod, const double4 *X4,
                                  const double4 *V4, const double4 *VR4, double *part,
                                  double *res, hipStream_t s) {
  const int nbe = (numels + NODE_BLOCK - 1) / NODE_BLOCK;
  const int nbk = (numnod + NODE_BLOCK - 1) / NODE_BLOCK;
  if (nbe + nbk > 0)
    hipLaunchKernelGGL(energy_partial_kernel<T>, dim3(nbe + nbk), dim3(NODE_BLOCK), 0, s, numels,
                       eint, eint2, vol0, emap, numnod, X4, V4, VR4, part, part + nbe, nbe);
  hipLaunchKernelGGL(energy_final_kernel, dim3(1), dim3(NODE_BLOCK), 0, s, part, nbe, part + nbe,
                     nbk, res);
  return hipGetLastError();
}

// Fixed-order sum of x[0..n): one partial per 256 values, then the partials.
__global__ __launch_bounds__(NODE_BLOCK) void sum_partial_kernel(const double *x, int n,
                                                                 double *part) {
  __shared__ double s[NODE_BLOCK];
  const int i = blockIdx.x * NODE_BLOCK + threadIdx.x;
  s[threadIdx.x] = i < n ? x[i] : 0.0;
  __syncthreads();
  for (int w = NODE_BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// part: (n + 255) / 256 doubles; res: two doubles (the sum, then zero)
hipError_t launch_sum(const double *x, int n, double *part, double *res, hipStream_t s) {
  const int nb = (n + NODE_BLOCK - 1) / NODE_BLOCK;
  if (nb > 0)
    hipLaunchKernelGGL(sum_partial_kernel, dim3(nb), dim3(NODE_BLOCK), 0, s, x, n, part);
  hipLaunchKernelGGL(energy_final_kernel, dim3(1), dim3(NODE_BLOCK), 0, s, part, nb, part, 0,
                     res);
  return hipGetLastError();
}

hipError_t launch_energy(bool fp32, int numels, const void *eint, const void *eint2,
                         const double *vol0, const int32_t *emap, int numnod, const double4 *X4,
                         const double4 *V4, const double4 *VR4, double *part, double *res,
                         hipStream_t s) {
  if (fp32)
    return launch_energy_t<float>(numels, (const float *)eint, (const float *)eint2, vol0, emap,
                                  numnod, X4, V4, VR4, part, res, s);
  return launch_energy_t<double>(numels, (const double *)eint, (const double *)eint2, vol0,
                                 emap, numnod, X4, V4, VR4, part, res, s);
}

// The engine's X(3,n), V(3,n) and, when VR4 is given, VR(3,n) (staged one after the
// other in xv) into the node records, keeping MS, the node word and IN in the fourth
// components.
__global__ __launch_bounds__(256) void pack_xv_kernel(int numnod, const double *xv, double4 *X4,
                                                      double4 *V4, double4 *VR4) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= numnod) return;
  const double *x = xv + 3 * (size_t)n, *v = xv + 3 * ((size_t)numnod + n);
  X4[n] = make_double4(x[0], x[1], x[2], X4[n].w);
  V4[n] = make_double4(v[0], v[1], v[2], V4[n].w);
  if (VR4) {
    const double *w = xv + 3 * (2 * (size_t)numnod + n);
    VR4[n] = make_double4(w[0], w[1], w[2], VR4[n].w);
  }
}

hipError_t launch_pack_xv(int numnod, const double *xv, double4 *X4, double4 *V4, double4 *VR4,
                          hipStream_t s) {
  const int nb = (numnod + 255) / 256;
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(pack_xv_kernel, dim3(nb), dim3(256), 0, s, numnod, xv, X4, V4, VR4);
  return hipGetLastError();
}

}  // namespace orad

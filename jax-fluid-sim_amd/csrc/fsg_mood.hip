// MOOD on gfx950: troubled-cell detection with wavefront-ballot compaction,
// the 64-subcell finite-volume recompute of troubled elements, the
// neighbours' mean correction, and the positivity limiter.
//
// Restates hybrid/limiter.py:11-343 and hybrid/subcell.py:442-950 of the
// reference for affine elements (every subcell table is then
// element-independent; physical subcell area vectors are |det J| J^-T times
// the reference ones, subcell volumes |det J| times the reference ones).
//
// Work mapping:
//   detect      one 64-lane wave per element: lanes over nodes and sample
//               points, wave reductions for the sensor energies and the PAD
//               test; then one lane per element for the ballot compaction.
//   subcell     one 64-lane workgroup per troubled element, lane = subcell,
//               persistent over the compacted list; every face flux is
//               computed by each subcell that receives it (no atomics).
//   correct     one wave per element; elements with no troubled neighbour
//               return at once.
//   positivity  one wave per element, samples over lanes, bisection on
//               the pressure theta only where theta = upper fails.
//
// A subface flux between a troubled element and a neighbour is computed
// twice (once by each side's kernel) from the same operands; contraction is
// disabled in this file so both copies round identically and the update
// stays conservative to round-off.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/fluid_gfx950.h"
#include "fsg_context.h"
#include "fsg_kernels.h"

namespace fsgmood {

using fsg::Phys;
using fsg::StageDiag;

constexpr int NSUB = 64;
constexpr int MAXN = 64;      // nodes per element (hexahedron p=3)
constexpr int MAXB = 3;       // element blocks

template <typename real>
struct TypeT {
  int N, Q, NS, NI, NB;
  real wsum;                   // sum of the reference volume weights
  real volsum;                 // sum of the reference subcell volumes
  int KS;                      // ceil((N + NS) / 64): samples per lane
  real lebesgue;               // max over samples k of sum_j |phi_j(x_k)| (>= 1)
  const real* mw;       // (N): mw[j] = sum_k wq[k] samples[j * NS + k], k < Q
  const real* p1proj;   // (N, N) transposed: P1[i][j] = p1proj[j * N + i]
  const real* samples;  // (N, NS): basis j at sample point k is samples[j * NS + k]
  const real* wq;
  const real* P;        // (64, N)
  const real* PT;       // (N, 64)
  const real* RT;       // (64, N): R[i][t] = RT[t * N + i]
  const real* subvol;
  const int32_t* int_o;
  const int32_t* int_n;
  const real* int_area;
  const real* int_dc;
  const int32_t* int_ptr;
  const int32_t* int_list;
  const int32_t* bnd_face;
  const int32_t* bnd_gather;
  const int32_t* tgt_ptr;
  const int32_t* tgt_list;
};

template <typename real>
struct BlockT {
  int type, max_faces;
  int64_t count, offset;
  const real* jinv;
  const real* detw;
  const int32_t* elem_faces;
  int32_t* flags;
  real* sensor;
};

template <typename real>
struct GroupT {
  int face_type, sub, owner_block, neighbor_block, K, Qf, NFo, NFn, TSO, TSN;
  const int32_t* owner_elem;
  const int32_t* neighbor_elem;
  const real* geom;
  const real* fw;
  const int32_t* owner_table;
  const real* owner_tables;
  const int32_t* owner_sup;
  const int32_t* neighbor_table;
  const real* neighbor_tables;
  const int32_t* neighbor_sup;
  const int32_t* bflags;
  const int32_t* pair_table;
  const int32_t* pair_ids;
  const real* pair_geom;
  const real* basis_o;
  const real* basis_n;
};

template <typename real>
struct Args {
  const TypeT<real>* types;
  const BlockT<real>* blocks;
  const GroupT<real>* groups;
  int nblocks;
  int64_t boff[MAXB + 1];   // global index of each block's first element, then the total
  int btype[MAXB];          // element type of each block
  const real* curr[MAXB];   // states the subcell update starts from
  const real* base[MAXB];   // stage base
  const real* inv[MAXB];    // inviscid rhs (stage)
  const real* visc[MAXB];   // viscous rhs / additional rhs (may be null)
  const real* det_src[MAXB];// detection source when not a stage
  real* work[MAXB];         // euler / DG candidate, corrected in place
  real* out[MAXB];          // positivity output
  const real* dte[MAXB];    // per-element dt or null
  real dt, alpha, beta;
  int32_t* counter;
  int32_t* list;
  real thr, dfloor, pfloor, rel;
  int recon, sensor_on, stage;
  StageDiag* diag;
};

template <typename real> __device__ __forceinline__ bool isnum(real x) { return x - x == real(0); }

// Wave-wide reductions (64 lanes).
template <typename real>
__device__ __forceinline__ real wsum(real x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
// min / max that propagate NaN (numpy semantics).
template <typename real>
__device__ __forceinline__ real nmin(real a, real b) { return (a != a || b != b) ? a + b : (a < b ? a : b); }
template <typename real>
__device__ __forceinline__ real nmax(real a, real b) { return (a != a || b != b) ? a + b : (a > b ? a : b); }
template <typename real>
__device__ __forceinline__ real wmin(real x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = nmin(x, __shfl_xor(x, o, 64));
  return x;
}
template <typename real>
__device__ __forceinline__ real wmax(real x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = nmax(x, __shfl_xor(x, o, 64));
  return x;
}
__device__ __forceinline__ bool wany(bool b) { return __ballot(b) != 0ull; }
// Reductions over groups of L lanes (L = 64: the whole wave; L = 32: two
// elements per wave).  With the inactive lanes of a group holding 0, the
// L = 32 reduction gives the same value as the 64-lane one (its first step
// would add or take the maximum with exact zeros).
template <int L, typename real>
__device__ __forceinline__ real gsum(real x) {
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
template <int L, typename real>
__device__ __forceinline__ real gmax_num(real x) {
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  return x;
}
template <int L>
__device__ __forceinline__ bool gany(bool b, int group) {
  const unsigned long long m = __ballot(b);
  if constexpr (L == 64) return m != 0ull;
  else return ((m >> (group * L)) & ((1ull << L) - 1)) != 0ull;
}
// Wave maximum that ignores NaN (fmax): one instruction per step.  Used for
// the deviation bounds of the sampling shortcuts, whose decision is false
// anyway when an input is NaN or infinite (the mean, and with it the bounds,
// are then NaN: every comparison with the floors fails).
template <typename real>
__device__ __forceinline__ real wmax_num(real x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  return x;
}

template <typename real>
__device__ __forceinline__ int locate(const Args<real>& a, int64_t g, int64_t* e) {
  int b = 0;
  while (b + 1 < a.nblocks && g >= a.boff[b + 1]) ++b;
  *e = g - a.boff[b];
  return b;
}

// dst[k] = src[k] for k < n, lane l of a group of S lanes taking k = l, l + S,
// ...  The loads go out in batches of CH, all issued before the first is
// waited on: a loop that stores each value before it loads the next waits
// one memory round trip per S values.  Indices past n load src[n - 1] (n > 0)
// and are not stored.
template <int S, int CH, typename real>
__device__ __forceinline__ void load_rows(real* dst, const real* src, int n, int lane) {
  for (int k0 = 0; k0 < n; k0 += S * CH) {
    real t[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int k = k0 + lane + S * i;
      t[i] = src[k < n ? k : n - 1];
    }
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int k = k0 + lane + S * i;
      if (k < n) dst[k] = t[i];
    }
  }
}

// hybrid/solver.py:83-98 pressure of a conserved state.
template <typename real>
__device__ __forceinline__ real pressure(const real* U, real gamma) {
  real safe = U[0] > real(1e-14) ? U[0] : real(1e-14);
  real ke = real(0.5) * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / safe;
  return (gamma - real(1)) * (U[4] - ke);
}

// The clamps of hybrid/subcell.py:648-663, 727-740, 823-837.
template <int V, typename real>
__device__ __forceinline__ void admissible(real* U) {
  U[0] = U[0] > real(1e-14) ? U[0] : real(1e-14);
  real ke = real(0.5) * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0];
  real emin = ke + real(1e-12);
  U[4] = U[4] > emin ? U[4] : emin;
#pragma unroll
  for (int s = 5; s < V; ++s) U[s] = U[s] > real(0) ? U[s] : real(0);
}

template <int V, typename real>
__device__ __forceinline__ void numflux(const real* UL, const real* UR, const real* n,
                                        const Phys<real>& ph, real* F) {
  if (ph.rusanov)
    fsg::rusanov<V, real>(UL, UR, n, ph.gamma, F);
  else
    fsg::hllc<V, real>(UL, UR, n, ph.gamma, F);
}

// Physical area vector |det J| J^-T a_ref; jinv[r*3+d] = d xi_r / d x_d.
template <typename real>
__device__ __forceinline__ void phys_area(const real* jinv, real det, const real* aref, real* a) {
#pragma unroll
  for (int d = 0; d < 3; ++d)
    a[d] = det * (jinv[0 * 3 + d] * aref[0] + jinv[1 * 3 + d] * aref[1] + jinv[2 * 3 + d] * aref[2]);
}

// Sum_m row[m] U[m, v] for one element row-major (N, V) in global memory.
template <int V, typename real>
__device__ __forceinline__ void row_dot(const real* row, const real* U, int N, real* out) {
#pragma unroll
  for (int v = 0; v < V; ++v) out[v] = real(0);
  for (int m = 0; m < N; ++m) {
    const real c = row[m];
#pragma unroll
    for (int v = 0; v < V; ++v) out[v] += c * U[m * V + v];
  }
}

// Flux of subface pair k of face i in group g (hybrid/subcell.py:748-853):
// a troubled side contributes its subcell average, the other its DG
// polynomial at the subface centre; boundary faces take the ghost state of
// the unclamped owner trace; both traces are clamped; flux times area.
template <int V, typename real>
__device__ void pair_flux(const Args<real>& a, const Phys<real>& ph, const GroupT<real>& G,
                          int64_t i, int k, real* F) {
  const BlockT<real>& bo = a.blocks[G.owner_block];
  const TypeT<real>& to = a.types[bo.type];
  const int64_t oe = G.owner_elem[i];
  const int tab = G.pair_table[i];
  const int32_t* ids = G.pair_ids + ((int64_t)tab * G.K + k) * 2;
  const real* pg = G.pair_geom + ((int64_t)tab * G.K + k) * 4;
  const real* Uo = a.curr[G.owner_block] + oe * to.N * V;
  real UL[V], UR[V];
  if (bo.flags[oe])
    row_dot<V, real>(to.P + (int64_t)to.bnd_gather[ids[0]] * to.N, Uo, to.N, UL);
  else
    row_dot<V, real>(G.basis_o + ((int64_t)tab * G.K + k) * to.N, Uo, to.N, UL);
  // normal and area from the side named by the table
  real av[3];
  const real aref[3] = {pg[1], pg[2], pg[3]};
  if (pg[0] == real(0)) {
    phys_area(bo.jinv + oe * 9, bo.detw[oe], aref, av);
  } else {
    const BlockT<real>& bn = a.blocks[G.neighbor_block];
    const int64_t ne = G.neighbor_elem[i];
    phys_area(bn.jinv + ne * 9, bn.detw[ne], aref, av);
  }
  const real area = sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
  const real n[3] = {av[0] / area, av[1] / area, av[2] / area};
  if (G.neighbor_block >= 0) {
    const BlockT<real>& bn = a.blocks[G.neighbor_block];
    const TypeT<real>& tn = a.types[bn.type];
    const int64_t ne = G.neighbor_elem[i];
    const real* Un = a.curr[G.neighbor_block] + ne * tn.N * V;
    if (bn.flags[ne])
      row_dot<V, real>(tn.P + (int64_t)tn.bnd_gather[ids[1]] * tn.N, Un, tn.N, UR);
    else
      row_dot<V, real>(G.basis_n + ((int64_t)tab * G.K + k) * tn.N, Un, tn.N, UR);
  } else {
    fsg::ghost_state<V, real>(UL, n, G.bflags ? G.bflags[i] : 0, ph, UR);
  }
  admissible<V, real>(UL);
  admissible<V, real>(UR);
  numflux<V, real>(UL, UR, n, ph, F);
#pragma unroll
  for (int v = 0; v < V; ++v) F[v] *= area;
}

// DG face flux of face i without the anchor, with the trace clamps
// (hybrid/subcell.py:690-746), at face quadrature point q.
template <int V, typename real>
__device__ void dg_point_flux(const Args<real>& a, const Phys<real>& ph, const GroupT<real>& G,
                              int64_t i, int q, real* F) {
  const BlockT<real>& bo = a.blocks[G.owner_block];
  const TypeT<real>& to = a.types[bo.type];
  const int64_t oe = G.owner_elem[i];
  const real* Uo = a.curr[G.owner_block] + oe * to.N * V;
  const int tid = G.owner_table[i];
  const real* phi = G.owner_tables + (int64_t)tid * G.TSO + (int64_t)q * G.NFo;
  const int32_t* sup = G.owner_sup + (int64_t)tid * G.NFo;
  real UL[V], UR[V];
#pragma unroll
  for (int v = 0; v < V; ++v) UL[v] = real(0);
  for (int m = 0; m < G.NFo; ++m) {
    const real c = phi[m];
    const real* u = Uo + (int64_t)sup[m] * V;
#pragma unroll
    for (int v = 0; v < V; ++v) UL[v] += c * u[v];
  }
  const real n[3] = {G.geom[i * 4 + 0], G.geom[i * 4 + 1], G.geom[i * 4 + 2]};
  if (G.neighbor_block >= 0) {
    const BlockT<real>& bn = a.blocks[G.neighbor_block];
    const TypeT<real>& tn = a.types[bn.type];
    const int64_t ne = G.neighbor_elem[i];
    const real* Un = a.curr[G.neighbor_block] + ne * tn.N * V;
    const int nid = G.neighbor_table[i];
    const real* phn = G.neighbor_tables + (int64_t)nid * G.TSN + (int64_t)q * G.NFn;
    const int32_t* supn = G.neighbor_sup + (int64_t)nid * G.NFn;
#pragma unroll
    for (int v = 0; v < V; ++v) UR[v] = real(0);
    for (int m = 0; m < G.NFn; ++m) {
      const real c = phn[m];
      const real* u = Un + (int64_t)supn[m] * V;
#pragma unroll
      for (int v = 0; v < V; ++v) UR[v] += c * u[v];
    }
  } else {
    fsg::ghost_state<V, real>(UL, n, G.bflags ? G.bflags[i] : 0, ph, UR);
  }
  admissible<V, real>(UL);
  admissible<V, real>(UR);
  numflux<V, real>(UL, UR, n, ph, F);
  const real w = G.fw[q] * G.geom[i * 4 + 3];
#pragma unroll
  for (int v = 0; v < V; ++v) F[v] *= w;
}

// ---------------------------------------------------------------------------
// Detection (hybrid/limiter.py:150-185, 285-300).  stage != 0: the detected
// state is alpha base + beta (curr + dt (inv + visc)), and curr + dt (inv +
// visc) is written to work; otherwise the detected state is det_src.  FP32
// instantiations are held to six waves per SIMD (a few bytes of register
// spill per lane; the extra waves keep more element loads in flight: 1.27 ->
// 1.00 ms per reacting-mesh launch).
template <int V, typename real, int EPW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(real) == 4 ? 6 : 1)))
void detect_kernel(Args<real> a, Phys<real> ph, int64_t total) {
  // EPW elements per wave, LG = 64 / EPW lanes each (EPW = 2 needs N <= 32)
  constexpr int LG = 64 / EPW, NE = MAXN / EPW;
  __shared__ real sU[4 * EPW][NE][V];
  __shared__ real sm[4 * EPW][NE][2];
  __shared__ real sW[4 * EPW][NE];   // the type's basis-weight products mw
  const int wl = threadIdx.x & 63, hh = wl / LG, lane = wl - hh * LG;
  const int w = (threadIdx.x >> 6) * EPW + hh;
  const int64_t g0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * EPW;
  if (g0 >= total) return;  // whole wave exits together; no block barriers below
  const bool live = g0 + hh < total;
  const int64_t g = live ? g0 + hh : total - 1;   // a spare group repeats the last element
  int64_t e;
  const int b = locate(a, g, &e);
  const BlockT<real>& B = a.blocks[b];
  const TypeT<real>& T = a.types[a.btype[b]];
  const int N = T.N;
  {
    // element values in node-major order, one coalesced pass over N V
    real* su = &sU[w][0][0];
    const int64_t o = e * N * V;
    const real mwl = lane < N ? T.mw[lane] : real(0);
    if (a.stage && !a.inv[b]) {
      // work already holds the Euler candidate curr + dt rhs (formed by the
      // residual's own epilogue): the stage candidate is alpha base + beta work
      const real* wrk = a.work[b] + o;
      const real* bas = a.base[b] + o;
      constexpr int CH = 4;
      const int n = N * V;
      for (int k0 = 0; k0 < n; k0 += LG * CH) {
        real rw[CH], rb[CH];
#pragma unroll
        for (int i = 0; i < CH; ++i) {
          const int k = k0 + lane + LG * i, kk = k < n ? k : n - 1;
          rw[i] = wrk[kk];
          rb[i] = bas[kk];
        }
#pragma unroll
        for (int i = 0; i < CH; ++i) {
          const int k = k0 + lane + LG * i;
          if (k < n) su[k] = a.alpha * rb[i] + a.beta * rw[i];
        }
      }
    } else if (a.stage) {
      const real h = a.dte[b] ? a.dte[b][e] : a.dt;
      const real* inv = a.inv[b] + o;
      const real* vis = a.visc[b] ? a.visc[b] + o : nullptr;
      const real* cur = a.curr[b] + o;
      const real* bas = a.base[b] + o;
      real* wrk = a.work[b] + o;
      // batches of CH indices per lane, every load of a batch in flight at once
      constexpr int CH = 4;
      const int n = N * V;
      for (int k0 = 0; k0 < n; k0 += LG * CH) {
        real ri[CH], rv[CH], rc[CH], rb[CH];
#pragma unroll
        for (int i = 0; i < CH; ++i) {
          const int k = k0 + lane + LG * i, kk = k < n ? k : n - 1;
          ri[i] = inv[kk];
          rc[i] = cur[kk];
          rb[i] = bas[kk];
          if (vis) rv[i] = vis[kk];
        }
#pragma unroll
        for (int i = 0; i < CH; ++i) {
          const int k = k0 + lane + LG * i;
          if (k < n) {
            real r = ri[i];
            if (vis) r = r + rv[i];
            const real eu = rc[i] + h * r;
            if (live) wrk[k] = eu;
            su[k] = a.alpha * rb[i] + a.beta * eu;
          }
        }
      }
    } else {
      load_rows<LG, 8>(su, a.det_src[b] + o, N * V, lane);
    }
    if (lane < N) sW[w][lane] = mwl;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (lane < N) {
    real u[V];
#pragma unroll
    for (int v = 0; v < V; ++v) u[v] = sU[w][lane][v];
    fsg::Prim<V, real> pr = fsg::c2p<V, real>(u, ph.gamma);
    sm[w][lane][0] = pr.rho;
    sm[w][lane][1] = pr.p;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  // Sensor: non-P1 energy of density and pressure at the nodes.
  real ratio = real(0);
  if (a.sensor_on) {
    real hi0 = 0, hi1 = 0, sc0 = 0, sc1 = 0;
    if (lane < N) {
      real l0 = 0, l1 = 0;
#pragma unroll 8
      for (int j = 0; j < N; ++j) {
        const real c = T.p1proj[j * N + lane];   // transposed table: P1[lane][j], coalesced
        l0 += c * sm[w][j][0];
        l1 += c * sm[w][j][1];
      }
      const real m0 = sm[w][lane][0], m1 = sm[w][lane][1];
      hi0 = (m0 - l0) * (m0 - l0);
      hi1 = (m1 - l1) * (m1 - l1);
      sc0 = m0 * m0;
      sc1 = m1 * m1;
    }
    hi0 = gsum<LG>(hi0);
    hi1 = gsum<LG>(hi1);
    sc0 = gsum<LG>(sc0);
    sc1 = gsum<LG>(sc1);
    const real r0 = hi0 / (sc0 > real(1e-12) ? sc0 : real(1e-12));
    const real r1 = hi1 / (sc1 > real(1e-12) ? sc1 : real(1e-12));
    ratio = nmax(r0, r1);
  }
  // PAD shortcut, as in the positivity limiter: every sample is
  // mean + sum_j phi_j(x_k) (U_j - mean), so with finite, moderate nodal
  // values and the Lebesgue-constant bounds on rho and p at or above the
  // floors (less a rounding allowance), no sample, and not the mean (a
  // positive-weight average of the volume samples; p is concave), can fail:
  // the sampled check below would find nothing.
  bool pad_done = false;
  {
    real mpart = real(0);
    if (lane < 5)
#pragma unroll 8
      for (int j = 0; j < N; ++j) mpart += sW[w][j] * sU[w][j][lane];
    real mv[5], D[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) mv[v] = __shfl(mpart, hh * LG + v, 64) / T.wsum;
    const real big = sizeof(real) == 4 ? real(1e15) : real(1e150);
    bool fine = true;
    if (lane < N)
#pragma unroll
      for (int v = 0; v < V; ++v) fine = fine && isnum(sU[w][lane][v]) && fabs(sU[w][lane][v]) <= big;
#pragma unroll
    for (int v = 0; v < 5; ++v) D[v] = gmax_num<LG>(lane < N ? fabs(sU[w][lane][v] - mv[v]) : real(0));
    const real L = T.lebesgue;
    const real eps = sizeof(real) == 4 ? real(1.2e-7) : real(2.3e-16);
    const real rnd = real(16) * real(N + 2) * eps * L;
    real bnd[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) bnd[v] = L * D[v] + rnd * (fabs(mv[v]) + D[v]);
    const real rho_lb = mv[0] - bnd[0];
    real m2 = real(0);
#pragma unroll
    for (int d = 1; d < 4; ++d) {
      const real mu = fabs(mv[d]) + bnd[d];
      m2 += mu * mu;
    }
    const real ke_ub = rho_lb > real(0) ? real(0.5) * m2 / rho_lb : real(0);
    const real p_lb = (ph.gamma - real(1)) * (mv[4] - bnd[4] - ke_ub);
    const real p_tol = real(16) * eps * (ph.gamma - real(1)) * (fabs(mv[4]) + bnd[4] + ke_ub);
    pad_done = !gany<LG>(!fine, hh) && rho_lb > real(0) && rho_lb >= a.dfloor &&
               p_lb - p_tol >= a.pfloor;
  }
  // PAD over nodes, volume points, face sample points and the mean.
  bool bad = false;
  real msum[V];
#pragma unroll
  for (int v = 0; v < V; ++v) msum[v] = real(0);
  auto check = [&](const real* u) {
    real W[V];
    fsg::c2p_full<V, real>(u, ph.gamma, W);
    bool fin = true;
#pragma unroll
    for (int v = 0; v < V; ++v) fin = fin && isnum(W[v]);
    return !fin || W[0] < a.dfloor || W[4] < a.pfloor;
  };
  if (!pad_done) {
  if (lane < N) bad = bad || check(sU[w][lane]);
  for (int k0 = 0; k0 < T.NS; k0 += LG) {
    const int k = k0 + lane;
    if (k < T.NS) {
      real u[V];
#pragma unroll
      for (int v = 0; v < V; ++v) u[v] = real(0);
      for (int j = 0; j < N; ++j) {
        const real c = T.samples[(int64_t)j * T.NS + k];
#pragma unroll
        for (int v = 0; v < V; ++v) u[v] += c * sU[w][j][v];
      }
      bad = bad || check(u);
      if (k < T.Q) {
#pragma unroll
        for (int v = 0; v < V; ++v) msum[v] += T.wq[k] * u[v];
      }
    }
  }
  real mean[V];
#pragma unroll
  for (int v = 0; v < V; ++v) mean[v] = gsum<LG>(msum[v]) / T.wsum;
  bad = bad || check(mean);
  }
  const bool troubled = gany<LG>(bad, hh) || ratio > a.thr;
  if (lane == 0 && live) {
    B.flags[e] = troubled ? 1 : 0;
    if (B.sensor) B.sensor[e] = ratio;
  }
}

// Ballot compaction of the flags into the troubled list (order across waves
// is arbitrary; every consumer is order-independent).
template <typename real>
__global__ __launch_bounds__(256) void compact_kernel(Args<real> a, int64_t total) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool f = false;
  if (g < total) {
    int64_t e;
    const int b = locate(a, g, &e);
    f = a.blocks[b].flags[e] != 0;
  }
  const uint64_t mask = __ballot(f);
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (lane == 0 && mask) base = atomicAdd(a.counter, (int)__popcll(mask));
  base = __shfl(base, 0, 64);
  if (f) {
    const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                               __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
    a.list[base + rank] = (int32_t)g;
  }
}

__global__ void zero_kernel(int32_t* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0;
}

// ---------------------------------------------------------------------------
// Subcell FV recompute of troubled elements (hybrid/subcell.py:855-950 for
// the troubled side): lane s owns subcell s.
template <int V, typename real>
__global__ __launch_bounds__(64) void subcell_kernel(Args<real> a, Phys<real> ph) {
  __shared__ real sU[MAXN][V];
  __shared__ real sS[NSUB][V];       // subcell averages, then updated averages
  __shared__ real sG[NSUB][3][V];    // Green-Gauss gradients
  const int s = threadIdx.x;
  const int count = *a.counter;
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    int64_t e;
    const int b = locate(a, (int64_t)a.list[item], &e);
    const BlockT<real>& B = a.blocks[b];
    const TypeT<real>& T = a.types[a.btype[b]];
    const int N = T.N;
    const real* U = a.curr[b] + e * N * V;
    for (int k = s; k < N * V; k += 64) sU[k / V][k % V] = U[k];
    __syncthreads();
    real jinv[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) jinv[k] = B.jinv[e * 9 + k];
    const real det = B.detw[e];
    // J from J^-1 (3x3 inverse by cofactors): dc_phys = J dc_ref.
    real J[9];
    {
      const real* m = jinv;
      const real c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8],
                 c02 = m[3] * m[7] - m[4] * m[6];
      const real id = real(1) / (m[0] * c00 + m[1] * c01 + m[2] * c02);
      J[0] = c00 * id;
      J[1] = (m[2] * m[7] - m[1] * m[8]) * id;
      J[2] = (m[1] * m[5] - m[2] * m[4]) * id;
      J[3] = c01 * id;
      J[4] = (m[0] * m[8] - m[2] * m[6]) * id;
      J[5] = (m[2] * m[3] - m[0] * m[5]) * id;
      J[6] = c02 * id;
      J[7] = (m[1] * m[6] - m[0] * m[7]) * id;
      J[8] = (m[0] * m[4] - m[1] * m[3]) * id;
    }
    // dg_to_subcells (hybrid/subcell.py:399-403)
    real sub[V];
#pragma unroll
    for (int v = 0; v < V; ++v) sub[v] = real(0);
    for (int m = 0; m < N; ++m) {
      const real c = T.PT[m * NSUB + s];
#pragma unroll
      for (int v = 0; v < V; ++v) sub[v] += c * sU[m][v];
    }
#pragma unroll
    for (int v = 0; v < V; ++v) sS[s][v] = sub[v];
    const real vol = det * T.subvol[s];
    __syncthreads();
    const int f0 = T.int_ptr[s], f1 = T.int_ptr[s + 1];
    if (a.recon != 0) {
      // Green-Gauss subcell gradients (hybrid/subcell.py:557-590).
      real g[3][V];
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int v = 0; v < V; ++v) g[d][v] = real(0);
      for (int k = f0; k < f1; ++k) {
        const int f = T.int_list[k];
        const int o = T.int_o[f], nb = T.int_n[f];
        real av[3];
        phys_area(jinv, det, T.int_area + f * 3, av);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const real fd = real(0.5) * av[d];
#pragma unroll
          for (int v = 0; v < V; ++v) g[d][v] += fd * (sS[nb][v] - sS[o][v]);
        }
      }
      const real vd = vol > real(1e-14) ? vol : real(1e-14);
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int v = 0; v < V; ++v) sG[s][d][v] = g[d][v] / vd;
    }
    __syncthreads();
    // Internal fluxes (hybrid/subcell.py:541-688).
    real res[V];
#pragma unroll
    for (int v = 0; v < V; ++v) res[v] = real(0);
    for (int k = f0; k < f1; ++k) {
      const int f = T.int_list[k];
      const int o = T.int_o[f], nb = T.int_n[f];
      real av[3];
      phys_area(jinv, det, T.int_area + f * 3, av);
      const real area = sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
      const real n[3] = {av[0] / area, av[1] / area, av[2] / area};
      real UL[V], UR[V];
      if (a.recon == 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
          UL[v] = sS[o][v];
          UR[v] = sS[nb][v];
        }
      } else {
        const real* dr = T.int_dc + f * 3;
        real dc[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) dc[d] = J[d * 3 + 0] * dr[0] + J[d * 3 + 1] * dr[1] + J[d * 3 + 2] * dr[2];
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const real uo = sS[o][v], un = sS[nb][v];
          const real dU = un - uo;
          real duo = real(0), dun = real(0);
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            duo = duo + sG[o][d][v] * dc[d];
            dun = dun - sG[nb][d][v] * dc[d];
          }
          if (a.recon == 1) {
            const real eps = real(1e-12);
            const real ro = dU / (duo + (duo >= real(0) ? eps : -eps));
            const real rn = (-dU) / (dun + (dun >= real(0) ? eps : -eps));
            real po = ro > real(0) ? (ro * ro + real(2) * ro) / (ro * ro + ro + real(2)) : real(0);
            real pn = rn > real(0) ? (rn * rn + real(2) * rn) / (rn * rn + rn + real(2)) : real(0);
            po = po < real(0) ? real(0) : (po > real(1) ? real(1) : po);
            pn = pn < real(0) ? real(0) : (pn > real(1) ? real(1) : pn);
            UL[v] = uo + real(0.5) * po * duo;
            UR[v] = un + real(0.5) * pn * dun;
          } else {
            const real so = duo, sn = -dun;
            const real Ph = real(0.5) * (uo + un) + real(0.125) * (so - sn);
            const real Po = uo + real(0.5) * so;
            const real Pn = un - real(0.5) * sn;
            const real e = real(1e-12);
            const real sec = dU;
            const real bh = (sn - so) * (sn - so) + real(0.1) * sec * sec;
            const real bo = (sec - so) * (sec - so) + real(0.1) * sec * sec;
            const real bn = (sec - sn) * (sec - sn) + real(0.1) * sec * sec;
            const real bc = sec * sec;
            const real tau = fabs(bc - bh);
            const real q0 = tau / (bh + e), q1 = tau / (bo + e), q2 = tau / (bc + e),
                       q3 = tau / (bn + e);
            const real a0 = real(0.70) * (real(1) + q0 * q0);
            const real a1 = real(0.28) * (real(1) + q1 * q1);
            const real a2 = real(0.02) * (real(1) + q2 * q2);
            const real sL = a0 + a1 + a2;
            UL[v] = (a0 / sL) * Ph + (a1 / sL) * Po + (a2 / sL) * uo;
            const real b1 = real(0.28) * (real(1) + q3 * q3);
            const real sR = a0 + b1 + a2;
            UR[v] = (a0 / sR) * Ph + (b1 / sR) * Pn + (a2 / sR) * un;
          }
        }
      }
      admissible<V, real>(UL);
      admissible<V, real>(UR);
      real F[V];
      numflux<V, real>(UL, UR, n, ph, F);
      const real sg = (s == o) ? real(-1) : real(1);
#pragma unroll
      for (int v = 0; v < V; ++v) res[v] += sg * (F[v] * area);
    }
    // External subfaces received by this subcell.
    for (int k = T.tgt_ptr[s]; k < T.tgt_ptr[s + 1]; ++k) {
      const int j = T.tgt_list[k];
      const int lf = T.bnd_face[j];
      for (int fe = 0; fe < B.max_faces; ++fe) {
        const int32_t code = B.elem_faces[(e * B.max_faces + fe) * 2];
        if (code < 0) continue;
        if (((code >> 16) & 0xff) != lf) continue;
        const int side = (code >> 24) & 1;
        const GroupT<real>& G = a.groups[code & 0xffff];
        const int64_t i = B.elem_faces[(e * B.max_faces + fe) * 2 + 1];
        const int tab = G.pair_table[i];
        const int32_t* ids = G.pair_ids + (int64_t)tab * G.K * 2 + side;
        // Each lane walks its own matching pairs (in increasing order); the
        // flux evaluations of all lanes then run together, one round per
        // match rather than one per pair index any lane matches.
        int kk = 0;
        while (true) {
          while (kk < G.K && ids[kk * 2] != j) ++kk;
          if (kk >= G.K) break;
          real F[V];
          pair_flux<V, real>(a, ph, G, i, kk, F);
          const real sg = side == 0 ? real(-1) : real(1);
#pragma unroll
          for (int v = 0; v < V; ++v) res[v] += sg * F[v];
          ++kk;
        }
      }
    }
    // FV update, additional rhs (hybrid/subcell.py:876-886).
    const real h = a.dte[b] ? a.dte[b][e] : a.dt;
    real fv[V];
#pragma unroll
    for (int v = 0; v < V; ++v) fv[v] = sub[v] + h * res[v] / vol;
    if (a.visc[b]) {
      const real* Vv = a.visc[b] + e * N * V;
      real pv[V];
      row_dot<V, real>(T.P + s * N, Vv, N, pv);   // (row read once per lane)
#pragma unroll
      for (int v = 0; v < V; ++v) fv[v] = fv[v] + h * pv[v];
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < V; ++v) sS[s][v] = fv[v];
    __syncthreads();
    // subcells_to_dg with the Zhang-Shu bound (hybrid/subcell.py:406-439).
    real raw[V];
    if (s < N) {
#pragma unroll
      for (int v = 0; v < V; ++v) raw[v] = real(0);
      for (int t = 0; t < NSUB; ++t) {
        const real c = T.RT[t * N + s];
#pragma unroll
        for (int v = 0; v < V; ++v) raw[v] += c * sS[t][v];
      }
    }
    const real vsum = det * T.volsum;
    const real ev = vsum > real(1e-14) ? vsum : real(1e-14);
    real* out = a.work[b] + e * N * V;
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const real mean = wsum(fv[v] * vol) / ev;
      const real smin = wmin(fv[v]), smax = wmax(fv[v]);
      const real dmin = wmin(s < N ? raw[v] : mean);
      const real dmax = wmax(s < N ? raw[v] : mean);
      const real eps = real(1e-12);
      const real tmax = dmax > mean + eps ? (smax - mean) / (dmax - mean + eps) : real(1);
      const real tmin = dmin < mean - eps ? (mean - smin) / (mean - dmin + eps) : real(1);
      real th = nmin(tmax, tmin);
      th = th < real(0) ? real(0) : (th > real(1) ? real(1) : th);
      if (s < N) out[s * V + v] = mean + th * (raw[v] - mean);
    }
    __syncthreads();
  }
}

// Mean correction of non-troubled neighbours (hybrid/subcell.py:891-944).
// Each lane first tests one element (not troubled, with a troubled face
// neighbour); the wave then corrects the elements it found one at a time,
// with the face points spread over the lanes.  Nothing to do without
// troubled elements.
template <int V, typename real>
__global__ __launch_bounds__(256) void correct_kernel(Args<real> a, Phys<real> ph, int64_t total) {
  if (*a.counter == 0) return;
  const int lane = threadIdx.x & 63;
  const int64_t g0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
  bool need = false;
  if (g0 + lane < total) {
    int64_t e;
    const int b = locate(a, g0 + lane, &e);
    const BlockT<real>& B = a.blocks[b];
    if (!B.flags[e]) {
      for (int fe = 0; fe < B.max_faces && !need; ++fe) {
        const int32_t code = B.elem_faces[(e * B.max_faces + fe) * 2];
        if (code < 0) continue;
        const GroupT<real>& G = a.groups[code & 0xffff];
        if (G.neighbor_block < 0) continue;       // the boundary side is never troubled
        const int side = (code >> 24) & 1;
        const int64_t i = B.elem_faces[(e * B.max_faces + fe) * 2 + 1];
        const BlockT<real>& Bo = a.blocks[side == 0 ? G.neighbor_block : G.owner_block];
        const int64_t other = side == 0 ? G.neighbor_elem[i] : G.owner_elem[i];
        need = Bo.flags[other] != 0;
      }
    }
  }
  unsigned long long todo = __ballot(need);
  while (todo) {
    const int l = __builtin_ctzll(todo);
    todo &= todo - 1;
    int64_t e;
    const int b = locate(a, g0 + l, &e);
    const BlockT<real>& B = a.blocks[b];
    const TypeT<real>& T = a.types[a.btype[b]];
    real corr[V];
#pragma unroll
    for (int v = 0; v < V; ++v) corr[v] = real(0);
    const real h = a.dte[b] ? a.dte[b][e] : a.dt;
    const real evol = B.detw[e] * T.volsum;
    for (int fe = 0; fe < B.max_faces; ++fe) {
      const int32_t code = B.elem_faces[(e * B.max_faces + fe) * 2];
      if (code < 0) continue;
      const GroupT<real>& G = a.groups[code & 0xffff];
      if (G.neighbor_block < 0) continue;
      const int side = (code >> 24) & 1;
      const int64_t i = B.elem_faces[(e * B.max_faces + fe) * 2 + 1];
      const BlockT<real>& Bo = a.blocks[side == 0 ? G.neighbor_block : G.owner_block];
      const int64_t other = side == 0 ? G.neighbor_elem[i] : G.owner_elem[i];
      if (!Bo.flags[other]) continue;
      real nf[V], of[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        nf[v] = real(0);
        of[v] = real(0);
      }
      if (lane < G.K) pair_flux<V, real>(a, ph, G, i, lane, nf);
      if (lane < G.Qf) dg_point_flux<V, real>(a, ph, G, i, lane, of);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const real d = wsum(nf[v]) - wsum(of[v]);
        corr[v] += (side == 0 ? -h : h) * d / evol;
      }
    }
    real* out = a.work[b] + e * T.N * V;
    for (int k = lane; k < T.N * V; k += 64) out[k] = out[k] + corr[k % V];
  }
}

// ---------------------------------------------------------------------------
// Positivity (hybrid/limiter.py:69-147) of one element held in LDS, by one
// wave.  The volume mean is sum_j mw_j U_j / sum_k w_k (the quadrature of the
// interpolant, with the basis-weight products summed once per type); lane l
// holds the deviations of samples l, l + 64, ... (nodes, then the volume and
// face sample points) in registers, KS of them.  Writes the limited state
// back to sU.  sW holds the type's basis-weight products mw.
template <int V, int KS, typename real>
__device__ void positivity_wave(const Args<real>& a, const Phys<real>& ph, const TypeT<real>& T,
                                real (*sU)[V], real* sM, const real* sW, int lane) {
  const int N = T.N, NT = N + T.NS;
  {
    // the mean: lane l sums variable l % VP over the nodes j = l / VP (mod
    // 64 / VP), then the partial sums of a variable are added across lanes
    constexpr int VP = V <= 8 ? 8 : 16, G = 64 / VP;
    static_assert(V <= VP, "variables per lane group");
    const int v = lane % VP, part = lane / VP;
    real m = real(0);
    if (v < V)
      for (int j = part; j < N; j += G) m += sW[j] * sU[j][v];
#pragma unroll
    for (int o = VP; o < 64; o <<= 1) m += __shfl_xor(m, o, 64);
    if (part == 0 && v < V) sM[v] = m / T.wsum;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  real mean[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) mean[v] = sM[v];
  real mfull[V];
#pragma unroll
  for (int v = 0; v < V; ++v) mfull[v] = sM[v];
  fsg::Prim<V, real> wm = fsg::c2p<V, real>(mfull, ph.gamma);
  const real rf = nmax(a.dfloor, a.rel * wm.rho);
  const real pf = nmax(a.pfloor, a.rel * wm.p);
  // Shortcut: every sample is mean + sum_j phi_j(x_k) (U_j - mean) (the
  // basis sums to one), so |u_k - mean| <= L max_j |U_j - mean| with L the
  // type's Lebesgue constant over its samples.  When the resulting lower
  // bounds on rho and p (less a generous allowance for the rounding of the
  // sample sums and of the pressure) clear the floors, the full evaluation
  // would find theta_rho = theta_p = 1; the limited state is then
  // mean + 1 (U - mean), as below, without sampling.
  {
    real D[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const real dev = lane < N ? fabs(sU[lane][v] - mean[v]) : real(0);
      D[v] = wmax_num(dev);
    }
    const real L = T.lebesgue;
    const real eps = sizeof(real) == 4 ? real(1.2e-7) : real(2.3e-16);
    const real rnd = real(16) * real(N + 2) * eps * L;
    real b[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) b[v] = L * D[v] + rnd * (fabs(mean[v]) + D[v]);
    const real rho_lb = mean[0] - b[0];
    const real E_lb = mean[4] - b[4];
    real m2 = real(0);
#pragma unroll
    for (int d = 1; d < 4; ++d) {
      const real mu = fabs(mean[d]) + b[d];
      m2 += mu * mu;
    }
    const real ke_ub = rho_lb > real(0) ? real(0.5) * m2 / rho_lb : real(0);
    const real p_lb = (ph.gamma - real(1)) * (E_lb - ke_ub);
    const real p_tol = real(16) * eps * (ph.gamma - real(1)) * (fabs(mean[4]) + b[4] + ke_ub);
    const bool safe = rho_lb > real(0) && rho_lb >= rf && p_lb - p_tol >= pf;
    if (safe) {
      for (int k = lane; k < N * V; k += 64) {
        const int i = k / V, v = k % V;
        sU[i][v] = sM[v] + real(1) * (sU[i][v] - sM[v]);
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      return;
    }
  }
  // sample deviations (first five variables) and the density minimum; the
  // node loop is outermost so each node's values are read from LDS once for
  // all of a lane's samples (the sum over nodes keeps its order per sample)
  real d[KS][5];
#pragma unroll
  for (int i = 0; i < KS; ++i)
#pragma unroll
    for (int v = 0; v < 5; ++v) d[i][v] = real(0);
  for (int j = 0; j < N; ++j) {
    real uj[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) uj[v] = sU[j][v];
    const real* cj = T.samples + (int64_t)j * T.NS;
#pragma unroll
    for (int i = 0; i < KS; ++i) {
      const int k = lane + 64 * i;
      if (k >= N && k < NT) {
        const real c = cj[k - N];
#pragma unroll
        for (int v = 0; v < 5; ++v) d[i][v] += c * uj[v];
      }
    }
  }
  real rmin = __builtin_huge_val();
#pragma unroll
  for (int i = 0; i < KS; ++i) {
    const int k = lane + 64 * i;
    real u[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) u[v] = k < N ? sU[k < N ? k : 0][v] : (k < NT ? d[i][v] : mean[v]);
    if (k < NT) rmin = nmin(rmin, u[0]);
#pragma unroll
    for (int v = 0; v < 5; ++v) d[i][v] = u[v] - mean[v];
  }
  rmin = wmin(rmin);
  const real den = mean[0] - rmin;
  real th = rmin < rf ? (mean[0] - rf) / (den > real(1e-30) ? den : real(1e-30)) : real(1);
  const real upper = th < real(0) ? real(0) : (th > real(1) ? real(1) : th);
  auto ok = [&](real t) {
    bool good = true;
#pragma unroll
    for (int i = 0; i < KS; ++i) {
      real u[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) u[v] = mean[v] + t * d[i][v];
      good = good && (lane + 64 * i >= NT || pressure(u, ph.gamma) >= pf);
    }
    return !wany(!good);
  };
  real theta = upper;
  if (!ok(upper)) {
    real lo = real(0), hi = upper;
    for (int it = 0; it < 30; ++it) {
      const real mid = real(0.5) * (lo + hi);
      if (ok(mid))
        lo = mid;
      else
        hi = mid;
    }
    theta = lo;
  }
  for (int k = lane; k < N * V; k += 64) {
    const int i = k / V, v = k % V;
    sU[i][v] = sM[v] + theta * (sU[i][v] - sM[v]);
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

constexpr int MAXNT = 64 + 640;   // nodes + sample points of one element

// mode 0: out = positivity(src); mode 1 (stage): out = positivity(alpha base
// + beta positivity(work)) with the node admissibility reductions.
template <int V, int KS, typename real>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(sizeof(real) == 4 && KS <= 5 ? 8 : (sizeof(real) == 4 && KS == 8 ? 6 : 1))))
void positivity_kernel(Args<real> a, Phys<real> ph, int mode, int64_t g0) {
  __shared__ real sU[MAXN][V];
  __shared__ real sM[V];
  __shared__ real sW[MAXN];
  const int lane = threadIdx.x;
  int64_t e;
  const int b = locate(a, g0 + (int64_t)blockIdx.x, &e);
  const TypeT<real>& T = a.types[a.btype[b]];
  const int N = T.N;
  // mode 0: limit det_src; 1: limit work, blend with base, limit again,
  // diagnostics; 2: limit det_src, diagnostics
  const real* src = mode == 1 ? a.work[b] : a.det_src[b];
  const real mwl = lane < N ? T.mw[lane] : real(0);
  load_rows<64, 8>(&sU[0][0], src + e * N * V, N * V, lane);
  if (lane < N) sW[lane] = mwl;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  positivity_wave<V, KS, real>(a, ph, T, sU, sM, sW, lane);
  if (mode == 1) {
    const real* bs = a.base[b] + e * N * V;
    real* su = &sU[0][0];
    constexpr int CH = 8;
    const int n = N * V;
    for (int k0 = 0; k0 < n; k0 += 64 * CH) {
      real t[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const int k = k0 + lane + 64 * i;
        t[i] = bs[k < n ? k : n - 1];
      }
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const int k = k0 + lane + 64 * i;
        if (k < n) su[k] = a.alpha * t[i] + a.beta * su[k];
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    positivity_wave<V, KS, real>(a, ph, T, sU, sM, sW, lane);
  }
  real* out = a.out[b] + e * N * V;
  for (int k = lane; k < N * V; k += 64) out[k] = sU[k / V][k % V];
  if (mode != 0 && a.diag) {
    // hybrid/time_integration.py:291-312 over the nodes
    real rmin = __builtin_huge_val(), pmin = __builtin_huge_val();
    bool nonfinite = false;
    if (lane < N) {
      real W[V];
      fsg::c2p_full<V, real>(sU[lane], ph.gamma, W);
#pragma unroll
      for (int v = 0; v < V; ++v) nonfinite = nonfinite || !isnum(sU[lane][v]);
      rmin = W[0];
      pmin = W[4];
    }
    rmin = wmin(rmin);
    pmin = wmin(pmin);
    const bool nf = wany(nonfinite);
    if (lane == 0) {
      // read here, not when the wave starts: with one wave per element, half a
      // million waves reading one address up front cost more than they hid
      const fsg::DiagSeen seen = fsg::diag_seen(a.diag);
      fsg::atomic_min_nan<real>(&a.diag->rho_min, (double)rmin, seen.rho);
      fsg::atomic_min_nan<real>(&a.diag->p_min, (double)pmin, seen.p);
      if (nf) atomicAdd(&a.diag->nonfinite, 1u);
    }
  }
}

// HybridModalShockLimiter (hybrid/limiter.py:186-264), one wave per element:
// the troubled-cell sensor; on a troubled element the P1 projection of every
// variable with the element mean restored, or the mean itself where that
// projection is non-finite or its density / pressure at the samples (nodes,
// volume and face points) falls below p0_factor x the floors or p0_rel x the
// mean's value.  Untroubled elements are copied.  Writes a.work.
template <int V, typename real>
__global__ __launch_bounds__(64) void modal_kernel(Args<real> a, Phys<real> ph, real p0_factor,
                                                   real p0_rel, int64_t g0) {
  __shared__ real sU[MAXN][V];
  __shared__ real sL[MAXN][V];
  __shared__ real sm[MAXN][2];
  __shared__ real sMean[2][V];
  const int lane = threadIdx.x;
  int64_t e;
  const int b = locate(a, g0 + (int64_t)blockIdx.x, &e);
  const TypeT<real>& T = a.types[a.btype[b]];
  const int N = T.N;
  load_rows<64, 8>(&sU[0][0], a.det_src[b] + e * N * V, N * V, lane);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (lane < N) {
    fsg::Prim<V, real> pr = fsg::c2p<V, real>(sU[lane], ph.gamma);
    sm[lane][0] = pr.rho;
    sm[lane][1] = pr.p;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  // the sensor of the detection kernel (hybrid/limiter.py:164-185)
  real hi0 = 0, hi1 = 0, sc0 = 0, sc1 = 0;
  if (lane < N) {
    real l0 = 0, l1 = 0;
    for (int j = 0; j < N; ++j) {
      const real c = T.p1proj[j * N + lane];
      l0 += c * sm[j][0];
      l1 += c * sm[j][1];
    }
    const real m0 = sm[lane][0], m1 = sm[lane][1];
    hi0 = (m0 - l0) * (m0 - l0);
    hi1 = (m1 - l1) * (m1 - l1);
    sc0 = m0 * m0;
    sc1 = m1 * m1;
  }
  hi0 = gsum<64>(hi0);
  hi1 = gsum<64>(hi1);
  sc0 = gsum<64>(sc0);
  sc1 = gsum<64>(sc1);
  const real r0 = hi0 / (sc0 > real(1e-12) ? sc0 : real(1e-12));
  const real r1 = hi1 / (sc1 > real(1e-12) ? sc1 : real(1e-12));
  const bool troubled = a.sensor_on && nmax(r0, r1) > a.thr;
  real* out = a.work[b] + e * N * V;
  if (!troubled) {
    for (int k = lane; k < N * V; k += 64) out[k] = sU[k / V][k % V];
    return;
  }
  if (lane < N)
    for (int v = 0; v < V; ++v) {
      real acc = 0;
      for (int j = 0; j < N; ++j) acc += T.p1proj[j * N + lane] * sU[j][v];
      sL[lane][v] = acc;
    }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (lane < V) {
    real mo = 0, ml = 0;
    for (int j = 0; j < N; ++j) {
      mo += T.mw[j] * sU[j][lane];
      ml += T.mw[j] * sL[j][lane];
    }
    sMean[0][lane] = mo / T.wsum;
    sMean[1][lane] = ml / T.wsum;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (lane < N)
    for (int v = 0; v < V; ++v) sL[lane][v] += sMean[0][v] - sMean[1][v];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  real rmin = __builtin_huge_val(), pmin = __builtin_huge_val();
  bool fin = true;
  for (int k = lane; k < N + T.NS; k += 64) {
    real u[V];
    if (k < N) {
#pragma unroll
      for (int v = 0; v < V; ++v) u[v] = sL[k][v];
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) u[v] = real(0);
      for (int j = 0; j < N; ++j) {
        const real c = T.samples[(int64_t)j * T.NS + (k - N)];
#pragma unroll
        for (int v = 0; v < V; ++v) u[v] += c * sL[j][v];
      }
    }
#pragma unroll
    for (int v = 0; v < V; ++v) fin = fin && isnum(u[v]);
    real W[V];
    fsg::c2p_full<V, real>(u, ph.gamma, W);
    rmin = nmin(rmin, W[0]);
    pmin = nmin(pmin, W[4]);
  }
  rmin = wmin(rmin);
  pmin = wmin(pmin);
  fin = !wany(!fin);
  real Wm[V];
  fsg::c2p_full<V, real>(sMean[0], ph.gamma, Wm);
  const real md = Wm[0] > a.dfloor ? Wm[0] : a.dfloor;
  const real mp = Wm[4] > a.pfloor ? Wm[4] : a.pfloor;
  const bool severe = !fin || rmin < p0_factor * a.dfloor || pmin < p0_factor * a.pfloor ||
                      rmin < p0_rel * md || pmin < p0_rel * mp;
  for (int k = lane; k < N * V; k += 64) out[k] = severe ? sMean[0][k % V] : sL[k / V][k % V];
}

}  // namespace fsgmood

// ---------------------------------------------------------------------------
struct fsg_mood {
  fsg_context* ctx;
  fsg_mood_params params;
  int dtype;
  int64_t total;
  int32_t* counter;
  int32_t* list;
  void* d_types = nullptr;
  void* d_blocks = nullptr;
  void* d_groups = nullptr;
  std::vector<void*> d_mw;      // per-type node weights of the volume mean
  std::vector<int> block_ks;
  std::vector<int> block_type;  // element type per block    // samples per lane of the positivity kernel, per block
  int max_n = 0;                // largest node count of the element types in use
  std::vector<int64_t> offsets;
  std::vector<int32_t*> flags;
};

namespace {

using namespace fsgmood;

template <typename real>
int upload(fsg_mood* m, const fsg_mood_type* types, const fsg_mood_block* blocks,
           const fsg_mood_group* groups) {
  fsg_context* c = m->ctx;
  std::vector<TypeT<real>> tt(3);
  memset(tt.data(), 0, sizeof(TypeT<real>) * 3);
  std::vector<bool> used(3, false);
  int type_ks[3] = {1, 1, 1};
  for (const auto& b : c->blocks) used[b.element_type] = true;
  for (int t = 0; t < 3; ++t) {
    if (!used[t]) continue;
    const fsg_mood_type& s = types[t];
    if (s.N > MAXN || s.N + s.NS > MAXNT || s.N <= 0)
      return fsg_fail(-20, "MOOD tables of type %d exceed the kernel limits (N %d, NS %d)", t,
                      s.N, s.NS);
    TypeT<real>& d = tt[t];
    d.N = s.N;
    d.Q = s.Q;
    d.NS = s.NS;
    d.NI = s.NI;
    d.NB = s.NB;
    d.p1proj = static_cast<const real*>(s.p1proj);
    d.samples = static_cast<const real*>(s.samples);
    d.wq = static_cast<const real*>(s.wq);
    d.P = static_cast<const real*>(s.P);
    d.PT = static_cast<const real*>(s.PT);
    d.RT = static_cast<const real*>(s.RT);
    d.subvol = static_cast<const real*>(s.subvol);
    d.int_o = s.int_o;
    d.int_n = s.int_n;
    d.int_area = static_cast<const real*>(s.int_area);
    d.int_dc = static_cast<const real*>(s.int_dc);
    d.int_ptr = s.int_ptr;
    d.int_list = s.int_list;
    d.bnd_face = s.bnd_face;
    d.bnd_gather = s.bnd_gather;
    d.tgt_ptr = s.tgt_ptr;
    d.tgt_list = s.tgt_list;
    // host copies of the weight sums (the tables are device memory)
    std::vector<real> w(s.Q), v(NSUB);
    if (hipMemcpy(w.data(), s.wq, sizeof(real) * s.Q, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(v.data(), s.subvol, sizeof(real) * NSUB, hipMemcpyDeviceToHost) != hipSuccess)
      return fsg_fail(-21, "cannot read the MOOD weight tables");
    real ws = 0, vs = 0;
    for (real x : w) ws += x;
    for (real x : v) vs += x;
    d.wsum = ws;
    d.volsum = vs;
    // mw[j] = sum_k w_k phi_j(x_k) over the volume points, summed in double
    std::vector<real> smp((size_t)s.N * s.NS), mw(s.N);
    if (s.Q > s.NS ||
        hipMemcpy(smp.data(), s.samples, sizeof(real) * smp.size(), hipMemcpyDeviceToHost) !=
            hipSuccess)
      return fsg_fail(-21, "cannot read the MOOD sample tables");
    for (int j = 0; j < s.N; ++j) {
      double acc = 0.0;
      for (int k = 0; k < s.Q; ++k) acc += (double)w[k] * (double)smp[(size_t)j * s.NS + k];
      mw[j] = (real)acc;
    }
    // Lebesgue constant of the samples (nodes: 1), rounded up
    double leb = 1.0;
    for (int k = 0; k < s.NS; ++k) {
      double acc = 0.0;
      for (int j = 0; j < s.N; ++j) acc += fabs((double)smp[(size_t)j * s.NS + k]);
      leb = acc > leb ? acc : leb;
    }
    d.lebesgue = (real)(leb * (1.0 + 1e-6));
    void* dmw = nullptr;
    if (hipMalloc(&dmw, sizeof(real) * s.N) != hipSuccess)
      return fsg_fail(-23, "cannot allocate the MOOD mean weights");
    m->d_mw.push_back(dmw);
    if (hipMemcpy(dmw, mw.data(), sizeof(real) * s.N, hipMemcpyHostToDevice) != hipSuccess)
      return fsg_fail(-21, "cannot upload the MOOD mean weights");
    d.mw = static_cast<const real*>(dmw);
    d.KS = (s.N + s.NS + 63) / 64;
    type_ks[t] = d.KS;
    if (s.N > m->max_n) m->max_n = s.N;
  }
  const int nb = (int)c->blocks.size();
  std::vector<BlockT<real>> bb(nb);
  int64_t off = 0;
  for (int b = 0; b < nb; ++b) {
    const fsg_block_desc& cb = c->blocks[b];
    BlockT<real>& d = bb[b];
    d.type = cb.element_type;
    d.max_faces = blocks[b].max_faces;
    d.count = cb.count;
    d.offset = off;
    d.jinv = static_cast<const real*>(cb.jinv);
    d.detw = static_cast<const real*>(cb.detw);
    d.elem_faces = blocks[b].elem_faces;
    d.flags = blocks[b].flags;
    d.sensor = static_cast<real*>(blocks[b].sensor);
    m->offsets.push_back(off);
    m->block_ks.push_back(type_ks[cb.element_type]);
    m->block_type.push_back(cb.element_type);
    m->flags.push_back(blocks[b].flags);
    off += cb.count;
  }
  m->total = off;
  const int ng = (int)c->groups.size();
  std::vector<GroupT<real>> gg(ng);
  const int p = c->params.degree;
  for (int g = 0; g < ng; ++g) {
    const fsg_group_desc& s = c->groups[g];
    GroupT<real>& d = gg[g];
    memset(&d, 0, sizeof d);
    d.face_type = s.face_type;
    d.sub = s.sub;
    d.owner_block = s.owner_block;
    d.neighbor_block = s.neighbor_block;
    d.K = groups[g].K;
    d.Qf = fsg::n_face_q(s.face_type, p);
    d.NFo = fsg::n_face_nodes(s.face_type, s.sub, p);
    d.NFn = s.neighbor_block >= 0 ? fsg::n_face_nodes(s.face_type, 0, p) : 0;
    const int No = fsg::n_nodes(c->blocks[s.owner_block].element_type, p);
    d.TSO = 2 * d.Qf * d.NFo + d.Qf * No * 3;
    if (s.neighbor_block >= 0) {
      const int Nn = fsg::n_nodes(c->blocks[s.neighbor_block].element_type, p);
      d.TSN = 2 * d.Qf * d.NFn + d.Qf * Nn * 3;
    }
    if (d.K > 64 || d.Qf > 64)
      return fsg_fail(-22, "group %d: %d subface pairs / %d points exceed one wave", g, d.K, d.Qf);
    d.owner_elem = s.owner_elem;
    d.neighbor_elem = s.neighbor_elem;
    d.geom = static_cast<const real*>(s.geom);
    d.fw = static_cast<const real*>(s.face_weights);
    d.owner_table = s.owner_table;
    d.owner_tables = static_cast<const real*>(s.owner_tables);
    d.owner_sup = s.owner_sup;
    d.neighbor_table = s.neighbor_table;
    d.neighbor_tables = static_cast<const real*>(s.neighbor_tables);
    d.neighbor_sup = s.neighbor_sup;
    d.bflags = s.bflags;
    d.pair_table = groups[g].pair_table;
    d.pair_ids = groups[g].pair_ids;
    d.pair_geom = static_cast<const real*>(groups[g].pair_geom);
    d.basis_o = static_cast<const real*>(groups[g].basis_o);
    d.basis_n = static_cast<const real*>(groups[g].basis_n);
  }
  size_t st = sizeof(TypeT<real>) * 3, sb = sizeof(BlockT<real>) * nb,
         sg = sizeof(GroupT<real>) * (ng > 0 ? ng : 1);
  if (hipMalloc(&m->d_types, st) != hipSuccess || hipMalloc(&m->d_blocks, sb) != hipSuccess ||
      hipMalloc(&m->d_groups, sg) != hipSuccess)
    return fsg_fail(-23, "cannot allocate the MOOD descriptor tables");
  if (hipMemcpy(m->d_types, tt.data(), st, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_blocks, bb.data(), sb, hipMemcpyHostToDevice) != hipSuccess ||
      (ng && hipMemcpy(m->d_groups, gg.data(), sizeof(GroupT<real>) * ng,
                       hipMemcpyHostToDevice) != hipSuccess))
    return fsg_fail(-23, "cannot upload the MOOD descriptor tables");
  return 0;
}

template <typename real>
Args<real> base_args(fsg_mood* m) {
  Args<real> a;
  memset(&a, 0, sizeof a);
  a.types = static_cast<const TypeT<real>*>(m->d_types);
  a.blocks = static_cast<const BlockT<real>*>(m->d_blocks);
  a.groups = static_cast<const GroupT<real>*>(m->d_groups);
  a.nblocks = (int)m->ctx->blocks.size();
  for (int b = 0; b < a.nblocks; ++b) {
    a.boff[b] = m->offsets[b];
    a.btype[b] = m->block_type[b];
  }
  a.boff[a.nblocks] = m->total;
  a.counter = m->counter;
  a.list = m->list;
  a.thr = real(m->params.threshold);
  a.dfloor = real(m->params.density_floor);
  a.pfloor = real(m->params.pressure_floor);
  a.rel = real(m->params.relative_floor);
  a.recon = m->params.reconstruction;
  a.sensor_on = m->params.sensor_on;
  return a;
}

template <typename real>
Phys<real> phys_of(fsg_mood* m, const fsg_boundary* bnd) {
  const fsg_params& p = m->ctx->params;
  Phys<real> ph;
  memset(&ph, 0, sizeof ph);
  ph.gamma = real(p.gamma);
  ph.gas_constant = real(p.gas_constant);
  ph.mu = real(p.dynamic_viscosity);
  ph.kappa = real(p.thermal_conductivity);
  ph.eta = real(p.br2_penalty);
  ph.rusanov = p.inviscid_flux == 1;
  if (bnd) {
    ph.has_boundary = bnd->has_boundary;
    ph.has_wall_temperature = bnd->has_wall_temperature;
    ph.wall_temperature = real(bnd->wall_temperature);
    for (int k = 0; k < 16; ++k) ph.farfield[k] = real(bnd->farfield[k]);
  }
  return ph;
}

template <typename T>
void set_ptrs(T* dst, const void* const* src, int n) {
  for (int b = 0; b < n; ++b) dst[b] = src ? (T)(src[b]) : nullptr;
}

int launched(const char* what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fsg_fail(-1, "%s launch failed: %s", what, hipGetErrorString(e));
}

constexpr unsigned SUBCELL_GRID = 8192;

// detect (+ compaction)
template <int V, typename real>
int run_detect(fsg_mood* m, Args<real>& a, const Phys<real>& ph, hipStream_t s) {
  const unsigned g256 = (unsigned)((m->total + 255) / 256);
  if (m->max_n <= MAXN / 2) {   // two elements per wave (p <= 2)
    const unsigned g8 = (unsigned)((m->total + 7) / 8);
    hipLaunchKernelGGL((detect_kernel<V, real, 2>), dim3(g8), dim3(256), 0, s, a, ph, m->total);
  } else {
    const unsigned g4 = (unsigned)((m->total + 3) / 4);
    hipLaunchKernelGGL((detect_kernel<V, real, 1>), dim3(g4), dim3(256), 0, s, a, ph, m->total);
  }
  if (int rc = launched("MOOD detect")) return rc;
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, m->counter);
  hipLaunchKernelGGL((compact_kernel<real>), dim3(g256), dim3(256), 0, s, a, m->total);
  return launched("MOOD compaction");
}

// One launch per element block, each with the sample count per lane of its
// element type (hexahedra and wedges need 5 at p = 2 where tetrahedra need
// 8: the deviation registers, and with them the occupancy, follow the block).
template <int V, typename real>
void launch_positivity(fsg_mood* m, const Args<real>& a, const Phys<real>& ph, int mode,
                       hipStream_t s) {
  for (size_t k = 0; k < m->offsets.size(); ++k) {
    const int64_t count = (k + 1 < m->offsets.size() ? m->offsets[k + 1] : m->total) - m->offsets[k];
    if (count <= 0) continue;
    const dim3 g((unsigned)count), b(64);
    const int64_t g0 = m->offsets[k];
    const int ks = m->block_ks[k];
    if (ks <= 2)
      hipLaunchKernelGGL((positivity_kernel<V, 2, real>), g, b, 0, s, a, ph, mode, g0);
    else if (ks <= 4)
      hipLaunchKernelGGL((positivity_kernel<V, 4, real>), g, b, 0, s, a, ph, mode, g0);
    else if (ks <= 5)
      hipLaunchKernelGGL((positivity_kernel<V, 5, real>), g, b, 0, s, a, ph, mode, g0);
    else if (ks <= 8)
      hipLaunchKernelGGL((positivity_kernel<V, 8, real>), g, b, 0, s, a, ph, mode, g0);
    else
      hipLaunchKernelGGL((positivity_kernel<V, 11, real>), g, b, 0, s, a, ph, mode, g0);
  }
}

template <int V, typename real>
int run_recompute_core(fsg_mood* m, Args<real>& a, const Phys<real>& ph, hipStream_t s) {
  const unsigned g256 = (unsigned)((m->total + 255) / 256);
  const unsigned gs = (unsigned)(m->total < SUBCELL_GRID ? m->total : SUBCELL_GRID);
  hipLaunchKernelGGL((subcell_kernel<V, real>), dim3(gs), dim3(64), 0, s, a, ph);
  if (int rc = launched("MOOD subcell")) return rc;
  hipLaunchKernelGGL((correct_kernel<V, real>), dim3(g256), dim3(256), 0, s, a, ph, m->total);
  return launched("MOOD correction");
}

template <int V, typename real>
int troubled_impl(fsg_mood* m, const void* const* U, hipStream_t s) {
  Args<real> a = base_args<real>(m);
  set_ptrs(a.det_src, U, a.nblocks);
  Phys<real> ph = phys_of<real>(m, nullptr);
  return run_detect<V, real>(m, a, ph, s);
}

template <int V, typename real>
int modal_impl(fsg_mood* m, const void* const* U, void* const* out, double p0_factor,
               double p0_rel, fsg_stage_diag* diag, hipStream_t s) {
  Args<real> a = base_args<real>(m);
  set_ptrs(a.det_src, U, a.nblocks);
  set_ptrs(a.work, (const void* const*)out, a.nblocks);
  Phys<real> ph = phys_of<real>(m, nullptr);
  if (a.sensor_on) {
    for (size_t k = 0; k < m->offsets.size(); ++k) {
      const int64_t count =
          (k + 1 < m->offsets.size() ? m->offsets[k + 1] : m->total) - m->offsets[k];
      if (count <= 0) continue;
      hipLaunchKernelGGL((modal_kernel<V, real>), dim3((unsigned)count), dim3(64), 0, s, a, ph,
                         real(p0_factor), real(p0_rel), m->offsets[k]);
    }
    if (int rc = launched("modal limiter")) return rc;
    set_ptrs(a.det_src, (const void* const*)out, a.nblocks);
  }
  // then HybridPositivityLimiter (hybrid/limiter.py:264), with the
  // admissibility of the result into diag when given
  set_ptrs(a.out, (const void* const*)out, a.nblocks);
  a.diag = reinterpret_cast<StageDiag*>(diag);
  launch_positivity<V, real>(m, a, ph, diag ? 2 : 0, s);
  return launched("positivity");
}

template <int V, typename real>
int positivity_impl(fsg_mood* m, const void* const* U, void* const* out, hipStream_t s) {
  Args<real> a = base_args<real>(m);
  set_ptrs(a.det_src, U, a.nblocks);
  set_ptrs(a.out, (const void* const*)out, a.nblocks);
  Phys<real> ph = phys_of<real>(m, nullptr);
  launch_positivity<V, real>(m, a, ph, 0, s);
  return launched("positivity");
}

template <int V, typename real>
int recompute_impl(fsg_mood* m, const void* const* states, const void* const* cand,
                   const void* const* detect, const void* const* visc, double dt,
                   const void* const* dt_elem, int use_flags, const fsg_boundary* bnd,
                   void* const* out, hipStream_t s) {
  Args<real> a = base_args<real>(m);
  const int nb = a.nblocks;
  Phys<real> ph = phys_of<real>(m, bnd);
  set_ptrs(a.curr, states, nb);
  set_ptrs(a.visc, visc, nb);
  set_ptrs(a.dte, dt_elem, nb);
  a.dt = real(dt);
  // the corrected candidate is assembled in out, then limited in place
  for (int b = 0; b < nb; ++b) {
    const fsg_block_desc& cb = m->ctx->blocks[b];
    const size_t bytes = sizeof(real) * (size_t)cb.count *
                         fsg::n_nodes(cb.element_type, m->ctx->params.degree) * m->ctx->params.nvars;
    if (out[b] != cand[b] &&
        hipMemcpyAsync(out[b], cand[b], bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return fsg_fail(-1, "copy of the DG candidate failed");
    a.work[b] = static_cast<real*>(out[b]);
    a.out[b] = static_cast<real*>(out[b]);
    a.det_src[b] = static_cast<const real*>(detect ? detect[b] : cand[b]);
  }
  int rc;
  if (!use_flags) {
    if ((rc = run_detect<V, real>(m, a, ph, s))) return rc;
  } else {
    const unsigned g256 = (unsigned)((m->total + 255) / 256);
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, m->counter);
    hipLaunchKernelGGL((compact_kernel<real>), dim3(g256), dim3(256), 0, s, a, m->total);
    if ((rc = launched("MOOD compaction"))) return rc;
  }
  if ((rc = run_recompute_core<V, real>(m, a, ph, s))) return rc;
  for (int b = 0; b < nb; ++b) a.det_src[b] = a.work[b];
  launch_positivity<V, real>(m, a, ph, 0, s);
  return launched("positivity");
}

template <int V, typename real>
int stage_impl(fsg_mood* m, const void* const* base, const void* const* curr,
               const void* const* inv, const void* const* visc, void* const* work,
               void* const* next, double dt, const void* const* dt_elem, double alpha,
               double beta, const fsg_boundary* bnd, fsg_stage_diag* diag, hipStream_t s) {
  Args<real> a = base_args<real>(m);
  const int nb = a.nblocks;
  Phys<real> ph = phys_of<real>(m, bnd);
  set_ptrs(a.base, base, nb);
  set_ptrs(a.curr, curr, nb);
  set_ptrs(a.inv, inv, nb);
  set_ptrs(a.visc, visc, nb);
  set_ptrs(a.dte, dt_elem, nb);
  set_ptrs(a.work, (const void* const*)work, nb);
  set_ptrs(a.out, (const void* const*)next, nb);
  a.dt = real(dt);
  a.alpha = real(alpha);
  a.beta = real(beta);
  a.stage = 1;
  a.diag = reinterpret_cast<StageDiag*>(diag);
  int rc;
  if ((rc = run_detect<V, real>(m, a, ph, s))) return rc;
  if ((rc = run_recompute_core<V, real>(m, a, ph, s))) return rc;
  launch_positivity<V, real>(m, a, ph, 1, s);
  return launched("positivity");
}

// The stage from one residual pass (total = inv + visc evaluated together):
// detection on the total, then the viscous rhs on the troubled elements only
// (the subcell recompute reads it nowhere else).
template <int V, typename real>
int stage_fused_impl(fsg_mood* m, const void* const* base, const void* const* curr,
                     const void* const* total, void* const* visc, void* const* work,
                     void* const* next, double dt, const void* const* dt_elem, double alpha,
                     double beta, const fsg_boundary* bnd, fsg_stage_diag* diag, hipStream_t s) {
  Args<real> a = base_args<real>(m);
  const int nb = a.nblocks;
  Phys<real> ph = phys_of<real>(m, bnd);
  set_ptrs(a.base, base, nb);
  set_ptrs(a.curr, curr, nb);
  set_ptrs(a.inv, total, nb);
  set_ptrs(a.dte, dt_elem, nb);
  set_ptrs(a.work, (const void* const*)work, nb);
  set_ptrs(a.out, (const void* const*)next, nb);
  a.dt = real(dt);
  a.alpha = real(alpha);
  a.beta = real(beta);
  a.stage = 1;
  a.diag = reinterpret_cast<StageDiag*>(diag);
  int rc;
  if ((rc = run_detect<V, real>(m, a, ph, s))) return rc;
  if (visc) {
    const int32_t* mask[MAXB];
    for (int b = 0; b < nb; ++b) mask[b] = m->flags[b];
    if ((rc = fsg_residual_masked(m->ctx, curr, visc, FSG_MODE_RHS, FSG_TERM_VISCOUS, mask, bnd,
                                  s)))
      return rc;
    set_ptrs(a.visc, (const void* const*)visc, nb);
  }
  if ((rc = run_recompute_core<V, real>(m, a, ph, s))) return rc;
  launch_positivity<V, real>(m, a, ph, 1, s);
  return launched("positivity");
}

#define FSG_MOOD_DISPATCH(call, args)                                        \
  do {                                                                       \
    const int nv = m->ctx->params.nvars;                                     \
    if (m->dtype == FSG_F64) {                                               \
      if (nv == 5) return call<5, double> args;                              \
      if (nv == 16) return call<16, double> args;                            \
    } else {                                                                 \
      if (nv == 5) return call<5, float> args;                               \
      if (nv == 16) return call<16, float> args;                             \
    }                                                                        \
    return fsg_fail(-4, "MOOD: unsupported nvars %d", nv);                   \
  } while (0)

}  // namespace

extern "C" {

int fsg_mood_create(fsg_context* ctx, const fsg_mood_params* params, const fsg_mood_type* types,
                    const fsg_mood_block* blocks, const fsg_mood_group* groups, int32_t* counter,
                    int32_t* list, fsg_mood** out) {
  if (!ctx || !params || !types || !blocks || !counter || !list || !out)
    return fsg_fail(-10, "null argument");
  if (ctx->blocks.size() > (size_t)MAXB) return fsg_fail(-11, "too many element blocks");
  if (!ctx->groups.empty() && !groups) return fsg_fail(-10, "null group tables");
  fsg_mood* m = new fsg_mood();
  m->ctx = ctx;
  m->params = *params;
  m->dtype = ctx->params.dtype;
  m->counter = counter;
  m->list = list;
  int rc = m->dtype == FSG_F64 ? upload<double>(m, types, blocks, groups)
                               : upload<float>(m, types, blocks, groups);
  if (rc) {
    fsg_mood_destroy(m);
    return rc;
  }
  *out = m;
  return 0;
}

void fsg_mood_destroy(fsg_mood* m) {
  if (!m) return;
  (void)hipFree(m->d_types);
  (void)hipFree(m->d_blocks);
  (void)hipFree(m->d_groups);
  for (void* p : m->d_mw) (void)hipFree(p);
  delete m;
}

int fsg_mood_troubled(fsg_mood* m, const void* const* U, void* stream) {
  if (!m || !U) return fsg_fail(-10, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  FSG_MOOD_DISPATCH(troubled_impl, (m, U, s));
}

int fsg_mood_positivity(fsg_mood* m, const void* const* U, void* const* out, void* stream) {
  if (!m || !U || !out) return fsg_fail(-10, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  FSG_MOOD_DISPATCH(positivity_impl, (m, U, out, s));
}

int fsg_mood_modal(fsg_mood* m, const void* const* U, void* const* out, double p0_factor,
                   double p0_relative, fsg_stage_diag* diag, void* stream) {
  if (!m || !U || !out) return fsg_fail(-10, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  FSG_MOOD_DISPATCH(modal_impl, (m, U, out, p0_factor, p0_relative, diag, s));
}

int fsg_mood_recompute(fsg_mood* m, const void* const* states, const void* const* dg_candidate,
                       const void* const* detect, const void* const* visc, double dt,
                       const void* const* dt_elem, int32_t use_flags,
                       const fsg_boundary* boundary, void* const* out, void* stream) {
  if (!m || !states || !dg_candidate || !out) return fsg_fail(-10, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  FSG_MOOD_DISPATCH(recompute_impl, (m, states, dg_candidate, detect, visc, dt, dt_elem,
                                     use_flags, boundary, out, s));
}

int fsg_mood_stage(fsg_mood* m, const void* const* base, const void* const* curr,
                   const void* const* inv, const void* const* visc, void* const* work,
                   void* const* next, double dt, const void* const* dt_elem, double alpha,
                   double beta, const fsg_boundary* boundary, fsg_stage_diag* diag,
                   void* stream) {
  if (!m || !base || !curr || !inv || !work || !next) return fsg_fail(-10, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  FSG_MOOD_DISPATCH(stage_impl, (m, base, curr, inv, visc, work, next, dt, dt_elem, alpha,
                                 beta, boundary, diag, s));
}

int fsg_mood_stage_fused(fsg_mood* m, const void* const* base, const void* const* curr,
                         const void* const* total, void* const* visc, void* const* work,
                         void* const* next, double dt, const void* const* dt_elem, double alpha,
                         double beta, const fsg_boundary* boundary, fsg_stage_diag* diag,
                         void* stream) {
  if (!m || !base || !curr || !work || !next) return fsg_fail(-10, "null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  FSG_MOOD_DISPATCH(stage_fused_impl, (m, base, curr, total, visc, work, next, dt, dt_elem,
                                       alpha, beta, boundary, diag, s));
}

int fsg_mood_count(fsg_mood* m, int64_t* count, void* stream) {
  if (!m || !count) return fsg_fail(-10, "null argument");
  int32_t c = 0;
  if (hipMemcpyAsync(&c, m->counter, sizeof c, hipMemcpyDeviceToHost,
                     static_cast<hipStream_t>(stream)) != hipSuccess ||
      hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess)
    return fsg_fail(-1, "cannot read the troubled count");
  *count = c;
  return 0;
}

}  // extern "C"

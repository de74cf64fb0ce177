// Hybrid DG residual kernels for gfx950.
//
// One RK stage (or one residual/rhs evaluation) is two launches per element
// type / face group:
//
//  1. face_kernel: one wavefront per face.  It gathers both cells' nodal
//     states into LDS, evaluates anchored Euler traces and HLLC
//     (hybrid/spatial.py:142-232), the BR2 jump, lift right-hand sides and
//     face gradients (hybrid/navier_stokes.py:138-188, 236-323), and writes
//     each side's surface contribution and lift right-hand side on that
//     side's face-support nodes only (the basis of a node off the face is
//     zero there).
//  2. element_kernel: one workgroup per element.  It gathers its face-side
//     records (a fixed CSR, no atomics), forms the BR2 lift, evaluates the
//     Euler and viscous volume integrals (hybrid/spatial.py:47-72,
//     hybrid/navier_stokes.py:215-234), applies the sliver-scaled inverse
//     mass matrix (hybrid/geometry.py:181-192) and, in stage mode, the SSP-RK
//     combination and admissibility reductions of
//     hybrid/time_integration.py:264-312, 404-439.
#pragma once
#include "fsg_lanes.h"
#include "fsg_mfma.h"
#include "fsg_physics.h"

namespace fsg {

struct StageDiag {         // device-side admissibility diagnostics
  double rho_min;
  double p_min;
  unsigned int nonfinite;
  unsigned int pad;
};

template <typename real>
struct ElemArgs {
  const real* U;          // (E, N, V) current state
  const real* Ubase;      // (E, N, V) stage base state (stage mode)
  real* out;              // (E, N, V) residual / rhs / next state
  const real* Bt;         // (N, Q)   basis at volume points, transposed
  const real* Dt;         // (N, 3, Q) reference gradients, q fastest
  const real* D;          // (Q, 3, N) reference gradients, node fastest
  const real* wq;         // (Q)
  const real* minv;       // (N, N) reference inverse mass
  const real* minvT;      // (N, N) its transpose
  const real* sf1d;       // hexahedra p >= 2: B1, D1, M1^-1 (or null: dense operators)
  const real* jinv;       // (E, 9)  [r*3+d] = d xi_r / d x_d
  const real* scale;      // (E) min(1, V/1e-4) / |det J|
  const real* detw;       // (E) |det J|
  const real* width;      // (E) WALE filter width (read only with WALE on)
  const int64_t* rec_ptr; // (E+1)
  const int64_t* rec_off; // face-output offsets (reals)
  const int32_t* rec_sup; // offsets into the support-index array
  const int32_t* rec_nf;
  const int32_t* support;
  const real* face_out;
  const real* dt_elem;    // (E) per-element dt or nullptr
  real dt, alpha, beta;
  int mode;               // 0 residual, 1 rhs, 2 stage
  int terms;
  StageDiag* diag;
  const int32_t* mask;    // (E) evaluate only elements with mask != 0, or null: all
  int64_t count;          // elements of the block
};

template <typename real>
struct FaceArgs {
  const real* Uo;          // owner block state
  const real* Un;          // neighbour block state
  const int32_t* owner_elem;
  const int32_t* neighbor_elem;
  const real* geom;        // (F, 4) normal, area
  const int32_t* owner_table;
  const int32_t* neighbor_table;
  const real* owner_tables;
  const real* neighbor_tables;
  const int32_t* owner_sup;     // (tables, FO)
  const int32_t* neighbor_sup;  // (tables, FN)
  const real* fw;          // (Qf) normalised face weights x reference measure
  const int32_t* bflags;   // (F) boundary flags
  const int32_t* owner_sf;    // hexahedral quad sides: per-table sum-factorisation map, or null
  const int32_t* neighbor_sf;
  const real* face_sf1d;      // 1-D factors at the face Gauss points
  const real* jinv_o;
  const real* scale_o;
  const real* jinv_n;
  const real* scale_n;
  const real* width_o;     // WALE filter widths of the owner / neighbour block
  const real* width_n;
  real* out_o;             // (F, FO, V+12)
  real* out_n;             // (F, FN, V+12)
  int64_t count;
  int terms;
  const int32_t* mask_o;   // owner / neighbour element masks: evaluate only faces
  const int32_t* mask_n;   // with a marked side, or null: all faces
};

// The two minima of the stage diagnostics as they stand, read coherently
// (agent scope: a plain load may see a stale cached line).  A stored minimum
// only moves down while a launch runs (NaN sits below every number), so a
// value that does not beat an earlier read cannot beat the current minimum:
// kernels read the minima when they start, where the read's latency hides
// behind their work, and hand them to atomic_min_nan at the end.
struct DiagSeen {
  double rho, p;
};
__device__ __forceinline__ DiagSeen diag_seen(StageDiag* d) {
  DiagSeen s;
  s.rho = __longlong_as_double(static_cast<long long>(__hip_atomic_load(
      reinterpret_cast<unsigned long long*>(&d->rho_min), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
  s.p = __longlong_as_double(static_cast<long long>(__hip_atomic_load(
      reinterpret_cast<unsigned long long*>(&d->p_min), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
  return s;
}

// NaN-propagating atomic minimum, starting from `seen`, an earlier coherent
// read of *addr: almost every caller finds the minimum already below its
// value and returns without a read-modify-write.
template <typename real>
__device__ __forceinline__ void atomic_min_nan(double* addr, double v, double seen) {
  unsigned long long* p = reinterpret_cast<unsigned long long*>(addr);
  unsigned long long old = static_cast<unsigned long long>(__double_as_longlong(seen));
  unsigned long long assumed;
  do {
    assumed = old;
    double cur = __longlong_as_double(static_cast<long long>(assumed));
    bool take = (v != v) ? (cur == cur) : (v < cur);
    if (!take) return;
    old = atomicCAS(p, assumed, static_cast<unsigned long long>(__double_as_longlong(v)));
  } while (old != assumed);
}

__device__ __forceinline__ double nan_min(double a, double b) {
  if (a != a || b != b) return __longlong_as_double(0x7ff8000000000000LL);
  return a < b ? a : b;
}

// Lanes per node in the element kernel's divergence: the largest power of
// two DS <= 64 with N DS <= NT.
__host__ __device__ constexpr int div_split(int N, int NT) {
  int d = 1;
  while (d < 64 && N * d * 2 <= NT) d *= 2;
  return d;
}

// The items (elements or faces) a workgroup evaluates.  A full pass has one
// item per workgroup, blockIdx.x.  A restricted pass (a mask is given: the
// viscous rhs of MOOD-troubled elements only; a separate instantiation, so
// that the full pass keeps its registers) runs on a capped grid that
// strides over chunks of 64 items: each lane tests one item, a ballot
// collects the marked ones and the workgroup evaluates those in turn.  With
// nothing marked such a pass costs one mask read per item instead of one
// workgroup launch per item.  Every wave of a workgroup forms the same
// ballot, so the items are uniform across the workgroup (its barriers are
// reached by all of its waves).
constexpr int64_t RESTRICTED_GRID = 2048;

__host__ __device__ constexpr int64_t pass_grid(int64_t count, bool restricted) {
  return restricted ? ((count + 63) / 64 < RESTRICTED_GRID ? (count + 63) / 64 : RESTRICTED_GRID)
                    : count;
}

template <bool RESTRICTED, class Marked, class Body>
__device__ __forceinline__ void for_each_item(int64_t count, Marked marked, Body body) {
  if constexpr (!RESTRICTED) {
    body((int64_t)blockIdx.x);
  } else {
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * 64;
    for (int64_t c0 = (int64_t)blockIdx.x * 64; c0 < count; c0 += stride) {
      const int64_t i = c0 + lane;
      uint64_t bits = __ballot(i < count && marked(i));
      while (bits) {
        const int j = __builtin_ctzll(bits);
        bits &= bits - 1;
        body(c0 + j);
        // the next item's staging must not overwrite LDS that a slower wave
        // still reads at the end of this one
        __syncthreads();
      }
    }
  }
}

__device__ __forceinline__ int64_t readlane64(int64_t x, int j) {
  const unsigned lo = __builtin_amdgcn_readlane((int)(uint32_t)x, j);
  const unsigned hi = __builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), j);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Face-side records of element e: surface terms into sR [N][V] and, with the
// viscous terms, lift right-hand sides into sb [N][12]; record by record, a
// barrier after each (records of one element share edge and corner nodes),
// so the sums keep the record order.  The element's record descriptors are
// read 64 at a time (lane j of every wave holds record j, broadcast with
// readlane) and the values of the next record are loaded while the current
// one is added: a record costs about one memory round trip instead of three
// dependent ones.  A thread holds RI items of a record, enough for a face of
// the element; items past RI NT (none for such faces) are read directly.
template <int NT, int RI, int V, typename real>
__device__ __forceinline__ void add_face_records(const ElemArgs<real>& a, int64_t e, bool visc,
                                                 real* sR, real* sb, int tid) {
  constexpr int W = V + 12;
  const int width = visc ? W : V;
  const int64_t r0 = a.rec_ptr[e], r1 = a.rec_ptr[e + 1];
  const int lt = tid & 63;
  for (int64_t rb = r0; rb < r1; rb += 64) {
    const int nr = (int)(r1 - rb < 64 ? r1 - rb : 64);
    int64_t moff = 0;
    int msup = 0, mnf = 0;
    if (lt < nr) {
      moff = a.rec_off[rb + lt];
      msup = a.rec_sup[rb + lt];
      mnf = a.rec_nf[rb + lt];
    }
    auto fetch = [&](int j, real (&v)[RI], int (&nd)[RI], int (&cl)[RI]) {
      const real* rec = a.face_out + readlane64(moff, j);
      const int32_t* sup = a.support + __builtin_amdgcn_readlane(msup, j);
      const int cnt = __builtin_amdgcn_readlane(mnf, j) * width;
#pragma unroll
      for (int i = 0; i < RI; ++i) {
        const int k = tid + NT * i;
        nd[i] = -1;
        cl[i] = 0;
        v[i] = real(0);
        if (k < cnt) {
          const int kk = k / width, c = k - kk * width;
          nd[i] = sup[kk];
          cl[i] = c;
          v[i] = rec[kk * W + c];
        }
      }
    };
    auto add = [&](int j, const real (&v)[RI], const int (&nd)[RI], const int (&cl)[RI]) {
#pragma unroll
      for (int i = 0; i < RI; ++i) {
        if (nd[i] < 0) continue;
        if (cl[i] < V)
          sR[nd[i] * V + cl[i]] += v[i];
        else
          sb[nd[i] * 12 + (cl[i] - V)] += v[i];
      }
      const int cnt = __builtin_amdgcn_readlane(mnf, j) * width;
      if (cnt > RI * NT) {
        const real* rec = a.face_out + readlane64(moff, j);
        const int32_t* sup = a.support + __builtin_amdgcn_readlane(msup, j);
        for (int k = RI * NT + tid; k < cnt; k += NT) {
          const int kk = k / width, c = k - kk * width;
          const int node = sup[kk];
          const real val = rec[kk * W + c];
          if (c < V)
            sR[node * V + c] += val;
          else
            sb[node * 12 + (c - V)] += val;
        }
      }
    };
    real va[RI], vb[RI];
    int na[RI], nb[RI], ca[RI], cb[RI];
    fetch(0, va, na, ca);
    for (int j = 0; j < nr; j += 2) {
      if (j + 1 < nr) fetch(j + 1, vb, nb, cb);
      add(j, va, na, ca);
      __syncthreads();
      if (j + 1 < nr) {
        if (j + 2 < nr) fetch(j + 2, va, na, ca);
        add(j + 1, vb, nb, cb);
        __syncthreads();
      }
    }
  }
}

// The dense element kernels keep one element per workgroup also in a
// restricted pass (an early exit on unmarked elements): a restricted
// instantiation of their body changed the full pass's register allocation
// (119 -> 130 VGPRs for tetrahedra in FP32, below four waves per SIMD).
template <int T, int P, int V, typename real, int NT>
__global__ __launch_bounds__(NT) void element_kernel(ElemArgs<real> a, Phys<real> ph) {
  constexpr int N = n_nodes(T, P);
  constexpr int Q = n_vol_q(T, P);
  constexpr int W = V + 12;
  __shared__ real sU[N * V];
  __shared__ real sR[N * V];
  __shared__ real sv[N * 4];
  __shared__ real sb[N * 12];
  __shared__ real sL[N * 12];
  // per point: the pulled-back flow fluxes [3][5]; with species also w (J_r .
  // u) [3], w (J_r . u_a) [3] and the species interpolants [V - 5].  The
  // row stride is odd, so that the lanes of a wave (one point each) write
  // and read distinct banks (21 + 11 = 32 words would put all on one).
  constexpr int SW0 = V > 5 ? 15 + 6 + (V - 5) : 3 * V;
  constexpr int SW = SW0 | 1;
  __shared__ real sF[Q * SW];
  __shared__ double red[2][NT / 64];
  __shared__ unsigned int redn[NT / 64];

  const int tid = threadIdx.x;
  const int64_t e = blockIdx.x;
  if (a.mask && a.mask[e] == 0) return;   // whole workgroup: no barrier passed yet
  DiagSeen seen{};
  if (a.mode == 2 && a.diag && tid == 0) seen = diag_seen(a.diag);
  const bool euler = a.terms & TERM_EULER;
  const bool visc = a.terms & TERM_VISCOUS;
  const real gamma = ph.gamma;

  const real* Ue = a.U + e * (N * V);
  {
    // every load of the element state in flight before the first is used
    constexpr int IU = (N * V + NT - 1) / NT;
    real t[IU];
#pragma unroll
    for (int i = 0; i < IU; ++i) {
      const int k = tid + NT * i;
      t[i] = Ue[k < N * V ? k : N * V - 1];
    }
#pragma unroll
    for (int i = 0; i < IU; ++i) {
      const int k = tid + NT * i;
      if (k < N * V) {
        sU[k] = t[i];
        sR[k] = real(0);
      }
    }
  }
  for (int k = tid; k < N * 12; k += NT) sb[k] = real(0);
  __syncthreads();
  if (visc)
    for (int i = tid; i < N; i += NT)
      viscous_variables<V, real>(sU + i * V, gamma, ph.gas_constant, sv + i * 4);

  // Face-side records: surface terms into sR, lift right-hand sides into sb.
  add_face_records<NT, ((P + 1) * (P + 1) * W + NT - 1) / NT, V, real>(a, e, visc, sR, sb, tid);

  const real s = a.scale[e];
  const real dw = a.detw[e];
  real J[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) J[k] = a.jinv[e * 9 + k];

  // BR2 global lift r = M_phys^-1 b  (hybrid/navier_stokes.py:157-186): two
  // threads per node, six columns each, so each row of M^-1 is read once per
  // thread rather than once per output.
  if (visc) {
    for (int k = tid; k < N * 2; k += NT) {
      const int i = k >> 1, c0 = (k & 1) * 6;
      real acc[6] = {0, 0, 0, 0, 0, 0};
      for (int j = 0; j < N; ++j) {
        const real m = a.minv[i * N + j];
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] += m * sb[j * 12 + c0 + c];
      }
#pragma unroll
      for (int c = 0; c < 6; ++c) sL[i * 12 + c0 + c] = s * acc[c];
    }
  }
  __syncthreads();

  // Volume quadrature points.  The flux differences dF are held for the five
  // flow columns only; a species column's Euler flux is U_s u_d, so its
  // pulled-back difference is formed from the contravariant velocities at
  // the end, and its viscous flux is zero (the plain interpolant Up, used by
  // the viscous flux, needs the flow columns only).
  constexpr int VF = 5;
  for (int q = tid; q < Q; q += NT) {
    real dF[3 * VF];
#pragma unroll
    for (int k = 0; k < 3 * VF; ++k) dF[k] = real(0);
    real Ua[V], Up[VF];
#pragma unroll
    for (int v = 0; v < V; ++v) Ua[v] = real(0);
#pragma unroll
    for (int v = 0; v < VF; ++v) Up[v] = real(0);
#pragma unroll 2
    for (int i = 0; i < N; ++i) {
      real b = a.Bt[i * Q + q];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        real u = sU[i * V + v];
        if (v < VF) Up[v] += b * u;
        Ua[v] += b * (u - sU[v]);
      }
    }
    real uq[3] = {0, 0, 0}, ua[3] = {0, 0, 0};
    if (euler) {
      // Anchored interpolation and baseline flux (hybrid/spatial.py:41-70).
#pragma unroll
      for (int v = 0; v < V; ++v) Ua[v] += sU[v];
      Prim<VF, real> wq = c2p<VF, real>(Ua, gamma);
      Prim<VF, real> wa = c2p<VF, real>(sU, gamma);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        real Fq[VF], Fa[VF];
        axis_flux<VF, real>(Ua, wq, d, Fq);
        axis_flux<VF, real>(sU, wa, d, Fa);
#pragma unroll
        for (int v = 0; v < VF; ++v) dF[d * VF + v] = Fq[v] - Fa[v];
        uq[d] = wq.vel[d];
        ua[d] = wa.vel[d];
      }
    }
    if (visc) {
      real G[12], g[12];
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        G[k] = real(0);
        g[k] = real(0);
      }
#pragma unroll 2
      for (int i = 0; i < N; ++i) {
        real dx = a.Dt[(i * 3 + 0) * Q + q];
        real dy = a.Dt[(i * 3 + 1) * Q + q];
        real dz = a.Dt[(i * 3 + 2) * Q + q];
        real b = a.Bt[i * Q + q];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          real vm = sv[i * 4 + m];
          G[m * 3 + 0] += dx * vm;
          G[m * 3 + 1] += dy * vm;
          G[m * 3 + 2] += dz * vm;
        }
#pragma unroll
        for (int c = 0; c < 12; ++c) g[c] += b * sL[i * 12 + c];
      }
      real raw[12];
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          raw[m * 3 + d] = G[m * 3 + 0] * J[0 * 3 + d] + G[m * 3 + 1] * J[1 * 3 + d] +
                           G[m * 3 + 2] * J[2 * 3 + d];
          g[m * 3 + d] += raw[m * 3 + d];
        }
      // WALE from the resolved gradient, the lift excluded
      // (hybrid/navier_stokes.py:220-290)
      real mu = ph.mu, kappa = ph.kappa;
      if (ph.wale) {
        const real mt = wale_viscosity<real>(raw, Up[0], a.width[e], ph.wale_constant);
        mu += mt;
        kappa += mt * ph.cp_over_prt;
      }
      real Fv[3 * VF];
      viscous_flux<VF, real>(Up, g, gamma, mu, kappa, Fv);
#pragma unroll
      for (int k = 0; k < 3 * VF; ++k) dF[k] -= Fv[k];
    }
    // Pull back to reference directions with the quadrature weight.
    const real w = a.wq[q] * dw;
    real* f = sF + q * SW;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
      for (int v = 0; v < VF; ++v)
        f[r * VF + v] = w * (J[r * 3 + 0] * dF[0 * VF + v] + J[r * 3 + 1] * dF[1 * VF + v] +
                             J[r * 3 + 2] * dF[2 * VF + v]);
      if constexpr (V > VF) {
        f[15 + r] = w * (J[r * 3 + 0] * uq[0] + J[r * 3 + 1] * uq[1] + J[r * 3 + 2] * uq[2]);
        f[18 + r] = w * (J[r * 3 + 0] * ua[0] + J[r * 3 + 1] * ua[1] + J[r * 3 + 2] * ua[2]);
      }
    }
    if constexpr (V > VF)
#pragma unroll
      for (int v = VF; v < V; ++v) f[21 + (v - VF)] = Ua[v];
  }
  __syncthreads();

  // R_i += sum_q sum_r dphi_i/dxi_r F^_r: DS adjacent lanes per node (DS a
  // power of two with N DS <= NT: 2 for 64 nodes, 8 for the 10-node
  // tetrahedron), each over every DS-th point and all variables (one read of
  // D per point and direction), partial sums combined by lane shuffles.  A
  // species column's flux difference w (U_s (J_r . u) - a_s (J_r . u_a)) is
  // rank two: its divergence is sum_q c_q U_s(q) - a_s sum_q c'_q with
  // c_q = sum_r D_qri w (J_r . u), c'_q likewise with the anchor velocity.
  {
    constexpr int DS = div_split(N, NT);
    const int i = tid / DS, h = tid - i * DS;
    real acc[V], acc_a = real(0);
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = real(0);
    // The operator entries D[q][r][i] of a chunk of CH points are loaded
    // together (one memory round trip per chunk instead of one per point).
    constexpr int NQ = (Q + DS - 1) / DS, CH = NQ < 8 ? NQ : 8;
    if (i < N)
#pragma unroll 1
      for (int k0 = 0; k0 < NQ; k0 += CH) {
        real dc[CH][3];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const int q = h + (k0 + c) * DS;
#pragma unroll
          for (int r = 0; r < 3; ++r) dc[c][r] = q < Q ? a.D[(q * 3 + r) * N + i] : real(0);
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const int q = h + (k0 + c) * DS;
          if (q >= Q) break;
          const real* f = sF + q * SW;
          const real* d = dc[c];
#pragma unroll
          for (int r = 0; r < 3; ++r) {
#pragma unroll
            for (int v = 0; v < VF; ++v) acc[v] += d[r] * f[r * VF + v];
          }
          if constexpr (V > VF) {
            const real cc = d[0] * f[15] + d[1] * f[16] + d[2] * f[17];
            acc_a += d[0] * f[18] + d[1] * f[19] + d[2] * f[20];
#pragma unroll
            for (int v = VF; v < V; ++v) acc[v] += cc * f[21 + (v - VF)];
          }
        }
      }
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = group_sum<DS>(acc[v]);
    if constexpr (V > VF) acc_a = group_sum<DS>(acc_a);
    if (i < N && h == 0) {
#pragma unroll
      for (int v = 0; v < VF; ++v) sR[i * V + v] += acc[v];
#pragma unroll
      for (int v = VF; v < V; ++v) sR[i * V + v] += acc[v] - sU[v] * acc_a;
    }
  }
  __syncthreads();

  real* out = a.out + e * (N * V);
  if (a.mode == 0) {
    for (int k = tid; k < N * V; k += NT) out[k] = sR[k];
    return;
  }
  const real dte = a.dt_elem ? a.dt_elem[e] : a.dt;
  // rhs = scale M^-1 R.  With at least two threads per node, one thread per
  // (node, variable); otherwise one thread per node over all variables (one
  // read of each M^-1 row).
  auto finish = [&](int k, real acc) {
    const real rhs = s * acc;
    if (a.mode == 1) {
      out[k] = rhs;
    } else {
      // without a base state (the Euler candidate U + dt rhs of a MOOD stage)
      // no base term is read
      const real nxt = (a.Ubase ? a.alpha * a.Ubase[e * (N * V) + k] : real(0)) +
                       a.beta * (sU[k] + dte * rhs);
      out[k] = nxt;
      sF[k] = nxt;
    }
  };
  if constexpr (2 * N <= NT) {
    for (int k = tid; k < N * V; k += NT) {
      const int i = k / V, v = k - i * V;
      real acc = real(0);
      for (int j = 0; j < N; ++j) acc += a.minv[i * N + j] * sR[j * V + v];
      finish(k, acc);
    }
  } else {
    for (int i = tid; i < N; i += NT) {
      real acc[V];
#pragma unroll
      for (int v = 0; v < V; ++v) acc[v] = real(0);
      for (int j = 0; j < N; ++j) {
        const real m = a.minv[i * N + j];
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] += m * sR[j * V + v];
      }
#pragma unroll
      for (int v = 0; v < V; ++v) finish(i * V + v, acc[v]);
    }
  }
  if (a.mode != 2 || !a.diag) return;
  __syncthreads();

  // Admissibility at nodes (hybrid/time_integration.py:291-312).
  double rmn = __longlong_as_double(0x7ff0000000000000LL), pmn = rmn;
  unsigned int bad = 0;
  for (int i = tid; i < N; i += NT) {
    const real* u = sF + i * V;
    for (int v = 0; v < V; ++v) bad += finite(u[v]) ? 0u : 1u;
    real rho = u[0];
    real ke = real(0.5) * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / rmax(rho, real(1.0e-14));
    real p = (gamma - real(1)) * (u[4] - ke);
    rmn = nan_min(rmn, double(rho));
    pmn = nan_min(pmn, double(p));
  }
  for (int off = 32; off > 0; off >>= 1) {
    rmn = nan_min(rmn, __shfl_down(rmn, off, 64));
    pmn = nan_min(pmn, __shfl_down(pmn, off, 64));
    bad += __shfl_down(bad, off, 64);
  }
  const int lane = tid & 63, wv = tid >> 6;
  if (lane == 0) {
    red[0][wv] = rmn;
    red[1][wv] = pmn;
    redn[wv] = bad;
  }
  __syncthreads();
  if (tid == 0) {
    for (int k = 1; k < NT / 64; ++k) {
      rmn = nan_min(rmn, red[0][k]);
      pmn = nan_min(pmn, red[1][k]);
      bad += redn[k];
    }
    atomic_min_nan<real>(&a.diag->rho_min, rmn, seen.rho);
    atomic_min_nan<real>(&a.diag->p_min, pmn, seen.p);
    if (bad) atomicAdd(&a.diag->nonfinite, bad);
  }
}

// BR2 face gradient of one hexahedral side on a quad face (p >= 2) by sum
// factorisation: the face points are the Qf1 x Qf1 Gauss grid on one face
// plane of the element, so the raw reference gradient
//   d v / d xi_r (q) = sum_n d phi_n / d xi_r (x_q) v_n
// is a contraction along the face normal (l(plane) or l'(plane)), then along
// the two tangential axes (B1 or D1 at the face Gauss points).  meta holds
// [normal axis, plane 0/1, t1, t2, index along t1 (Qf), index along t2 (Qf)]
// for this trace table.  Writes sg[q * 12 + m * 3 + d] = J^-T grad + eta
// scale (local lift), as the dense phase does.  scratch: 2 N1^2 * 4 +
// 3 N1 Qf1 * 4 reals of LDS.
template <int P, typename real>
__device__ __forceinline__ void face_grad_hex_sf(const int32_t* __restrict__ meta,
                                                 const real* __restrict__ sf1, const real* sv,
                                                 const real* lop_base, const real* sbb,
                                                 const real* Jx, real sc, real eta, real* sg,
                                                 real* scratch, int tid) {
  constexpr int N1 = P + 1, QF1 = P == 2 ? 3 : 4, QF = QF1 * QF1, NF = N1 * N1;
  const int dn = meta[0], plane = meta[1], t1 = meta[2], t2 = meta[3];
  const int32_t* aq = meta + 4;
  const int32_t* bq = meta + 4 + QF;
  const real* Lg = sf1;
  const real* dLg = sf1 + QF1 * N1;
  const real* en = sf1 + 2 * QF1 * N1 + plane * 2 * N1;
  const real* den = en + N1;
  auto stride = [](int axis) { return axis == 0 ? 1 : (axis == 1 ? N1 : N1 * N1); };
  const int st1 = stride(t1), st2 = stride(t2), stn = stride(dn);
  real* Vn = scratch;                  // [i1][i2][m]
  real* Dn = Vn + NF * 4;
  real* A = Dn + NF * 4;               // [i1][b][m]
  real* B = A + N1 * QF1 * 4;
  real* C = B + N1 * QF1 * 4;
  for (int t = tid; t < NF * 4; t += 64) {          // normal contraction
    const int m = t & 3, i12 = t >> 2, i1 = i12 / N1, i2 = i12 - i1 * N1;
    const int base = i1 * st1 + i2 * st2;
    real vn = real(0), dv = real(0);
#pragma unroll
    for (int k = 0; k < N1; ++k) {
      const real x = sv[(base + k * stn) * 4 + m];
      vn += en[k] * x;
      dv += den[k] * x;
    }
    Vn[t] = vn;
    Dn[t] = dv;
  }
  __syncthreads();
  for (int t = tid; t < N1 * QF1 * 4; t += 64) {    // along t2
    const int m = t & 3, r = t >> 2, i1 = r / QF1, b = r - i1 * QF1;
    real av = real(0), bv = real(0), cv = real(0);
#pragma unroll
    for (int i2 = 0; i2 < N1; ++i2) {
      const int src = ((i1 * N1 + i2) << 2) + m;
      const real l = Lg[b * N1 + i2], dl = dLg[b * N1 + i2];
      av += l * Vn[src];
      bv += dl * Vn[src];
      cv += l * Dn[src];
    }
    A[t] = av;
    B[t] = bv;
    C[t] = cv;
  }
  __syncthreads();
  for (int t = tid; t < QF * 4; t += 64) {          // along t1, per (point, variable)
    const int q = t >> 2, m = t & 3, qa = aq[q], qb = bq[q];
    real g1 = real(0), g2 = real(0), gn = real(0);
#pragma unroll
    for (int i1 = 0; i1 < N1; ++i1) {
      const int src = ((i1 * QF1 + qb) << 2) + m;
      g1 += dLg[qa * N1 + i1] * A[src];
      g2 += Lg[qa * N1 + i1] * B[src];
      gn += Lg[qa * N1 + i1] * C[src];
    }
    real G[3];
    G[t1] = g1;
    G[t2] = g2;
    G[dn] = gn;
    const real* lop = lop_base + q * NF;
    real L[3] = {real(0), real(0), real(0)};
#pragma unroll
    for (int kk = 0; kk < NF; ++kk) {
      const real c = lop[kk];
#pragma unroll
      for (int d = 0; d < 3; ++d) L[d] += c * sbb[kk * 12 + m * 3 + d];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
      sg[q * 12 + m * 3 + d] =
          G[0] * Jx[0 * 3 + d] + G[1] * Jx[1 * 3 + d] + G[2] * Jx[2 * 3 + d] + eta * (sc * L[d]);
  }
}

// WALE: the variant that adds the WALE eddy viscosity of the resolved face
// gradient (a separate instantiation, so that the plain kernels keep their
// registers).
template <int FT, int TO, int TN, int SUB, int P, int V, typename real, bool WALE>
__device__ __forceinline__ void face_body(const FaceArgs<real>& a, const Phys<real>& ph,
                                          const int64_t f) {
  constexpr int NO = n_nodes(TO, P);
  constexpr int NN = TN >= 0 ? n_nodes(TN, P) : 1;
  constexpr int QF = n_face_q(FT, P);
  constexpr int FO = n_face_nodes(FT, SUB, P);
  constexpr int FN = TN >= 0 ? n_face_nodes(FT, 0, P) : 1;
  constexpr int TSO = QF * FO * 2 + QF * NO * 3;
  constexpr int TSN = QF * FN * 2 + QF * NN * 3;
  constexpr int W = V + 12;
  constexpr bool NBR = TN >= 0;
  // Unroll depth of the trace and face-gradient loops: full, except in FP64
  // with species, where full unrolling overflows the register file.
  constexpr bool WIDE = V > 5 && sizeof(real) == 8;
  constexpr int UO = WIDE ? 4 : FO, UN = WIDE ? 4 : FN;
  // the dense face-gradient loops are only a fallback on hexahedral quad
  // sides (sum factorisation below): keep them short there
  constexpr bool HQ = FT == QUAD && SUB == 0 && P >= 2;
  constexpr int UG = (WIDE || (HQ && TO == HEX)) ? 4 : NO;
  constexpr int UGN = (WIDE || (HQ && TN == HEX)) ? 4 : NN;

  // LDS: the staged states (anchor node 0, then the face-support nodes) and
  // the anchored traces are dead after the Euler flux (phase 1b); the face
  // gradients (phases 3-4) and the sum-factorisation scratch reuse them.
  constexpr int SO = FO + 1, SN = NBR ? FN + 1 : 1;
  // BR2 face gradients by sum factorisation on hexahedral quad sides
  constexpr bool SFO = FT == QUAD && SUB == 0 && TO == HEX && P >= 2;
  constexpr bool SFN = FT == QUAD && SUB == 0 && TN == HEX && P >= 2;
  constexpr int SFN1 = P + 1, SFQ1 = P == 2 ? 3 : 4;
  constexpr int SFS = (SFO || SFN) ? 2 * SFN1 * SFN1 * 4 + 3 * SFN1 * SFQ1 * 4 : 1;
  constexpr int NU = (SO + SN) * V > 2 * QF * 12 ? (SO + SN) * V : 2 * QF * 12;
  constexpr int NA = (NBR ? 2 : 1) * QF * V > SFS ? (NBR ? 2 : 1) * QF * V : SFS;
  __shared__ real sUU[NU], sAA[NA], svo[NO * 4], svn[NN * 4];
  __shared__ real sSo[QF * V], sSn[QF * V], sPo[QF * V], sPn[QF * V], sJ[QF * 4];
  __shared__ real sbo[FO * 12], sbn[FN * 12];
  __shared__ real sMt[WALE ? 2 * QF : 1]; // WALE eddy viscosity per side and point
  __shared__ real sB[2][6];              // anchor normal flux [5] and u.n per side
  real* const sUo = sUU;                 // [SO][V], slot 0 = node 0, slot 1 + k = support node k
  real* const sUn = sUU + SO * V;        // [SN][V]
  real* const sgo = sUU;                 // [QF][12] after phase 1b
  real* const sgn = sUU + QF * 12;
  real* const sAo = sAA;                 // [QF][V]
  real* const sAn = sAA + QF * V;
  real* const sfs = sAA;                 // after phase 1b

  const int tid = threadIdx.x;
  const bool euler = a.terms & TERM_EULER;
  const bool visc = a.terms & TERM_VISCOUS;
  const real gamma = ph.gamma;

  const int64_t eo = a.owner_elem[f];
  const int64_t en = NBR ? a.neighbor_elem[f] : 0;
  const real n[3] = {a.geom[f * 4 + 0], a.geom[f * 4 + 1], a.geom[f * 4 + 2]};
  const real area = a.geom[f * 4 + 3];
  const int flags = NBR ? 0 : a.bflags[f];
  const int to_id = a.owner_table[f];
  const real* tbo = a.owner_tables + (int64_t)to_id * TSO;
  const int32_t* supo = a.owner_sup + to_id * FO;
  const real* tbn = nullptr;
  const int32_t* supn = nullptr;
  if constexpr (NBR) {
    const int tn_id = a.neighbor_table[f];
    tbn = a.neighbor_tables + (int64_t)tn_id * TSN;
    supn = a.neighbor_sup + tn_id * FN;
  }

  // stage node 0 (the anchor) and the face-support nodes of each side; the
  // viscous variables of every node come from the flow columns directly.
  // Every load of both sides is issued before the first value is used: the
  // support indices, then the staged states and the viscous nodes' flow
  // columns (a loop that stores each value before loading the next waits
  // two dependent round trips per 64 values and side).
  {
    constexpr int IO = (SO * V + 63) / 64, IN = (SN * V + 63) / 64;
    constexpr int VO = (NO + 63) / 64, VN = NBR ? (NN + 63) / 64 : 1;
    const real* Ue = a.Uo + eo * (NO * V);
    const real* Un = a.Un + en * (NN * V);
    int ro[IO], rn[IN];
#pragma unroll
    for (int i = 0; i < IO; ++i) {
      const int k = tid + 64 * i < SO * V ? tid + 64 * i : SO * V - 1, sl = k / V;
      const int s = supo[sl > 0 ? sl - 1 : 0];
      ro[i] = (sl == 0 ? 0 : s) * V + (k - sl * V);
    }
    if constexpr (NBR) {
#pragma unroll
      for (int i = 0; i < IN; ++i) {
        const int k = tid + 64 * i < SN * V ? tid + 64 * i : SN * V - 1, sl = k / V;
        const int s = supn[sl > 0 ? sl - 1 : 0];
        rn[i] = (sl == 0 ? 0 : s) * V + (k - sl * V);
      }
    }
    real fo[VO][5], fn[VN][5];
    if (visc) {
#pragma unroll
      for (int j = 0; j < VO; ++j) {
        const int i = tid + 64 * j < NO ? tid + 64 * j : NO - 1;
#pragma unroll
        for (int v = 0; v < 5; ++v) fo[j][v] = Ue[i * V + v];
      }
      if constexpr (NBR) {
#pragma unroll
        for (int j = 0; j < VN; ++j) {
          const int i = tid + 64 * j < NN ? tid + 64 * j : NN - 1;
#pragma unroll
          for (int v = 0; v < 5; ++v) fn[j][v] = Un[i * V + v];
        }
      }
    }
    real uo[IO], un[IN];
#pragma unroll
    for (int i = 0; i < IO; ++i) uo[i] = Ue[ro[i]];
    if constexpr (NBR) {
#pragma unroll
      for (int i = 0; i < IN; ++i) un[i] = Un[rn[i]];
    }
#pragma unroll
    for (int i = 0; i < IO; ++i)
      if (tid + 64 * i < SO * V) sUo[tid + 64 * i] = uo[i];
    if constexpr (NBR) {
#pragma unroll
      for (int i = 0; i < IN; ++i)
        if (tid + 64 * i < SN * V) sUn[tid + 64 * i] = un[i];
    }
    if (visc) {
#pragma unroll
      for (int j = 0; j < VO; ++j)
        if (tid + 64 * j < NO) viscous_variables<5, real>(fo[j], gamma, ph.gas_constant, svo + (tid + 64 * j) * 4);
      if constexpr (NBR) {
#pragma unroll
        for (int j = 0; j < VN; ++j)
          if (tid + 64 * j < NN)
            viscous_variables<5, real>(fn[j], gamma, ph.gas_constant, svn + (tid + 64 * j) * 4);
      }
    }
  }
  __syncthreads();

  // The anchor's normal flux and normal velocity, a constant of each side,
  // once per face (owner on lane 0, neighbour on lane 1) for the Euler flux.
  if (euler && tid < (NBR ? 2 : 1)) {
    const real* u0 = tid ? sUn : sUo;
    Prim<5, real> wa = c2p<5, real>(u0, gamma);
    real B[5];
    normal_flux<5, real>(u0, wa, n, B);
#pragma unroll
    for (int v = 0; v < 5; ++v) sB[tid][v] = B[v];
    sB[tid][5] = wa.vel[0] * n[0] + wa.vel[1] * n[1] + wa.vel[2] * n[2];
  }

  // Phase 1a: traces at the face points: primal Up = sum_k phi_k u_k (for
  // the viscous flux) and anchored Ua = u_0 + sum_k phi_k (u_k - u_0) (for
  // the Euler flux).  With at most 16 face points, 16 support nodes and 16
  // variables each is a 16x16 tile product Phi (points x nodes) U (nodes x
  // variables) on the matrix cores (v_mfma 16x16x4, K = support nodes in
  // steps of 4; rows and columns past the sizes are zero); otherwise one lane
  // per (point, variable).  FP32 quadrilateral faces keep the vector form:
  // there the tile accumulators cost two waves per SIMD of occupancy for no
  // gain (hex-hex p=2: 1.37 ms either way), while the FP64 kernels (hex-hex
  // 3.32 -> 2.79 ms) and the FP32 triangles (hex-tet 1.08 -> 1.02 ms) gain.
  constexpr bool MT = QF <= 16 && FO <= 16 && FN <= 16 && V <= 16 &&
                      !(sizeof(real) == 4 && FT == QUAD);
  if constexpr (MT) {
    const int lr = tid & 15, lk = tid >> 4;
    auto traces = [&](const real* tb, const real* sUs, int NF, real* sA, real* sP) {
      const real u0 = lr < V ? sUs[lr] : real(0);
      auto ca = zero4<real>(), cp = zero4<real>();
#pragma unroll
      for (int k0 = 0; k0 < 16; k0 += 4) {
        if (k0 >= NF) break;
        const int k = k0 + lk;
        const bool in = k < NF;
        const real phi = (in && lr < QF) ? tb[lr * NF + k] : real(0);
        const real u = (in && lr < V) ? sUs[(k + 1) * V + lr] : real(0);
        ca = mma(phi, in ? u - u0 : real(0), ca);
        cp = mma(phi, u, cp);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int q = crow<real>(tid, j);
        if (q < QF && lr < V) {
          sA[q * V + lr] = ca[j] + u0;
          sP[q * V + lr] = cp[j];
        }
      }
    };
    traces(tbo, sUo, FO, sAo, sPo);
    if constexpr (NBR) {
      traces(tbn, sUn, FN, sAn, sPn);
    } else {
      for (int k = tid; k < QF * V; k += 64) sPn[k] = sPo[k];  // unused: the boundary flux is the owner's
    }
  } else
  for (int k = tid; k < QF * V; k += 64) {
    const int q = k / V, v = k - q * V;
    const real* phio = tbo + q * FO;
    const real u0 = sUo[v];
    real ua = real(0), up = real(0);
#pragma unroll UO
    for (int kk = 0; kk < FO; ++kk) {
      const real b = phio[kk];
      const real u = sUo[(kk + 1) * V + v];
      up += b * u;
      ua += b * (u - u0);
    }
    sAo[k] = ua + u0;
    sPo[k] = up;
    if constexpr (NBR) {
      const real* phin = tbn + q * FN;
      const real n0 = sUn[v];
      real ub = real(0), vp = real(0);
#pragma unroll UN
      for (int kk = 0; kk < FN; ++kk) {
        const real b = phin[kk];
        const real u = sUn[(kk + 1) * V + v];
        vp += b * u;
        ub += b * (u - n0);
      }
      sAn[k] = ub + n0;
      sPn[k] = vp;
    } else {
      sPn[k] = up;  // unused: the boundary flux is the owner's
    }
  }
  __syncthreads();

  // Phase 1b: the Euler numerical flux at each face point.  With species
  // (V > 5) the flow columns stay in registers and each species column is
  // streamed through LDS (numflux_split), which keeps the per-lane state to
  // the five flow variables.
  if constexpr (V > 5) {
    for (int q = tid; q < QF; q += 64) {
      const real w = a.fw[q] * area;
      const real* Ao = sAo + q * V;
      real UL[5], UR[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) UL[v] = Ao[v];
      // ghost species (ghost_state): rho_g * Y, Y = far (inlet) or U_s / safe rho
      real rho_g = real(0), safe_in = real(1);
      if constexpr (NBR) {
#pragma unroll
        for (int v = 0; v < 5; ++v) UR[v] = sAn[q * V + v];
      } else {
        ghost_state<5, real>(UL, n, flags, ph, UR);
        rho_g = UR[0];
        safe_in = rmax(UL[0], real(1.0e-14));
      }
      auto sl = [&](int sp) { return Ao[sp]; };
      auto sr = [&](int sp) -> real {
        if constexpr (NBR) {
          return sAn[q * V + sp];
        } else {
          if (!ph.has_boundary) return Ao[sp];
          const real Y = (flags & BND_INLET) ? ph.farfield[sp] : Ao[sp] / safe_in;
          return rho_g * Y;
        }
      };
      if (euler) {
        real F[5], Bo[5], Bn[5];
        FluxScalars<real> fs;
        numflux_split<V, real>(UL, UR, n, gamma, ph.rusanov, sl, sr, F, fs);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          Bo[v] = sB[0][v];
          Bn[v] = NBR ? sB[1][v] : real(0);
        }
        const real una = sB[0][5], unb = NBR ? sB[1][5] : real(0);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          sSo[q * V + v] = -w * (F[v] - Bo[v]);
          sSn[q * V + v] = NBR ? w * (F[v] - Bn[v]) : real(0);
        }
        // each species column's flux once, for both sides
        for (int sp = 5; sp < V; ++sp) {
          const real fsp = species_flux(fs, sl(sp), sr(sp));
          sSo[q * V + sp] = -w * (fsp - sUo[sp] * una);
          sSn[q * V + sp] = NBR ? w * (fsp - sUn[sp] * unb) : real(0);
        }
      } else {
        for (int v = 0; v < V; ++v) {
          sSo[q * V + v] = real(0);
          sSn[q * V + v] = real(0);
        }
      }
    }
  } else
  for (int q = tid; q < QF; q += 64) {
    const real w = a.fw[q] * area;
    real Ua_o[V], Ua_n[V];
#pragma unroll
    for (int v = 0; v < V; ++v) Ua_o[v] = sAo[q * V + v];
    if constexpr (NBR) {
#pragma unroll
      for (int v = 0; v < V; ++v) Ua_n[v] = sAn[q * V + v];
    } else {
      ghost_state<V, real>(Ua_o, n, flags, ph, Ua_n);
    }
    real So[V], Sn[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      So[v] = real(0);
      Sn[v] = real(0);
    }
    if (euler) {
      real Fs[V], Bo[V];
      if (ph.rusanov) rusanov<V, real>(Ua_o, Ua_n, n, gamma, Fs);
      else hllc<V, real>(Ua_o, Ua_n, n, gamma, Fs);
      static_assert(V == 5, "flow variables only in this branch");
#pragma unroll
      for (int v = 0; v < V; ++v) Bo[v] = sB[0][v];
#pragma unroll
      for (int v = 0; v < V; ++v) So[v] = -w * (Fs[v] - Bo[v]);
      if constexpr (NBR) {
        real Bn[V];
#pragma unroll
        for (int v = 0; v < V; ++v) Bn[v] = sB[1][v];
#pragma unroll
        for (int v = 0; v < V; ++v) Sn[v] = w * (Fs[v] - Bn[v]);
      }
    }
#pragma unroll
    for (int v = 0; v < V; ++v) {
      sSo[q * V + v] = So[v];
      sSn[q * V + v] = Sn[v];
    }
  }
  // Jump of the viscous variables at the face points, one lane per (point,
  // variable): [[v]]_m = v_m(neighbour trace) - v_m(owner trace), weighted.
  if (visc) {
    for (int k = tid; k < QF * 4; k += 64) {
      const int q = k >> 2, m = k & 3;
      const real w = a.fw[q] * area;
      const real* phio = tbo + q * FO;
      real vo = real(0), vn = real(0);
#pragma unroll UO
      for (int kk = 0; kk < FO; ++kk) vo += phio[kk] * svo[supo[kk] * 4 + m];
      if constexpr (NBR) {
        const real* phin = tbn + q * FN;
#pragma unroll UN
        for (int kk = 0; kk < FN; ++kk) vn += phin[kk] * svn[supn[kk] * 4 + m];
      } else {
        real g[4];
        ghost_viscous<V, real>(sPo + q * V, n, flags, ph, g);
        vn = g[m];
      }
      sJ[k] = w * (vn - vo);
    }
  }
  __syncthreads();

  if (visc) {
    // Phase 2: lift right-hand sides b = 1/2 sum_q w [[v]] (x) n phi.  Four
    // lanes per support node split the points; each table entry is read once.
    {
      const int kk = tid >> 2, part = tid & 3;
      real acc_o[4] = {0, 0, 0, 0}, acc_n[4] = {0, 0, 0, 0};
      for (int q = part; q < QF; q += 4) {
        if (kk < FO) {
          const real c = tbo[q * FO + kk];
#pragma unroll
          for (int m = 0; m < 4; ++m) acc_o[m] += c * sJ[q * 4 + m];
        }
        if constexpr (NBR) {
          if (kk < FN) {
            const real c = tbn[q * FN + kk];
#pragma unroll
            for (int m = 0; m < 4; ++m) acc_n[m] += c * sJ[q * 4 + m];
          }
        }
      }
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        acc_o[m] = group_sum<4>(acc_o[m]);
        acc_n[m] = group_sum<4>(acc_n[m]);
      }
      static_assert(FO <= 16 && FN <= 16, "four lanes per support node in one wave");
      if (part == 0 && kk < FO)
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int d = 0; d < 3; ++d) sbo[kk * 12 + m * 3 + d] = real(0.5) * acc_o[m] * n[d];
      if constexpr (NBR)
        if (part == 0 && kk < FN)
#pragma unroll
          for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int d = 0; d < 3; ++d) sbn[kk * 12 + m * 3 + d] = real(0.5) * acc_n[m] * n[d];
    }
    __syncthreads();

    // Phase 3: face gradients = raw polynomial gradient + eta * local lift.
    // Hexahedral quad sides (p >= 2) by sum factorisation when the tables are
    // given; otherwise the dense gradient tables.
    // FP32 dense: four lanes per face point split the element nodes and the
    // support nodes (the gradient table is read once per (point, node)); FP64
    // dense keeps one lane per (point, variable), which needs fewer registers.
    bool sf_o = false, sf_n = false;
    if constexpr ((SFO || SFN) && !WALE) {
      if (a.face_sf1d) {
        constexpr int MW = 4 + 2 * QF;
        if constexpr (SFO) {
          if (a.owner_sf) {
          real Jo[9];
#pragma unroll
          for (int k = 0; k < 9; ++k) Jo[k] = a.jinv_o[eo * 9 + k];
          face_grad_hex_sf<P, real>(a.owner_sf + to_id * MW, a.face_sf1d, svo, tbo + QF * FO, sbo,
                                    Jo, a.scale_o[eo], ph.eta, sgo, sfs, tid);
          sf_o = true;
          __syncthreads();
          }
        }
        if constexpr (SFN) {
          if (a.neighbor_sf) {
          real Jn[9];
#pragma unroll
          for (int k = 0; k < 9; ++k) Jn[k] = a.jinv_n[en * 9 + k];
          face_grad_hex_sf<P, real>(a.neighbor_sf + a.neighbor_table[f] * MW, a.face_sf1d, svn,
                                    tbn + QF * FN, sbn, Jn, a.scale_n[en], ph.eta, sgn, sfs, tid);
          sf_n = true;
          }
        }
      }
    }
    if constexpr (sizeof(real) == 4) {
      static_assert(QF <= 16 || QF == 25, "face points per wave");
      auto side = [&](const real* tb, const real* sv, const real* sbb, const real* Jx, real sc,
                      int NE, int NF, real* sg, const real* sP, real wd, real* smt) {
        for (int q0 = 0; q0 < QF; q0 += 16) {
          const int q = q0 + (tid >> 2), part = tid & 3;
          real G[12], L[12];
#pragma unroll
          for (int k = 0; k < 12; ++k) {
            G[k] = real(0);
            L[k] = real(0);
          }
          if (q < QF) {
            const real* dph = tb + QF * NF * 2 + q * NE * 3;
            for (int i = part; i < NE; i += 4) {
              const real d0 = dph[i * 3 + 0], d1 = dph[i * 3 + 1], d2 = dph[i * 3 + 2];
#pragma unroll
              for (int m = 0; m < 4; ++m) {
                const real vm = sv[i * 4 + m];
                G[m * 3 + 0] += d0 * vm;
                G[m * 3 + 1] += d1 * vm;
                G[m * 3 + 2] += d2 * vm;
              }
            }
            const real* lop = tb + QF * NF + q * NF;
            for (int kk = part; kk < NF; kk += 4) {
              const real c = lop[kk];
#pragma unroll
              for (int k = 0; k < 12; ++k) L[k] += c * sbb[kk * 12 + k];
            }
          }
#pragma unroll
          for (int k = 0; k < 12; ++k) {
            G[k] = group_sum<4>(G[k]);
            L[k] = group_sum<4>(L[k]);
          }
          if (q < QF && part == 0) {
            real raw[12];
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
              for (int d = 0; d < 3; ++d) {
                raw[m * 3 + d] = G[m * 3 + 0] * Jx[0 * 3 + d] + G[m * 3 + 1] * Jx[1 * 3 + d] +
                                 G[m * 3 + 2] * Jx[2 * 3 + d];
                sg[q * 12 + m * 3 + d] = raw[m * 3 + d] + ph.eta * (sc * L[m * 3 + d]);
              }
            // WALE from the resolved gradient, the lift excluded
            if constexpr (WALE) smt[q] = wale_viscosity<real>(raw, sP[q * V], wd, ph.wale_constant);
          }
        }
      };
      if (!sf_o) {
        real Jo[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) Jo[k] = a.jinv_o[eo * 9 + k];
        side(tbo, svo, sbo, Jo, a.scale_o[eo], NO, FO, sgo, sPo,
             WALE ? a.width_o[eo] : real(0), sMt);
      }
      if constexpr (NBR) {
        if (!sf_n) {
          real Jn[9];
#pragma unroll
          for (int k = 0; k < 9; ++k) Jn[k] = a.jinv_n[en * 9 + k];
          side(tbn, svn, sbn, Jn, a.scale_n[en], NN, FN, sgn, sPn,
               WALE ? a.width_n[en] : real(0), WALE ? sMt + QF : sMt);
        }
      }
    } else {
      const real so = a.scale_o[eo];
      real Jo[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) Jo[k] = a.jinv_o[eo * 9 + k];
      for (int k = sf_o ? QF * 4 : tid; k < QF * 4; k += 64) {
        int q = k >> 2, m = k & 3;
        const real* dph = tbo + QF * FO * 2 + q * NO * 3;
        real G0 = 0, G1 = 0, G2 = 0;
#pragma unroll UG
        for (int i = 0; i < NO; ++i) {
          real vm = svo[i * 4 + m];
          G0 += dph[i * 3 + 0] * vm;
          G1 += dph[i * 3 + 1] * vm;
          G2 += dph[i * 3 + 2] * vm;
        }
        const real* lop = tbo + QF * FO + q * FO;
        real L[3] = {0, 0, 0};
        for (int kk = 0; kk < FO; ++kk) {
          real c = lop[kk];
#pragma unroll
          for (int d = 0; d < 3; ++d) L[d] += c * sbo[kk * 12 + m * 3 + d];
        }
        real rd[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          rd[d] = G0 * Jo[0 * 3 + d] + G1 * Jo[1 * 3 + d] + G2 * Jo[2 * 3 + d];
          sgo[q * 12 + m * 3 + d] = rd[d] + ph.eta * (so * L[d]);
        }
        // WALE: the point's four lanes (one wave, same iteration) hand the
        // raw velocity gradient rows to the m = 0 lane
        if constexpr (WALE) {
          real r9[9];
#pragma unroll
          for (int mm = 0; mm < 3; ++mm)
#pragma unroll
            for (int d = 0; d < 3; ++d) r9[mm * 3 + d] = __shfl(rd[d], (tid & ~3) + mm, 64);
          if (m == 0) sMt[q] = wale_viscosity<real>(r9, sPo[q * V], a.width_o[eo], ph.wale_constant);
        }
      }
      if constexpr (NBR) {
        const real sn = a.scale_n[en];
        real Jn[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) Jn[k] = a.jinv_n[en * 9 + k];
        for (int k = sf_n ? QF * 4 : tid; k < QF * 4; k += 64) {
          int q = k >> 2, m = k & 3;
          const real* dph = tbn + QF * FN * 2 + q * NN * 3;
          real G0 = 0, G1 = 0, G2 = 0;
#pragma unroll UGN
          for (int i = 0; i < NN; ++i) {
            real vm = svn[i * 4 + m];
            G0 += dph[i * 3 + 0] * vm;
            G1 += dph[i * 3 + 1] * vm;
            G2 += dph[i * 3 + 2] * vm;
          }
          const real* lop = tbn + QF * FN + q * FN;
          real L[3] = {0, 0, 0};
          for (int kk = 0; kk < FN; ++kk) {
            real c = lop[kk];
#pragma unroll
            for (int d = 0; d < 3; ++d) L[d] += c * sbn[kk * 12 + m * 3 + d];
          }
          real rd[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            rd[d] = G0 * Jn[0 * 3 + d] + G1 * Jn[1 * 3 + d] + G2 * Jn[2 * 3 + d];
            sgn[q * 12 + m * 3 + d] = rd[d] + ph.eta * (sn * L[d]);
          }
          if constexpr (WALE) {
            real r9[9];
#pragma unroll
            for (int mm = 0; mm < 3; ++mm)
#pragma unroll
              for (int d = 0; d < 3; ++d) r9[mm * 3 + d] = __shfl(rd[d], (tid & ~3) + mm, 64);
            if (m == 0)
              sMt[QF + q] = wale_viscosity<real>(r9, sPn[q * V], a.width_n[en], ph.wale_constant);
          }
        }
      }
    }
    __syncthreads();

    // Phase 4: central viscous flux (owner flux only on the boundary), one
    // lane per (point, side): lane QF + q evaluates the neighbour's flux at
    // point q and hands its rows to lane q.  The viscous flux has momentum
    // and energy rows only (the mass and species rows are zero): the other
    // columns are left as they are.
    static_assert(2 * QF <= 64, "two face sides per point in one wave");
    {
      const int side = (NBR && tid >= QF) ? 1 : 0;
      const int q = tid - side * QF;
      real Fx[3 * V];
#pragma unroll
      for (int k = 0; k < 3 * V; ++k) Fx[k] = real(0);
      if (q < QF) {
        real mu = ph.mu, kappa = ph.kappa;
        if constexpr (WALE) {
          const real mt = sMt[side * QF + q];
          mu += mt;
          kappa += mt * ph.cp_over_prt;
        }
        viscous_flux<V, real>((side ? sPn : sPo) + q * V, (side ? sgn : sgo) + q * 12, gamma, mu,
                              kappa, Fx);
      }
      real Fo[3][5];
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int v = 1; v < 5; ++v) {
          Fo[d][v] = Fx[d * V + v];
          if constexpr (NBR) {
            const real fnb = __shfl(Fx[d * V + v], tid + QF, 64);
            Fo[d][v] = real(0.5) * (Fo[d][v] + fnb);
          }
        }
      if (tid < QF) {
        const real w = a.fw[q] * area;
        real fn[5];
#pragma unroll
        for (int v = 1; v < 5; ++v) fn[v] = Fo[0][v] * n[0] + Fo[1][v] * n[1] + Fo[2][v] * n[2];
        if (!NBR && ph.has_boundary && (flags & BND_WALL)) fn[4] = real(0);
#pragma unroll
        for (int v = 1; v < 5; ++v) {
          sSo[q * V + v] += w * fn[v];
          sSn[q * V + v] -= w * fn[v];
        }
      }
    }
    __syncthreads();
  }

  // Phase 5: project onto the face-support basis of each side, records
  // rec[k][v] = sum_q phi_k(x_q) S_q,v: with the tile sizes of phase 1a a
  // 16x16 tile product Phi^T (nodes x points) S (points x variables) on the
  // matrix cores, written straight to the records; otherwise four lanes per
  // support node split the points.
  if constexpr (MT) {
    const int lr = tid & 15, lk = tid >> 4;
    auto project = [&](const real* tb, const real* sS, int NF, real* out) {
      auto c = zero4<real>();
#pragma unroll
      for (int q0 = 0; q0 < 16; q0 += 4) {
        if (q0 >= QF) break;
        const int q = q0 + lk;
        const bool in = q < QF;
        const real phi = (in && lr < NF) ? tb[q * NF + lr] : real(0);
        const real sv = (in && lr < V) ? sS[q * V + lr] : real(0);
        c = mma(phi, sv, c);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int kk = crow<real>(tid, j);
        if (kk < NF && lr < V) out[kk * W + lr] = c[j];
      }
    };
    real* oo = a.out_o + f * (FO * W);
    project(tbo, sSo, FO, oo);
    if (visc)
      for (int k = tid; k < FO * 12; k += 64) oo[(k / 12) * W + V + (k % 12)] = sbo[k];
    if constexpr (NBR) {
      real* on = a.out_n + f * (FN * W);
      project(tbn, sSn, FN, on);
      if (visc)
        for (int k = tid; k < FN * 12; k += 64) on[(k / 12) * W + V + (k % 12)] = sbn[k];
    }
  } else
  {
    const int kk = tid >> 2, part = tid & 3;
    real ao[V], an[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      ao[v] = real(0);
      an[v] = real(0);
    }
    for (int q = part; q < QF; q += 4) {
      if (kk < FO) {
        const real c = tbo[q * FO + kk];
#pragma unroll
        for (int v = 0; v < V; ++v) ao[v] += c * sSo[q * V + v];
      }
      if constexpr (NBR) {
        if (kk < FN) {
          const real c = tbn[q * FN + kk];
#pragma unroll
          for (int v = 0; v < V; ++v) an[v] += c * sSn[q * V + v];
        }
      }
    }
#pragma unroll
    for (int v = 0; v < V; ++v) {
      ao[v] = group_sum<4>(ao[v]);
      an[v] = group_sum<4>(an[v]);
    }
    real* oo = a.out_o + f * (FO * W);
    if (part == 0 && kk < FO)
#pragma unroll
      for (int v = 0; v < V; ++v) oo[kk * W + v] = ao[v];
    if (visc)
      for (int k = tid; k < FO * 12; k += 64) oo[(k / 12) * W + V + (k % 12)] = sbo[k];
    if constexpr (NBR) {
      real* on = a.out_n + f * (FN * W);
      if (part == 0 && kk < FN)
#pragma unroll
        for (int v = 0; v < V; ++v) on[kk * W + v] = an[v];
      if (visc)
        for (int k = tid; k < FN * 12; k += 64) on[(k / 12) * W + V + (k % 12)] = sbn[k];
    }
  }
}

template <int FT, int TO, int TN, int SUB, int P, int V, typename real, bool WALE, bool RESTRICTED>
__global__ __launch_bounds__(64) void face_kernel(FaceArgs<real> a, Phys<real> ph) {
  constexpr bool NBR = TN >= 0;
  // a restricted pass evaluates the faces with a marked side
  for_each_item<RESTRICTED>(
      a.count,
      [&](int64_t f) {
        return a.mask_o[a.owner_elem[f]] != 0 || (NBR && a.mask_n[a.neighbor_elem[f]] != 0);
      },
      [&](int64_t f) { face_body<FT, TO, TN, SUB, P, V, real, WALE>(a, ph, f); });
}

}  // namespace fsg

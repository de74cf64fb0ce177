// Element kernel as batched dense contractions on the matrix cores.
//
// One workgroup (4 wavefronts) takes EB elements.  Every per-element
// operator of hybrid/spatial.py:47-72 and hybrid/navier_stokes.py:138-234 is
// a product of a reference matrix (shared by all elements of the type) with
// the batch's nodal data laid out as columns (element-major, variable-minor):
//
//   lift            L  = M^-1 b              (N x N)   x (N x 12 EB)
//   interpolation   Uq = B (U - a) + a       (Q x N)   x (N x V EB)
//   raw gradient    Gr = D_r v               (3Q x N)  x (N x 4 EB)
//   lift at points  gl = B L                 (Q x N)   x (N x 12 EB)
//   divergence      R += D^T F^              (N x 3Q)  x (3Q x V EB)
//   inverse mass    rhs = M^-1 R             (N x N)   x (N x V EB)
//
// run as 16x16x4 MFMA tiles (f32 for the FP32 path, f64 for the plain path),
// the reference matrices streamed from the L2-resident tables as A operands
// and the batch data from LDS as B operands.  The volume points are taken in
// chunks of 16: the chunk's interpolants go to LDS, the pointwise fluxes
// (anchored Euler flux difference minus the BR2 viscous flux, pulled back by
// J^-1 and weighted) are formed per (point, element), and the chunk's
// contribution to the divergence accumulates in registers.
#pragma once
#include "fsg_kernels.h"
#include "fsg_mfma.h"

namespace fsg {

// Elements per workgroup: as many as the LDS budget (160 KB) allows for
// hexahedra at p = 3 with five or sixteen variables.
template <int V, typename real>
__host__ __device__ constexpr int mfma_batch() {
  return V <= 5 ? (sizeof(real) == 4 ? 16 : 6) : (sizeof(real) == 4 ? 4 : 2);
}

template <int T, int P, int V, typename real>
struct MfmaShape {
  static constexpr int EB = mfma_batch<V, real>();
  static constexpr int N = n_nodes(T, P);
  static constexpr int Q = n_vol_q(T, P);
  static constexpr int NK = (N + 3) / 4 * 4;      // K extent over the nodes
  static constexpr int NR = (N + 15) / 16;        // row tiles over the nodes
  static constexpr int QC = (Q + 15) / 16;        // chunks of volume points
  static constexpr int CU = (EB * V + 15) / 16;   // column tiles of U / R
  static constexpr int CV = (EB * 4 + 15) / 16;   // ... of the viscous variables
  static constexpr int CL = (EB * 12 + 15) / 16;  // ... of the lifts
  static constexpr int WU = CU * 16, WV = CV * 16, WL = CL * 16;
  static constexpr int LDS = (NK * WU * 2 + WU + NK * WV + NK * WL + 16 * WU + 3 * 16 * WV +
                              16 * WL + 48 * WU + EB * 11) * (int)sizeof(real);
  static_assert(LDS <= 160 * 1024, "element batch exceeds the LDS budget");
};

template <int T, int P, int V, typename real>
__global__ __launch_bounds__(256) void element_mfma_kernel(ElemArgs<real> a, Phys<real> ph,
                                                           int64_t count) {
  using S = MfmaShape<T, P, V, real>;
  using v4 = typename Acc<real>::v4;
  constexpr int EB = S::EB, N = S::N, Q = S::Q, NK = S::NK, NR = S::NR, QC = S::QC;
  constexpr int CU = S::CU, CV = S::CV, CL = S::CL, WU = S::WU, WV = S::WV, WL = S::WL;
  constexpr int W = V + 12;

  __shared__ __align__(16) real smem[S::LDS / sizeof(real)];
  real* sU = smem;                                 // [NK][WU] nodal states
  real* sR = sU + NK * WU;                         // [NK][WU] surface terms, then R
  real* sA = sR + NK * WU;                         // [WU] anchors (node 0)
  real* sV = sA + WU;                              // [NK][WV] viscous variables
  real* sB = sV + NK * WV;                         // [NK][WL] lift rhs, then lift
  real* sQ = sB + NK * WL;                         // [16][WU] chunk interpolants
  real* sG = sQ + 16 * WU;                         // [3][16][WV] chunk raw gradients
  real* sGL = sG + 3 * 16 * WV;                    // [16][WL] chunk lift values
  real* sF = sGL + 16 * WL;                        // [48][WU] chunk fluxes, row = 3 q + r
  real* sJ = sF + 48 * WU;                         // [EB][9] J^-1
  real* sS = sJ + EB * 9;                          // [EB] inverse-mass scale
  real* sD = sS + EB;                              // [EB] |det J|

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t e0 = (int64_t)blockIdx.x * EB;
  DiagSeen seen{};
  if (a.mode == 2 && a.diag && lane == 0) seen = diag_seen(a.diag);
  const int ne = (int)(count - e0 < EB ? count - e0 : EB);
  if (a.mask) {
    // masked evaluation: a batch without a marked element has nothing to do
    // (the decision is uniform over the workgroup, before any barrier)
    bool any = false;
    for (int e = 0; e < ne; ++e) any |= a.mask[e0 + e] != 0;
    if (!any) return;
  }
  const bool euler = a.terms & TERM_EULER;
  const bool visc = a.terms & TERM_VISCOUS;
  const real gamma = ph.gamma;

  // -- load the batch --------------------------------------------------------
  for (int k = tid; k < NK * WU; k += 256) {
    sU[k] = real(0);
    sR[k] = real(0);
  }
  for (int k = tid; k < NK * WL; k += 256) sB[k] = real(0);
  for (int k = tid; k < NK * WV; k += 256) sV[k] = real(0);
  for (int k = tid; k < EB * 9; k += 256) {
    const int e = k / 9;
    sJ[k] = e < ne ? a.jinv[(e0 + e) * 9 + (k - e * 9)] : real(0);
  }
  for (int e = tid; e < EB; e += 256) {
    sS[e] = e < ne ? a.scale[e0 + e] : real(0);
    sD[e] = e < ne ? a.detw[e0 + e] : real(0);
  }
  __syncthreads();
  {
    const real* Ub = a.U + e0 * (N * V);
    for (int k = tid; k < ne * N * V; k += 256) {
      const int e = k / (N * V), r = k - e * (N * V), i = r / V, v = r - i * V;
      sU[i * WU + e * V + v] = Ub[k];
    }
  }
  __syncthreads();
  for (int c = tid; c < WU; c += 256) sA[c] = sU[c];
  if (visc)
    for (int k = tid; k < N * EB; k += 256) {
      const int i = k / EB, e = k - i * EB;
      real vv[4];
      viscous_variables<V, real>(sU + i * WU + e * V, gamma, ph.gas_constant, vv);
#pragma unroll
      for (int m = 0; m < 4; ++m) sV[i * WV + e * 4 + m] = e < ne ? vv[m] : real(0);
    }

  // -- face-side records: surface terms -> sR, lift right-hand sides -> sB --
  // Round r adds the r-th record of every element of the batch: within a
  // round the entries touch distinct (node, column) pairs, so the 256
  // threads issue their loads independently; rounds are ordered by barriers,
  // so each node's sum runs in record order (deterministic).
  {
    __shared__ int64_t sRec[EB + 1];
    __shared__ int sRounds;
    if (tid <= EB) sRec[tid] = a.rec_ptr[e0 + (tid < ne ? tid : ne)];
    __syncthreads();
    if (tid == 0) {
      int m = 0;
      for (int e = 0; e < ne; ++e) m = max(m, (int)(sRec[e + 1] - sRec[e]));
      sRounds = m;
    }
    __syncthreads();
    const int width = visc ? W : V;
    constexpr int MAXF = (P + 1) * (P + 1);          // support nodes of any face side
    const int per = MAXF * width;
    for (int r = 0; r < sRounds; ++r) {
      for (int idx = tid; idx < ne * per; idx += 256) {
        const int e = idx / per, k = idx - e * per;
        const int64_t rr = sRec[e] + r;
        if (rr >= sRec[e + 1]) continue;
        const int nf = a.rec_nf[rr];
        if (k >= nf * width) continue;
        const int kk = k / width, c = k - kk * width;
        const int node = a.support[a.rec_sup[rr] + kk];
        const real val = a.face_out[a.rec_off[rr] + kk * W + c];
        if (c < V)
          sR[node * WU + e * V + c] += val;
        else
          sB[node * WL + e * 12 + (c - V)] += val;
      }
      __syncthreads();
    }
  }

  // -- BR2 lift L = scale M^-1 b (hybrid/navier_stokes.py:157-186) ------------
  if (visc) {
    // wave w: row tile w / CGL, column tiles of group w % CGL (A shared by the group)
    constexpr int CGL = NR >= 4 ? 1 : 4 / NR, GL = (CL + CGL - 1) / CGL;
    const int rt = wave / CGL, cg = wave - rt * CGL;
    v4 acc[GL];
#pragma unroll
    for (int g = 0; g < GL; ++g) acc[g] = zero4<real>();
    if (rt < NR) {
#pragma unroll
      for (int k0 = 0; k0 < NK; k0 += 4) {
        const int i = rt * 16 + (lane & 15), k = k0 + (lane >> 4);
        const real av = (i < N && k < N) ? a.minvT[k * N + i] : real(0);
#pragma unroll
        for (int g = 0; g < GL; ++g) {
          const int ct = cg * GL + g;
          if (ct < CL) acc[g] = mma(av, sB[k * WL + ct * 16 + (lane & 15)], acc[g]);
        }
      }
    }
    __syncthreads();
    if (rt < NR) {
#pragma unroll
      for (int g = 0; g < GL; ++g) {
        const int ct = cg * GL + g;
        if (ct >= CL) continue;
        const int col = ct * 16 + (lane & 15), e = col / 12;
        const real s = e < EB ? sS[e] : real(0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int i = rt * 16 + crow<real>(lane, j);
          if (i < NK) sB[i * WL + col] = i < N ? s * acc[g][j] : real(0);
        }
      }
    }
    __syncthreads();
  }

  // -- volume points, in chunks of 16 ----------------------------------------
  // divergence / inverse-mass units: wave w -> row tile w / CGD, column group w % CGD
  constexpr int CGD = NR >= 4 ? 1 : 4 / NR, GD = (CU + CGD - 1) / CGD;
  const int drt = wave / CGD, dcg = wave - drt * CGD;
  v4 div[GD];
#pragma unroll
  for (int g = 0; g < GD; ++g) div[g] = zero4<real>();
  // chunk products against the basis rows (interpolation, lift at the points):
  // wave w takes tiles w, w + 4, ... of [CU interpolation | CL lift]
  constexpr int BTMAX = CU + CL, MB = (BTMAX + 3) / 4;

  for (int qc = 0; qc < QC; ++qc) {
    const int q = qc * 16 + (lane & 15);
    {
      const int BT = CU + (visc ? CL : 0);
      v4 acc[MB];
#pragma unroll
      for (int m = 0; m < MB; ++m) acc[m] = zero4<real>();
#pragma unroll
      for (int k0 = 0; k0 < NK; k0 += 4) {
        const int k = k0 + (lane >> 4);
        const real av = (q < Q && k < N) ? a.Bt[k * Q + q] : real(0);
#pragma unroll
        for (int m = 0; m < MB; ++m) {
          const int t = wave + 4 * m;
          if (t < BT) {
            real bv;
            if (t < CU) {
              const int col = t * 16 + (lane & 15);
              bv = sU[k * WU + col] - sA[col];
            } else {
              bv = sB[k * WL + (t - CU) * 16 + (lane & 15)];
            }
            acc[m] = mma(av, bv, acc[m]);
          }
        }
      }
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const int t = wave + 4 * m;
        if (t >= BT) continue;
        if (t < CU) {
          const int col = t * 16 + (lane & 15);
#pragma unroll
          for (int j = 0; j < 4; ++j) sQ[crow<real>(lane, j) * WU + col] = sA[col] + acc[m][j];
        } else {
          const int col = (t - CU) * 16 + (lane & 15);
#pragma unroll
          for (int j = 0; j < 4; ++j) sGL[crow<real>(lane, j) * WL + col] = acc[m][j];
        }
      }
    }
    if (visc && wave < 3) {
      // raw reference gradient along direction r = wave
      const int r = wave;
      v4 acc[CV];
#pragma unroll
      for (int c = 0; c < CV; ++c) acc[c] = zero4<real>();
#pragma unroll
      for (int k0 = 0; k0 < NK; k0 += 4) {
        const int k = k0 + (lane >> 4);
        const real av = (q < Q && k < N) ? a.Dt[(k * 3 + r) * Q + q] : real(0);
#pragma unroll
        for (int c = 0; c < CV; ++c) acc[c] = mma(av, sV[k * WV + c * 16 + (lane & 15)], acc[c]);
      }
#pragma unroll
      for (int c = 0; c < CV; ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          sG[(r * 16 + crow<real>(lane, j)) * WV + c * 16 + (lane & 15)] = acc[c][j];
    }
    __syncthreads();

    // pointwise fluxes per (point, element)
    for (int p = tid; p < 16 * EB; p += 256) {
      const int ql = p / EB, e = p - ql * EB, q = qc * 16 + ql;
      real Fh[3][V];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int v = 0; v < V; ++v) Fh[r][v] = real(0);
      if (q < Q && e < ne) {
        const real* J = sJ + e * 9;
        const real* Uq = sQ + ql * WU + e * V;
        const real* An = sA + e * V;
        real dF[3 * V];
#pragma unroll
        for (int k = 0; k < 3 * V; ++k) dF[k] = real(0);
        real Ua[V];
#pragma unroll
        for (int v = 0; v < V; ++v) Ua[v] = Uq[v];
        if (euler) {
          real Av[V];
#pragma unroll
          for (int v = 0; v < V; ++v) Av[v] = An[v];
          Prim<V, real> wq = c2p<V, real>(Ua, gamma);
          Prim<V, real> wa = c2p<V, real>(Av, gamma);
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            real Fq[V], Fa[V];
            axis_flux<V, real>(Ua, wq, d, Fq);
            axis_flux<V, real>(Av, wa, d, Fa);
#pragma unroll
            for (int v = 0; v < V; ++v) dF[d * V + v] = Fq[v] - Fa[v];
          }
        }
        if (visc) {
          real g[12];
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            const real G0 = sG[(0 * 16 + ql) * WV + e * 4 + m];
            const real G1 = sG[(1 * 16 + ql) * WV + e * 4 + m];
            const real G2 = sG[(2 * 16 + ql) * WV + e * 4 + m];
#pragma unroll
            for (int d = 0; d < 3; ++d)
              g[m * 3 + d] = sGL[ql * WL + e * 12 + m * 3 + d] +
                             (G0 * J[0 * 3 + d] + G1 * J[1 * 3 + d] + G2 * J[2 * 3 + d]);
          }
          real Fv[3 * V];
          viscous_flux<V, real>(Ua, g, gamma, ph.mu, ph.kappa, Fv);
#pragma unroll
          for (int k = 0; k < 3 * V; ++k) dF[k] -= Fv[k];
        }
        const real w = a.wq[q] * sD[e];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int v = 0; v < V; ++v)
            Fh[r][v] = w * (J[r * 3 + 0] * dF[0 * V + v] + J[r * 3 + 1] * dF[1 * V + v] +
                            J[r * 3 + 2] * dF[2 * V + v]);
      }
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int v = 0; v < V; ++v) sF[(ql * 3 + r) * WU + e * V + v] = Fh[r][v];
    }
    __syncthreads();

    // R += D^T F^ over this chunk: K = 48 = 16 points x 3 directions
    if (drt < NR) {
      const int i = drt * 16 + (lane & 15);
#pragma unroll
      for (int k0 = 0; k0 < 48; k0 += 4) {
        const int k = k0 + (lane >> 4), ql = k / 3, r = k - ql * 3, qq = qc * 16 + ql;
        const real av = (qq < Q && i < N) ? a.D[(qq * 3 + r) * N + i] : real(0);
#pragma unroll
        for (int g = 0; g < GD; ++g) {
          const int ct = dcg * GD + g;
          if (ct < CU) div[g] = mma(av, sF[k * WU + ct * 16 + (lane & 15)], div[g]);
        }
      }
    }
    __syncthreads();
  }

  // -- residual: R = volume terms + surface terms ----------------------------
  if (drt < NR) {
#pragma unroll
    for (int g = 0; g < GD; ++g) {
      const int ct = dcg * GD + g;
      if (ct >= CU) continue;
      const int col = ct * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = drt * 16 + crow<real>(lane, j);
        if (i < NK) sR[i * WU + col] += div[g][j];
      }
    }
  }
  __syncthreads();
  if (a.mode == 0) {
    real* ob = a.out + e0 * (N * V);
    for (int k = tid; k < ne * N * V; k += 256) {
      const int e = k / (N * V), r = k - e * (N * V), i = r / V, v = r - i * V;
      if (a.mask && a.mask[e0 + e] == 0) continue;
      ob[k] = sR[i * WU + e * V + v];
    }
    return;
  }

  // -- rhs = scale M^-1 R, then the stage epilogue -----------------------------
  if (drt < NR) {
#pragma unroll
    for (int g = 0; g < GD; ++g) div[g] = zero4<real>();
    const int i = drt * 16 + (lane & 15);
#pragma unroll
    for (int k0 = 0; k0 < NK; k0 += 4) {
      const int k = k0 + (lane >> 4);
      const real av = (i < N && k < N) ? a.minvT[k * N + i] : real(0);
#pragma unroll
      for (int g = 0; g < GD; ++g) {
        const int ct = dcg * GD + g;
        if (ct < CU) div[g] = mma(av, sR[k * WU + ct * 16 + (lane & 15)], div[g]);
      }
    }
  }
  __syncthreads();   // every wave has read sR: it now holds the new state
  if (drt < NR) {
#pragma unroll
    for (int g = 0; g < GD; ++g) {
      const int ct = dcg * GD + g;
      if (ct >= CU) continue;
      const int col = ct * 16 + (lane & 15), e = col / V, v = col - e * V;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = drt * 16 + crow<real>(lane, j);
        if (i >= N || e >= ne) continue;
        const real rhs = sS[e] * div[g][j];
        const int64_t gi = ((e0 + e) * N + i) * V + v;
        if (a.mode == 1) {
          if (!a.mask || a.mask[e0 + e] != 0) a.out[gi] = rhs;
        } else {
          const real dte = a.dt_elem ? a.dt_elem[e0 + e] : a.dt;
          const real nxt = (a.Ubase ? a.alpha * a.Ubase[gi] : real(0)) +
                           a.beta * (sU[i * WU + col] + dte * rhs);
          a.out[gi] = nxt;
          sR[i * WU + col] = nxt;
        }
      }
    }
  }
  if (a.mode != 2 || !a.diag) return;
  __syncthreads();

  // Admissibility at the nodes (hybrid/time_integration.py:291-312).
  double rmn = __longlong_as_double(0x7ff0000000000000LL), pmn = rmn;
  unsigned int bad = 0;
  for (int k = tid; k < N * ne; k += 256) {
    const int i = k / ne, e = k - i * ne;
    const real* u = sR + i * WU + e * V;
    for (int v = 0; v < V; ++v) bad += finite(u[v]) ? 0u : 1u;
    const real rho = u[0];
    const real ke = real(0.5) * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / rmax(rho, real(1.0e-14));
    const real p = (gamma - real(1)) * (u[4] - ke);
    rmn = nan_min(rmn, double(rho));
    pmn = nan_min(pmn, double(p));
  }
  for (int off = 32; off > 0; off >>= 1) {
    rmn = nan_min(rmn, __shfl_down(rmn, off, 64));
    pmn = nan_min(pmn, __shfl_down(pmn, off, 64));
    bad += __shfl_down(bad, off, 64);
  }
  if (lane == 0) {
    atomic_min_nan<real>(&a.diag->rho_min, rmn, seen.rho);
    atomic_min_nan<real>(&a.diag->p_min, pmn, seen.p);
    if (bad) atomicAdd(&a.diag->nonfinite, bad);
  }
}

}  // namespace fsg

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
  const int ne = (int)(count - e0 < EB ? count - e0 : EB);
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
  {
    const int grp = tid >> 4, gl = tid & 15;   // 16 lanes per element
    for (int e = grp; e < ne; e += 16) {
      const int64_t r0 = a.rec_ptr[e0 + e], r1 = a.rec_ptr[e0 + e + 1];
      const int width = visc ? W : V;
      for (int64_t r = r0; r < r1; ++r) {
        const real* rec = a.face_out + a.rec_off[r];
        const int32_t* sup = a.support + a.rec_sup[r];
        const int nf = a.rec_nf[r];
        for (int k = gl; k < nf * width; k += 16) {
          const int kk = k / width, c = k - kk * width;
          const int node = sup[kk];
          const real val = rec[kk * W + c];
          if (c < V)
            sR[node * WU + e * V + c] += val;
          else
            sB[node * WL + e * 12 + (c - V)] += val;
        }
      }
    }
  }
  __syncthreads();

  // -- BR2 lift L = scale M^-1 b (hybrid/navier_stokes.py:157-186) ------------
  if (visc) {
    constexpr int CT4 = (CL + 3) / 4;   // column tiles per wave: wave, wave + 4, ...
    v4 acc[NR * CT4];
#pragma unroll
    for (int rt = 0; rt < NR; ++rt)
#pragma unroll
      for (int ci = 0; ci < CT4; ++ci) {
        const int ct = wave + 4 * ci;
        v4 c = zero4<real>();
        if (ct < CL) {
          for (int k0 = 0; k0 < NK; k0 += 4) {
            const int i = rt * 16 + (lane & 15), k = k0 + (lane >> 4);
            const real av = (i < N && k < N) ? a.minvT[k * N + i] : real(0);
            const real bv = sB[(k0 + (lane >> 4)) * WL + ct * 16 + (lane & 15)];
            c = mma(av, bv, c);
          }
        }
        acc[rt * CT4 + ci] = c;
      }
    __syncthreads();
#pragma unroll
    for (int rt = 0; rt < NR; ++rt)
#pragma unroll
      for (int ci = 0; ci < CT4; ++ci) {
        const int ct = wave + 4 * ci;
        if (ct >= CL) continue;
        const v4 c = acc[rt * CT4 + ci];
        const int col = ct * 16 + (lane & 15), e = col / 12;
        const real s = e < EB ? sS[e] : real(0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int i = rt * 16 + crow<real>(lane, j);
          if (i < NK) sB[i * WL + col] = i < N ? s * c[j] : real(0);
        }
      }
    __syncthreads();
  }

  // -- volume points, in chunks of 16 ----------------------------------------
  constexpr int DT = (NR * CU + 3) / 4;   // divergence tiles per wave
  v4 div[DT];
#pragma unroll
  for (int t = 0; t < DT; ++t) div[t] = zero4<real>();

  for (int qc = 0; qc < QC; ++qc) {
    // interpolants: tiles [interp CU | grad 3 CV | lift CL], round-robin over waves
    const int ntile = CU + (visc ? 3 * CV + CL : 0);
    for (int t = wave; t < ntile; t += 4) {
      v4 c = zero4<real>();
      const int q = qc * 16 + (lane & 15);
      if (t < CU) {
        const int col = t * 16 + (lane & 15);
        for (int k0 = 0; k0 < NK; k0 += 4) {
          const int k = k0 + (lane >> 4);
          const real av = (q < Q && k < N) ? a.Bt[k * Q + q] : real(0);
          const int kb = k0 + (lane >> 4);
          const real bv = sU[kb * WU + col] - sA[col];
          c = mma(av, bv, c);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int ql = crow<real>(lane, j);
          sQ[ql * WU + col] = sA[col] + c[j];
        }
      } else if (t < CU + 3 * CV) {
        const int u = t - CU, r = u / CV, ct = u - r * CV;
        const int col = ct * 16 + (lane & 15);
        for (int k0 = 0; k0 < NK; k0 += 4) {
          const int k = k0 + (lane >> 4);
          const real av = (q < Q && k < N) ? a.Dt[(k * 3 + r) * Q + q] : real(0);
          const real bv = sV[(k0 + (lane >> 4)) * WV + col];
          c = mma(av, bv, c);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) sG[(r * 16 + crow<real>(lane, j)) * WV + col] = c[j];
      } else {
        const int ct = t - CU - 3 * CV;
        const int col = ct * 16 + (lane & 15);
        for (int k0 = 0; k0 < NK; k0 += 4) {
          const int k = k0 + (lane >> 4);
          const real av = (q < Q && k < N) ? a.Bt[k * Q + q] : real(0);
          const real bv = sB[(k0 + (lane >> 4)) * WL + col];
          c = mma(av, bv, c);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) sGL[(crow<real>(lane, j)) * WL + col] = c[j];
      }
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
#pragma unroll
    for (int t = 0; t < DT; ++t) {
      const int tile = wave + 4 * t;
      if (tile < NR * CU) {
        const int rt = tile / CU, ct = tile - rt * CU;
        const int i = rt * 16 + (lane & 15), col = ct * 16 + (lane & 15);
        v4 c = div[t];
        for (int k0 = 0; k0 < 48; k0 += 4) {
          const int k = k0 + (lane >> 4), ql = k / 3, r = k - ql * 3, q = qc * 16 + ql;
          const real av = (q < Q && i < N) ? a.D[(q * 3 + r) * N + i] : real(0);
          const real bv = sF[k * WU + col];
          c = mma(av, bv, c);
        }
        div[t] = c;
      }
    }
    __syncthreads();
  }

  // -- residual: R = volume terms + surface terms ----------------------------
#pragma unroll
  for (int t = 0; t < DT; ++t) {
    const int tile = wave + 4 * t;
    if (tile < NR * CU) {
      const int rt = tile / CU, ct = tile - rt * CU, col = ct * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = rt * 16 + crow<real>(lane, j);
        if (i < NK) sR[i * WU + col] += div[t][j];
      }
    }
  }
  __syncthreads();
  if (a.mode == 0) {
    real* ob = a.out + e0 * (N * V);
    for (int k = tid; k < ne * N * V; k += 256) {
      const int e = k / (N * V), r = k - e * (N * V), i = r / V, v = r - i * V;
      ob[k] = sR[i * WU + e * V + v];
    }
    return;
  }

  // -- rhs = scale M^-1 R, then the stage epilogue -----------------------------
#pragma unroll
  for (int t = 0; t < DT; ++t) {
    const int tile = wave + 4 * t;
    if (tile >= NR * CU) continue;
    const int rt = tile / CU, ct = tile - rt * CU;
    const int col = ct * 16 + (lane & 15);
    v4 c = zero4<real>();
    for (int k0 = 0; k0 < NK; k0 += 4) {
      const int i = rt * 16 + (lane & 15), k = k0 + (lane >> 4);
      const real av = (i < N && k < N) ? a.minvT[k * N + i] : real(0);
      const real bv = sR[(k0 + (lane >> 4)) * WU + col];
      c = mma(av, bv, c);
    }
    div[t] = c;
  }
  __syncthreads();   // every wave has read sR: it now holds the new state
#pragma unroll
  for (int t = 0; t < DT; ++t) {
    const int tile = wave + 4 * t;
    if (tile >= NR * CU) continue;
    const int rt = tile / CU, ct = tile - rt * CU;
    const int col = ct * 16 + (lane & 15), e = col / V, v = col - e * V;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = rt * 16 + crow<real>(lane, j);
      if (i >= N || e >= ne) continue;
      const real rhs = sS[e] * div[t][j];
      const int64_t g = ((e0 + e) * N + i) * V + v;
      if (a.mode == 1) {
        a.out[g] = rhs;
      } else {
        const real dte = a.dt_elem ? a.dt_elem[e0 + e] : a.dt;
        const real nxt = a.alpha * a.Ubase[g] + a.beta * (sU[i * WU + col] + dte * rhs);
        a.out[g] = nxt;
        sR[i * WU + col] = nxt;
      }
    }
  }
  if (a.mode != 2) return;
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
    atomic_min_nan<real>(&a.diag->rho_min, rmn);
    atomic_min_nan<real>(&a.diag->p_min, pmn);
    if (bad) atomicAdd(&a.diag->nonfinite, bad);
  }
}

}  // namespace fsg

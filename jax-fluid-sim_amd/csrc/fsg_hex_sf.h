// Element kernel for hexahedra at p >= 2 by sum factorisation.
//
// The hexahedral nodes and volume points are tensor products in the
// reference's own ordering (x fastest: hybrid/elements.py nodes and the
// Gauss rule of volume_quadrature), so every dense reference operator of
// the element kernel is a Kronecker product of 1-D ones:
//
//   B  = B1 (x) B1 (x) B1,  d/dxi_x = B1 (x) B1 (x) D1 (and cyclic),
//   M^-1 = M1^-1 (x) M1^-1 (x) M1^-1,
//
// with B1 (Q1 x N1) the line basis at the Gauss points, D1 its derivative and
// M1 = B1^T W1 B1.  Each dense contraction (N x Q per variable) becomes three
// passes of N1- or Q1-term sums through LDS: at p = 3 the element costs
// ~0.15 Mflop instead of the dense 0.85 Mflop.  The arithmetic per point (the
// anchored Euler flux difference, the BR2 viscous flux, the pull-back by
// J^-1 and the weights) is the dense kernel's; the sums run in a different
// order, so results agree to round-off, not bit for bit.  A uniform state
// still gives an exactly zero Euler residual: the interpolation acts on
// U - a, which is exactly zero.
//
// The volume state at the points is the anchored interpolant a + B (U - a)
// for both the Euler and the viscous flux (the dense kernel uses the plain
// interpolant B U for the viscous flux: the two differ by round-off).
#pragma once
#include <type_traits>

#include "fsg_kernels.h"

namespace fsg {

template <int P>
struct HexSF {
  static constexpr int N1 = P + 1;
  static constexpr int Q1 = P == 3 ? 5 : 3;   // Gauss points per direction
  static constexpr int N = N1 * N1 * N1;
  static constexpr int Q = Q1 * Q1 * Q1;
};

// 1-D tables, one block of reals: B1 (Q1, N1), D1 (Q1, N1), M1inv (N1, N1).
template <int P>
__host__ __device__ constexpr int hex_sf_table_size() {
  return 2 * HexSF<P>::Q1 * HexSF<P>::N1 + HexSF<P>::N1 * HexSF<P>::N1;
}

__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }

// One line ("pencil") of a 1-D pass: out[o * so] = sc * sum_i A(o, i) in[i * si]
// with A(o, i) = op[o * NI + i], or op[i * NO + o] when TR (transposed);
// NI inputs read once, NO outputs.
template <int NO, int NI, bool TR, typename real>
__device__ __forceinline__ void pencil(const real* op, const real* in, int si, real* out, int so,
                                       real sub = real(0), real sc = real(1)) {
  real x[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) x[i] = in[i * si] - sub;
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    real acc = real(0);
#pragma unroll
    for (int i = 0; i < NI; ++i) acc += op[TR ? i * NO + o : o * NI + i] * x[i];
    out[o * so] = sc * acc;
  }
}

// The same line contraction with the outputs kept in registers.
template <int NO, int NI, typename real>
__device__ __forceinline__ void pencil_vals(const real* op, const real* in, int si, real* out) {
  real x[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) x[i] = in[i * si];
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    real acc = real(0);
#pragma unroll
    for (int i = 0; i < NI; ++i) acc += op[o * NI + i] * x[i];
    out[o] = acc;
  }
}

template <int P, int V, typename real, int NT>
__device__ __forceinline__ void hex_sf_body(const ElemArgs<real>& a, const Phys<real>& ph,
                                            const int64_t e) {
  using S = HexSF<P>;
  constexpr int N1 = S::N1, Q1 = S::Q1, N = S::N, Q = S::Q;
  constexpr int W = V + 12;
  // column layouts of the interpolation passes
  constexpr int CI = V + 16;        // x pass with B1: U - a (V), lift (12), v (4)
  constexpr int CX = CI + 4;        // ... plus D1 along x of v (4)
  constexpr int CY = CI + 8;        // y pass: B1 of CI, D1 of v (d/dy), B1 of the x derivative
  constexpr int CZ = V + 12;        // z pass: U - a, g = J^-T grad v + lift
  constexpr int SX = N1 * N1 * Q1 * CX, SY = N1 * Q1 * Q1 * CY, SZ = Q * CZ;
  constexpr int SF = Q * 3 * V, STZ = N1 * Q1 * Q1 * 3 * V, STY = N1 * N1 * Q1 * 2 * V;
  // the transposed z pass runs in place in w2 (each thread reads its whole
  // line before writing the N1 <= Q1 results back into it); the transposed
  // y pass writes w1
  constexpr int W1S = cmax(cmax(SX, SZ), cmax(STY, N * 12));
  // w2 also holds the lift right-hand side / lift (sb), the viscous
  // variables (sv) and the element state (sU) until the x pass of the
  // interpolation has read them: they sit behind the first N * 12 entries,
  // which the inverse-mass passes of the lift use as their middle buffer;
  // the y pass overwrites them later.  The anchor (node 0) is kept apart and
  // the stage epilogue reads the state again from memory.
  constexpr int W2S = cmax(cmax(cmax(SY, SF), cmax(STZ, N * 12)), N * (28 + V));

  __shared__ real sR[N * V];
  __shared__ real sA[V];
  __shared__ real w1[W1S];
  __shared__ real w2[W2S];
  real* const sb = w2 + N * 12;      // [N][12], dead after the x pass
  real* const sv = w2 + N * 24;      // [N][4], dead after the x pass
  real* const sU = w2 + N * 28;      // [N][V], dead after the x pass
  __shared__ real sB1[Q1 * N1], sD1[Q1 * N1], sM1[N1 * N1];
  __shared__ double red[2][NT / 64];
  __shared__ unsigned int redn[NT / 64];

  const int tid = threadIdx.x;
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
        if (k < V) sA[k] = t[i];
        sR[k] = real(0);
      }
    }
  }
  for (int k = tid; k < N * 12; k += NT) sb[k] = real(0);
  for (int k = tid; k < hex_sf_table_size<P>(); k += NT) {
    const real t = a.sf1d[k];
    if (k < Q1 * N1) sB1[k] = t;
    else if (k < 2 * Q1 * N1) sD1[k - Q1 * N1] = t;
    else sM1[k - 2 * Q1 * N1] = t;
  }
  __syncthreads();
  if (visc)
    for (int i = tid; i < N; i += NT)
      viscous_variables<V, real>(sU + i * V, gamma, ph.gas_constant, sv + i * 4);

  // Face-side records: surface terms into sR, lift right-hand sides into sb.
  add_face_records<NT, (N1 * N1 * W + NT - 1) / NT, V, real>(a, e, visc, sR, sb, tid);
  __syncthreads();

  const real s = a.scale[e];
  const real dw = a.detw[e];
  real J[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) J[k] = a.jinv[e * 9 + k];

  // inverse mass by three 1-D passes: out = sc M^-1 in over C columns, one
  // thread per (line, column); C is a compile-time constant.
  auto minv3 = [&](const real* in, real* out, auto ctag, real sc) {
    constexpr int C = decltype(ctag)::value;
    for (int t = tid; t < N1 * N1 * C; t += NT) {    // x lines: (k, j)
      const int jk = t / C, c = t - jk * C;
      pencil<N1, N1, false, real>(sM1, in + jk * N1 * C + c, C, w1 + jk * N1 * C + c, C);
    }
    __syncthreads();
    for (int t = tid; t < N1 * N1 * C; t += NT) {    // y lines: (k, i)
      const int ki = t / C, c = t - ki * C, k = ki / N1, i = ki - k * N1;
      const int base = (k * N1 * N1 + i) * C + c;
      pencil<N1, N1, false, real>(sM1, w1 + base, N1 * C, w2 + base, N1 * C);
    }
    __syncthreads();
    for (int t = tid; t < N1 * N1 * C; t += NT) {    // z lines: (j, i)
      const int ij = t / C, c = t - ij * C;
      pencil<N1, N1, false, real>(sM1, w2 + ij * C + c, N1 * N1 * C, out + ij * C + c,
                                  N1 * N1 * C, real(0), sc);
    }
    __syncthreads();
  };

  // BR2 global lift r = s M^-1 b (hybrid/navier_stokes.py:157-186), in place.
  if (visc) minv3(sb, sb, std::integral_constant<int, 12>(), s);

  // Interpolation to the points by x, y, z passes, one thread per (line,
  // column) and each pass split by operator (B1 or D1) so that no thread
  // selects per term; the column counts are compile-time (all columns with
  // the viscous terms, the V state columns without).
  auto interpolate = [&](auto vtag) {
    constexpr bool VI = decltype(vtag)::value;
    constexpr int NB = VI ? CI : V;                  // x: B1 columns
    // x pass: (k, j, i) -> (k, j, a)
    for (int t = tid; t < N1 * N1 * NB; t += NT) {
      const int jk = t / NB, c = t - jk * NB;
      const real* src;
      int stride;
      real sub = real(0);
      if (c < V) { src = sU + c; stride = V; sub = sA[c]; }
      else if (c < V + 12) { src = sb + (c - V); stride = 12; }
      else { src = sv + (c - V - 12); stride = 4; }
      pencil<Q1, N1, false, real>(sB1, src + jk * N1 * stride, stride, w1 + jk * Q1 * CX + c, CX,
                                  sub);
    }
    if constexpr (VI)
      for (int t = tid; t < N1 * N1 * 4; t += NT) {
        const int jk = t >> 2, m = t & 3;
        pencil<Q1, N1, false, real>(sD1, sv + jk * N1 * 4 + m, 4, w1 + jk * Q1 * CX + CI + m, CX);
      }
    __syncthreads();
    // y pass: (k, j, a) -> (k, b, a); B1 on the CI plain columns and on the
    // x derivative (columns CI + 4.. from CI..), D1 on v (d/dy)
    constexpr int NBY = VI ? CI + 4 : V;
    for (int t = tid; t < N1 * Q1 * NBY; t += NT) {
      const int ka = t / NBY, c = t - ka * NBY, k = ka / Q1, qa = ka - k * Q1;
      const int dst = c < CI ? c : c + 4, src = c < CI ? c : CI + (c - CI);
      pencil<Q1, N1, false, real>(sB1, w1 + (k * N1 * Q1 + qa) * CX + src, Q1 * CX,
                                  w2 + (k * Q1 * Q1 + qa) * CY + dst, Q1 * CY);
    }
    if constexpr (VI)
      for (int t = tid; t < N1 * Q1 * 4; t += NT) {
        const int ka = t >> 2, m = t & 3, k = ka / Q1, qa = ka - k * Q1;
        pencil<Q1, N1, false, real>(sD1, w1 + (k * N1 * Q1 + qa) * CX + V + 12 + m, Q1 * CX,
                                    w2 + (k * Q1 * Q1 + qa) * CY + CI + m, Q1 * CY);
      }
    __syncthreads();
    // z pass: (k, b, a) -> (c, b, a); B1 on U - a.  With the viscous terms,
    // one thread per (line, m) also takes the lift (B1), dv_m/dx and dv_m/dy
    // (B1) and dv_m/dz (D1) along the line and stores the physical gradient
    // plus lift g = r + J^-T grad v per point, 12 columns instead of 24.
    for (int t = tid; t < Q1 * Q1 * V; t += NT) {
      const int ab = t / V, c = t - ab * V;
      pencil<Q1, N1, false, real>(sB1, w2 + ab * CY + c, Q1 * Q1 * CY, w1 + ab * CZ + c,
                                  Q1 * Q1 * CZ);
    }
    if constexpr (VI)
      for (int t = tid; t < Q1 * Q1 * 4; t += NT) {
        const int ab = t >> 2, m = t & 3;
        const real* in = w2 + ab * CY;
        real lf[3][Q1], gx[Q1], gy[Q1], gz[Q1];
#pragma unroll
        for (int d = 0; d < 3; ++d)
          pencil_vals<Q1, N1, real>(sB1, in + V + m * 3 + d, Q1 * Q1 * CY, lf[d]);
        pencil_vals<Q1, N1, real>(sB1, in + CI + 4 + m, Q1 * Q1 * CY, gx);
        pencil_vals<Q1, N1, real>(sB1, in + CI + m, Q1 * Q1 * CY, gy);
        pencil_vals<Q1, N1, real>(sD1, in + V + 12 + m, Q1 * Q1 * CY, gz);
#pragma unroll
        for (int c = 0; c < Q1; ++c)
#pragma unroll
          for (int d = 0; d < 3; ++d)
            w1[(c * Q1 * Q1 + ab) * CZ + V + m * 3 + d] =
                lf[d][c] + (gx[c] * J[0 * 3 + d] + gy[c] * J[1 * 3 + d] + gz[c] * J[2 * 3 + d]);
      }
    __syncthreads();
  };
  if (visc)
    interpolate(std::true_type());
  else
    interpolate(std::false_type());

  // Pointwise fluxes, pulled back to reference directions with the weight.
  // dF holds the five flow columns; a species column's Euler flux is U_s u_d
  // (its viscous flux is zero), so its pulled-back difference is formed from
  // the contravariant velocities.
  constexpr int VF = 5;
  for (int q = tid; q < Q; q += NT) {
    const real* z = w1 + q * CZ;
    real Ua[V];
#pragma unroll
    for (int v = 0; v < V; ++v) Ua[v] = z[v] + sA[v];
    real dF[3 * VF];
#pragma unroll
    for (int k = 0; k < 3 * VF; ++k) dF[k] = real(0);
    real uq[3] = {0, 0, 0}, ua[3] = {0, 0, 0};
    if (euler) {
      Prim<VF, real> wq = c2p<VF, real>(Ua, gamma);
      Prim<VF, real> wa = c2p<VF, real>(sA, gamma);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        real Fq[VF], Fa[VF];
        axis_flux<VF, real>(Ua, wq, d, Fq);
        axis_flux<VF, real>(sA, wa, d, Fa);
#pragma unroll
        for (int v = 0; v < VF; ++v) dF[d * VF + v] = Fq[v] - Fa[v];
        uq[d] = wq.vel[d];
        ua[d] = wa.vel[d];
      }
    }
    if (visc) {
      real g[12];
#pragma unroll
      for (int k = 0; k < 12; ++k) g[k] = z[V + k];
      real Fv[3 * VF];
      viscous_flux<VF, real>(Ua, g, gamma, ph.mu, ph.kappa, Fv);
#pragma unroll
      for (int k = 0; k < 3 * VF; ++k) dF[k] -= Fv[k];
    }
    const real w = a.wq[q] * dw;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
      for (int v = 0; v < VF; ++v)
        w2[q * 3 * V + r * V + v] =
            w * (J[r * 3 + 0] * dF[0 * VF + v] + J[r * 3 + 1] * dF[1 * VF + v] +
                 J[r * 3 + 2] * dF[2 * VF + v]);
      if constexpr (V > VF) {
        const real ur = J[r * 3 + 0] * uq[0] + J[r * 3 + 1] * uq[1] + J[r * 3 + 2] * uq[2];
        const real uar = J[r * 3 + 0] * ua[0] + J[r * 3 + 1] * ua[1] + J[r * 3 + 2] * ua[2];
#pragma unroll
        for (int v = VF; v < V; ++v) w2[q * 3 * V + r * V + v] = w * (Ua[v] * ur - sA[v] * uar);
      }
    }
  }
  __syncthreads();

  // Weak divergence R_i += sum_r (d phi_i / d xi_r)^T F^_r, transposed passes.
  // z: (c, b, a) -> (k, b, a), D1^T for r = z, B1^T otherwise
  for (int t = tid; t < Q1 * Q1 * 2 * V; t += NT) {
    const int ab = t / (2 * V), col = t - ab * (2 * V);
    pencil<N1, Q1, true, real>(sB1, w2 + ab * 3 * V + col, Q1 * Q1 * 3 * V, w2 + ab * 3 * V + col,
                               Q1 * Q1 * 3 * V);
  }
  for (int t = tid; t < Q1 * Q1 * V; t += NT) {
    const int ab = t / V, col = 2 * V + (t - ab * V);
    pencil<N1, Q1, true, real>(sD1, w2 + ab * 3 * V + col, Q1 * Q1 * 3 * V, w2 + ab * 3 * V + col,
                               Q1 * Q1 * 3 * V);
  }
  __syncthreads();
  // y: (k, b, a) -> (k, j, a); columns [0, V): D1^T of r = y plus B1^T of
  // r = z (both taken by B1^T in x next); [V, 2V): B1^T of r = x (D1^T in x)
  for (int t = tid; t < N1 * Q1 * V; t += NT) {
    const int ka = t / V, v = t - ka * V, k = ka / Q1, qa = ka - k * Q1;
    const real* in = w2 + (k * Q1 * Q1 + qa) * 3 * V;
    real ry[Q1], rz[Q1];
#pragma unroll
    for (int qb = 0; qb < Q1; ++qb) {
      ry[qb] = in[qb * Q1 * 3 * V + V + v];
      rz[qb] = in[qb * Q1 * 3 * V + 2 * V + v];
    }
    real* o = w1 + (k * N1 * Q1 + qa) * 2 * V + v;
#pragma unroll
    for (int j = 0; j < N1; ++j) {
      real acc = real(0);
#pragma unroll
      for (int qb = 0; qb < Q1; ++qb) acc += sD1[qb * N1 + j] * ry[qb] + sB1[qb * N1 + j] * rz[qb];
      o[j * Q1 * 2 * V] = acc;
    }
  }
  for (int t = tid; t < N1 * Q1 * V; t += NT) {
    const int ka = t / V, v = t - ka * V, k = ka / Q1, qa = ka - k * Q1;
    pencil<N1, Q1, true, real>(sB1, w2 + (k * Q1 * Q1 + qa) * 3 * V + v, Q1 * 3 * V,
                               w1 + (k * N1 * Q1 + qa) * 2 * V + V + v, Q1 * 2 * V);
  }
  __syncthreads();
  // x: (k, j, a) -> (k, j, i), added to the surface terms
  for (int t = tid; t < N1 * N1 * V; t += NT) {
    const int jk = t / V, v = t - jk * V;
    const real* in = w1 + jk * Q1 * 2 * V;
    real xb[Q1], xd[Q1];
#pragma unroll
    for (int qa = 0; qa < Q1; ++qa) {
      xb[qa] = in[qa * 2 * V + v];
      xd[qa] = in[qa * 2 * V + V + v];
    }
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      real acc = real(0);
#pragma unroll
      for (int qa = 0; qa < Q1; ++qa) acc += sB1[qa * N1 + i] * xb[qa] + sD1[qa * N1 + i] * xd[qa];
      sR[(jk * N1 + i) * V + v] += acc;
    }
  }
  __syncthreads();

  real* out = a.out + e * (N * V);
  if (a.mode == 0) {
    for (int k = tid; k < N * V; k += NT) out[k] = sR[k];
    return;
  }
  // rhs = scale M^-1 R (into sR), then the stage epilogue
  minv3(sR, sR, std::integral_constant<int, V>(), s);
  const real dte = a.dt_elem ? a.dt_elem[e] : a.dt;
  for (int k = tid; k < N * V; k += NT) {
    const real rhs = sR[k];
    if (a.mode == 1) {
      out[k] = rhs;
    } else {
      const real nxt = (a.Ubase ? a.alpha * a.Ubase[e * (N * V) + k] : real(0)) +
                       a.beta * (Ue[k] + dte * rhs);
      out[k] = nxt;
      sR[k] = nxt;
    }
  }
  if (a.mode != 2 || !a.diag) return;
  __syncthreads();

  // Admissibility at nodes (hybrid/time_integration.py:291-312).
  double rmn = __longlong_as_double(0x7ff0000000000000LL), pmn = rmn;
  unsigned int bad = 0;
  for (int i = tid; i < N; i += NT) {
    const real* u = sR + i * V;
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

template <int P, int V, typename real, int NT, bool RESTRICTED>
__global__ __launch_bounds__(NT) void hex_sf_kernel(ElemArgs<real> a, Phys<real> ph) {
  for_each_item<RESTRICTED>(a.count, [&](int64_t e) { return a.mask[e] != 0; },
                            [&](int64_t e) { hex_sf_body<P, V, real, NT>(a, ph, e); });
}

}  // namespace fsg

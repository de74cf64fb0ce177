// Pointwise compressible physics for the gfx950 DG kernels.
//
// Each function restates one function of the reference's hybrid/solver.py and
// hybrid/boundary.py (cited per function) for one quadrature point held in
// registers.  ``real`` is double on the plain path and float on the FP32 path.
//
// The conversion and Euler flux functions disable floating-point contraction:
// the residual subtracts F(a) from F(U) with U == a bitwise in uniform
// regions (the anchor of hybrid/spatial.py:41-70), which cancels exactly
// only if every inlined copy rounds identically.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsg {

enum { HEX = 0, WEDGE = 1, TET = 2 };
enum { QUAD = 0, TRI = 1 };
enum { TERM_EULER = 1, TERM_VISCOUS = 2 };
enum { BND_WALL = 1, BND_INLET = 2, BND_NOSLIP = 4, BND_SLIP = 8 };

__host__ __device__ constexpr int n_nodes(int t, int p) {
  return t == HEX ? (p + 1) * (p + 1) * (p + 1)
       : t == WEDGE ? (p + 1) * (p + 1) * (p + 2) / 2
                    : (p + 1) * (p + 2) * (p + 3) / 6;
}
// hybrid/elements.py:60-128
__host__ __device__ constexpr int n_vol_q(int t, int p) {
  return p == 1 ? (t == HEX ? 8 : t == WEDGE ? 6 : 4) : (p == 2 && t == HEX) ? 27 : 125;
}
// hybrid/geometry.py:48-89
__host__ __device__ constexpr int n_face_q(int ft, int p) {
  return ft == QUAD ? (p == 1 ? 4 : p == 2 ? 9 : 16) : (p == 1 ? 3 : p == 2 ? 16 : 25);
}
// Nodes lying on a face (its basis support): a quad face (or a triangle
// sub-face cut from one) carries (p+1)^2 nodes, a triangle (p+1)(p+2)/2.
__host__ __device__ constexpr int n_face_nodes(int ft, int sub, int p) {
  return (ft == QUAD || sub) ? (p + 1) * (p + 1) : (p + 1) * (p + 2) / 2;
}

template <typename real>
struct Phys {
  real gamma, gas_constant, mu, kappa, eta;
  int has_boundary;      // a boundary_state object was supplied
  int rusanov;           // inviscid_flux == "rusanov" (hybrid/solver.py:302-305)
  int has_wall_temperature;
  real wall_temperature;
  real farfield[16];     // primitive farfield [rho, u, v, w, p, Y...]
  int wale;              // WALE eddy viscosity on (hybrid/solver.py:317-406)
  real wale_constant;    // C_w
  real cp_over_prt;      // c_p / Pr_t: eddy conductivity per eddy viscosity
};

template <typename real> __device__ __forceinline__ real rmax(real a, real b) { return a > b ? a : b; }
template <typename real> __device__ __forceinline__ real rmin(real a, real b) { return a < b ? a : b; }
template <typename real> __device__ __forceinline__ bool finite(real x) { return x - x == real(0); }

// hybrid/solver.py:83-98 (density floored at 1e-14 for division only).
template <int V, typename real>
struct Prim {
  real rho, vel[3], p, safe;
};

template <int V, typename real>
__device__ __forceinline__ Prim<V, real> c2p(const real* U, real gamma) {
#pragma clang fp contract(off)
  Prim<V, real> w;
  w.rho = U[0];
  w.safe = rmax(U[0], real(1.0e-14));
  w.vel[0] = U[1] / w.safe;
  w.vel[1] = U[2] / w.safe;
  w.vel[2] = U[3] / w.safe;
  real ke = real(0.5) * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / w.safe;
  w.p = (gamma - real(1)) * (U[4] - ke);
  return w;
}

// hybrid/solver.py:63-81 from primitive [rho, vel, p, Y...].
template <int V, typename real>
__device__ __forceinline__ void p2c(const real* W, real gamma, real* U) {
#pragma clang fp contract(off)
  real rho = W[0];
  U[0] = rho;
  U[1] = rho * W[1];
  U[2] = rho * W[2];
  U[3] = rho * W[3];
  U[4] = W[4] / (gamma - real(1)) + real(0.5) * rho * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
#pragma unroll
  for (int s = 5; s < V; ++s) U[s] = rho * W[s];
}

template <int V, typename real>
__device__ __forceinline__ void c2p_full(const real* U, real gamma, real* W) {
#pragma clang fp contract(off)
  Prim<V, real> w = c2p<V, real>(U, gamma);
  W[0] = w.rho;
  W[1] = w.vel[0];
  W[2] = w.vel[1];
  W[3] = w.vel[2];
  W[4] = w.p;
#pragma unroll
  for (int s = 5; s < V; ++s) W[s] = U[s] / w.safe;
}

// hybrid/solver.py:100-122: F(U) . n
template <int V, typename real>
__device__ __forceinline__ void normal_flux(const real* U, const Prim<V, real>& w, const real* n,
                                            real* F) {
#pragma clang fp contract(off)
  real un = w.vel[0] * n[0] + w.vel[1] * n[1] + w.vel[2] * n[2];
  F[0] = w.rho * un;
  F[1] = w.rho * w.vel[0] * un + w.p * n[0];
  F[2] = w.rho * w.vel[1] * un + w.p * n[1];
  F[3] = w.rho * w.vel[2] * un + w.p * n[2];
  F[4] = un * (U[4] + w.p);
#pragma unroll
  for (int s = 5; s < V; ++s) F[s] = U[s] * un;
}

// hybrid/solver.py:124-130: Cartesian flux along axis d.
template <int V, typename real>
__device__ __forceinline__ void axis_flux(const real* U, const Prim<V, real>& w, int d, real* F) {
#pragma clang fp contract(off)
  real un = w.vel[d];
  F[0] = w.rho * un;
  F[1] = w.rho * w.vel[0] * un + (d == 0 ? w.p : real(0));
  F[2] = w.rho * w.vel[1] * un + (d == 1 ? w.p : real(0));
  F[3] = w.rho * w.vel[2] * un + (d == 2 ? w.p : real(0));
  F[4] = un * (U[4] + w.p);
#pragma unroll
  for (int s = 5; s < V; ++s) F[s] = U[s] * un;
}

// hybrid/solver.py:154-161
template <typename real>
__device__ __forceinline__ real safe_signed(real x) {
  const real tol = real(1.0e-14);
  return fabs(x) > tol ? x : (x >= real(0) ? tol : -tol);
}

// hybrid/solver.py:132-152
template <int V, typename real>
__device__ __forceinline__ void rusanov(const real* UL, const real* UR, const real* n, real gamma,
                                        real* F) {
#pragma clang fp contract(off)
  Prim<V, real> wl = c2p<V, real>(UL, gamma), wr = c2p<V, real>(UR, gamma);
  real unl = wl.vel[0] * n[0] + wl.vel[1] * n[1] + wl.vel[2] * n[2];
  real unr = wr.vel[0] * n[0] + wr.vel[1] * n[1] + wr.vel[2] * n[2];
  real rl = rmax(wl.rho, real(1.0e-14)), rr = rmax(wr.rho, real(1.0e-14));
  real pl = rmax(wl.p, real(0)), pr = rmax(wr.p, real(0));
  real cl = sqrt(gamma * pl / rl), cr = sqrt(gamma * pr / rr);
  real smax = rmax(fabs(unl) + cl, fabs(unr) + cr);
  real FL[V], FR[V];
  normal_flux<V, real>(UL, wl, n, FL);
  normal_flux<V, real>(UR, wr, n, FR);
#pragma unroll
  for (int v = 0; v < V; ++v) F[v] = real(0.5) * (FL[v] + FR[v]) - real(0.5) * smax * (UR[v] - UL[v]);
}

// hybrid/solver.py:209-243 then F* = F + S (U* - U).
template <int V, typename real>
__device__ __forceinline__ void hllc_side(const real* U, const Prim<V, real>& w, const real* n,
                                          real S, real SM, const real* F, real* Fs) {
#pragma clang fp contract(off)
  real un = w.vel[0] * n[0] + w.vel[1] * n[1] + w.vel[2] * n[2];
  real den = safe_signed(S - SM);
  real comp = (S - un) / den;
  real rs = w.rho * comp;
  real ps = w.p + w.rho * (S - un) * (SM - un);
  real Es = ((S - un) * U[4] - w.p * un + ps * SM) / den;
  real Us[V];
  Us[0] = rs;
#pragma unroll
  for (int d = 0; d < 3; ++d) Us[1 + d] = rs * (w.vel[d] + (SM - un) * n[d]);
  Us[4] = Es;
#pragma unroll
  for (int s = 5; s < V; ++s) Us[s] = U[s] * comp;
#pragma unroll
  for (int v = 0; v < V; ++v) Fs[v] = F[v] + S * (Us[v] - U[v]);
}

// hybrid/solver.py:245-305: HLLC with the pointwise Rusanov fallback and the
// bitwise identical-state shortcut.
template <int V, typename real>
__device__ __forceinline__ void hllc(const real* UL, const real* UR, const real* n, real gamma,
                                     real* F) {
#pragma clang fp contract(off)
  Prim<V, real> wl = c2p<V, real>(UL, gamma), wr = c2p<V, real>(UR, gamma);
  real FL[V];
  normal_flux<V, real>(UL, wl, n, FL);
  bool same = true;
#pragma unroll
  for (int v = 0; v < V; ++v) same = same && (UL[v] == UR[v]);
  if (same) {
#pragma unroll
    for (int v = 0; v < V; ++v) F[v] = FL[v];
    return;
  }
  real FR[V];
  normal_flux<V, real>(UR, wr, n, FR);
  real unl = wl.vel[0] * n[0] + wl.vel[1] * n[1] + wl.vel[2] * n[2];
  real unr = wr.vel[0] * n[0] + wr.vel[1] * n[1] + wr.vel[2] * n[2];
  real cl = sqrt(rmax(gamma * wl.p / rmax(wl.rho, real(1.0e-14)), real(0)));
  real cr = sqrt(rmax(gamma * wr.p / rmax(wr.rho, real(1.0e-14)), real(0)));
  real SL = rmin(unl - cl, unr - cr);
  real SR = rmax(unl + cl, unr + cr);
  real den = wl.rho * (SL - unl) - wr.rho * (SR - unr);
  real SM = (wr.p - wl.p + wl.rho * unl * (SL - unl) - wr.rho * unr * (SR - unr)) / safe_signed(den);
  // One star-state evaluation on the side that SM selects (lanes of a wave
  // that would take different sides then share one evaluation), then
  // H = F_L (SL >= 0), star state (SM >= 0 or SR > 0), or F_R.
  const bool lft = SM >= real(0);
  real Ux[V], Fx[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    Ux[v] = lft ? UL[v] : UR[v];
    Fx[v] = lft ? FL[v] : FR[v];
  }
  Prim<V, real> wx;
  wx.rho = lft ? wl.rho : wr.rho;
  wx.safe = lft ? wl.safe : wr.safe;
#pragma unroll
  for (int d = 0; d < 3; ++d) wx.vel[d] = lft ? wl.vel[d] : wr.vel[d];
  wx.p = lft ? wl.p : wr.p;
  real Hs[V];
  hllc_side<V, real>(Ux, wx, n, lft ? SL : SR, SM, Fx, Hs);
  const bool star = lft || SR > real(0);
  real H[V];
#pragma unroll
  for (int v = 0; v < V; ++v) H[v] = SL >= real(0) ? FL[v] : (star ? Hs[v] : FR[v]);
  bool ok = (wl.rho > real(0)) && (wr.rho > real(0)) && (wl.p > real(0)) && (wr.p > real(0));
#pragma unroll
  for (int v = 0; v < V; ++v) ok = ok && finite(H[v]);
  if (ok) {
#pragma unroll
    for (int v = 0; v < V; ++v) F[v] = H[v];
  } else {
    rusanov<V, real>(UL, UR, n, gamma, F);
  }
}

// The numerical flux of hllc / rusanov above with the species columns kept
// out of the state registers (states with many passive species).  The flow
// part (columns 0-4) goes to F; every species column follows from the
// scalars in FluxScalars and that column's two states alone
// (species_flux), by the same expressions as the all-column functions, so
// results are bit-identical.  sl(s) / sr(s) return the left / right state of
// species column s (s >= 5).
template <typename real>
struct FluxScalars {
  int mode;     // 0: F_L, 1: left star, 2: right star, 3: F_R, 4: Rusanov
  real unl, unr, S, comp, smax;
};

// The species column of each mode (F_L, left star, right star, F_R,
// Rusanov) as selects between the mode expressions, without branches: the
// lanes of a wave take different modes.
template <typename real>
__device__ __forceinline__ real species_flux(const FluxScalars<real>& f, real uL, real uR) {
#pragma clang fp contract(off)
  const real FL = uL * f.unl, FR = uR * f.unr;
  const bool right = f.mode == 2;
  const real u = right ? uR : uL, F = right ? FR : FL;
  const real star = F + f.S * (u * f.comp - u);
  const real rus = real(0.5) * (FL + FR) - real(0.5) * f.smax * (uR - uL);
  return f.mode == 0 ? FL : (f.mode == 3 ? FR : (f.mode == 4 ? rus : star));
}

template <int V, typename real>
__device__ __forceinline__ void rusanov_scalars(const real* UL, const real* UR, const real* n,
                                                real gamma, FluxScalars<real>& f) {
#pragma clang fp contract(off)
  Prim<5, real> wl = c2p<5, real>(UL, gamma), wr = c2p<5, real>(UR, gamma);
  real rl = rmax(wl.rho, real(1.0e-14)), rr = rmax(wr.rho, real(1.0e-14));
  real pl = rmax(wl.p, real(0)), pr = rmax(wr.p, real(0));
  real cl = sqrt(gamma * pl / rl), cr = sqrt(gamma * pr / rr);
  f.mode = 4;
  f.smax = rmax(fabs(f.unl) + cl, fabs(f.unr) + cr);
}

template <int V, typename real, class SL, class SR>
__device__ __forceinline__ void numflux_split(const real* UL, const real* UR, const real* n,
                                              real gamma, bool use_rusanov, SL sl, SR sr,
                                              real* F, FluxScalars<real>& f) {
#pragma clang fp contract(off)
  Prim<5, real> wl = c2p<5, real>(UL, gamma), wr = c2p<5, real>(UR, gamma);
  f.unl = wl.vel[0] * n[0] + wl.vel[1] * n[1] + wl.vel[2] * n[2];
  f.unr = wr.vel[0] * n[0] + wr.vel[1] * n[1] + wr.vel[2] * n[2];
  f.S = real(0);
  f.comp = real(0);
  f.smax = real(0);
  if (use_rusanov) {
    rusanov<5, real>(UL, UR, n, gamma, F);
    rusanov_scalars<V, real>(UL, UR, n, gamma, f);
    return;
  }
  real FL[5];
  normal_flux<5, real>(UL, wl, n, FL);
  bool same = true;
#pragma unroll
  for (int v = 0; v < 5; ++v) same = same && (UL[v] == UR[v]);
  // non-short-circuit accumulations (no branch per column)
  for (int s = 5; s < V; ++s) same = same & (sl(s) == sr(s));
  f.mode = 0;
  if (same) {
#pragma unroll
    for (int v = 0; v < 5; ++v) F[v] = FL[v];
    return;
  }
  real FR[5];
  normal_flux<5, real>(UR, wr, n, FR);
  real cl = sqrt(rmax(gamma * wl.p / rmax(wl.rho, real(1.0e-14)), real(0)));
  real cr = sqrt(rmax(gamma * wr.p / rmax(wr.rho, real(1.0e-14)), real(0)));
  real SLs = rmin(f.unl - cl, f.unr - cr);
  real SRs = rmax(f.unl + cl, f.unr + cr);
  real den = wl.rho * (SLs - f.unl) - wr.rho * (SRs - f.unr);
  real SM = (wr.p - wl.p + wl.rho * f.unl * (SLs - f.unl) - wr.rho * f.unr * (SRs - f.unr)) /
            safe_signed(den);
  real H[5];
  if (SLs >= real(0)) {
#pragma unroll
    for (int v = 0; v < 5; ++v) H[v] = FL[v];
  } else if (SM >= real(0)) {
    hllc_side<5, real>(UL, wl, n, SLs, SM, FL, H);
    f.mode = 1;
    f.S = SLs;
    f.comp = (SLs - f.unl) / safe_signed(SLs - SM);
  } else if (SRs > real(0)) {
    hllc_side<5, real>(UR, wr, n, SRs, SM, FR, H);
    f.mode = 2;
    f.S = SRs;
    f.comp = (SRs - f.unr) / safe_signed(SRs - SM);
  } else {
#pragma unroll
    for (int v = 0; v < 5; ++v) H[v] = FR[v];
    f.mode = 3;
  }
  bool ok = (wl.rho > real(0)) && (wr.rho > real(0)) && (wl.p > real(0)) && (wr.p > real(0));
#pragma unroll
  for (int v = 0; v < 5; ++v) ok = ok && finite(H[v]);
  for (int s = 5; s < V; ++s) ok = ok & finite(species_flux(f, sl(s), sr(s)));
  if (ok) {
#pragma unroll
    for (int v = 0; v < 5; ++v) F[v] = H[v];
  } else {
    rusanov<5, real>(UL, UR, n, gamma, F);
    rusanov_scalars<V, real>(UL, UR, n, gamma, f);
  }
}

// hybrid/solver.py:307-315: [u, v, w, T = p / (rho R)].
template <int V, typename real>
__device__ __forceinline__ void viscous_variables(const real* U, real gamma, real R, real* v) {
  Prim<V, real> w = c2p<V, real>(U, gamma);
  v[0] = w.vel[0];
  v[1] = w.vel[1];
  v[2] = w.vel[2];
  v[3] = w.p / (rmax(w.rho, real(1.0e-14)) * R);
}

// hybrid/solver.py:317-355: WALE eddy dynamic viscosity rho nu_t from the
// resolved velocity gradient g[i*3+d] = d u_i / d x_d, the filter width and
// the density.  With s = |S|^2 and t = |S^d|^2 (Frobenius norms of the
// strain rate and of the traceless symmetric part of g^2), s clipped to
// [0, 1e8] and t to [0, 1e10]: nu_t = (C_w D)^2 t^(3/2) / (s^(5/2) +
// t^(5/4) + 1e-30) times the Ducros sensor |w|^2 / (|w|^2 + (div u)^2 +
// 1e-12), clipped to [0, 50] (NaN to 0); rho floored at 0 (NaN kept).
template <typename real>
__device__ __forceinline__ real wale_viscosity(const real* g, real rho, real width, real cw) {
  real g2[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      g2[i * 3 + j] = g[i * 3 + 0] * g[0 * 3 + j] + g[i * 3 + 1] * g[1 * 3 + j] +
                      g[i * 3 + 2] * g[2 * 3 + j];
  const real third = (g2[0] + g2[4] + g2[8]) / real(3);
  real sn = real(0), sd = real(0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const real sij = real(0.5) * (g[i * 3 + j] + g[j * 3 + i]);
      const real dij = real(0.5) * (g2[i * 3 + j] + g2[j * 3 + i]) - (i == j ? third : real(0));
      sn += sij * sij;
      sd += dij * dij;
    }
  const real div = g[0] + g[4] + g[8];
  const real cx = g[7] - g[5], cy = g[2] - g[6], cz = g[3] - g[1];
  const real vort = cx * cx + cy * cy + cz * cz;
  const real ducros = vort / (vort + div * div + real(1.0e-12));
  // clips that keep NaN (both comparisons false)
  const real ss = sn < real(0) ? real(0) : (sn > real(1.0e8) ? real(1.0e8) : sn);
  const real st = sd < real(0) ? real(0) : (sd > real(1.0e10) ? real(1.0e10) : sd);
  // t^(3/2) = t sqrt(t), t^(5/4) = t sqrt(sqrt(t)), s^(5/2) = s^2 sqrt(s)
  const real rs = sqrt(ss), rt = sqrt(st);
  const real cwd = cw * width;
  real nu = (cwd * cwd * (st * rt) / (ss * ss * rs + st * sqrt(rt) + real(1.0e-30))) * ducros;
  nu = nu > real(50) ? real(50) : (nu >= real(0) ? nu : real(0));
  const real r = rho < real(0) ? real(0) : rho;
  return r * nu;
}

// hybrid/solver.py:357-406: grad[m*3+d] = d v_m / d x_d; mu and kappa
// include the eddy viscosity and conductivity when WALE is on.
// Writes Fv[d*V + v]; only the momentum and energy columns are non-zero.
template <int V, typename real>
__device__ __forceinline__ void viscous_flux(const real* U, const real* grad, real gamma, real mu,
                                             real kappa, real* Fv) {
  Prim<V, real> w = c2p<V, real>(U, gamma);
  real div = grad[0] + grad[4] + grad[8];
  real tau[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int d = 0; d < 3; ++d)
      tau[i][d] = mu * (grad[i * 3 + d] + grad[d * 3 + i] - (i == d ? real(2.0 / 3.0) * div : real(0)));
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    Fv[d * V + 0] = real(0);
#pragma unroll
    for (int i = 0; i < 3; ++i) Fv[d * V + 1 + i] = tau[i][d];
    Fv[d * V + 4] = w.vel[0] * tau[0][d] + w.vel[1] * tau[1][d] + w.vel[2] * tau[2][d] +
                    kappa * grad[9 + d];
#pragma unroll
    for (int s = 5; s < V; ++s) Fv[d * V + s] = real(0);
  }
}

// hybrid/boundary.py:54-72: Euler ghost state for a boundary face point.
template <int V, typename real>
__device__ __forceinline__ void ghost_state(const real* Uin, const real* n, int flags,
                                            const Phys<real>& ph, real* Ug) {
  if (!ph.has_boundary) {
#pragma unroll
    for (int v = 0; v < V; ++v) Ug[v] = Uin[v];
    return;
  }
  real W[V];
  c2p_full<V, real>(Uin, ph.gamma, W);
  if (flags & BND_WALL) {
    real un = W[1] * n[0] + W[2] * n[1] + W[3] * n[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) W[1 + d] = W[1 + d] - real(2) * un * n[d];
  }
  if (flags & BND_INLET) {
#pragma unroll
    for (int v = 0; v < V; ++v) W[v] = ph.farfield[v];
  }
  p2c<V, real>(W, ph.gamma, Ug);
}

// hybrid/boundary.py:74-102 (and hybrid/navier_stokes.py:80-89 when no
// boundary object is given): ghost viscous variables.
template <int V, typename real>
__device__ __forceinline__ void ghost_viscous(const real* Uin, const real* n, int flags,
                                              const Phys<real>& ph, real* vg) {
  viscous_variables<V, real>(Uin, ph.gamma, ph.gas_constant, vg);
  if (!ph.has_boundary) return;
  // The slip ghost replaces the no-slip one when a tag is in both sets: it
  // reflects the inside velocity and keeps the inside temperature.
  if (flags & BND_SLIP) {
    real un = vg[0] * n[0] + vg[1] * n[1] + vg[2] * n[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) vg[d] = vg[d] - real(2) * un * n[d];
  } else if (flags & BND_NOSLIP) {
    vg[0] = -vg[0];
    vg[1] = -vg[1];
    vg[2] = -vg[2];
    if (ph.has_wall_temperature) vg[3] = real(2) * ph.wall_temperature - vg[3];
  }
  if (flags & BND_INLET) {
    real Uf[V];
    p2c<V, real>(ph.farfield, ph.gamma, Uf);
    viscous_variables<V, real>(Uf, ph.gamma, ph.gas_constant, vg);
  }
}

}  // namespace fsg

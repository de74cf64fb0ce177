// Point-implicit 11-species air chemistry on gfx950 (FP64, one cell per lane).
//
// Restates hybrid/chemistry/integrator.py:31-104 (backward Euler on the
// species with a fixed Newton iteration count, non-negativity clip,
// temperature recovery at constant internal energy, total-mass rescale),
// hybrid/chemistry/mechanism.py:159-287 (mass-action rates with detailed
// balance, mixture energy, damped T(e) Newton) and
// hybrid/chemistry/coupling.py:116-177 (per-node coupling to the flow state).
// The reference obtains the Newton Jacobian with jax.jacobian; here it is the
// analytic derivative of the same expression, including JAX's half-gradient
// at max(x, y) ties (rho_s = 0 after the clip).  The 11x11 system is solved
// by LU with partial pivoting held in registers (fully unrolled; the pivot
// swap is a select over rows, so no register array is indexed dynamically).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "../../include/fluid_gfx950.h"
#include "fsg_park11.h"

namespace fsgchem {

constexpr int NS = 11;
constexpr int NR = 22;
constexpr double R_UNIV = 8.31446261815324;
constexpr double P_STD = 101325.0;

// Mechanism table layout (doubles), written by the host:
//   [0, 11)              molar masses
//   [11, 11 + 17*11)     NASA-7 per species: t_low, t_mid, t_high, low[7], high[7]
//   then per reaction 6: A, b, theta, third_body, fwd_mode, rev_mode
//   then nu' (22x11), nu'' (22x11), alpha (22x11)
//   then per reaction 3: ln A, sum_s (nu''_s - nu'_s), sum_s M_s (nu''_s - nu'_s)
//   then per species 18: 1/M, R/M, 1/t_low, 1/t_high, and per fit range
//   (low, high) a1/2, a2/3, a3/4, a4/5 (h/RT) and a2/2, a3/3, a4/4 (s/R)
constexpr int OFF_NASA = NS;
constexpr int OFF_RX = OFF_NASA + 17 * NS;
constexpr int OFF_NU1 = OFF_RX + 6 * NR;
constexpr int OFF_NU2 = OFF_NU1 + NR * NS;
constexpr int OFF_ALPHA = OFF_NU2 + NR * NS;
constexpr int OFF_DERIVED = OFF_ALPHA + NR * NS;
constexpr int OFF_SPEC = OFF_DERIVED + 3 * NR;
constexpr int SPEC_W = 18;
// last entry: the fingerprint of the table's stoichiometry and third-body
// data (hybrid/chemistry/air11_park.py stoichiometry_fingerprint)
constexpr int OFF_FP = OFF_SPEC + SPEC_W * NS;
constexpr int TABLE_SIZE = OFF_FP + 1;
static_assert(park11::NR == NR && park11::NS == NS, "compiled-in mechanism shape");

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double clampd(double x, double lo, double hi) { return dmin(dmax(x, lo), hi); }
// d max(x, y) / dx with the half gradient at a tie.
__device__ __forceinline__ double dmax_dx(double x, double y) {
  return x > y ? 1.0 : (x == y ? 0.5 : 0.0);
}

// NASA-7 (hybrid/chemistry/nasa.py:47-119) with T clamped to [t_low, t_high]
// inside the polynomial, in Horner form with the divided coefficients of the
// species' derived block (a1/2, a2/3, ... written by the host, so no
// division by 3 or 5 per evaluation) and 1/t given: the reference's power
// sums to rounding.  The coefficients are read with scalar (wave-uniform)
// loads: each fit range is evaluated when some lane of the wave lies in it
// (a uniform branch; the lanes of a wave almost always share a range) and
// every lane selects its own.
struct Thermo {
  double cp, h;  // cp/R, h/(RT)
};

__device__ __forceinline__ double clamped_inverse(const double* c, const double* sp, double t,
                                                  double T, double iT) {
  return t == T ? iT : (t == c[0] ? sp[2] : sp[3]);
}

__device__ __forceinline__ void nasa_range(const double* a, const double* d, double t, double it,
                                           double& cp, double& h) {
  cp = fma(t, fma(t, fma(t, fma(t, a[4], a[3]), a[2]), a[1]), a[0]);
  h = fma(a[5], it, fma(t, fma(t, fma(t, fma(t, d[3], d[2]), d[1]), d[0]), a[0]));
}

__device__ __forceinline__ Thermo nasa(const double* c, const double* sp, double T, double iT) {
  const double t = clampd(T, c[0], c[2]);
  const double it = clamped_inverse(c, sp, t, T, iT);
  const bool lo = t < c[1];
  double cpl = 0.0, hl = 0.0, cph = 0.0, hh = 0.0;
  if (__any(lo)) nasa_range(c + 3, sp + 4, t, it, cpl, hl);
  if (__any(!lo)) nasa_range(c + 10, sp + 11, t, it, cph, hh);
  return Thermo{lo ? cpl : cph, lo ? hl : hh};
}

// g = h/RT - s/R of one species at tT > 0 with ln(tT) and 1/tT given (the
// logarithm of the clamped temperature otherwise).
__device__ __forceinline__ double gibbs_range(const double* a, const double* d, double t, double lt,
                                              double it) {
  const double h = fma(a[5], it, fma(t, fma(t, fma(t, fma(t, d[3], d[2]), d[1]), d[0]), a[0]));
  const double sr = fma(a[0], lt, fma(t, fma(t, fma(t, fma(t, d[6], d[5]), d[4]), a[1]), a[6]));
  return h - sr;
}

__device__ __forceinline__ double gibbs(const double* c, const double* sp, double tT, double lT,
                                        double iT) {
  const double t = clampd(tT, c[0], c[2]);
  const double lt = t == tT ? lT : log(t);
  const double it = clamped_inverse(c, sp, t, tT, iT);
  const bool lo = t < c[1];
  double gl = 0.0, gh = 0.0;
  if (__any(lo)) gl = gibbs_range(c + 3, sp + 4, t, lt, it);
  if (__any(!lo)) gh = gibbs_range(c + 10, sp + 11, t, lt, it);
  return lo ? gl : gh;
}

// The table pointer offset by a zero the compiler cannot see through:
// addresses formed from it inside a loop stay inside the loop (scalar loads
// with immediate offsets from one base) instead of being hoisted as a
// hundred 64-bit sums that spill.  The offset is wave-uniform ("s").
__device__ __forceinline__ const double* table_launder(const double* p) {
  int zero = 0;
  __asm__ volatile("" : "+s"(zero));
  return p + zero;
}

__device__ __forceinline__ const double* species_block(const double* tab, int s) {
  return tab + OFF_SPEC + SPEC_W * s;
}

// e_mix(T) with h = h/RT * R T / M_s (T unclamped outside), as
// hybrid/chemistry/mechanism.py:222-260.
__device__ __forceinline__ double mixture_energy(const double* tab, const double* Y, double T) {
  const double iT = 1.0 / T;
  double e = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double* sp = species_block(tab, s);
    Thermo th = nasa(tab + OFF_NASA + 17 * s, sp, T, iT);
    e = fma(Y[s], (th.h - 1.0) * (sp[1] * T), e);
  }
  return e;
}

// The damped Newton T(e) of hybrid/chemistry/mechanism.py:262-287, stopped
// once a step moves T by at most 1e-13 T (further steps of the fixed count
// would change T below that) or after max_iters.  Used for the coupler's
// initial temperature (hybrid/chemistry/coupling.py:135-154, 8 steps there),
// the integrator's temperature after each Newton iteration
// (hybrid/chemistry/integrator.py:88-95, 10 steps there) and the error
// check; the fused react_kernel and the split prepare_kernel therefore agree
// bit for bit.
__device__ __forceinline__ double recover_temperature_converged(const double* tab,
                                                                const double* rho,
                                                                double e_target, double T,
                                                                int max_iters) {
  double tot = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) tot += rho[s];
  tot = dmax(tot, 1.0e-14);
  const double itot = 1.0 / tot;
  double Y[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) Y[s] = dmax(rho[s], 0.0) * itot;
#pragma unroll 1
  for (int it = 0; it < max_iters; ++it) {
    const double* tb = table_launder(tab);
    const double Ts = clampd(T, 50.0, 50000.0);
    const double iTs = 1.0 / Ts;
    double e = 0.0, cv = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double* sp = species_block(tb, s);
      Thermo th = nasa(tb + OFF_NASA + 17 * s, sp, Ts, iTs);
      e = fma(Y[s], (th.h - 1.0) * (sp[1] * Ts), e);
      cv = fma(Y[s], (th.cp - 1.0) * sp[1], cv);
    }
    double dT = -(e - e_target) / dmax(cv, 1.0e-5);
    const double Tn = clampd(T + clampd(dT, -0.5 * T, 0.5 * T), 100.0, 50000.0);
    const bool done = fabs(Tn - T) <= 1.0e-13 * T;
    T = Tn;
    if (done) break;
  }
  return T;
}

// Net mass production rates w and their Jacobian d w / d rho at T = Tv,
// the integrator's case (hybrid/chemistry/mechanism.py:159-220, the
// derivative jax.jacobian takes, with its half gradient at max(x, y) ties).
// With T = Tv every temperature mode is tT = max(T, 10).  The arithmetic is
// rearranged (values equal to rounding): k_f fwd and k_r rev are each one
// exponential of ln A + b ln tT - theta / tT + lf (and of that minus ln K_c
// plus lr), with the clamps of the reference on the same arguments
// (ln K_c >= ln 1e-300 for k_r = k_f / max(K_c, 1e-300)); the Gibbs
// energies share log tT; species that take no part in a reaction are
// skipped (the stoichiometry is uniform across the wave: scalar branches);
// the Jacobian's column sums come from the reactions' mass balances sum_s
// M_s (nu''_s - nu'_s); the mass-sum correction w - Y sum(w) is
// differentiated with one reciprocal of the total.  PARK: the stoichiometry
// and third-body data are the compiled-in Park-11 constants (fsg_park11.h)
// and the reaction loop is unrolled, each reaction's absent species folded
// away at compile time (same arithmetic); otherwise they are read from the
// table, one reaction per loop iteration.
template <bool PARK>
__device__ __forceinline__ void rates(const double* __restrict__ tab, const double* rho, double T,
                                      double* w, double (*dw)[NS]) {
  const double* mw = tab;
  // C_s is formed again where a third body needs it (not held)
  double dC[NS], lc[NS], dlc[NS], g[NS];
  const double tT = dmax(T, 10.0);
  const double lT = log(tT), iT = 1.0 / tT;
  const double lP = log(P_STD / (R_UNIV * tT));
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double* sp = species_block(tab, s);
    const double x = rho[s] * sp[0];
    const double C = dmax(x, 0.0);
    dC[s] = dmax_dx(x, 0.0) * sp[0];
    const double cs = dmax(C, 1.0e-100);
    lc[s] = log(cs);
    dlc[s] = dmax_dx(C, 1.0e-100) / cs;
    g[s] = gibbs(tab + OFF_NASA + 17 * s, sp, tT, lT, iT);
  }
  double om[NS], csum[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    om[s] = 0.0;
    csum[s] = 0.0;
#pragma unroll
    for (int j = 0; j < NS; ++j) dw[s][j] = 0.0;
  }
  auto reaction = [&](int r) {
    const double* tb = table_launder(tab);
    const double* rx = tb + OFF_RX + 6 * r;
    const double b = rx[1], th = rx[2];
    const bool third = PARK ? park11::THIRD[r] : rx[3] != 0.0;
    const double* n1 = tb + OFF_NU1 + r * NS;
    const double* n2 = tb + OFF_NU2 + r * NS;
    const double* alt = tb + OFF_ALPHA + r * NS;
    auto nu1 = [&](int k) { return PARK ? park11::NU1[r][k] : n1[k]; };
    auto nu2 = [&](int k) { return PARK ? park11::NU2[r][k] : n2[k]; };
    auto al = [&](int k) { return PARK ? park11::ALPHA[r][k] : alt[k]; };
    const double* dr = tb + OFF_DERIVED + 3 * r;
    const double lnA = dr[0], dn = dr[1], mr = dr[2];
    double dG = 0.0, lf = 0.0, lr = 0.0, Meff = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double a1 = nu1(s), a2 = nu2(s);
      if (a1 != 0.0 || a2 != 0.0) {
        dG = fma(a2 - a1, g[s], dG);
        lf = fma(a1, lc[s], lf);
        lr = fma(a2, lc[s], lr);
      }
      if (third && al(s) != 0.0)
        Meff = fma(al(s), dmax(rho[s] * tb[OFF_SPEC + SPEC_W * s], 0.0), Meff);
    }
    const double lkf = fma(-th, iT, fma(b, lT, lnA));
    const double lKc = dmax(fma(dn, lP, -clampd(dG, -700.0, 700.0)), -690.7755278982137);
    const double fwdk = exp(lkf + clampd(lf, -700.0, 700.0));
    const double revk = exp(lkf - lKc + clampd(lr, -700.0, 700.0));
    const double cf = fabs(lf) < 700.0 ? 1.0 : (fabs(lf) == 700.0 ? 0.5 : 0.0);
    const double cr = fabs(lr) < 700.0 ? 1.0 : (fabs(lr) == 700.0 ? 0.5 : 0.0);
    const double col = third ? Meff : 1.0;
    const double net = fwdk - revk;
    const double q = net * col;
    const double fc = col * fwdk * cf, rc = col * revk * cr;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double d = nu2(s) - nu1(s);
      if (d != 0.0) om[s] = fma(d, q, om[s]);
    }
    // column j of this reaction's d q / d rho, added to every row it feeds
    // (no per-reaction vector held)
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const double a1 = nu1(j), a2 = nu2(j);
      const double dcol = third ? al(j) : 0.0;
      if (a1 != 0.0 || a2 != 0.0 || dcol != 0.0) {
        const double dq = fma(fma(fc, a1, -(rc * a2)), dlc[j], dcol * net) * dC[j];
        csum[j] = fma(mr, dq, csum[j]);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          const double d = nu2(s) - nu1(s);
          if (d != 0.0) dw[s][j] = fma(tb[s] * d, dq, dw[s][j]);
        }
      }
    }
  };
  if constexpr (PARK) {
#pragma unroll
    for (int r = 0; r < NR; ++r) reaction(r);
  } else {
#pragma unroll 1
    for (int r = 0; r < NR; ++r) reaction(r);
  }
  // w = M omega, then the exact mass-sum correction w - Y sum(w)
  double srho = 0.0, S = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    srho += rho[s];
    S = fma(om[s], mw[s], S);
  }
  const double tot = dmax(srho, 1.0e-14), dtot = dmax_dx(srho, 1.0e-14);
  const double itot = 1.0 / tot;
  const double k2 = dtot * (itot * itot);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double rp = dmax(rho[s], 0.0);
    const double Y = rp * itot;
    w[s] = fma(om[s], mw[s], -(Y * S));
    // d Y_s / d rho_j = delta_sj [rho_s > 0] / tot - rho_s dtot / tot^2
    const double off = rp * k2;
    const double diag = dmax_dx(rho[s], 0.0) * itot;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const double dY = (s == j ? diag : 0.0) - off;
      dw[s][j] = fma(-Y, csum[j], fma(-dY, S, dw[s][j]));
    }
  }
}

// w = net mass production rates and the diagonal d w_s / d rho_s of their
// Jacobian at T = Tv (hybrid/chemistry/mechanism.py:159-220), for the
// integrator's error check.  The arithmetic of rates() rearranged for
// speed (values equal to rounding): with T = Tv every temperature mode is
// tT = max(T, 10); k_f fwd and k_r rev are each one exponential of
// ln A + b ln tT - theta / tT + lf (and of that minus ln K_c plus lr), the
// clamps of rates() applied to the same arguments.  The stoichiometry and
// third-body efficiencies are the compiled-in Park-11 constants
// (fsg_park11.h; the caller checks the table's fingerprint against them),
// so the reaction loop is unrolled with each reaction's absent species
// folded away at compile time; the rate constants, thermodynamics and
// molar masses come from the table.
__device__ __forceinline__ void rates_diag(const double* __restrict__ tab, const double* rho,
                                           double T, double* w, double* dd) {
  const double* mw = tab;
  // C_s and d C_s / d rho_s are formed again where used (not held)
  double lc[NS], dlc[NS], g[NS];
  const double tT = dmax(T, 10.0);
  const double lT = log(tT), iT = 1.0 / tT;
  const double lP = log(P_STD / (R_UNIV * tT));
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double* sp = species_block(tab, s);
    const double C = dmax(rho[s] * sp[0], 0.0);
    const double cs = dmax(C, 1.0e-100);
    lc[s] = log(cs);
    dlc[s] = dmax_dx(C, 1.0e-100) / cs;
    g[s] = gibbs(tab + OFF_NASA + 17 * s, sp, tT, lT, iT);
  }
  // the diagonal only scales the error estimate: its accumulators and
  // factors are FP32 (a non-finite diagonal is caught by the caller)
  double om[NS];
  float dom[NS], csum[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    om[s] = 0.0;
    dom[s] = csum[s] = 0.0f;
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const double* tb = table_launder(tab);
    const double b = tb[OFF_RX + 6 * r + 1], th = tb[OFF_RX + 6 * r + 2];
    const bool third = park11::THIRD[r];
    const double lnA = tb[OFF_DERIVED + 3 * r + 0], dn = tb[OFF_DERIVED + 3 * r + 1];
    const double mr = tb[OFF_DERIVED + 3 * r + 2];
    double dG = 0.0, lf = 0.0, lr = 0.0, Meff = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double a1 = park11::NU1[r][s], a2 = park11::NU2[r][s];
      if (a1 != 0.0 || a2 != 0.0) {
        dG = fma(a2 - a1, g[s], dG);
        lf = fma(a1, lc[s], lf);
        lr = fma(a2, lc[s], lr);
      }
      if (third && park11::ALPHA[r][s] != 0.0)
        Meff = fma(park11::ALPHA[r][s], dmax(rho[s] * tb[OFF_SPEC + SPEC_W * s], 0.0), Meff);
    }
    const double lkf = fma(-th, iT, fma(b, lT, lnA));
    // ln K_c = -clamp(dG) + dn ln(P / R T); k_r = k_f / max(K_c, 1e-300)
    const double lKc = dmax(fma(dn, lP, -clampd(dG, -700.0, 700.0)), -690.7755278982137);
    const double fwdk = exp(lkf + clampd(lf, -700.0, 700.0));
    const double revk = exp(lkf - lKc + clampd(lr, -700.0, 700.0));
    const double cf = fabs(lf) < 700.0 ? 1.0 : (fabs(lf) == 700.0 ? 0.5 : 0.0);
    const double cr = fabs(lr) < 700.0 ? 1.0 : (fabs(lr) == 700.0 ? 0.5 : 0.0);
    const double col = third ? Meff : 1.0;
    const double q = (fwdk - revk) * col;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double a1 = park11::NU1[r][s], a2 = park11::NU2[r][s];
      const double dcol = third ? park11::ALPHA[r][s] : 0.0;
      if (a1 != 0.0 || a2 != 0.0 || dcol != 0.0) {
        const double m = tb[OFF_SPEC + SPEC_W * s];
        const double dC = dmax_dx(rho[s] * m, 0.0) * m;
        const float dq = (float)(fma(col * fma(fwdk * cf, a1, -(revk * cr * a2)), dlc[s],
                                     dcol * (fwdk - revk)) * dC);
        const double d = a2 - a1;
        om[s] = fma(d, q, om[s]);
        dom[s] = fmaf((float)d, dq, dom[s]);
        csum[s] = fmaf((float)mr, dq, csum[s]);
      }
    }
  }
  double srho = 0.0, S = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    srho += rho[s];
    S += om[s] * mw[s];
  }
  const double tot = dmax(srho, 1.0e-14), dtot = dmax_dx(srho, 1.0e-14);
  // w = M omega - Y sum(M omega) and its diagonal derivative (the mass-sum
  // correction differentiated as in rates())
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double Y = dmax(rho[s], 0.0) / tot;
    w[s] = om[s] * mw[s] - Y * S;
    const double dY = dmax_dx(rho[s], 0.0) / tot - dmax(rho[s], 0.0) / (tot * tot) * dtot;
    dd[s] = mw[s] * (double)dom[s] - dY * S - Y * (double)csum[s];
  }
}

// Solve A x = b (A overwritten) by LU with partial pivoting, all unrolled
// (jnp.linalg.solve, hybrid/chemistry/integrator.py:81): the pivot swap is a
// select over rows (no register array indexed at run time), each column's
// multipliers use one reciprocal of its pivot.
__device__ __forceinline__ void solve(double (*A)[NS], double* b) {
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    int piv = k;
    double best = fabs(A[k][k]);
#pragma unroll
    for (int i = k + 1; i < NS; ++i) {
      double v = fabs(A[i][k]);
      if (v > best) {
        best = v;
        piv = i;
      }
    }
#pragma unroll
    for (int i = k + 1; i < NS; ++i) {
      bool sw = (piv == i);
#pragma unroll
      for (int j = k; j < NS; ++j) {
        double t = A[k][j];
        A[k][j] = sw ? A[i][j] : t;
        A[i][j] = sw ? t : A[i][j];
      }
      double t = b[k];
      b[k] = sw ? b[i] : t;
      b[i] = sw ? t : b[i];
    }
    const double inv = 1.0 / A[k][k];
#pragma unroll
    for (int i = k + 1; i < NS; ++i) {
      double f = A[i][k] * inv;
#pragma unroll
      for (int j = k + 1; j < NS; ++j) A[i][j] = fma(-f, A[k][j], A[i][j]);
      b[i] = fma(-f, b[k], b[i]);
    }
  }
#pragma unroll
  for (int k = NS - 1; k >= 0; --k) {
    double acc = b[k];
#pragma unroll
    for (int j = k + 1; j < NS; ++j) acc = fma(-A[k][j], b[j], acc);
    b[k] = acc / A[k][k];
  }
}

// hybrid/chemistry/integrator.py:31-104: backward Euler on the species with
// a fixed Newton count, non-negativity clip, temperature recovery at the
// initial internal energy after each iteration (unless held), total-mass
// rescale.  The reference evaluates its rates at (T, Tv) with Tv = T here.
template <bool PARK>
__device__ __forceinline__ double integrate(const double* tab, double* rho, double T, double dt, int iters,
                            int constant_temperature) {
  double rho0[NS], Y0[NS];
  double total = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    rho0[s] = dmax(rho[s], 1.0e-20);
    total += rho0[s];
  }
  {
    double tot = dmax(total, 1.0e-14);
#pragma unroll
    for (int s = 0; s < NS; ++s) Y0[s] = dmax(rho0[s], 0.0) / tot;
  }
  const double e_init = mixture_energy(tab, Y0, T);
  double r[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) r[s] = rho0[s];
  double Tk = T;
  for (int it = 0; it < iters; ++it) {
    double w[NS];
    double J[NS][NS];
    rates<PARK>(tab, r, Tk, w, J);
    double rhs[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      rhs[s] = -fma(-dt, w[s], r[s] - rho0[s]);
#pragma unroll
      for (int j = 0; j < NS; ++j) J[s][j] = fma(-dt, J[s][j], s == j ? 1.0 : 0.0);
    }
    solve(J, rhs);
#pragma unroll
    for (int s = 0; s < NS; ++s) r[s] = dmax(r[s] + rhs[s], 0.0);
    if (!constant_temperature) Tk = recover_temperature_converged(tab, r, e_init, Tk, 10);
  }
  double sr = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) sr += r[s];
#pragma unroll
  for (int s = 0; s < NS; ++s) rho[s] = r[s] * (total / sr);
  return Tk;
}

// The integrator kernels come in two instantiations, launched together: PARK
// (the compiled-in Park-11 stoichiometry) and the table-driven one; each
// leaves at once unless the table's stoichiometry fingerprint selects it.
template <bool PARK>
__device__ __forceinline__ bool selected(const double* tab) {
  return (tab[OFF_FP] == park11::FINGERPRINT) == PARK;
}

template <bool PARK>
__global__ __launch_bounds__(64) void integrate_kernel(const double* __restrict__ tab,
                                                       const double* __restrict__ rho_in,
                                                       const double* __restrict__ T_in,
                                                       const double* __restrict__ dt_cell,
                                                       double dt, int iters, int const_T,
                                                       double* __restrict__ rho_out,
                                                       double* __restrict__ T_out, int64_t n) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n || !selected<PARK>(tab)) return;
  double r[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) r[s] = rho_in[c * NS + s];
  double h = dt_cell ? dt_cell[c] : dt;
  double T = integrate<PARK>(tab, r, T_in[c], h, iters, const_T);
#pragma unroll
  for (int s = 0; s < NS; ++s) rho_out[c * NS + s] = r[s];
  if (T_out) T_out[c] = T;
}

// hybrid/chemistry/coupling.py:116-177 on a (cells, 16) state, in place.
template <bool PARK>
__global__ __launch_bounds__(64) void react_kernel(const double* __restrict__ tab,
                                                   double* __restrict__ U, int64_t n,
                                                   int64_t nodes_per_elem,
                                                   const double* __restrict__ dt_elem, double dt,
                                                   fsg_chem_params p) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n || !selected<PARK>(tab)) return;
  double* u = U + c * (5 + NS);
  double h = dt_elem ? dt_elem[c / nodes_per_elem] : dt;
  double rt = dmax(u[0], 1.0e-14);
  double rs[NS], sum = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    rs[s] = dmax(u[5 + s], 0.0);
    sum += rs[s];
  }
  double scale = rt / dmax(sum, 1.0e-20);
#pragma unroll
  for (int s = 0; s < NS; ++s) rs[s] *= scale;
  double ek = 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / (rt * rt);
  // the dimensionless and dimensional couplers differ in their scalings
  // only: one inlined integrator serves both
  const bool dl = p.dimensionless;
  const double eth_s = dmax(u[4] / rt - ek, dl ? 1.0e-4 : 1.0e3);
  const double eth = dl ? eth_s * p.e_ref : eth_s;
  if (dl) {
#pragma unroll
    for (int s = 0; s < NS; ++s) rs[s] = rs[s] * p.rho_ref;
  }
  const double Tg = dl ? dmax((p.gamma - 1.0) * eth_s * p.T_ref, 200.0) : 1000.0;
  const double T0 = recover_temperature_converged(tab, rs, eth, Tg, 8);
  integrate<PARK>(tab, rs, T0, dl ? h * p.t_ref : h, p.newton_iters, 0);
  if (dl) {
#pragma unroll
    for (int s = 0; s < NS; ++s) rs[s] = rs[s] / p.rho_ref;
  }
  double s2 = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    rs[s] = dmax(rs[s], 0.0);
    s2 += rs[s];
  }
  double f = rt / dmax(s2, 1.0e-20);
#pragma unroll
  for (int s = 0; s < NS; ++s) u[5 + s] = rs[s] * f;
}

// The coupler's steps before and after its integrator call
// (hybrid/chemistry/coupling.py:116-154 and :165-177), split out so another
// substep (the learned surrogate) can run in between.  Flow states of either
// precision; chemistry in FP64.  prepare_kernel is held to three waves per
// SIMD: its 36 B of register spill per lane cost less than the third wave
// gains on the latency of the temperature recovery (683 -> 562 us per launch
// on the reacting mesh's blocks).
template <typename ureal>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
void prepare_kernel(const double* __restrict__ tab,
                                                     const ureal* __restrict__ U, int64_t n,
                                                     fsg_chem_params p, double* __restrict__ rho_out,
                                                     double* __restrict__ T_out) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const ureal* u = U + c * (5 + NS);
  const double rt = dmax((double)u[0], 1.0e-14);
  double rs[NS], sum = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    rs[s] = dmax((double)u[5 + s], 0.0);
    sum += rs[s];
  }
  const double scale = rt / dmax(sum, 1.0e-20);
#pragma unroll
  for (int s = 0; s < NS; ++s) rs[s] *= scale;
  const double m0 = u[1], m1 = u[2], m2 = u[3];
  const double ek = 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / (rt * rt);
  double T0;
  if (p.dimensionless) {
    const double eth_s = dmax((double)u[4] / rt - ek, 1.0e-4);
#pragma unroll
    for (int s = 0; s < NS; ++s) rs[s] *= p.rho_ref;
    const double Tg = dmax((p.gamma - 1.0) * eth_s * p.T_ref, 200.0);
    T0 = recover_temperature_converged(tab, rs, eth_s * p.e_ref, Tg, 8);
  } else {
    const double eth = dmax((double)u[4] / rt - ek, 1.0e3);
    T0 = recover_temperature_converged(tab, rs, eth, 1000.0, 8);
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) rho_out[c * NS + s] = rs[s];
  T_out[c] = T0;
}

template <typename ureal>
__global__ __launch_bounds__(64) void finish_kernel(ureal* __restrict__ U, int64_t n,
                                                    fsg_chem_params p,
                                                    const double* __restrict__ rho_new) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  ureal* u = U + c * (5 + NS);
  const double rt = dmax((double)u[0], 1.0e-14);
  double rs[NS], s2 = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double r = rho_new[c * NS + s];
    if (p.dimensionless) r = r / p.rho_ref;
    rs[s] = dmax(r, 0.0);
    s2 += rs[s];
  }
  const double f = rt / dmax(s2, 1.0e-20);
#pragma unroll
  for (int s = 0; s < NS; ++s) u[5 + s] = (ureal)(rs[s] * f);
}

// The integrator's judgement of a predicted substep: the backward-Euler
// residual R = rho_p - rho_0 - dt w(rho_p, T_p) of the scheme
// hybrid/chemistry/integrator.py:31-104 solves (rho_0 = max(rho_in, 1e-20);
// T_p recovered from rho_p at the initial internal energy unless the
// temperature is held), scaled by the Jacobian diagonal into the one-step
// error estimate |R_s| / max(|1 - dt dw_s/drho_s|, 1) (a stiff, decaying
// species' residual is divided down; no residual is amplified).  A cell whose estimate
// exceeds rtol |rho_p,s - rho_0,s| + atol sum(rho_0) for any species is
// flagged (flag |= 1) for the integrator.  Held to three waves per SIMD:
// the 216 B of register spill per lane cost less than the third wave gains
// (1.75 -> 1.66 ms per reacting-mesh launch).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
void check_kernel(const double* __restrict__ tab,
                                                   const double* __restrict__ rho_in,
                                                   const double* __restrict__ T_in,
                                                   const double* __restrict__ dt_cell, double dt,
                                                   int const_T, const double* __restrict__ rho_p,
                                                   double rtol, double atol,
                                                   int32_t* __restrict__ flag,
                                                   double* __restrict__ err_out, int64_t n) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  // the rates use the compiled-in Park-11 stoichiometry: a table built for
  // other stoichiometry or third-body data is not judged here, its cells go
  // to the integrator
  if (tab[OFF_FP] != park11::FINGERPRINT) {
    flag[c] = 1;
    if (err_out) err_out[c] = __builtin_huge_val();
    return;
  }
  // rho_0 is read again after the rates rather than held across them
  double rp[NS], Y0[NS], total = 0.0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    Y0[s] = dmax(rho_in[c * NS + s], 1.0e-20);
    rp[s] = rho_p[c * NS + s];
    total += Y0[s];
  }
  double T = T_in[c];
  if (!const_T) {
    const double tot = dmax(total, 1.0e-14);
#pragma unroll
    for (int s = 0; s < NS; ++s) Y0[s] = Y0[s] / tot;
    T = recover_temperature_converged(tab, rp, mixture_energy(tab, Y0, T), T, 10);
  }
  const double h = dt_cell ? dt_cell[c] : dt;
  double w[NS], dd[NS];
  rates_diag(tab, rp, T, w, dd);
  double worst = 0.0;
  bool bad = false;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const double r0 = dmax(rho_in[c * NS + s], 1.0e-20);
    const double R = rp[s] - r0 - h * w[s];
    const double a = fabs(1.0 - h * dd[s]);
    const double est = fabs(R) / (a - a == 0.0 ? dmax(a, 1.0) : 1.0);
    const double tol = rtol * fabs(rp[s] - r0) + atol * total;
    bad = bad || !(est <= tol);
    worst = dmax(worst, est / dmax(total, 1.0e-300));
  }
  if (bad) flag[c] = 1;
  if (err_out) err_out[c] = worst;
}

// Integrator on the flagged cells of a batch only (the surrogate's
// fallback), without a host round trip: flagged cell ids are compacted with
// wavefront ballots into list[0..*count), then a bounded grid strides over
// the list entries (and returns at once when the list is empty).  Cells are
// independent, so the list order (set by the atomics) does not change any
// result.
__global__ void zero_count_kernel(int32_t* count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *count = 0;
}

__global__ __launch_bounds__(256) void compact_flags_kernel(const int32_t* __restrict__ flag,
                                                            int64_t n, int32_t* __restrict__ list,
                                                            int32_t* __restrict__ count) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool f = c < n && flag[c] != 0;
  const uint64_t mask = __ballot(f);
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (lane == 0 && mask) base = atomicAdd(count, (int)__popcll(mask));
  base = __shfl(base, 0, 64);
  if (f) {
    const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                               __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
    list[base + rank] = (int32_t)c;
  }
}

template <bool PARK>
__global__ __launch_bounds__(64) void integrate_list_kernel(
    const double* __restrict__ tab, const double* __restrict__ rho_in,
    const double* __restrict__ T_in, const double* __restrict__ dt_cell, double dt, int iters,
    int const_T, const int32_t* __restrict__ list, const int32_t* __restrict__ count,
    double* __restrict__ rho_out, double* __restrict__ T_out) {
  if (!selected<PARK>(tab)) return;
  const int64_t m = *count;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = list[i];
    double r[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) r[s] = rho_in[c * NS + s];
    const double h = dt_cell ? dt_cell[c] : dt;
    const double T = integrate<PARK>(tab, r, T_in[c], h, iters, const_T);
#pragma unroll
    for (int s = 0; s < NS; ++s) rho_out[c * NS + s] = r[s];
    if (T_out) T_out[c] = T;
  }
}

}  // namespace fsgchem

// Scratch of fsg_chem_integrate_flagged_noalloc: the compacted list of
// flagged cells (FSG_NOALLOC_MAX_CELLS int32) followed by its counter.  One
// copy per GPU, held until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns on each device.
static __device__ __align__(256) int32_t scratch[FSG_NOALLOC_MAX_CELLS + 64];

static thread_local std::string g_chem_error;

static int chem_fail(int code, const char* msg) {
  g_chem_error = msg;
  return code;
}

extern "C" {

int32_t fsg_chem_table_size(void) { return fsgchem::TABLE_SIZE; }

int fsg_chem_integrate(const double* table, const double* rho_in, const double* T_in,
                       const double* dt_cell, double dt, int32_t newton_iters,
                       int32_t constant_temperature, double* rho_out, double* T_out, int64_t n,
                       void* stream) {
  if (!table || !rho_in || !T_in || !rho_out) return chem_fail(-10, "null argument");
  if (n <= 0) return 0;
  if (newton_iters < 0) return chem_fail(-11, "negative iteration count");
  unsigned grid = (unsigned)((n + 63) / 64);
  hipLaunchKernelGGL(fsgchem::integrate_kernel<true>, dim3(grid), dim3(64), 0,
                     static_cast<hipStream_t>(stream), table, rho_in, T_in, dt_cell, dt,
                     newton_iters, constant_temperature, rho_out, T_out, n);
  hipLaunchKernelGGL(fsgchem::integrate_kernel<false>, dim3(grid), dim3(64), 0,
                     static_cast<hipStream_t>(stream), table, rho_in, T_in, dt_cell, dt,
                     newton_iters, constant_temperature, rho_out, T_out, n);
  return hipGetLastError() == hipSuccess ? 0 : chem_fail(-1, "chemistry launch failed");
}

int fsg_chem_react(const double* table, double* U, int64_t ncells, int64_t nodes_per_elem,
                   const double* dt_elem, double dt, const fsg_chem_params* params,
                   void* stream) {
  if (!table || !U || !params) return chem_fail(-10, "null argument");
  if (ncells <= 0) return 0;
  if (nodes_per_elem <= 0) return chem_fail(-12, "nodes_per_elem must be positive");
  unsigned grid = (unsigned)((ncells + 63) / 64);
  hipLaunchKernelGGL(fsgchem::react_kernel<true>, dim3(grid), dim3(64), 0,
                     static_cast<hipStream_t>(stream), table, U, ncells, nodes_per_elem, dt_elem,
                     dt, *params);
  hipLaunchKernelGGL(fsgchem::react_kernel<false>, dim3(grid), dim3(64), 0,
                     static_cast<hipStream_t>(stream), table, U, ncells, nodes_per_elem, dt_elem,
                     dt, *params);
  return hipGetLastError() == hipSuccess ? 0 : chem_fail(-1, "chemistry launch failed");
}

int fsg_chem_integrate_flagged(const double* table, const double* rho_in, const double* T_in,
                               const double* dt_cell, double dt, int32_t newton_iters,
                               int32_t constant_temperature, const int32_t* flag, int32_t* list,
                               int32_t* count, double* rho_out, double* T_out, int64_t n,
                               void* stream) {
  if (!table || !rho_in || !T_in || !flag || !list || !count || !rho_out)
    return chem_fail(-10, "null argument");
  if (n <= 0) return 0;
  if (n > INT32_MAX) return chem_fail(-12, "batch too large for int32 cell ids");
  if (newton_iters < 0) return chem_fail(-11, "negative iteration count");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(fsgchem::zero_count_kernel, dim3(1), dim3(64), 0, s, count);
  hipLaunchKernelGGL(fsgchem::compact_flags_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256),
                     0, s, flag, n, list, count);
  // a bounded grid strides over the list: 8192 64-lane workgroups exceed the
  // integrator's resident capacity (one wave per SIMD at its register count)
  // and leave at once when nothing is flagged
  const int64_t blocks = (n + 63) / 64 < 8192 ? (n + 63) / 64 : 8192;
  hipLaunchKernelGGL(fsgchem::integrate_list_kernel<true>, dim3((unsigned)blocks), dim3(64),
                     0, s, table,
                     rho_in, T_in, dt_cell, dt, newton_iters, constant_temperature, list, count,
                     rho_out, T_out);
  hipLaunchKernelGGL(fsgchem::integrate_list_kernel<false>, dim3((unsigned)blocks), dim3(64),
                     0, s, table,
                     rho_in, T_in, dt_cell, dt, newton_iters, constant_temperature, list, count,
                     rho_out, T_out);
  return hipGetLastError() == hipSuccess ? 0 : chem_fail(-1, "chemistry launch failed");
}

int fsg_chem_integrate_flagged_noalloc(const double* table, const double* rho_in,
                                       const double* T_in, const double* dt_cell, double dt,
                                       int32_t newton_iters, int32_t constant_temperature,
                                       const int32_t* flag, double* rho_out, double* T_out,
                                       int64_t n, void* stream) {
  if (!table || !rho_in || !T_in || !flag || !rho_out) return chem_fail(-10, "null argument");
  if (n <= 0) return 0;
  if (n > FSG_NOALLOC_MAX_CELLS)
    return chem_fail(-13, "batch exceeds FSG_NOALLOC_MAX_CELLS cells of the static scratch");
  // the scratch address of the current device, looked up once per device
  static std::mutex lock;
  static int32_t* base[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
    return chem_fail(-1, "no current HIP device");
  int32_t* ptr = nullptr;
  {
    std::lock_guard<std::mutex> g(lock);
    if (!base[dev]) {
      void* addr = nullptr;
      if (hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)) != hipSuccess || !addr)
        return chem_fail(-1, "hipGetSymbolAddress of the chemistry scratch failed");
      base[dev] = static_cast<int32_t*>(addr);
    }
    ptr = base[dev];
  }
  return fsg_chem_integrate_flagged(table, rho_in, T_in, dt_cell, dt, newton_iters,
                                    constant_temperature, flag, ptr, ptr + FSG_NOALLOC_MAX_CELLS,
                                    rho_out, T_out, n, stream);
}

int fsg_chem_check(const double* table, const double* rho_in, const double* T_in,
                   const double* dt_cell, double dt, int32_t constant_temperature,
                   const double* rho_pred, double rtol, double atol, int32_t* flag,
                   double* err_out, int64_t n, void* stream) {
  if (!table || !rho_in || !T_in || !rho_pred || !flag) return chem_fail(-10, "null argument");
  if (n <= 0) return 0;
  hipLaunchKernelGGL(fsgchem::check_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0,
                     static_cast<hipStream_t>(stream), table, rho_in, T_in, dt_cell, dt,
                     constant_temperature, rho_pred, rtol, atol, flag, err_out, n);
  return hipGetLastError() == hipSuccess ? 0 : chem_fail(-1, "chemistry launch failed");
}

int fsg_chem_prepare(const double* table, const void* U, int32_t dtype, int64_t ncells,
                     const fsg_chem_params* params, double* rho_out, double* T_out,
                     void* stream) {
  if (!table || !U || !params || !rho_out || !T_out) return chem_fail(-10, "null argument");
  if (ncells <= 0) return 0;
  unsigned grid = (unsigned)((ncells + 63) / 64);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == FSG_F64)
    hipLaunchKernelGGL(fsgchem::prepare_kernel<double>, dim3(grid), dim3(64), 0, s, table,
                       static_cast<const double*>(U), ncells, *params, rho_out, T_out);
  else
    hipLaunchKernelGGL(fsgchem::prepare_kernel<float>, dim3(grid), dim3(64), 0, s, table,
                       static_cast<const float*>(U), ncells, *params, rho_out, T_out);
  return hipGetLastError() == hipSuccess ? 0 : chem_fail(-1, "chemistry launch failed");
}

int fsg_chem_finish(void* U, int32_t dtype, int64_t ncells, const fsg_chem_params* params,
                    const double* rho_new, void* stream) {
  if (!U || !params || !rho_new) return chem_fail(-10, "null argument");
  if (ncells <= 0) return 0;
  unsigned grid = (unsigned)((ncells + 63) / 64);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == FSG_F64)
    hipLaunchKernelGGL(fsgchem::finish_kernel<double>, dim3(grid), dim3(64), 0, s,
                       static_cast<double*>(U), ncells, *params, rho_new);
  else
    hipLaunchKernelGGL(fsgchem::finish_kernel<float>, dim3(grid), dim3(64), 0, s,
                       static_cast<float*>(U), ncells, *params, rho_new);
  return hipGetLastError() == hipSuccess ? 0 : chem_fail(-1, "chemistry launch failed");
}

const char* fsg_chem_last_error(void) { return g_chem_error.c_str(); }

}  // extern "C"

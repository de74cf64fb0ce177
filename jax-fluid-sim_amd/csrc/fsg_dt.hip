// Time-step control on the device: the CFL estimate of
// hybrid/time_integration.py:93-169 reduced to one dt, the element-local dt
// of :171-262 (neighbour smoothing and sliver scaling), the FP64
// conservation sums, and the accept / reject bookkeeping of the step loop in
// run_hybrid_sphere.py:265-346, so that a run needs one host
// synchronisation per batch of steps instead of one per step.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/fluid_gfx950.h"
#include "fsg_context.h"
#include "fsg_physics.h"

struct fsg_dt {
  fsg_context* ctx;
  std::vector<fsg_dt_block> blocks;
  std::vector<fsg_dt_conn> conns;
  double* work;
  int64_t work_len;
};

namespace fsgdt {

constexpr int WG = 256;

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

struct RateParams {
  double gamma, mu, kappa, cp, pen, k;
  int wale;            // add the WALE eddy viscosity (hybrid/time_integration.py:119-148)
  double cw, cp_prt;   // C_w and c_p / Pr_t
};

// Per-block tables of the WALE term: basis and reference gradients at the
// volume points, J^-1 and the filter widths (context dtype).
template <typename real>
struct WaleTables {
  const real* Bt;     // (N, Q)
  const real* Dt;     // (N, 3, Q)
  const real* jinv;   // (E, 9)
  const real* width;  // (E)
  int Q;
};

// max over the volume points of the WALE viscosity of the resolved velocity
// gradient (the physical basis gradients applied to the nodal velocities,
// density interpolated to the point), in FP64.  One wave per element, at
// most 64 nodes: lane i holds node i's density and velocity and every lane
// reads them by shuffles for its points.
template <typename real>
__device__ __forceinline__ double wale_max(const real* U, int N, int V, int64_t e,
                                           const WaleTables<real>& w, double cw, int lane) {
  double rho = 0.0, u0 = 0.0, u1 = 0.0, u2 = 0.0;
  if (lane < N) {
    const real* u = U + (int64_t)lane * V;
    rho = double(u[0]);
    const double safe = fmax(rho, 1.0e-14);
    u0 = double(u[1]) / safe;
    u1 = double(u[2]) / safe;
    u2 = double(u[3]) / safe;
  }
  double J[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) J[k] = double(w.jinv[e * 9 + k]);
  const double width = double(w.width[e]);
  double mx = 0.0;
  for (int q0 = 0; q0 < w.Q; q0 += 64) {
    const int q = q0 + lane;
    const int qc = q < w.Q ? q : w.Q - 1;
    double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, rq = 0.0;
    for (int i = 0; i < N; ++i) {
      const double ri = __shfl(rho, i, 64);
      const double vi[3] = {__shfl(u0, i, 64), __shfl(u1, i, 64), __shfl(u2, i, 64)};
      const double b = double(w.Bt[(int64_t)i * w.Q + qc]);
      const double d[3] = {double(w.Dt[((int64_t)i * 3 + 0) * w.Q + qc]),
                           double(w.Dt[((int64_t)i * 3 + 1) * w.Q + qc]),
                           double(w.Dt[((int64_t)i * 3 + 2) * w.Q + qc])};
      rq += b * ri;
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int r = 0; r < 3; ++r) G[m * 3 + r] += d[r] * vi[m];
    }
    double g[9];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        g[m * 3 + c] = G[m * 3 + 0] * J[0 * 3 + c] + G[m * 3 + 1] * J[1 * 3 + c] +
                       G[m * 3 + 2] * J[2 * 3 + c];
    const double mt = fsg::wale_viscosity<double>(g, rq, width, cw);
    if (q < w.Q) mx = mt != mt ? mt : (mx != mx ? mx : fmax(mx, mt));
  }
  // NaN-propagating wave maximum
  bool nan = mx != mx;
  nan = __any(nan);
  mx = wave_max(nan ? 0.0 : mx);
  return nan ? __longlong_as_double(0x7ff8000000000000LL) : mx;
}

// Stability rate of one element (one wave per element, lanes over nodes):
// k max_i(|u_i| + c_i) / L, plus 2 eta k^2 max(4/3 mu, kappa / cp) / rho_mean
// / max(L, 0.01)^2 when viscous.  Densities and pressures are floored at
// 1e-14 as in hybrid/time_integration.py:104-108.
template <typename real>
__device__ __forceinline__ double element_rate(const real* U, int N, int V, double L,
                                               const RateParams& q, int lane, int64_t e,
                                               const WaleTables<real>& w) {
  double smax = 0.0, rsum = 0.0;
  bool nan = false;
  for (int i = lane; i < N; i += 64) {
    const real* u = U + (int64_t)i * V;
    const double rho_raw = double(u[0]);
    const double m0 = double(u[1]), m1 = double(u[2]), m2 = double(u[3]), E = double(u[4]);
    const double safe = fmax(rho_raw, 1.0e-14);
    const double v0 = m0 / safe, v1 = m1 / safe, v2 = m2 / safe;
    const double ke = 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) / safe;
    const double p = (q.gamma - 1.0) * (E - ke);
    const double rho = fmax(rho_raw, 1.0e-14), pf = fmax(p, 1.0e-14);
    const double s = sqrt(v0 * v0 + v1 * v1 + v2 * v2) + sqrt(q.gamma * pf / rho);
    nan |= s != s || rho != rho;
    smax = fmax(smax, s);
    rsum += rho;
  }
  smax = wave_max(smax);
  rsum = wave_sum(rsum);
  nan = __any(nan);
  double rate = q.k * smax / L;
  if (q.mu > 0.0 || q.wale) {
    const double rmin = fmax(rsum / N, 0.1);
    double mu = q.mu, kappa = q.kappa;
    if (q.wale) {
      const double mt = wale_max<real>(U, N, V, e, w, q.cw, lane);
      mu += mt;
      kappa += mt * q.cp_prt;
    }
    const double diff = fmax((4.0 / 3.0) * mu / rmin, kappa / (rmin * q.cp));
    const double lv = fmax(L, 0.01);
    rate = rate + 2.0 * q.pen * q.k * q.k * diff / (lv * lv);
  }
  const double qnan = __longlong_as_double(0x7ff8000000000000LL);
  return nan ? qnan : rate;
}

// Global maximum rate: NaN and +inf count as 1e7 (the nan_to_num of
// :166); non-negative doubles order like their bit patterns.  A bounded grid
// strides over the elements (one wave per element at a time), each
// workgroup reduces its waves' maxima in LDS and issues one atomic: one
// atomic per element wave on a single address serialised the launch.
constexpr int RATE_GRID = 4096;

template <typename real>
__global__ __launch_bounds__(WG) void max_rate_kernel(const real* __restrict__ U, int64_t E,
                                                      int N, int V, const double* __restrict__ L,
                                                      RateParams q, WaleTables<real> w,
                                                      unsigned long long* __restrict__ out) {
  __shared__ unsigned long long sm[WG / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long mb = 0ull;
  for (int64_t e = (int64_t)blockIdx.x * (WG / 64) + wave; e < E;
       e += (int64_t)gridDim.x * (WG / 64)) {
    double r = element_rate<real>(U + e * N * V, N, V, L[e], q, lane, e, w);
    if (r != r || r > 1.0e300) r = 1.0e7;
    const unsigned long long rb = (unsigned long long)__double_as_longlong(r);
    mb = rb > mb ? rb : mb;
  }
  if (lane == 0) sm[wave] = mb;
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 1; k < WG / 64; ++k) mb = sm[k] > mb ? sm[k] : mb;
    atomicMax(out, mb);
  }
}

__global__ void zero_u64_kernel(unsigned long long* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0ull;
}

// dt = clip(cfl / max(rate, 1e-12), 1e-7, 1e-3) (:167-169); with a step
// state, dt also becomes its CFL dt when the accepted-step count is a
// multiple of the refresh interval (run_hybrid_sphere.py:283-285).
__global__ void finish_dt_kernel(const unsigned long long* __restrict__ maxr, double cfl,
                                 double* __restrict__ dt_out, fsg_step_state* st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double r = __longlong_as_double((long long)*maxr);
  const double dt = fmin(fmax(cfl / fmax(r, 1.0e-12), 1.0e-7), 1.0e-3);
  if (dt_out) *dt_out = dt;
  if (st && (st->have_dt == 0 || st->step % st->refresh_interval == 0)) {
    st->cfl_dt = dt;
    st->have_dt = 1;
  }
}

// Element-local dt = cfl / max(rate, 1e-12) (:220), in FP64.
template <typename real>
__global__ __launch_bounds__(WG) void local_rate_kernel(const real* __restrict__ U, int64_t E,
                                                        int N, int V, const double* __restrict__ L,
                                                        RateParams q, WaleTables<real> w,
                                                        double cfl, double* __restrict__ dt) {
  const int lane = threadIdx.x & 63;
  const int64_t e = (int64_t)blockIdx.x * (WG / 64) + (threadIdx.x >> 6);
  if (e >= E) return;
  const double r = element_rate<real>(U + e * N * V, N, V, L[e], q, lane, e, w);
  if (lane == 0) dt[e] = r != r ? r : cfl / fmax(r, 1.0e-12);
}

// One connection group of the smoothing (:223-230): both sides read before
// either is updated, then dt = min(dt, 1.3 dt_other) on each side.
__global__ void smooth_gather_kernel(const double* __restrict__ dto, const double* __restrict__ dtn,
                                     const int32_t* __restrict__ oc, const int32_t* __restrict__ nc,
                                     int64_t n, double* __restrict__ tmp) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  tmp[f] = dto[oc[f]];
  tmp[n + f] = dtn[nc[f]];
}

__device__ __forceinline__ void atomic_min_pos(double* a, double v) {
  atomicMin(reinterpret_cast<unsigned long long*>(a), (unsigned long long)__double_as_longlong(v));
}

__global__ void smooth_apply_kernel(double* __restrict__ dto, double* __restrict__ dtn,
                                    const int32_t* __restrict__ oc, const int32_t* __restrict__ nc,
                                    int64_t n, const double* __restrict__ tmp) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  atomic_min_pos(dto + oc[f], 1.3 * tmp[n + f]);
  atomic_min_pos(dtn + nc[f], 1.3 * tmp[f]);
}

// Sliver scaling (:232-240), an optional retry scale, and the conversion to
// the stage dtype; dt64 (may be null) receives the FP64 value.
template <typename real>
__global__ void finish_local_kernel(const double* __restrict__ dt, const double* __restrict__ sliver,
                                    int64_t E, const double* __restrict__ factor,
                                    real* __restrict__ out, double* __restrict__ dt64) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double v = dt[e] * sliver[e];
  if (factor) v *= *factor;
  out[e] = real(v);
  if (dt64) dt64[e] = v;
}

// Conservation sums: per element sum_i (int phi_i) |det J| U_i, partial
// sums per workgroup in a fixed order, then a fixed-order final sum.
template <typename real>
__global__ __launch_bounds__(WG) void totals_partial_kernel(const real* __restrict__ U, int64_t E,
                                                            int N, int V,
                                                            const double* __restrict__ nw,
                                                            const double* __restrict__ detw,
                                                            double* __restrict__ part) {
  __shared__ double s[WG];
  const int64_t base = (int64_t)blockIdx.x * WG;
  const int64_t e = base + threadIdx.x;
  for (int v = 0; v < V; ++v) {
    double acc = 0.0;
    if (e < E) {
      const real* u = U + e * N * V + v;
      for (int i = 0; i < N; ++i) acc += nw[i] * double(u[(int64_t)i * V]);
      acc *= detw[e];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = WG / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * V + v] = s[0];
    __syncthreads();
  }
}

__global__ void totals_final_kernel(const double* __restrict__ part, int64_t nparts, int V,
                                    double* __restrict__ out) {
  const int v = threadIdx.x;
  if (v >= V) return;
  double acc = 0.0;
  for (int64_t k = 0; k < nparts; ++k) acc += part[k * V + v];
  out[v] = acc;
}

// Step begin: dt = min(retry * cfl_dt, target - t) for a global step, or
// retry * dt_LTS per element; zero when the run is finished or failed.
__global__ void step_begin_kernel(fsg_step_state* st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const bool done = st->status != 0 || st->time >= st->target_time ||
                    st->step >= st->maximum_steps;
  st->active = done ? 0 : 1;
  const double dt = done ? 0.0 : fmin(st->retry * st->cfl_dt, st->target_time - st->time);
  st->dt = dt;
  st->factor = done ? 0.0 : st->retry;
}

template <typename real>
__global__ void fill_dt_kernel(const fsg_step_state* __restrict__ st, int64_t E,
                               real* __restrict__ out, double* __restrict__ dt64) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  out[e] = real(st->dt);
  if (dt64) dt64[e] = st->dt;
}

// Step decision from the stage diagnostics (run_hybrid_sphere.py:301-346).
__global__ void step_decide_kernel(fsg_step_state* st, const fsg_stage_diag* d, double rho_floor,
                                   double p_floor, double dt_taken) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  st->accept = 0;
  if (!st->active) return;
  const bool valid = d->nonfinite == 0 && d->rho_min >= rho_floor && d->p_min >= p_floor;
  st->attempts += 1;
  if (valid) {
    st->accept = 1;
    const double dt = dt_taken >= 0.0 ? dt_taken : st->dt;
    st->time += dt;
    st->step += 1;
    st->consecutive = 0;
    st->retry = fmin(1.0, 1.25 * st->retry);
    st->min_dt = fmin(st->min_dt, dt);
    const double tol = 32.0 * 1.1920928955078125e-07 * fmax(fabs(st->target_time), 1.0e-12);
    if (st->target_time - st->time <= tol) st->time = st->target_time;
  } else {
    st->rejections += 1;
    st->consecutive += 1;
    st->retry *= 0.5;
    st->last_rho_min = d->rho_min;
    st->last_p_min = d->p_min;
    if (st->consecutive >= st->maximum_rejections) st->status = 1;
  }
}

// states <- candidate where the step was accepted.
template <typename real>
__global__ void commit_kernel(const fsg_step_state* __restrict__ st, const real* __restrict__ cand,
                              real* __restrict__ U, int64_t n) {
  if (!st->accept) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    U[i] = cand[i];
}

// min over the elements of the local dt (the dt reported for an LTS step).
__global__ void min_dt_kernel(const double* __restrict__ dt, int64_t E,
                              unsigned long long* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v = e < E ? dt[e] : 1.0e300;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  if (lane == 0 && e - lane < E) atomicMin(out, (unsigned long long)__double_as_longlong(v));
}

__global__ void set_u64_kernel(unsigned long long* p, unsigned long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void lts_dt_kernel(fsg_step_state* st, const unsigned long long* mn) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  st->dt = st->active ? __longlong_as_double((long long)*mn) : 0.0;
}

}  // namespace fsgdt

static thread_local std::string g_dt_error;

static int dt_fail(int code, const char* msg) {
  g_dt_error = msg;
  return code;
}

static int dt_launched() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return dt_fail(-1, hipGetErrorString(e));
  return 0;
}

namespace {

using fsgdt::RateParams;

RateParams rate_params(const fsg_context* c) {
  RateParams q;
  q.gamma = c->params.gamma;
  q.mu = c->params.dynamic_viscosity;
  q.kappa = c->params.thermal_conductivity;
  q.cp = c->params.gamma * c->params.gas_constant / (c->params.gamma - 1.0);
  q.pen = c->params.br2_penalty;
  q.k = 2.0 * c->params.degree + 1.0;
  q.wale = c->params.use_wale;
  q.cw = c->params.wale_constant;
  q.cp_prt = c->params.use_wale ? q.cp / c->params.turbulent_prandtl : 0.0;
  return q;
}

int vol_points_of(int type, int p) {
  if (p == 1) return type == 0 ? 8 : (type == 1 ? 6 : 4);
  return (p == 2 && type == 0) ? 27 : 125;
}

template <typename real>
fsgdt::WaleTables<real> wale_tables(const fsg_context* c, const fsg_block_desc& bd) {
  fsgdt::WaleTables<real> w;
  w.Bt = static_cast<const real*>(bd.Bt);
  w.Dt = static_cast<const real*>(bd.Dt);
  w.jinv = static_cast<const real*>(bd.jinv);
  w.width = static_cast<const real*>(bd.width);
  w.Q = vol_points_of(bd.element_type, c->params.degree);
  return w;
}

int nodes_of(int type, int p) {
  if (type == 0) return (p + 1) * (p + 1) * (p + 1);
  if (type == 1) return (p + 1) * (p + 1) * (p + 2) / 2;
  return (p + 1) * (p + 2) * (p + 3) / 6;
}

unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

// Scratch layout in fsg_dt::work: [0] max-rate bits, [1] min-dt bits,
// [8, 8 + sum E) element dt, then smoothing temporaries (2 x max group
// size), then the totals' partial sums.
int64_t elem_base(const fsg_dt*) { return 8; }

}  // namespace

extern "C" {

int fsg_dt_create(fsg_context* ctx, const fsg_dt_block* blocks, int32_t nconn,
                  const fsg_dt_conn* conns, double* work, int64_t work_len, fsg_dt** out) {
  if (!ctx || !blocks || !out || !work || (nconn > 0 && !conns))
    return dt_fail(-10, "null argument");
  const int64_t need = fsg_dt_work_size(ctx, nconn, conns);
  if (work_len < need) return dt_fail(-12, "work buffer smaller than fsg_dt_work_size");
  for (size_t b = 0; b < ctx->blocks.size(); ++b)
    if (!blocks[b].length || !blocks[b].sliver || !blocks[b].node_weights || !blocks[b].detw)
      return dt_fail(-10, "null table in a dt block");
  fsg_dt* d = new fsg_dt;
  d->ctx = ctx;
  d->blocks.assign(blocks, blocks + ctx->blocks.size());
  d->conns.assign(conns, conns + (nconn > 0 ? nconn : 0));
  d->work = work;
  d->work_len = work_len;
  *out = d;
  return 0;
}

int64_t fsg_dt_work_size(const fsg_context* ctx, int32_t nconn, const fsg_dt_conn* conns) {
  if (!ctx) return -1;
  int64_t E = 0, parts = 0;
  for (const fsg_block_desc& b : ctx->blocks) {
    E += b.count;
    parts += (b.count + fsgdt::WG - 1) / fsgdt::WG;
  }
  int64_t mc = 0;
  for (int32_t k = 0; k < nconn && conns; ++k) mc = conns[k].count > mc ? conns[k].count : mc;
  return 8 + E + 2 * mc + parts * ctx->params.nvars + 8;
}

void fsg_dt_destroy(fsg_dt* d) { delete d; }

int fsg_stable_dt(fsg_dt* d, const void* const* U, double cfl, double* dt_out,
                  fsg_step_state* state, void* stream) {
  if (!d || !U || (!dt_out && !state)) return dt_fail(-10, "null argument");
  const fsg_context* c = d->ctx;
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned long long* mx = reinterpret_cast<unsigned long long*>(d->work);
  hipLaunchKernelGGL(fsgdt::zero_u64_kernel, dim3(1), dim3(64), 0, s, mx);
  const RateParams q = rate_params(c);
  const int V = c->params.nvars, P = c->params.degree;
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    const fsg_block_desc& bd = c->blocks[b];
    if (!U[b]) return dt_fail(-10, "null state block");
    if (bd.count == 0) continue;
    const int N = nodes_of(bd.element_type, P);
    const unsigned g = std::min(blocks_for(bd.count, fsgdt::WG / 64), (unsigned)fsgdt::RATE_GRID);
    if (c->params.dtype == FSG_F64)
      hipLaunchKernelGGL(fsgdt::max_rate_kernel<double>, dim3(g), dim3(fsgdt::WG), 0, s,
                         static_cast<const double*>(U[b]), bd.count, N, V, d->blocks[b].length,
                         q, wale_tables<double>(c, bd), mx);
    else
      hipLaunchKernelGGL(fsgdt::max_rate_kernel<float>, dim3(g), dim3(fsgdt::WG), 0, s,
                         static_cast<const float*>(U[b]), bd.count, N, V, d->blocks[b].length,
                         q, wale_tables<float>(c, bd), mx);
  }
  hipLaunchKernelGGL(fsgdt::finish_dt_kernel, dim3(1), dim3(64), 0, s, mx, cfl, dt_out, state);
  return dt_launched();
}

int fsg_local_dt(fsg_dt* d, const void* const* U, double cfl, void* const* dt_elem,
                 double* const* dt_elem64, const double* factor, void* stream) {
  if (!d || !U || !dt_elem) return dt_fail(-10, "null argument");
  const fsg_context* c = d->ctx;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const RateParams q = rate_params(c);
  const int V = c->params.nvars, P = c->params.degree;
  std::vector<double*> dt(c->blocks.size());
  int64_t off = elem_base(d);
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    dt[b] = d->work + off;
    off += c->blocks[b].count;
  }
  double* tmp = d->work + off;
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    const fsg_block_desc& bd = c->blocks[b];
    if (!U[b] || !dt_elem[b]) return dt_fail(-10, "null state or dt block");
    if (bd.count == 0) continue;
    const int N = nodes_of(bd.element_type, P);
    const unsigned g = blocks_for(bd.count, fsgdt::WG / 64);
    if (c->params.dtype == FSG_F64)
      hipLaunchKernelGGL(fsgdt::local_rate_kernel<double>, dim3(g), dim3(fsgdt::WG), 0, s,
                         static_cast<const double*>(U[b]), bd.count, N, V, d->blocks[b].length,
                         q, wale_tables<double>(c, bd), cfl, dt[b]);
    else
      hipLaunchKernelGGL(fsgdt::local_rate_kernel<float>, dim3(g), dim3(fsgdt::WG), 0, s,
                         static_cast<const float*>(U[b]), bd.count, N, V, d->blocks[b].length,
                         q, wale_tables<float>(c, bd), cfl, dt[b]);
  }
  for (int round = 0; round < 2; ++round)
    for (const fsg_dt_conn& k : d->conns) {
      if (k.count == 0) continue;
      const unsigned g = blocks_for(k.count, 256);
      hipLaunchKernelGGL(fsgdt::smooth_gather_kernel, dim3(g), dim3(256), 0, s,
                         dt[k.owner_block], dt[k.neighbor_block], k.owner_cells,
                         k.neighbor_cells, k.count, tmp);
      hipLaunchKernelGGL(fsgdt::smooth_apply_kernel, dim3(g), dim3(256), 0, s,
                         dt[k.owner_block], dt[k.neighbor_block], k.owner_cells,
                         k.neighbor_cells, k.count, tmp);
    }
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    const int64_t E = c->blocks[b].count;
    if (E == 0) continue;
    double* d64 = dt_elem64 ? dt_elem64[b] : nullptr;
    if (c->params.dtype == FSG_F64)
      hipLaunchKernelGGL(fsgdt::finish_local_kernel<double>, dim3(blocks_for(E, 256)), dim3(256),
                         0, s, dt[b], d->blocks[b].sliver, E, factor,
                         static_cast<double*>(dt_elem[b]), d64);
    else
      hipLaunchKernelGGL(fsgdt::finish_local_kernel<float>, dim3(blocks_for(E, 256)), dim3(256),
                         0, s, dt[b], d->blocks[b].sliver, E, factor,
                         static_cast<float*>(dt_elem[b]), d64);
  }
  return dt_launched();
}

int fsg_conserved_totals(fsg_dt* d, const void* const* U, double* out, void* stream) {
  if (!d || !U || !out) return dt_fail(-10, "null argument");
  const fsg_context* c = d->ctx;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int V = c->params.nvars, P = c->params.degree;
  int64_t E = 0, mc = 0;
  for (const fsg_block_desc& b : c->blocks) E += b.count;
  for (const fsg_dt_conn& k : d->conns) mc = k.count > mc ? k.count : mc;
  double* part = d->work + elem_base(d) + E + 2 * mc;
  int64_t nparts = 0;
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    const fsg_block_desc& bd = c->blocks[b];
    if (!U[b]) return dt_fail(-10, "null state block");
    if (bd.count == 0) continue;
    const int N = nodes_of(bd.element_type, P);
    const unsigned g = blocks_for(bd.count, fsgdt::WG);
    if (c->params.dtype == FSG_F64)
      hipLaunchKernelGGL(fsgdt::totals_partial_kernel<double>, dim3(g), dim3(fsgdt::WG), 0, s,
                         static_cast<const double*>(U[b]), bd.count, N, V,
                         d->blocks[b].node_weights, d->blocks[b].detw, part + nparts * V);
    else
      hipLaunchKernelGGL(fsgdt::totals_partial_kernel<float>, dim3(g), dim3(fsgdt::WG), 0, s,
                         static_cast<const float*>(U[b]), bd.count, N, V,
                         d->blocks[b].node_weights, d->blocks[b].detw, part + nparts * V);
    nparts += g;
  }
  hipLaunchKernelGGL(fsgdt::totals_final_kernel, dim3(1), dim3(64), 0, s, part, nparts, V, out);
  return dt_launched();
}

int fsg_step_begin(fsg_dt* d, fsg_step_state* state, void* const* dt_elem,
                   double* const* dt_elem64, int32_t lts, const void* const* U, double cfl,
                   void* stream) {
  if (!d || !state || !dt_elem) return dt_fail(-10, "null argument");
  const fsg_context* c = d->ctx;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(fsgdt::step_begin_kernel, dim3(1), dim3(64), 0, s, state);
  if (lts) {
    // retry x local dt per element (state->factor is 0 once the run is over)
    if (!U) return dt_fail(-10, "null state for the local dt");
    int rc = fsg_local_dt(d, U, cfl, dt_elem, dt_elem64, &state->factor, stream);
    if (rc) return rc;
    unsigned long long* mn = reinterpret_cast<unsigned long long*>(d->work + 1);
    hipLaunchKernelGGL(fsgdt::set_u64_kernel, dim3(1), dim3(64), 0, s, mn,
                       (unsigned long long)0x7fefffffffffffffULL);
    for (size_t b = 0; b < c->blocks.size(); ++b) {
      const int64_t E = c->blocks[b].count;
      if (E == 0) continue;
      if (!dt_elem64 || !dt_elem64[b]) return dt_fail(-10, "LTS steps need the FP64 dt arrays");
      hipLaunchKernelGGL(fsgdt::min_dt_kernel, dim3(blocks_for(E, 256)), dim3(256), 0, s,
                         dt_elem64[b], E, mn);
    }
    hipLaunchKernelGGL(fsgdt::lts_dt_kernel, dim3(1), dim3(64), 0, s, state, mn);
    return dt_launched();
  }
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    const int64_t E = c->blocks[b].count;
    if (E == 0) continue;
    if (!dt_elem[b]) return dt_fail(-10, "null dt block");
    double* d64 = dt_elem64 ? dt_elem64[b] : nullptr;
    if (c->params.dtype == FSG_F64)
      hipLaunchKernelGGL(fsgdt::fill_dt_kernel<double>, dim3(blocks_for(E, 256)), dim3(256), 0,
                         s, state, E, static_cast<double*>(dt_elem[b]), d64);
    else
      hipLaunchKernelGGL(fsgdt::fill_dt_kernel<float>, dim3(blocks_for(E, 256)), dim3(256), 0,
                         s, state, E, static_cast<float*>(dt_elem[b]), d64);
  }
  return dt_launched();
}

int fsg_step_commit(fsg_dt* d, fsg_step_state* state, const fsg_stage_diag* diag,
                    double density_floor, double pressure_floor, const void* const* candidate,
                    void* const* states, void* stream) {
  if (!d || !state || !diag || !candidate || !states) return dt_fail(-10, "null argument");
  const fsg_context* c = d->ctx;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(fsgdt::step_decide_kernel, dim3(1), dim3(64), 0, s, state, diag,
                     density_floor, pressure_floor, -1.0);
  const int V = c->params.nvars, P = c->params.degree;
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    const fsg_block_desc& bd = c->blocks[b];
    if (bd.count == 0) continue;
    if (!candidate[b] || !states[b]) return dt_fail(-10, "null state block");
    const int64_t n = bd.count * nodes_of(bd.element_type, P) * V;
    const unsigned g = (unsigned)(n / 256 + 1 < 4096 ? n / 256 + 1 : 4096);
    if (c->params.dtype == FSG_F64)
      hipLaunchKernelGGL(fsgdt::commit_kernel<double>, dim3(g), dim3(256), 0, s, state,
                         static_cast<const double*>(candidate[b]), static_cast<double*>(states[b]),
                         n);
    else
      hipLaunchKernelGGL(fsgdt::commit_kernel<float>, dim3(g), dim3(256), 0, s, state,
                         static_cast<const float*>(candidate[b]), static_cast<float*>(states[b]),
                         n);
  }
  return dt_launched();
}

const char* fsg_dt_last_error(void) { return g_dt_error.c_str(); }

}  // extern "C"

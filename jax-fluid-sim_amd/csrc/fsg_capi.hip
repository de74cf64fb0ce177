// C ABI over the DG residual / RK-stage kernels (declared in
// include/fluid_gfx950.h).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/fluid_gfx950.h"
#include "fsg_context.h"
#include "fsg_dispatch.h"

namespace fsg {
template <int P, int V, typename real>
int launch_element(int, int64_t, const ElemArgs<real>&, const Phys<real>&, hipStream_t, int);
template <int P, int V, typename real>
int launch_face(int, int, int, int, int64_t, const FaceArgs<real>&, const Phys<real>&, hipStream_t);
}  // namespace fsg

namespace fsg {
__global__ void diag_reset_kernel(StageDiag* d) {
  if (threadIdx.x == 0) {
    d->rho_min = __longlong_as_double(0x7ff0000000000000LL);
    d->p_min = __longlong_as_double(0x7ff0000000000000LL);
    d->nonfinite = 0u;
    d->pad = 0u;
  }
}
}  // namespace fsg

static thread_local std::string g_error;

int fsg_fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}
#define fail fsg_fail



namespace {
// Records the start event of a timed launch; returns the pair index or -1.
long timing_begin(fsg_context* c, int kind, hipStream_t s, bool masked) {
  if (!c->timing || (c->timing == 2 && masked)) return -1;
  size_t i = c->used;
  if (2 * (i + 1) > c->pool.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1;
    c->pool.push_back(a);
    c->pool.push_back(b);
    c->kinds.push_back(kind);
  }
  c->kinds[i] = kind;
  // timing is best effort: a failed record only loses that sample (the
  // launch it brackets reports its own errors)
  (void)hipEventRecord(c->pool[2 * i], s);
  c->used = i + 1;
  return (long)i;
}
void timing_end(fsg_context* c, long i, hipStream_t s) {
  if (i >= 0) (void)hipEventRecord(c->pool[2 * i + 1], s);
}
}  // namespace

namespace {

using namespace fsg;

template <typename real>
Phys<real> make_phys(const fsg_params& p, const fsg_boundary* b) {
  Phys<real> ph;
  memset(&ph, 0, sizeof ph);
  ph.gamma = real(p.gamma);
  ph.gas_constant = real(p.gas_constant);
  ph.mu = real(p.dynamic_viscosity);
  ph.kappa = real(p.thermal_conductivity);
  ph.eta = real(p.br2_penalty);
  ph.rusanov = p.inviscid_flux == 1;
  ph.wale = p.use_wale != 0;
  ph.wale_constant = real(p.wale_constant);
  ph.cp_over_prt = real(p.gamma * p.gas_constant / (p.gamma - 1.0) / p.turbulent_prandtl);
  if (b) {
    ph.has_boundary = b->has_boundary;
    ph.has_wall_temperature = b->has_wall_temperature;
    ph.wall_temperature = real(b->wall_temperature);
    for (int k = 0; k < 16; ++k) ph.farfield[k] = real(b->farfield[k]);
  }
  return ph;
}

template <int P, int V, typename real>
int run_faces(fsg_context* c, const void* const* U, int terms, const int32_t* const* mask,
              const Phys<real>& ph, hipStream_t s) {
  real* fo = static_cast<real*>(c->face_out);
  for (const fsg_group_desc& g : c->groups) {
    const fsg_block_desc& bo = c->blocks[g.owner_block];
    FaceArgs<real> a;
    memset(&a, 0, sizeof a);
    a.Uo = static_cast<const real*>(U[g.owner_block]);
    a.owner_elem = g.owner_elem;
    a.geom = static_cast<const real*>(g.geom);
    a.owner_table = g.owner_table;
    a.owner_tables = static_cast<const real*>(g.owner_tables);
    a.owner_sup = g.owner_sup;
    a.fw = static_cast<const real*>(g.face_weights);
    a.bflags = g.bflags;
    // WALE needs the raw face gradient apart from the lift: dense gradients
    if (c->hex_sum_fact != 0 && g.face_sf1d && !c->params.use_wale) {
      a.owner_sf = g.owner_sf;
      a.neighbor_sf = g.neighbor_sf;
      a.face_sf1d = static_cast<const real*>(g.face_sf1d);
    }
    a.jinv_o = static_cast<const real*>(bo.jinv);
    a.scale_o = static_cast<const real*>(bo.scale);
    a.width_o = static_cast<const real*>(bo.width);
    a.out_o = fo + g.out_owner;
    if (mask) a.mask_o = mask[g.owner_block];
    int tn = -1;
    if (g.neighbor_block >= 0) {
      const fsg_block_desc& bn = c->blocks[g.neighbor_block];
      tn = bn.element_type;
      a.Un = static_cast<const real*>(U[g.neighbor_block]);
      a.neighbor_elem = g.neighbor_elem;
      a.neighbor_table = g.neighbor_table;
      a.neighbor_tables = static_cast<const real*>(g.neighbor_tables);
      a.neighbor_sup = g.neighbor_sup;
      a.jinv_n = static_cast<const real*>(bn.jinv);
      a.scale_n = static_cast<const real*>(bn.scale);
      a.width_n = static_cast<const real*>(bn.width);
      a.out_n = fo + g.out_neighbor;
      if (mask) a.mask_n = mask[g.neighbor_block];
    }
    a.count = g.count;
    a.terms = terms;
    long ev = timing_begin(c, int(&g - c->groups.data()), s, mask != nullptr);
    int rc = launch_face<P, V, real>(g.face_type, bo.element_type, tn, g.sub, g.count, a, ph, s);
    timing_end(c, ev, s);
    if (rc) return fail(rc, "face launch failed (type %d/%d/%d sub %d): %s", g.face_type,
                        bo.element_type, tn, g.sub, hipGetErrorString(hipGetLastError()));
  }
  return 0;
}

template <int P, int V, typename real>
int run_elements(fsg_context* c, const void* const* U, const void* const* base, void* const* out,
                 int mode, int terms, double dt, const void* const* dt_elem, double alpha,
                 double beta, fsg_stage_diag* diag, const int32_t* const* mask,
                 const Phys<real>& ph, hipStream_t s) {
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    const fsg_block_desc& d = c->blocks[b];
    ElemArgs<real> a;
    memset(&a, 0, sizeof a);
    a.U = static_cast<const real*>(U[b]);
    a.Ubase = base ? static_cast<const real*>(base[b]) : nullptr;
    a.out = static_cast<real*>(out[b]);
    a.Bt = static_cast<const real*>(d.Bt);
    a.Dt = static_cast<const real*>(d.Dt);
    a.D = static_cast<const real*>(d.D);
    a.wq = static_cast<const real*>(d.wq);
    a.minv = static_cast<const real*>(d.minv);
    a.minvT = static_cast<const real*>(d.minvT);
    // WALE runs the vector element kernel (raw gradient apart from the lift)
    const bool wale = c->params.use_wale != 0;
    a.sf1d = c->hex_sum_fact != 0 && !wale ? static_cast<const real*>(d.sf1d) : nullptr;
    a.width = static_cast<const real*>(d.width);
    a.jinv = static_cast<const real*>(d.jinv);
    a.scale = static_cast<const real*>(d.scale);
    a.detw = static_cast<const real*>(d.detw);
    a.rec_ptr = d.rec_ptr;
    a.rec_off = d.rec_off;
    a.rec_sup = d.rec_sup;
    a.rec_nf = d.rec_nf;
    a.support = c->support;
    a.face_out = static_cast<const real*>(c->face_out);
    a.dt_elem = dt_elem ? static_cast<const real*>(dt_elem[b]) : nullptr;
    a.dt = real(dt);
    a.alpha = real(alpha);
    a.beta = real(beta);
    a.mode = mode;
    a.terms = terms;
    a.diag = reinterpret_cast<StageDiag*>(diag);
    a.mask = mask ? mask[b] : nullptr;
    a.count = d.count;
    long ev = timing_begin(c, int(c->groups.size() + b), s, mask != nullptr);
    const int mfma = c->element_mfma > 0 && !wale ? 1 : 0;
    int rc = launch_element<P, V, real>(d.element_type, d.count, a, ph, s, mfma);
    timing_end(c, ev, s);
    if (rc) return fail(rc, "element launch failed (type %d): %s", d.element_type,
                        hipGetErrorString(hipGetLastError()));
  }
  return 0;
}

template <int P, int V, typename real>
int run(fsg_context* c, const void* const* U, const void* const* base, void* const* out, int mode,
        int terms, double dt, const void* const* dt_elem, double alpha, double beta,
        const fsg_boundary* bnd, fsg_stage_diag* diag, const int32_t* const* mask,
        hipStream_t s) {
  Phys<real> ph = make_phys<real>(c->params, bnd);
  int rc = run_faces<P, V, real>(c, U, terms, mask, ph, s);
  if (rc) return rc;
  return run_elements<P, V, real>(c, U, base, out, mode, terms, dt, dt_elem, alpha, beta, diag,
                                  mask, ph, s);
}

int dispatch(fsg_context* c, const void* const* U, const void* const* base, void* const* out,
             int mode, int terms, double dt, const void* const* dt_elem, double alpha,
             double beta, const fsg_boundary* bnd, fsg_stage_diag* diag,
             const int32_t* const* mask, hipStream_t s) {
  const fsg_params& p = c->params;
  const bool f64 = p.dtype == FSG_F64;
#define FSG_RUN(P, V)                                                                          \
  if (p.degree == P && p.nvars == V)                                                          \
    return f64 ? run<P, V, double>(c, U, base, out, mode, terms, dt, dt_elem, alpha, beta, bnd, \
                                   diag, mask, s)                                              \
               : run<P, V, float>(c, U, base, out, mode, terms, dt, dt_elem, alpha, beta, bnd,  \
                                  diag, mask, s);
  FSG_RUN(1, 5)
  FSG_RUN(2, 5)
  FSG_RUN(3, 5)
  FSG_RUN(2, 16)
#undef FSG_RUN
  return fail(-4, "unsupported degree %d / nvars %d", p.degree, p.nvars);
}

}  // namespace

extern "C" {

int fsg_create(const fsg_params* params, int32_t nblocks, const fsg_block_desc* blocks,
               int32_t ngroups, const fsg_group_desc* groups, const int32_t* support,
               void* face_out, fsg_context** out) {
  if (!params || !out) return fail(-10, "null argument");
  *out = nullptr;
  const fsg_params& p = *params;
  if (p.degree < 1 || p.degree > 3) return fail(-11, "degree %d not in 1..3", p.degree);
  if (!(p.nvars == 5 || (p.nvars == 16 && p.degree == 2)))
    return fail(-12, "nvars %d unsupported at degree %d (5, or 16 at degree 2)", p.nvars, p.degree);
  if (p.dtype != FSG_F64 && p.dtype != FSG_F32) return fail(-13, "dtype %d", p.dtype);
  if (nblocks < 1 || nblocks > 3) return fail(-14, "nblocks %d", nblocks);
  for (int b = 0; b < nblocks; ++b)
    if (blocks[b].element_type < 0 || blocks[b].element_type > 2 || blocks[b].count < 0)
      return fail(-15, "block %d invalid", b);
  if (p.use_wale != 0 && p.use_wale != 1) return fail(-18, "use_wale %d not 0 or 1", p.use_wale);
  if (p.use_wale) {
    if (!(p.wale_constant >= 0.0) || !(p.turbulent_prandtl > 0.0))
      return fail(-18, "WALE constant %g / turbulent Prandtl number %g", p.wale_constant,
                  p.turbulent_prandtl);
    for (int b = 0; b < nblocks; ++b)
      if (!blocks[b].width && blocks[b].count > 0)
        return fail(-18, "WALE needs the filter widths of block %d", b);
  }
  for (int g = 0; g < ngroups; ++g) {
    const fsg_group_desc& d = groups[g];
    if (d.owner_block < 0 || d.owner_block >= nblocks || d.neighbor_block >= nblocks ||
        d.face_type < 0 || d.face_type > 1 || d.sub < 0 || d.sub > 1 || d.count < 0)
      return fail(-16, "group %d invalid", g);
  }
  fsg_context* c = new fsg_context;
  c->params = p;
  c->blocks.assign(blocks, blocks + nblocks);
  c->groups.assign(groups, groups + ngroups);
  c->support = support;
  c->face_out = face_out;
  *out = c;
  return 0;
}

int fsg_residual(fsg_context* ctx, const void* const* U, void* const* out, int32_t mode,
                 int32_t terms, const fsg_boundary* boundary, void* stream) {
  if (!ctx || !U || !out) return fail(-10, "null argument");
  if (mode != FSG_MODE_RESIDUAL && mode != FSG_MODE_RHS) return fail(-17, "mode %d", mode);
  return dispatch(ctx, U, nullptr, out, mode, terms, 0.0, nullptr, 0.0, 0.0, boundary, nullptr,
                  nullptr, static_cast<hipStream_t>(stream));
}

int fsg_residual_masked(fsg_context* ctx, const void* const* U, void* const* out, int32_t mode,
                        int32_t terms, const int32_t* const* mask, const fsg_boundary* boundary,
                        void* stream) {
  if (!ctx || !U || !out || !mask) return fail(-10, "null argument");
  if (mode != FSG_MODE_RESIDUAL && mode != FSG_MODE_RHS) return fail(-17, "mode %d", mode);
  for (size_t b = 0; b < ctx->blocks.size(); ++b)
    if (!mask[b] && ctx->blocks[b].count > 0) return fail(-10, "null mask of block %d", (int)b);
  return dispatch(ctx, U, nullptr, out, mode, terms, 0.0, nullptr, 0.0, 0.0, boundary, nullptr,
                  mask, static_cast<hipStream_t>(stream));
}

int fsg_rk_stage(fsg_context* ctx, const void* const* base, const void* const* curr,
                 void* const* next, double dt, const void* const* dt_elem, double alpha,
                 double beta, int32_t terms, const fsg_boundary* boundary, fsg_stage_diag* diag,
                 void* stream) {
  // base may be null (next = beta (curr + dt rhs), no base term read) and
  // diag may be null (no admissibility reduction)
  if (!ctx || !curr || !next) return fail(-10, "null argument");
  return dispatch(ctx, curr, base, next, 2, terms, dt, dt_elem, alpha, beta, boundary, diag,
                  nullptr, static_cast<hipStream_t>(stream));
}

int fsg_diag_reset(fsg_stage_diag* diag, void* stream) {
  if (!diag) return fail(-10, "null argument");
  hipLaunchKernelGGL(fsg::diag_reset_kernel, dim3(1), dim3(64), 0,
                     static_cast<hipStream_t>(stream), reinterpret_cast<fsg::StageDiag*>(diag));
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fail(-1, "diag reset: %s", hipGetErrorString(e));
}

int fsg_set_option(fsg_context* ctx, int32_t option, int32_t value) {
  if (!ctx) return fail(-10, "null argument");
  if (option == FSG_OPT_ELEMENT_MFMA) {
    ctx->element_mfma = value < 0 ? -1 : (value != 0);
    return 0;
  }
  if (option == FSG_OPT_HEX_SUM_FACTORISATION) {
    ctx->hex_sum_fact = value < 0 ? -1 : (value != 0);
    return 0;
  }
  return fail(-13, "unknown option %d", option);
}

int fsg_timing(fsg_context* ctx, int32_t enable) {
  if (!ctx) return fail(-10, "null argument");
  ctx->timing = enable == 2 ? 2 : (enable != 0 ? 1 : 0);
  ctx->used = 0;
  return 0;
}

int fsg_timing_read(fsg_context* ctx, double* group_ms, double* block_ms, int64_t* group_launches,
                    int64_t* block_launches) {
  if (!ctx) return fail(-10, "null argument");
  const size_t ng = ctx->groups.size(), nb = ctx->blocks.size();
  for (size_t k = 0; k < ng; ++k) {
    if (group_ms) group_ms[k] = 0.0;
    if (group_launches) group_launches[k] = 0;
  }
  for (size_t k = 0; k < nb; ++k) {
    if (block_ms) block_ms[k] = 0.0;
    if (block_launches) block_launches[k] = 0;
  }
  for (size_t i = 0; i < ctx->used; ++i) {
    if (hipEventSynchronize(ctx->pool[2 * i + 1]) != hipSuccess)
      return fail(-1, "event synchronize failed");
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->pool[2 * i], ctx->pool[2 * i + 1]) != hipSuccess)
      return fail(-1, "event elapsed time failed");
    int k = ctx->kinds[i];
    if (k < (int)ng) {
      if (group_ms) group_ms[k] += ms;
      if (group_launches) group_launches[k] += 1;
    } else {
      if (block_ms) block_ms[k - ng] += ms;
      if (block_launches) block_launches[k - ng] += 1;
    }
  }
  ctx->used = 0;
  return 0;
}

void fsg_destroy(fsg_context* ctx) {
  if (!ctx) return;
  for (hipEvent_t e : ctx->pool) (void)hipEventDestroy(e);
  delete ctx;
}

const char* fsg_last_error(void) { return g_error.c_str(); }

int32_t fsg_abi_version(void) { return FSG_ABI_VERSION; }

}  // extern "C"

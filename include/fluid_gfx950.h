/* C ABI of the MI355X-native hybrid DG hot path (libfluid_gfx950.so).
 *
 * All array arguments are device pointers owned by the caller (torch tensors
 * on ROCm today; XLA buffers when the entry points are registered as
 * jax.ffi custom calls).  A context keeps raw pointers to the caller's
 * immutable tables and never allocates state memory.  No entry point
 * synchronises the stream; diagnostics are written to device memory.
 * Every entry point returns 0 on success or a negative status; the text of
 * the last error is available from fsg_last_error().
 *
 * Precision: dtype FSG_F64 runs every kernel in double (the plain path);
 * FSG_F32 runs the same kernels in float (the FP32 path).  The state and
 * table buffers must be of the context's dtype.
 */
#ifndef FLUID_GFX950_H
#define FLUID_GFX950_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSG_F64 0
#define FSG_F32 1

#define FSG_MODE_RESIDUAL 0 /* weak-form residual before M^-1 */
#define FSG_MODE_RHS 1      /* M^-1 residual */

#define FSG_TERM_EULER 1
#define FSG_TERM_VISCOUS 2

/* Boundary flags per boundary face (the tag classes of
 * hybrid/boundary.py:22-46). */
#define FSG_BND_WALL 1
#define FSG_BND_INLET 2
#define FSG_BND_NOSLIP 4
#define FSG_BND_SLIP 8

typedef struct fsg_context fsg_context;

/* Physics of HybridCompressibleSolver (hybrid/solver.py:20-53) and the BR2
 * penalty of HybridNavierStokesOperator (hybrid/navier_stokes.py:10-16).
 * use_wale adds the WALE eddy viscosity of hybrid/solver.py:317-406 to mu
 * (and mu_t c_p / Pr_t to kappa) in the viscous term, evaluated from the
 * resolved polynomial gradient with each block's filter widths; it needs
 * fsg_block_desc.width on every block and runs the vector element kernel
 * and the dense face gradients on every block. */
typedef struct {
  double gamma;
  double gas_constant;
  double dynamic_viscosity;
  double thermal_conductivity;
  double br2_penalty;
  int32_t inviscid_flux; /* 0 = hllc, 1 = rusanov */
  int32_t nvars;         /* 5 + number of species (5 or 16) */
  int32_t degree;        /* 1, 2 or 3 */
  int32_t dtype;         /* FSG_F64 or FSG_F32 */
  int32_t use_wale;      /* 0 or 1 */
  int32_t reserved;
  double wale_constant;     /* C_w (0.5 in the reference) */
  double turbulent_prandtl; /* Pr_t (0.9 in the reference) */
} fsg_params;

/* One element block (hybrid/mesh.py:122-139 + hybrid/geometry.py:27-36 for
 * affine cells).  Reference tables are shared by the block; per-element
 * metrics are J^-1 (9 per element, [r*3+d] = d xi_r / d x_d), the
 * sliver-scaled inverse-mass factor min(1, V/1e-4)/|det J| and |det J|.  The
 * CSR (rec_*) lists each element's face-side records in the face-output
 * buffer. */
typedef struct {
  int32_t element_type; /* 0 hexahedron, 1 wedge, 2 tetra */
  int32_t reserved;
  int64_t count;
  const void* Bt;    /* (N, Q) */
  const void* Dt;    /* (N, 3, Q) */
  const void* D;     /* (Q, 3, N) */
  const void* wq;    /* (Q) */
  const void* minv;  /* (N, N) reference inverse mass */
  const void* jinv;  /* (E, 9) */
  const void* scale; /* (E) */
  const void* detw;  /* (E) */
  const int64_t* rec_ptr; /* (E + 1) */
  const int64_t* rec_off; /* (R) offsets in reals into face_out */
  const int32_t* rec_sup; /* (R) offsets into the support-index array */
  const int32_t* rec_nf;  /* (R) support size */
  const void* minvT; /* (N, N) transposed reference inverse mass */
  const void* sf1d;  /* hexahedra at p >= 2: 1-D factors B1 (Q1, N1), D1 (Q1, N1),
                        M1^-1 (N1, N1) of the volume operators, or null */
  const void* width; /* (E) WALE filter width V^(1/3) / (p + 1)
                        (hybrid/navier_stokes.py:17-25); may be null
                        without use_wale */
} fsg_block_desc;

/* One face group: (face type, owner block, neighbour block) as in
 * hybrid/spatial.py:20-39, further split by sub-faces (a quad face of the
 * owner cut into two triangles).  Trace tables hold, per distinct set of
 * reference points: face-support basis (Qf, NF), lift operator
 * (Phi M_ref^-1 restricted to the support, (Qf, NF)) and reference
 * gradients (Qf, N, 3). */
typedef struct {
  int32_t face_type;      /* 0 quad, 1 triangle */
  int32_t owner_block;
  int32_t neighbor_block; /* -1: boundary */
  int32_t sub;
  int64_t count;
  const int32_t* owner_elem;
  const int32_t* neighbor_elem;
  const void* geom; /* (F, 4): unit normal owner->neighbour, area */
  const int32_t* owner_table;
  const int32_t* neighbor_table;
  const void* owner_tables;
  const void* neighbor_tables;
  const int32_t* owner_sup;
  const int32_t* neighbor_sup;
  const void* face_weights; /* (Qf) normalized face quadrature weights */
  const int32_t* bflags;    /* (F) FSG_BND_* flags (boundary groups) */
  int64_t out_owner;        /* offset (reals) of the owner records in face_out */
  int64_t out_neighbor;
  /* Hexahedral quad sides at p >= 2 (or null): per trace table 4 + 2 Qf
   * int32 [normal axis, face plane 0/1, tangential axes t1 < t2, the point's
   * index along t1 (Qf), along t2 (Qf)], and the 1-D factors [B1 (Qf1, N1),
   * D1 (Qf1, N1), l(0), l'(0), l(1), l'(1) (N1 each)] at the Qf1 face Gauss
   * points, for the BR2 face gradients by sum factorisation. */
  const int32_t* owner_sf;
  const int32_t* neighbor_sf;
  const void* face_sf1d;
} fsg_group_desc;

/* Boundary object (FarfieldWallBoundary, hybrid/boundary.py:6-102).
 * has_boundary = 0 reproduces boundary_state=None (U_ghost = U_inside). */
typedef struct {
  int32_t has_boundary;
  int32_t has_wall_temperature;
  double wall_temperature;
  double farfield[16]; /* primitive [rho, u, v, w, p, Y...] */
} fsg_boundary;

/* Admissibility diagnostics written by a stage (device memory), as
 * returned by hybrid/time_integration.py:291-312: valid == (nonfinite == 0
 * && rho_min >= floor && p_min >= floor). */
typedef struct {
  double rho_min;
  double p_min;
  uint32_t nonfinite;
  uint32_t reserved;
} fsg_stage_diag;

/* Builds a context over caller-owned tables.  face_out is a device scratch
 * buffer of the context dtype sized by the caller (sum over groups of
 * count * (NF_owner + NF_neighbor) * (nvars + 12) reals).
 * Replaces the set-up of HybridEulerOperator / HybridNavierStokesOperator
 * (hybrid/spatial.py:11-18, hybrid/navier_stokes.py:10-25). */
int fsg_create(const fsg_params* params, int32_t nblocks, const fsg_block_desc* blocks,
               int32_t ngroups, const fsg_group_desc* groups, const int32_t* support,
               void* face_out, fsg_context** out);

/* Residual or rhs of the listed terms for every block.
 * mode FSG_MODE_RESIDUAL, terms EULER     = HybridEulerOperator.residual (hybrid/spatial.py:137)
 * mode FSG_MODE_RHS,      terms EULER     = inviscid_rhs (hybrid/navier_stokes.py:345)
 * mode FSG_MODE_RESIDUAL, terms VISCOUS   = viscous_residual (hybrid/navier_stokes.py:210)
 * mode FSG_MODE_RHS,      terms VISCOUS   = viscous_rhs (hybrid/navier_stokes.py:326)
 * mode FSG_MODE_RHS,      terms both      = rhs (hybrid/navier_stokes.py:338)
 * U[b] and out[b] are (E_b, N, nvars) arrays of block b. */
int fsg_residual(fsg_context* ctx, const void* const* U, void* const* out, int32_t mode,
                 int32_t terms, const fsg_boundary* boundary, void* stream);

/* fsg_residual restricted to marked elements: mask[b] is an (E_b) int32
 * array per block; the faces with a marked side and the marked elements are
 * evaluated, and out[b] holds the result on the elements with mask != 0
 * (equal to fsg_residual there: bitwise in FP64, to rounding in FP32, where
 * the restricted kernels may contract multiply-adds differently; other
 * elements' values are unspecified).  Serves the MOOD stage, whose viscous rhs is read only on
 * the troubled elements (hybrid/subcell.py:876-886). */
int fsg_residual_masked(fsg_context* ctx, const void* const* U, void* const* out, int32_t mode,
                        int32_t terms, const int32_t* const* mask, const fsg_boundary* boundary,
                        void* stream);

/* One SSP-RK stage fused into the residual epilogue
 * (hybrid/time_integration.py:404-439 without a limiter):
 *   next = alpha * base + beta * (curr + dt * rhs(curr)),
 * followed by the node admissibility reductions into diag (device).  dt is
 * a scalar unless dt_elem is non-null (one dt per element, the LTS dict).
 * next may alias curr or base.  base may be null (no base term: next =
 * beta * (curr + dt * rhs), the Euler candidate of a MOOD stage with
 * alpha = 0, beta = 1) and diag may be null (no admissibility reductions). */
int fsg_rk_stage(fsg_context* ctx, const void* const* base, const void* const* curr,
                 void* const* next, double dt, const void* const* dt_elem, double alpha,
                 double beta, int32_t terms, const fsg_boundary* boundary,
                 fsg_stage_diag* diag, void* stream);

/* Sets diag to (+inf, +inf, 0) on the stream (a fill kernel, capture-safe). */
int fsg_diag_reset(fsg_stage_diag* diag, void* stream);

/* Per-launch timing with HIP events recorded on the launch stream around each
 * face-group and element-block kernel (no counterpart in the reference; used
 * by the benchmark to report a kernel's own duration).  fsg_timing(ctx, 1)
 * starts a fresh recording of every launch, fsg_timing(ctx, 2) of the
 * launches of full (unmasked) residual passes only, fsg_timing(ctx, 0) stops
 * recording; fsg_timing_read waits for the recorded events,
 * writes the summed milliseconds and launch counts per group / block (arrays
 * of ngroups / nblocks entries, any may be null) and clears the record. */
int fsg_timing(fsg_context* ctx, int32_t enable);
int fsg_timing_read(fsg_context* ctx, double* group_ms, double* block_ms,
                    int64_t* group_launches, int64_t* block_launches);

/* Context options (no counterpart in the reference).
 * FSG_OPT_ELEMENT_MFMA: 1 runs the element operators as batched matrix-core
 * products, 0 or -1 (default) as one workgroup per element on the vector
 * units (the faster of the two on MI355X for the p = 3 hexahedron in both
 * precisions); both compute the same discrete operator. */
#define FSG_OPT_ELEMENT_MFMA 1
/* FSG_OPT_HEX_SUM_FACTORISATION: 1 or -1 (default) evaluates the hexahedral
 * volume operators at p >= 2 by sum factorisation (their 1-D factors, when
 * the block carries sf1d), 0 by the dense operators; same discrete operator,
 * round-off apart. */
#define FSG_OPT_HEX_SUM_FACTORISATION 2
int fsg_set_option(fsg_context* ctx, int32_t option, int32_t value);

void fsg_destroy(fsg_context* ctx);
const char* fsg_last_error(void);

/* C = A B for 16x16 row-major matrices through the 16x16x4 matrix-core tiles
 * the element kernel uses (no counterpart in the reference: a device check
 * of the operand / accumulator lane layout). */
int fsg_mfma_selftest(int32_t dtype, const void* A, const void* B, void* C, void* stream);
/* Version of the structure layouts and entry points of this header
 * (FSG_ABI_VERSION); a caller built against another version must not bind. */
#define FSG_ABI_VERSION 2
int32_t fsg_abi_version(void);

/* ---- XLA custom-call target -------------------------------------------------
 * fsg_xla_rhs is the untyped GPU custom call
 *   void f(hipStream_t stream, void** buffers, const char* opaque, size_t len)
 * for jax.ffi.register_ffi_target(name, capsule, platform="ROCM", api_version=0),
 * replacing HybridNavierStokesOperator.rhs / residual at the reference's call
 * sites (run_hybrid_sphere.py:200-216; hybrid/navier_stokes.py:338).  opaque:
 * "ctx=<hex>;bnd=<hex or 0>;mode=<0|1>;terms=<1..3>;nblocks=<n>"; buffers =
 * [U_0 .. U_{n-1}, out_0 .. out_{n-1}].  Errors go to fsg_last_error() (the
 * signature has no status).  fsg_xla_parse decodes an opaque string. */
typedef struct {
  fsg_context* ctx;
  const fsg_boundary* boundary;
  int32_t mode, terms, nblocks, reserved;
} fsg_xla_rhs_config;
int fsg_xla_parse(const char* opaque, size_t len, fsg_xla_rhs_config* cfg);
void fsg_xla_rhs(void* stream, void** buffers, const char* opaque, size_t len);

/* ---- Time-step control ------------------------------------------------------
 * The CFL estimate HybridSSPRK3.stable_dt (hybrid/time_integration.py:93-169,
 * without WALE), the element-local dt local_dt (:171-262) and the
 * accept / reject loop of run_hybrid_sphere.py:265-346 on the device, plus
 * the FP64 conservation sums (sum over elements of int U dV).  A dt context
 * extends a residual context (same blocks). */
typedef struct fsg_dt fsg_dt;

/* Per element block (device arrays, FP64): */
typedef struct {
  const double* length;       /* (E) max(2V / sum of face areas, V^(1/3) / (4 (p+1))) */
  const double* sliver;       /* (E) min(1, V / 1e-4) */
  const double* detw;         /* (E) |det J| */
  const double* node_weights; /* (N) integral of each basis function over the reference element */
} fsg_dt_block;

/* One group of interior faces of the local-dt smoothing, in the order of
 * hybrid/time_integration.py:27-47 (face type, owner block, neighbour
 * block). */
typedef struct {
  int32_t owner_block;
  int32_t neighbor_block;
  int64_t count;
  const int32_t* owner_cells;
  const int32_t* neighbor_cells;
} fsg_dt_conn;

/* Step-loop state in device memory (the caller initialises it): dt =
 * min(retry * cfl_dt, target_time - time); an accepted step advances time
 * and step and grows retry by 1.25 up to 1; a rejected one halves retry;
 * maximum_rejections consecutive rejections set status = 1, after which
 * (as after target_time or maximum_steps) steps do nothing.  cfl_dt is
 * refreshed by fsg_stable_dt when step % refresh_interval == 0. */
typedef struct {
  double cfl_dt, retry, time, target_time, dt, factor, min_dt, last_rho_min, last_p_min;
  int64_t step, maximum_steps, attempts, rejections;
  int32_t consecutive, maximum_rejections, status, active, accept, have_dt, refresh_interval,
      reserved;
} fsg_step_state;

/* work: device FP64 scratch of fsg_dt_work_size() entries, owned by the caller. */
int64_t fsg_dt_work_size(const fsg_context* ctx, int32_t nconn, const fsg_dt_conn* conns);
int fsg_dt_create(fsg_context* ctx, const fsg_dt_block* blocks, int32_t nconn,
                  const fsg_dt_conn* conns, double* work, int64_t work_len, fsg_dt** out);
void fsg_dt_destroy(fsg_dt* d);

/* HybridSSPRK3.stable_dt (hybrid/time_integration.py:93-169): the clipped
 * CFL dt into dt_out (device double, may be null), and into state->cfl_dt
 * when state is given and its refresh is due. */
int fsg_stable_dt(fsg_dt* d, const void* const* U, double cfl, double* dt_out,
                  fsg_step_state* state, void* stream);

/* HybridSSPRK3.local_dt (hybrid/time_integration.py:171-262): per-element
 * dt into dt_elem[b] (context dtype, (E_b)) and dt_elem64[b] (FP64, may be
 * null), multiplied by *factor (device double, may be null). */
int fsg_local_dt(fsg_dt* d, const void* const* U, double cfl, void* const* dt_elem,
                 double* const* dt_elem64, const double* factor, void* stream);

/* sum over all elements of int_K U dV per variable, in FP64 (out: nvars
 * device doubles), reduced in a fixed order (deterministic). */
int fsg_conserved_totals(fsg_dt* d, const void* const* U, double* out, void* stream);

/* Step loop on the device (run_hybrid_sphere.py:265-346): begin fills the
 * per-element dt of the attempt (context dtype, and FP64 for the chemistry)
 * from the state, or the retry-scaled local dt when lts != 0; the caller
 * then runs the stages (fsg_rk_stage / fsg_mood_stage_fused with those dt
 * arrays) and the chemistry on the candidate; commit reads the stage
 * diagnostics, updates the state and copies the candidate into states when
 * the step is accepted. */
int fsg_step_begin(fsg_dt* d, fsg_step_state* state, void* const* dt_elem,
                   double* const* dt_elem64, int32_t lts, const void* const* U, double cfl,
                   void* stream);
int fsg_step_commit(fsg_dt* d, fsg_step_state* state, const fsg_stage_diag* diag,
                    double density_floor, double pressure_floor, const void* const* candidate,
                    void* const* states, void* stream);
const char* fsg_dt_last_error(void);

/* ---- MOOD: troubled-cell detection, subcell FV fallback, positivity --------
 * HybridMOODSubcellLimiter (hybrid/limiter.py:267-343) with its
 * HybridTroubledCellSensor (:150-185), HybridPositivityLimiter (:11-147) and
 * HybridSubcellFV (hybrid/subcell.py:442-950), for affine elements.  A MOOD
 * context extends a residual context (same blocks, face groups, dtype). */
typedef struct fsg_mood fsg_mood;

/* Element-independent tables of one element type (all reals in the context
 * dtype; built by hybrid/mood.py). */
typedef struct {
  int32_t N;   /* nodes */
  int32_t Q;   /* volume points */
  int32_t NS;  /* sample points: the Q volume points, then the face sample points */
  int32_t NI;  /* internal subcell faces */
  int32_t NB;  /* boundary subfaces */
  int32_t reserved;
  const void* p1proj;       /* (N, N) P1 projector at the nodes, transposed (sensor) */
  const void* samples;      /* (N, NS) basis at the sample points, point fastest */
  const void* wq;           /* (Q) reference volume weights */
  const void* P;            /* (64, N) subcell averages of the basis */
  const void* PT;           /* (N, 64) the same, subcell fastest */
  const void* RT;           /* (64, N) least-squares reconstruction R (N, 64), transposed */
  const void* subvol;       /* (64) reference subcell volumes */
  const int32_t* int_o;     /* (NI) owner subcell */
  const int32_t* int_n;     /* (NI) neighbour subcell */
  const void* int_area;     /* (NI, 3) reference area vector, owner -> neighbour */
  const void* int_dc;       /* (NI, 3) reference centre offset neighbour - owner */
  const int32_t* int_ptr;   /* (65) per-subcell list of internal faces */
  const int32_t* int_list;  /* (2 NI) */
  const int32_t* bnd_face;  /* (NB) parent local face of each boundary subface */
  const int32_t* bnd_gather;/* (NB) subcell read as the subface's FV trace */
  const int32_t* tgt_ptr;   /* (65) per-subcell list of boundary subfaces it receives */
  const int32_t* tgt_list;
} fsg_mood_type;

/* Per element block: each element's face entries, max_faces pairs of int32
 * (group | local_face << 16 | side << 24, face index); group -1 = unused. */
typedef struct {
  int32_t max_faces;
  int32_t reserved;
  const int32_t* elem_faces; /* (E, max_faces, 2) */
  int32_t* flags;            /* (E) troubled flags, written by the MOOD calls */
  void* sensor;              /* (E) sensor values written by detection (may be null) */
} fsg_mood_block;

/* Per face group: K subface pairs per face (hybrid/subcell.py:473-533),
 * de-duplicated into tables: pair_ids (tables, K, 2) = owner / neighbour
 * boundary-subface ids, pair_geom (tables, K, 4) = side giving the normal
 * (0 owner, 1 neighbour) and its reference area vector oriented owner ->
 * neighbour, basis_o / basis_n (tables, K, N) = basis at the subface
 * centres on each side. */
typedef struct {
  int32_t K;
  int32_t reserved;
  const int32_t* pair_table; /* (F) */
  const int32_t* pair_ids;
  const void* pair_geom;
  const void* basis_o;
  const void* basis_n;       /* null on boundary groups */
} fsg_mood_group;

typedef struct {
  double threshold;        /* sensor threshold (hybrid/limiter.py:150) */
  double density_floor;    /* positivity floors (hybrid/limiter.py:11-26) */
  double pressure_floor;
  double relative_floor;
  int32_t reconstruction;  /* 0 constant, 1 tvd, 2 weno (hybrid/subcell.py:541-665) */
  int32_t sensor_on;       /* degree >= 2 */
} fsg_mood_params;

/* types[t] for t = 0 hexahedron, 1 wedge, 2 tetra (entries of absent types
 * are ignored); blocks / groups in the order of the residual context.
 * counter (1 int32) and list (sum of block sizes int32) are device scratch. */
int fsg_mood_create(fsg_context* ctx, const fsg_mood_params* params, const fsg_mood_type* types,
                    const fsg_mood_block* blocks, const fsg_mood_group* groups, int32_t* counter,
                    int32_t* list, fsg_mood** out);
void fsg_mood_destroy(fsg_mood* m);

/* HybridMOODSubcellLimiter.troubled (hybrid/limiter.py:285-300): sensor or
 * PAD failure of U, into the blocks' flags. */
int fsg_mood_troubled(fsg_mood* m, const void* const* U, void* stream);

/* HybridPositivityLimiter.__call__ (hybrid/limiter.py:141-147): out may
 * alias U. */
int fsg_mood_positivity(fsg_mood* m, const void* const* U, void* const* out, void* stream);

/* HybridModalShockLimiter.__call__ (hybrid/limiter.py:186-264): on the
 * elements the troubled-cell sensor flags (degree >= 2), the P1 projection
 * with the element mean restored, or the mean where that projection's
 * density or pressure at the samples drops below p0_factor x the floors or
 * p0_relative x the mean's (or is non-finite); then the positivity limiter.
 * out may alias U.  diag (may be null) receives the admissibility of the
 * result (accumulated: reset it first with fsg_diag_reset). */
int fsg_mood_modal(fsg_mood* m, const void* const* U, void* const* out, double p0_factor,
                   double p0_relative, fsg_stage_diag* diag, void* stream);

/* HybridMOODSubcellLimiter.recompute (hybrid/limiter.py:302-338): detection
 * on `detect` (or on dg_candidate when null) unless use_flags != 0 (then
 * the blocks' flags are taken as given), the subcell FV recompute of the
 * troubled elements from `states` with the neighbours' mean correction, then
 * positivity.  visc (may be null) is the additional rhs; out may alias
 * dg_candidate. */
int fsg_mood_recompute(fsg_mood* m, const void* const* states, const void* const* dg_candidate,
                       const void* const* detect, const void* const* visc, double dt,
                       const void* const* dt_elem, int32_t use_flags,
                       const fsg_boundary* boundary, void* const* out, void* stream);

/* The MOOD branch of HybridSSPRK3.rk_stage (hybrid/time_integration.py:
 * 404-439) after the residuals: euler = curr + dt (inv + visc), detection on
 * alpha base + beta euler, recompute, positivity, blend with base,
 * positivity, admissibility into diag.  work holds euler (may alias next). */
int fsg_mood_stage(fsg_mood* m, const void* const* base, const void* const* curr,
                   const void* const* inv, const void* const* visc, void* const* work,
                   void* const* next, double dt, const void* const* dt_elem, double alpha,
                   double beta, const fsg_boundary* boundary, fsg_stage_diag* diag,
                   void* stream);

/* fsg_mood_stage from one residual pass: total = rhs(curr) (inviscid and
 * viscous terms together, fsg_residual mode RHS) replaces inv + visc; after
 * detection the viscous rhs of curr is evaluated on the troubled elements
 * only (fsg_residual_masked, into visc, which must hold (E_b, N, nvars)
 * arrays; null when the viscosity is zero) and feeds their subcell
 * recompute.  Same stage as hybrid/time_integration.py:404-439 up to the
 * rounding of inv + visc.  total may be null: work then already holds the
 * Euler candidate curr + dt * rhs(curr) (fsg_rk_stage with a null base,
 * alpha 0, beta 1, the same dt), which detection reads instead of total and
 * curr. */
int fsg_mood_stage_fused(fsg_mood* m, const void* const* base, const void* const* curr,
                         const void* const* total, void* const* visc, void* const* work,
                         void* const* next, double dt, const void* const* dt_elem, double alpha,
                         double beta, const fsg_boundary* boundary, fsg_stage_diag* diag,
                         void* stream);

/* Number of troubled elements found by the last detection (device counter,
 * read synchronously on the stream). */
int fsg_mood_count(fsg_mood* m, int64_t* count, void* stream);

/* ---- Park 11-species air chemistry (FP64) ---------------------------------
 * table: device buffer of fsg_chem_table_size() doubles holding the
 * mechanism (molar masses, NASA-7 fits, Arrhenius and temperature modes,
 * nu', nu'', third-body efficiencies, derived per-reaction and per-species
 * values, the stoichiometry fingerprint; layout in fsg_chem.hip), built from
 * Air11ParkMechanism (hybrid/chemistry/air11_park.py:20-318). */
typedef struct {
  int32_t newton_iters;  /* 12 in ReactingFlowCoupler (hybrid/chemistry/coupling.py:30) */
  int32_t dimensionless; /* non-dimensional flow scales (coupling.py:55-67, 135-154) */
  double gamma;
  double T_ref;
  double rho_ref;
  double e_ref;
  double t_ref;
} fsg_chem_params;

int32_t fsg_chem_table_size(void);
/* The integrator entry points below use the Park-11 stoichiometry and
 * third-body data compiled into the library (fsg_park11.h) when the table's
 * last entry (the stoichiometry fingerprint) matches them, and read them
 * from the table otherwise (same results). */

/* PointImplicitIntegrator.step over n cells
 * (hybrid/chemistry/integrator.py:31-104): rho_in/rho_out (n, 11), T_in (n),
 * dt per cell (dt_cell, may be null) or the scalar dt; T_out (n) may be null. */
int fsg_chem_integrate(const double* table, const double* rho_in, const double* T_in,
                       const double* dt_cell, double dt, int32_t newton_iters,
                       int32_t constant_temperature, double* rho_out, double* T_out,
                       int64_t n, void* stream);

/* ReactingFlowCoupler.react_block (hybrid/chemistry/coupling.py:116-197) in
 * place on U (ncells, 16); dt_elem (ncells / nodes_per_elem) may be null. */
int fsg_chem_react(const double* table, double* U, int64_t ncells, int64_t nodes_per_elem,
                   const double* dt_elem, double dt, const fsg_chem_params* params,
                   void* stream);

/* fsg_chem_integrate on the cells with flag[c] != 0 only: rho_out and T_out
 * of the other cells are left as they are (the surrogate's predictions).
 * list (n int32) and count (1 int32) are caller scratch; *count ends as the
 * number of flagged cells.  No host synchronisation.  Same per-cell result as
 * fsg_chem_integrate (integrator.py:9-104). */
int fsg_chem_integrate_flagged(const double* table, const double* rho_in, const double* T_in,
                               const double* dt_cell, double dt, int32_t newton_iters,
                               int32_t constant_temperature, const int32_t* flag, int32_t* list,
                               int32_t* count, double* rho_out, double* T_out, int64_t n,
                               void* stream);

/* fsg_chem_integrate_flagged without caller scratch and without any device
 * allocation: the flagged-cell list and its counter live in a static device
 * array of the library (one per GPU, FSG_NOALLOC_MAX_CELLS + 64 int32 =
 * 8 MiB, kept until the process ends).  No counterpart in the reference: for
 * callers that cannot allocate (code under HIP graph capture, serving loops).
 * n above FSG_NOALLOC_MAX_CELLS is rejected (error -13).  The counter is
 * cleared by a fill kernel at each call; calls that overlap on one GPU (two
 * streams, or two host threads) share the array and race.  Same per-cell
 * result as fsg_chem_integrate_flagged. */
#define FSG_NOALLOC_MAX_CELLS (1 << 21)
int fsg_chem_integrate_flagged_noalloc(const double* table, const double* rho_in,
                                       const double* T_in, const double* dt_cell, double dt,
                                       int32_t newton_iters, int32_t constant_temperature,
                                       const int32_t* flag, double* rho_out, double* T_out,
                                       int64_t n, void* stream);

/* The integrator's judgement of a predicted substep (the surrogate's error
 * check): the backward-Euler residual R = rho_pred - rho_0 - dt w(rho_pred, T)
 * of PointImplicitIntegrator.step (hybrid/chemistry/integrator.py:31-104;
 * rho_0 = max(rho_in, 1e-20); T recovered from rho_pred at the initial
 * internal energy unless constant_temperature), one rate evaluation without
 * a Newton solve, scaled by the Jacobian diagonal into the error estimate
 * |R_s| / max(|1 - dt dw_s / drho_s|, 1).  flag[c] is set to 1 where the estimate
 * exceeds rtol |rho_pred,s - rho_0,s| + atol sum(rho_0) for any species (and
 * left as it is elsewhere); err_out (may be null) receives the largest
 * estimate relative to sum(rho_0).  The rates use the Park-11 stoichiometry
 * and third-body efficiencies compiled into the library (fsg_park11.h); a
 * table whose last entry (the stoichiometry fingerprint) differs from the
 * compiled one is not judged: every cell is flagged (err_out = +inf). */
int fsg_chem_check(const double* table, const double* rho_in, const double* T_in,
                   const double* dt_cell, double dt, int32_t constant_temperature,
                   const double* rho_pred, double rtol, double atol, int32_t* flag,
                   double* err_out, int64_t n, void* stream);

/* The coupler's steps around its integrator call, for another substep in
 * between (the surrogate): prepare = species clip and rescale to rho, thermal
 * energy, T0 recovery (hybrid/chemistry/coupling.py:116-154) giving the
 * integrator's rho_s (n, 11, kg/m^3) and T0 (n, K); finish = convert back,
 * clip, rescale to rho and write the species (:165-177).  U is (n, 16) of
 * dtype FSG_F64 or FSG_F32; chemistry in FP64. */
int fsg_chem_prepare(const double* table, const void* U, int32_t dtype, int64_t ncells,
                     const fsg_chem_params* params, double* rho_out, double* T_out,
                     void* stream);
int fsg_chem_finish(void* U, int32_t dtype, int64_t ncells, const fsg_chem_params* params,
                    const double* rho_new, void* stream);

const char* fsg_chem_last_error(void);

/* ---- Chemistry surrogate (MLP 14 -> 128 -> 128 -> 128 -> 11) --------------
 * No counterpart in the reference: a learned replacement for the stiff
 * substep of PointImplicitIntegrator.step (hybrid/chemistry/integrator.py:31),
 * trained on that integrator's own results.  All training buffers are FP32,
 * row-major; W is (N out, K in) like torch.nn.Linear. */

/* Z = X W^T + b (M x N), H = silu(Z) (or Z when silu_on == 0). */
int fsg_mlp_linear_forward(const float* X, const float* W, const float* b, float* Z, float* H,
                           int32_t M, int32_t N, int32_t K, int32_t silu_on, void* stream);
/* Backward of the layer above given dH (M x N, overwritten with dZ):
 * dW = dZ^T X (N x K), db = column sums of dZ (N), dX = dZ W (M x K; may be
 * null for the first layer).  The weight and bias gradients are summed over
 * batch chunks into work (fsg_mlp_backward_work_size(M, N, K) floats, caller
 * scratch) and reduced in a fixed order, so results are deterministic.
 * N <= 256. */
int fsg_mlp_linear_backward(float* dH, const float* Z, const float* X, const float* W, float* dW,
                            float* db, float* dX, float* work, int32_t M, int32_t N, int32_t K,
                            int32_t silu_on, void* stream);
int64_t fsg_mlp_backward_work_size(int32_t M, int32_t N, int32_t K);
/* g = scale * (O - T) over n entries (scale = 2 / n for the mean squared error). */
int fsg_mlp_mse_grad(const float* O, const float* T, float* g, int64_t n, float scale,
                     void* stream);
/* One Adam step (Kingma & Ba 2015, bias-corrected; step counts from 1). */
int fsg_mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                 float b2, float eps, int32_t step, void* stream);

/* Inference on BF16 MFMA: packed_weights holds fsg_mlp_packed_weight_count()
 * bf16 values in the lane order of the matrix-core A operand (written by
 * hybrid/chemistry/surrogate.py), biases 3*128 + 16 floats, consts
 * fsg_mlp_consts_size() bytes (feature/target normalisation, training hull,
 * projection tolerance, species data).  For each of n cells (rho_in (n, 11),
 * T_in (n), dt per cell or scalar) writes the predicted species densities,
 * re-projected onto the input's N, O, charge and mass content, into rho_out;
 * cells outside the hull or whose projection misses the tolerance get
 * fallback[c] = 1 and rho_out = rho_in, to be advanced by fsg_chem_integrate.
 * rho_out must not overlap rho_in (the raw network outputs are staged in
 * rho_out before the projection reads rho_in; error -11).  A block holds 16
 * waves of 64 cells: blocks <= 0 launches ceil(n / 1024) blocks, a positive
 * value caps the grid (the waves then stride over the cells). */
int32_t fsg_mlp_packed_weight_count(void);
int32_t fsg_mlp_consts_size(void);
int fsg_mlp_infer(const void* packed_weights, const float* biases, const void* consts,
                  const double* rho_in, const double* T_in, const double* dt_cell, double dt,
                  double* rho_out, int32_t* fallback, int64_t n, int32_t blocks, void* stream);
const char* fsg_mlp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

/*
 * zhenmode.h -- C ABI of the MI355X mode-split step (libzhenmode_hip.so).
 *
 * The library replaces the compiled entry points that the reference builds in
 * make_solver_global / _compile_solver
 * (src/zhenmode/model/solver/factory.py:30-68, 721-799).  A handle holds one
 * frozen parameter set (the reference's FDPhysParams, solver/state.py:46-187)
 * on the GPU, exactly as the reference's jitted closures hold theirs.
 *
 * Layout: 3-D fields are (nz, ny, nx) with i (longitude) fastest, 2-D fields
 * (ny, nx).  The reference stores (nx, ny, nz); the Python wrapper presents the
 * same buffers to users as (nx, ny, nz) views, so no copy happens per step.
 * All device pointers are in the compute dtype chosen at zm_create (float32 or
 * float64).  Plain pointers and sizes only; no torch types cross the ABI.
 *
 * Errors: every call returns ZM_OK (0) or a nonzero code; zm_last_error gives
 * the message (per handle, or per thread for a failed zm_create when the
 * handle is NULL).  As in the reference, a step never reports a numerical
 * failure: callers classify the state (runtime/monitor.py:20-30).
 * Threading: one handle per stream; calls on a handle must be serialised.
 */
#ifndef ZHENMODE_H
#define ZHENMODE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZM_OK 0
#define ZM_EINVAL 1   /* invalid configuration or argument (ValueError) */
#define ZM_EHIP 2     /* HIP runtime failure */
#define ZM_ENOMEM 3   /* device allocation failure */

typedef struct zm_handle zm_handle;

/* Frozen parameters of the legacy mode-split step, derived on the host in
 * float64 (fd_metrics.py:9-148, factory.py:397-695).  Array pointers are host
 * pointers and are copied at zm_create. */
typedef struct zm_config {
  int32_t nx, ny, nz;
  int32_t dtype_bytes;          /* 4 = float32, 8 = float64 */
  double dt;                    /* baroclinic step [s] */
  double dt_bt;                 /* barotropic substep dt / n_subcyc */
  int32_t n_subcyc;             /* barotropic substeps (factory.py:540-542) */
  int32_t n_nu;                 /* nu_h substeps per half-step (step.py:99) */
  int32_t conv_nsub;            /* convective substeps (factory.py:565-566) */
  int32_t adv_nsub;             /* must be 1 (factory.py:567) */
  int32_t polar_cap_rows, polar_cap_taper;
  double nu_h, nu_v, kappa_h, kappa_v, kappa_conv, nu_bi, kappa_bi;
  double T_ref, S_ref, r_bot;
  double lambda_bulk, restore_coef_S, ice_freeze_temp_c, ice_salt_flux;
  double dy, H_sw, dz_surface, dz_bnd_top, dz_bnd_bot, d2z_h0_top, d2z_h0_bot;
  const double* cos_lat;        /* [ny] */
  const double* dx_row;         /* [ny] dx of each latitude row */
  const double* f_row;          /* [ny] Coriolis parameter of each row */
  const double* dz_node;        /* [nz] */
  const double* dz_iface;       /* [nz-1] */
  const double* dz;             /* [nz-1] */
  const double* dz_norm;        /* [nz-1] */
  const double* d2z_hm;         /* [nz-2] */
  const double* d2z_hp;         /* [nz-2] */
  const double* d2z_denom;      /* [nz-2] */
  const double* dz_denom_interior; /* [nz-2] */
  const double* cap_weights;    /* [polar_cap_rows + polar_cap_taper] */
  const int32_t* dealias_mask;  /* [nx] 1 = longitude wavenumber kept (factory.py:529-533) */
  const int32_t* nwet;          /* [ny*nx] wet levels per column */
  const double* wet_mask;       /* [ny*nx] */
  const double* tau_x;          /* [ny*nx] default forcing */
  const double* tau_y;
  const double* q_heat;
  const double* t_atm;
  const double* s_ref_surf;     /* SSS restoring target */
  const double* coastal_restore_coef;
  const double* coastal_restore_T;
  const double* coastal_bulk_lambda;
  const double* kappa_h_2d;     /* kappa_h + coastal kappa_h */
  const double* kappa_v_2d;     /* kappa_v + coastal kappa_v */
  /* production tracer/mixing options of the reference factory (factory.py:49-51, 649-653) */
  int32_t conservative_kv;      /* interface-flux vertical tracer diffusion (physics/vertical.py:45-53) */
  int32_t localize_conv;        /* convective flux needs its own interface unstable (tendencies.py:149-165) */
  int32_t fct_adv;              /* minmod-limited tracer face values (transport.py:199-245, 266-275) */
  int32_t monotone_adv;         /* donor-cell tracer face values; fct_adv takes precedence */
  /* column-divergence projection of the advecting velocity of the second
   * tracer residual (factory.py:49-53, 368-386; dynamics/projection.py:64-127) */
  int32_t project_adv_vel;
  int32_t projection_niter;            /* CG iteration cap per solve (>= 1) */
  int32_t projection_preconditioner;   /* 0 none, 1 Jacobi */
  int32_t projection_max_refinements;  /* 0..2 transport-checked re-solves */
  double projection_rtol;              /* CG relative tolerance, dtype floor applied */
  double projection_transport_rtol;    /* max(1e-10, 32 eps(dtype)) */
} zm_config;

/* Prognostic state (reference JaxStateG(u, v, T, S, eta, ice),
 * solver/state.py:10): device pointers. */
typedef struct zm_state {
  void* u;
  void* v;
  void* T;
  void* S;
  void* eta;
  void* ice;
} zm_state;

/* Stage selectors for zm_stage. */
#define ZM_STAGE_LINEAR 0      /* _linear_half_step, step.py:76-181 */
#define ZM_STAGE_NONLINEAR 1   /* _explicit_full_step, step.py:249-320 */
#define ZM_STAGE_BAROTROPIC 2  /* mode-split block, step.py:393-428 */
#define ZM_STAGE_TRACER_RES 3  /* _compute_tracer_residual -> out.T, out.S */
#define ZM_STAGE_MOMENTUM_RES 4 /* _compute_momentum_residual + r_bot u bm (the drag strip of
                                  step.py:273-277) -> out.u, out.v */

/* Replaces make_solver_global's parameter freezing and jit compilation
 * (factory.py:30-718).  Returns ZM_EINVAL with a message for an invalid or
 * unsupported configuration. */
int zm_create(const zm_config* cfg, zm_handle** out);
int zm_destroy(zm_handle* h);
const char* zm_last_error(const zm_handle* h);
int32_t zm_version(void);

/* step(state) -> state (factory.py:724-726).  Reads *in, writes *out (which may
 * alias *in); nsteps > 1 continues in place on *out. */
int zm_step(zm_handle* h, const zm_state* in, const zm_state* out, int32_t nsteps, void* stream);

/* step_dyn(state, tau_x, tau_y, q_heat, T_atm_3d) -> state (factory.py:735-750):
 * one step with runtime 2-D forcing (device pointers, (ny, nx); t_atm may be
 * NULL to keep the frozen field). */
int zm_step_dyn(zm_handle* h, const zm_state* in, const zm_state* out, const void* tau_x,
                const void* tau_y, const void* q_heat, const void* t_atm, void* stream);

/* diagnostics(state) -> (rho, p, w) (factory.py:752-760); 3-D outputs. */
int zm_diagnostics(zm_handle* h, const zm_state* in, void* rho, void* pressure, void* w, void* stream);

/* terms_fn(state) -> _tracer_terms(state, params) (factory.py:763-764,
 * dynamics/tendencies.py:74-92): terms = device array [6][nz][ny][nx] of the
 * handle's dtype receiving the temperature's advection, horizontal diffusion,
 * vertical diffusion, convection (one pass with kappa_conv, no sub-cycle), GM
 * and Redi tendencies (the last two zero: the closures are not supported), each
 * masked to wet nodes.  Uses the handle's workspace (calls on one handle are
 * serialised, as for every entry point). */
int zm_tracer_terms(zm_handle* h, const zm_state* in, void* terms, void* stream);

/* One stage of the step (ZM_STAGE_*), for stage-level parity checks. */
int zm_stage(zm_handle* h, int32_t stage, const zm_state* in, const zm_state* out, void* stream);

/* Use hipGraph capture for zm_step/zm_step_dyn (default 1). */
int zm_set_graph(zm_handle* h, int32_t enabled);

/* Average duration [ms] of one launch of a step kernel, timed with HIP events
 * around `reps` back-to-back launches on the handle's own stream.
 * which: 0 = nu_h Laplacian substep (one substep per launch), 1 = tracer
 * residual, 2 = momentum residual, 3 = x de-alias, 4 = one barotropic
 * substep (2 launches), 5 = temporally blocked nu_h launch
 * (zm_nu_substeps_per_launch substeps), 6 = plain copy of one 3-D field
 * (bandwidth calibration), 7 = linear-step finish, 8 = convection,
 * 9 = Fz column scan, 10 = unit D1 pair, 11 = momentum advection,
 * 12 = final masks, 14 = x de-alias by the four-step DFT (3 uses the mixed-radix FFT when nx
 * has no prime factor above 5). */
int zm_time_kernel(zm_handle* h, int32_t which, int32_t reps, double* ms_per_launch);

/* Number of nu_h substeps one blocked launch performs (1 = unblocked). */
int32_t zm_nu_substeps_per_launch(const zm_handle* h);

/* ---- implicit vertical mixing and its learned surrogate -----------------
 * No reference function: the reference's vertical diffusion and convection
 * are explicit.  The mixing operator is the backward-Euler counterpart of its
 * flux-form interface operator (_diffusive_interface_flux /
 * _iface_flux_divergence, numerics/vertical.py:296-332) with the linear EOS
 * (physics/eos.py:13-18):
 *   kappa_{k+1/2} = kappa_v + kappa_conv [rho'_k > rho'_{k+1}] on wet interfaces,
 *   (I - dt D_kappa) X_new = X for X = T, S; zero-flux ends; dry levels held.
 * The handle supplies the grid (wet levels, dz_node, dz_iface, T_ref, S_ref). */

/* Thomas solve per column (one column per lane, float64 elimination).
 * T, S, T_out, S_out: 3-D device fields in the handle's dtype; outputs must not
 * alias inputs.  only_columns: optional int32 [ny*nx] device mask; when given,
 * only columns with a nonzero entry are solved (others untouched).  The first
 * call on a handle allocates its float64 elimination scratch (3 nz ny nx). */
int zm_vertical_mix(zm_handle* h, const void* T, const void* S, void* T_out, void* S_out, double dt,
                    double kappa_v, double kappa_conv, const int32_t* only_columns, void* stream);

/* Allocation-free variant of zm_vertical_mix (same arguments, same results bit
 * for bit).  No counterpart in the reference: it exists for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The elimination values go to a static
 * 32 MiB float64 device array held by every GPU the library runs on until the
 * process ends, and the grid is solved in ranges of
 * floor(4 Mi / (3 nz) / 256) * 256 columns, one launch per range, in stream
 * order (27,904 columns per range at nz = 50).  Levels beyond what one
 * 256-column range fits (nz > 5461) are rejected with ZM_EINVAL.  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other threads) race and must be serialised by the caller. */
int zm_vertical_mix_noalloc(zm_handle* h, const void* T, const void* S, void* T_out, void* S_out, double dt,
                            double kappa_v, double kappa_conv, const int32_t* only_columns, void* stream);

/* Surrogate input rows X[c][0..255] (bfloat16, c = j*nx + i, rows >= ny*nx,
 * rows beyond ny*nx zero): T, S, u, v on nz levels (0 below the seafloor),
 * N^2 at each level's lower interface, q_heat, the SSS-restoring salt flux,
 * |tau|, zero padding; each feature f normalised as (x - mean[f]) * inv_std[f].
 * tau_x / tau_y / q_heat may be NULL for the handle's frozen forcing. nz <= 50. */
int zm_column_features(zm_handle* h, const zm_state* in, const void* tau_x, const void* tau_y,
                       const void* q_heat, const float* mean, const float* inv_std, void* X, int64_t rows,
                       void* stream);

/* Apply surrogate outputs Y (float32 rows of ldy >= 2 nz: nz T then nz S
 * values, normalised).  Per column and tracer X, with e_k the explicit
 * increment of the mixing operator over dt (dt D_kappa X, kappa as in
 * zm_vertical_mix) and s = max_k |e_k| the column's mixing scale:
 * dX_k = e_k + s (y_k out_std_k + out_mean_k) on wet levels, shifted so that
 * the float64 column integral sum dz_node dX is zero (heat and salt conserved
 * exactly), X_new = X + dX into T_out/S_out.  Columns whose implicit residual
 * max|X_new - dt D(X_new) - X| exceeds tol_T / tol_S get flags[c] = 1 and are
 * counted into *n_flagged (device int32, accumulated); the caller passes the
 * same flags to zm_vertical_mix to replace them by the Thomas solution.
 * nz <= 50.  No reference counterpart (the reference has no learned part). */
int zm_mlp_apply(zm_handle* h, const zm_state* in, const float* Y, int32_t ldy, const float* out_mean,
                 const float* out_std, double dt, double kappa_v, double kappa_conv, double tol_T,
                 double tol_S, void* T_out, void* S_out, int32_t* flags, int32_t* n_flagged, void* stream);

/* Training targets of the surrogate, as float32 rows Y[c][0..ldy) (row
 * c = j nx + i; nz T values, nz S values, zero padding; ldy >= 2 nz): per
 * column and tracer, (X_thomas_k - X_k - e_k) / s with e_k, s as in
 * zm_mlp_apply, 0 where s = 0 and on dry levels.  T_thomas / S_thomas hold the
 * zm_vertical_mix solution of the same state with the same dt, kappa_v,
 * kappa_conv.  No reference counterpart (the reference has no learned part). */
int zm_mix_targets(zm_handle* h, const zm_state* in, const void* T_thomas, const void* S_thomas, double dt,
                   double kappa_v, double kappa_conv, float* Y, int32_t ldy, void* stream);

/* The surrogate step in one launch on a float32 handle: features as
 * zm_column_features (in_mean / in_inv_std), the four layers
 * 256 -> 256 -> 256 -> 256 -> 128 (weights[l]: bfloat16 [N][256] row-major,
 * biases[l]: float32 [N]; GELU after the first three) on MFMA, and the apply
 * step of zm_mlp_apply, with the activations kept on chip.  Results equal
 * zm_column_features -> zm_gemm_bf16_nt x 4 -> zm_mlp_apply bit for bit.
 * tau_x / tau_y / q_heat may be NULL for the handle's frozen forcing.  nz <= 50.
 * No reference counterpart (the reference has no learned part). */
int zm_mlp_mix(zm_handle* h, const zm_state* in, const void* tau_x, const void* tau_y, const void* q_heat,
               const float* in_mean, const float* in_inv_std, const void* const* weights, const float* const* biases,
               const float* out_mean, const float* out_std, double dt, double kappa_v, double kappa_conv,
               double tol_T, double tol_S, void* T_out, void* S_out, int32_t* flags, int32_t* n_flagged,
               void* stream);

/* C = A B^T, A (M x K) and B (N x K) bfloat16 row-major device arrays, float32
 * accumulation on MFMA.  M and N multiples of 128, K a multiple of 32.
 * epilogue 0: C_f32 = acc; 1: C_f32 = acc + bias[n];
 * 2: Z = acc + bias[n] -> C_f32 (optional), C_bf16 = gelu(Z) (M x N) and
 *    CT_bf16 = its transpose (N x M, optional);
 * 3: C_bf16 = acc * gelu'(aux[m][n]) and CT_bf16 = its transpose (optional). */
int zm_gemm_bf16_nt(int64_t M, int64_t N, int64_t K, const void* A, const void* B, int32_t epilogue,
                    const float* bias, const float* aux, float* C_f32, void* C_bf16, void* CT_bf16,
                    void* stream);

/* Mean-squared-error gradient over the first n_out of ld entries per row:
 * dY = 2 (Y - target) / (rows n_out) as bfloat16 rows (dY) and columns (dYT),
 * *loss_sum += sum (Y - target)^2 (device float). */
int zm_mse_grad(int64_t rows, int32_t ld, int32_t n_out, const float* Y, const float* target, void* dY,
                void* dYT, float* loss_sum, void* stream);

/* out[r] = sum_c in[r][c] over a rows x cols bfloat16 matrix (float32 sums). */
int zm_rowsum_bf16(int64_t rows, int64_t cols, const void* in, float* out, void* stream);

/* out (cols x rows) = transpose of in (rows x cols), bfloat16. */
int zm_transpose_bf16(int64_t rows, int64_t cols, const void* in, void* out, void* stream);

/* Adam step t (>= 1) on float32 master weights w (n_out x n_in) with gradient
 * gw and moments m, v; writes the bfloat16 copies w_bf (n_out x n_in) and
 * wT_bf (n_in x n_out).  b/gb/bm/bv (length n_out) update the bias when b is
 * not NULL. */
int zm_adam(int32_t n_out, int32_t n_in, float* w, const float* gw, float* m, float* v, void* w_bf,
            void* wT_bf, float* b, const float* gb, float* bm, float* bv, float lr, float beta1, float beta2,
            float eps, int32_t t, void* stream);

/* ---- run-time kernels: forcing, monitor, budget --------------------------- */

/* Seasonal forcing blend of interp_monthly_field_jit
 * (inputs/forcing/seasonal.py:86-103): stack holds 12 months of nfields 2-D
 * fields, [12][nfields][ny][nx] in the handle's dtype; out [nfields][ny][nx]
 * = (1 - wa - wb) month(mi) + wa month(mi+1) + wb month(mi-1) for the 30-day
 * calendar position of `day` (wa, wb from blend_days, 0 <= blend_days <= 30). */
int zm_seasonal_forcing(zm_handle* h, const void* stack, int32_t nfields, double day, double blend_days,
                        void* out, void* stream);

/* Device-resident monitor status: classify_state (runtime/monitor.py:20-30)
 * of the last classified state, running peaks, and the first-rejection
 * record of the run loop (runtime/run_loop.py:53-96): once a state fails,
 * later steps leave the status untouched.  failure: 0 accepted, 1 non-finite,
 * 3 velocity above the limit, 4 surface height above the limit. */
typedef struct zm_monitor_status {
  double max_u, max_velocity, max_eta;
  double peak_u, peak_velocity, peak_eta;
  int32_t finite, failure;
  int64_t first_failure_step; /* -1 while every state was accepted */
  int64_t last_step;
} zm_monitor_status;

/* status: device pointer to one zm_monitor_status. */
int zm_monitor_reset(zm_handle* h, zm_monitor_status* status, void* stream);
/* Classify state number `step` (>= 0, increasing) on the device; no host sync. */
int zm_monitor(zm_handle* h, const zm_state* state, int64_t step, double velocity_limit, double eta_limit,
               int32_t inclusive, zm_monitor_status* status, void* stream);

/* Float64 compensated budget sums (diagnostics/snapshot.py:67-128) into the
 * device array sums[9]: wet volume (area x dz_node), sum T vol, sum S vol, wet
 * surface area, sum T_top area, sum S_top area, total area, sum eta area over
 * wet columns, 0. */
int zm_budget(zm_handle* h, const zm_state* state, double* sums, void* stream);

/* ---- shadow-step stage ledger --------------------------------------------
 * Replaces make_budget_step(params) (diagnostics/budgets.py:244-270) and its
 * _StageRecorder (:69-225): one step, identical to zm_step_dyn, that also
 * writes the float64 audit of the step into the device array ledger
 * [ZM_LEDGER_SIZE].  Metrics are (heat J, salt kg, displacement m3); stages
 * are (linear_diffusion, sponge, nonlinear, free_surface, polar_cap_and_masks,
 * dynamic_ice); sources are (prescribed_heat, bulk_heat, coastal_bulk_heat,
 * temperature_restore, salinity_restore, prescribed_brine, sponge,
 * ice_atmosphere_heat, ice_brine); processes are (advection, convection, gm,
 * redi).  Entries the legacy mode-split step never books (sponge, dynamic
 * ice, GM/Redi, split bottom drag, transport consistency) are zero.
 * tau_x / tau_y / q_heat / t_atm may be NULL for the handle's frozen forcing.
 * out must not alias in. */
#define ZM_LEDGER_OBSERVED 0                          /* [3] */
#define ZM_LEDGER_STAGES 3                            /* [6][3] */
#define ZM_LEDGER_SOURCES 21                          /* [9][3] */
#define ZM_LEDGER_DECOMPOSITION 48                    /* [3] observed - sum(stages) */
#define ZM_LEDGER_BUDGET 51                           /* [3] observed - sum(sources) */
#define ZM_LEDGER_ABS_DECOMPOSITION 54                /* [3] */
#define ZM_LEDGER_ABS_BUDGET 57                       /* [3] */
#define ZM_LEDGER_CHANGE_SCALE 60                     /* [3] sum of |stage changes| */
#define ZM_LEDGER_PROCESSES 63                        /* [4][3] */
#define ZM_LEDGER_PROCESS_SCALE 75                    /* [3] */
#define ZM_LEDGER_ADVECTION_BOUNDARY 78               /* [3] surface advective flux */
#define ZM_LEDGER_ADVECTION_BOUNDARY_RESIDUAL 81      /* [3] advection - boundary */
#define ZM_LEDGER_ABS_ADVECTION_BOUNDARY_RESIDUAL 84  /* [3] */
#define ZM_LEDGER_NONLINEAR_RESIDUAL 87               /* [3] nonlinear stage - sources[:6] - processes */
#define ZM_LEDGER_ABS_NONLINEAR_RESIDUAL 90           /* [3] */
#define ZM_LEDGER_SURFACE_DISPLACEMENT 93             /* [3] eta * surface tracer change */
#define ZM_LEDGER_PROJECTION_NORMS 96                 /* [2] surface transport norm^2 before / after */
#define ZM_LEDGER_PROJECTION_MAX 98                   /* worst sqrt(after / before) */
#define ZM_LEDGER_TRANSPORT_CONSISTENCY 99            /* [4] */
#define ZM_LEDGER_TRANSPORT_AUDITED 103
#define ZM_LEDGER_DRAG_ENERGY 104
#define ZM_LEDGER_DRAG_FACE_CHANGE 105
#define ZM_LEDGER_DRAG_HALVES 106
#define ZM_LEDGER_SIZE 107
int zm_budget_step(zm_handle* h, const zm_state* in, const zm_state* out, const void* tau_x, const void* tau_y,
                   const void* q_heat, const void* t_atm, double* ledger, void* stream);

/* ---- column-divergence projection ----------------------------------------
 * Replaces _project_column_divergence(u, v, p, dt) (dynamics/projection.py:64-127):
 * (u_out, v_out) = (u, v) - dt g grad_c3d(psi) with psi the preconditioned CG
 * solution of the area-weighted column-divergence Poisson problem, then up to
 * projection_max_refinements re-solves while the surface transport norm^2
 * exceeds transport_rtol^2 times that of (u, v).  Float64 solver arithmetic
 * in both dtypes; u, v, u_out, v_out are 3-D fields in the handle's dtype
 * (in place allowed).  iterations: optional device int32[3] receiving the CG
 * iterations of the first solve and of each refinement (0 = not run). */
int zm_project_velocity(zm_handle* h, const void* u, const void* v, void* u_out, void* v_out, double dt,
                        int32_t* iterations, void* stream);
/* The projection operator A(psi) = -column_divergence(grad_c3d(psi)) * area
 * (projection.py:85-87) as nine taps per column, computed on the host from
 * the grid fields of cfg (nwet, wet_mask, dx_row, dy, cos_lat, dz_node):
 *   A(psi)[c] = sum_t taps[t][c] psi[c + offset_t], t = centre, east, west,
 *   east+2, west+2, north, south, north+2, south+2 (x periodic; taps reaching
 *   past the first or last row are zero), c = j*nx + i.
 * diagonal receives _column_projection_diagonal (projection.py:10-50), the
 * Jacobi scaling.  taps [9*ny*nx] and diagonal [ny*nx] are host arrays; either
 * may be NULL.  No device work. */
int zm_projection_operator(const zm_config* cfg, double* taps, double* diagonal);
/* CG iterations of the last projection issued on the handle (by a step with
 * project_adv_vel or by zm_project_velocity) into device int32[3]. */
int zm_projection_iterations(zm_handle* h, int32_t* iterations, void* stream);

/* ---- tiles, autotuner and batched ensembles ------------------------------
 * No reference counterpart: the reference steps one grid per jitted call. */

/* Launch shapes of a handle: nu_block = nu_h substeps per strip launch
 * (1..4 in float32, 1..4 in float64), seg = levels per thread of the
 * column-marching kernels (1..nz).  Results do not depend on either. */
/* Barotropic sub-cycle as one persistent launch (enabled = 1; needs its
 * (ny + 1) / 2 blocks of 1024 threads to fit one per CU) or as one launch per
 * substep (enabled = 0, the default); results are bit-identical.  No counterpart in the reference:
 * it selects how the sub-cycle of _free_surface_step_fd
 * (dynamics/barotropic.py:86-254, driven by timestepping/step.py:410-418) is
 * launched.  Clears the handle's captured graphs.  ZM_EINVAL when enabling a
 * grid whose blocks would not all be resident. */
int zm_set_barotropic_persistent(zm_handle* h, int32_t enabled);

/* Number of bounded waits of the persistent sub-cycle that gave up since the
 * handle was created (a block waited ~1e6 polls for a neighbour); nonzero
 * means the steps since then are invalid.  Synchronises the handle's stream.
 * No counterpart in the reference. */
int zm_barotropic_timeouts(zm_handle* h, int32_t* count);

int zm_set_tiles(zm_handle* h, int32_t nu_block, int32_t seg);
int zm_get_tiles(const zm_handle* h, int32_t* nu_block, int32_t* seg);

/* Times `reps` eagerly issued steps (no graph capture) from a copy of *sample
 * for every (nu_block, seg) candidate up to the handle's current nu_block on
 * this device, keeps the fastest on the handle and reports it; two scratch
 * states are allocated for the duration of the call; *sample is not modified.
 * Results of later steps do not depend on the choice. */
int zm_autotune(zm_handle* h, const zm_state* sample, int32_t reps, int32_t* nu_block, int32_t* seg,
                double* best_ms);

/* An ensemble of float32 handles with equal barotropic substep counts and
 * nx <= 1024 (sizes may differ).  One zm_ensemble_step advances every member
 * one step as a single hipGraph launch: each member's 3-D stages, then the
 * barotropic sub-cycle as one batched launch per substep over a device member
 * table, then each member's final masks; results equal each member's own
 * zm_step bit for bit.  Members keep their own workspaces, so a member handle
 * must not be stepped on its own while an ensemble step is in flight. */
typedef struct zm_ensemble zm_ensemble;
int zm_ensemble_create(zm_handle* const* members, int32_t n_members, zm_ensemble** out);
/* in[n_members], out[n_members]; out[i] may alias in[i] */
int zm_ensemble_step(zm_ensemble* e, const zm_state* in, const zm_state* out, void* stream);
int zm_ensemble_set_graph(zm_ensemble* e, int32_t enabled);
int32_t zm_ensemble_size(const zm_ensemble* e);
int zm_ensemble_destroy(zm_ensemble* e);

#ifdef __cplusplus
}
#endif

#endif /* ZHENMODE_H */

/*
 * cavlbm.h — C-ABI of the MI355X (gfx950) thermal LBM cavity time-stepper.
 *
 * Drop-in boundary for the reference's uniform-grid solver type
 * `crate::uniform::solver::CavitySolver` (bbrangeo/cfd_cavity,
 * src/uniform/solver.rs:35-104) and its diagnostics (src/uniform/post.rs).
 * The reference has no FFI of its own; each entry point below names the Rust
 * item it replaces.  Conventions:
 *   - every call returns 0 on success or a negative CAVLBM_E_* code; the
 *     message is available from cavlbm_last_error(handle) (or
 *     cavlbm_last_error(NULL) for a failed cavlbm_create);
 *   - host arrays are copied in or out, never retained;
 *   - a handle is not re-entrant (like `&mut self`); separate handles are
 *     independent and each owns its own HIP stream;
 *   - per-node arrays use k = iy*nx + ix (solver.rs:162); population arrays
 *     are k-major [k][9] in the direction order given below;
 *   - non-finite values are reported by cavlbm_diagnostics, never an error.
 *
 * Direction order: i = 0..8 with c_i = (0,0) (1,0) (1,1) (0,1) (-1,1) (-1,0)
 * (-1,-1) (0,-1) (1,-1); opp(i) = i+4 for i = 1..4.
 */
#ifndef CAVLBM_H
#define CAVLBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CAVLBM_ABI_VERSION 1

enum {
  CAVLBM_OK = 0,
  CAVLBM_E_PARAM = -1,       /* bad argument; the reference assert!s (solver.rs:130-154) */
  CAVLBM_E_SIZE = -2,        /* array length mismatch (solver.rs:236-240) */
  CAVLBM_E_HIP = -3,         /* HIP runtime failure */
  CAVLBM_E_UNSUPPORTED = -4, /* feature not on this path (dynamic SGS) */
  CAVLBM_E_NODEVICE = -5,    /* no usable gfx950 device */
  CAVLBM_E_IO = -6           /* snapshot file error (the reference exits 6) */
};

/* Population storage in HBM.  Arithmetic is FP64 for FP64 storage and FP32
 * otherwise; reductions are always FP64. */
enum {
  CAVLBM_STORAGE_FP64 = 0,
  CAVLBM_STORAGE_FP32 = 1,
  CAVLBM_STORAGE_FP16_OFFSET = 2 /* f stored as half(f - o_i), o_i the weights w_i rounded
                                    to multiples of 2^-22 that sum to 1; g as
                                    half(g - w_i T_j) with T_j the writing node's
                                    temperature, kept in one FP32 plane after the 18 half
                                    planes; both rounded stochastically, with random words
                                    hashed from the node index and the stored state */
};

enum { CAVLBM_FIELD_RHO = 0, CAVLBM_FIELD_UX = 1, CAVLBM_FIELD_UY = 2,
       CAVLBM_FIELD_T = 3, CAVLBM_FIELD_NU_T = 4, CAVLBM_FIELD_FY_NODE = 5 };
enum { CAVLBM_POPS_F = 0, CAVLBM_POPS_G = 1 };

/* Phase-split entry points, the crate-private phases of step()
 * (solver.rs:287-291) and the test-only apply_bounce_back (solver.rs:920). */
enum {
  CAVLBM_PHASE_COLLIDE = 0,
  CAVLBM_PHASE_STREAM = 1,
  CAVLBM_PHASE_THERMAL_BC = 2,
  CAVLBM_PHASE_UPDATE_MACROSCOPIC = 3,
  CAVLBM_PHASE_ENFORCE_WALL_VELOCITY = 4,
  CAVLBM_PHASE_BOUNCE_BACK = 5
};

typedef struct cavlbm_solver cavlbm_solver;

/* The twelve arguments of CavitySolver::new (solver.rs:108) plus the device
 * options.  Fill with cavlbm_params_default() first. */
typedef struct {
  uint64_t nx, ny;
  double ra, pr, cs, pr_t, utarget;
  uint64_t seed;
  uint64_t jobs; /* recorded for parity with the reference; the GPU ignores it */
  double alpha_mult, nu_start_mult;
  int32_t has_nu_override; /* Option<f64> nu_override */
  int32_t storage;         /* CAVLBM_STORAGE_* */
  double nu_override;
  int32_t geq_order;       /* 2: quadratic g equilibrium (default), 1: linear */
  int32_t device;          /* HIP device ordinal */
} cavlbm_params;

/* Values of post::* and main.rs's progress diagnostics at the current state. */
typedef struct {
  double nusselt_hot;           /* post.rs:68-78 */
  double nusselt_cold;          /* post.rs:100-111 */
  double nusselt_hot_wall_link; /* post.rs:86-96 */
  double nusselt_volume;        /* post.rs:116-132 */
  double total_mass;            /* solver.rs:1012-1014 */
  double total_f_mass;          /* solver.rs:1018-1020 */
  double max_uy;                /* post.rs:153-155 */
  double max_u;                 /* main.rs:192-196 */
  double nu_t_max;              /* main.rs:170-173 */
  double y_plus_hot;            /* post.rs:190-199 */
  int32_t all_finite;           /* post.rs:406-411 */
  int32_t reserved;
  uint64_t time_step;           /* completed steps */
} cavlbm_diag;

/* Lattice-unit calibration of CavitySolver::new (solver.rs:123-156, :194). */
typedef struct {
  double lz, lx, nu, alpha, alpha_phys, beta_eff, tau, tau_t, cs, pr_t,
      nu_start_mult, t_ref;
  uint64_t spin_up_steps;
} cavlbm_derived;

/* Defaults of src/cli.rs:94-130 (64x256, Ra 1e6, Pr 0.71, Cs 0.10, Pr_t 0.9,
 * utarget 0.02, seed 42, jobs 1) with FP64 storage on device 0. */
int cavlbm_params_default(cavlbm_params *p);

/* CavitySolver::new (solver.rs:108-210): calibration, asserts (as
 * CAVLBM_E_PARAM), linear-T + xorshift64* noise init, f = feq, g = geq. */
int cavlbm_create(const cavlbm_params *p, cavlbm_solver **out);
/* Drop of CavitySolver. */
int cavlbm_destroy(cavlbm_solver *s);

/* CavitySolver::step (solver.rs:279-293). */
int cavlbm_step(cavlbm_solver *s);
/* n consecutive step() calls (the loop of main.rs:164-165 without I/O).
 * After the buoyancy spin-up the steady steps are replayed from captured HIP
 * graphs; with CAVLBM_GRAPHS=0 in the environment (read once per process)
 * every solver, batch and octree handle launches them directly instead, with
 * the same results bit for bit. */
int cavlbm_run(cavlbm_solver *s, uint64_t nsteps);
/* Blocks until queued work on the handle's stream has finished. */
int cavlbm_synchronize(cavlbm_solver *s);

/* CavitySolver::init_from_fields (solver.rs:232-262): rho, ux, uy, T of n
 * = nx*ny values each; f = feq, g = geq, wall u = 0, time_step = 0. */
int cavlbm_init_from_fields(cavlbm_solver *s, const double *rho,
                            const double *ux, const double *uy,
                            const double *t, uint64_t n);

/* Read or overwrite one per-node field (the public fields lbm.rho, lbm.ux,
 * lbm.uy, lbm.temperature, nu_t and the crate-private fy_node). */
int cavlbm_get_field(cavlbm_solver *s, int32_t field, double *out, uint64_t n);
int cavlbm_set_field(cavlbm_solver *s, int32_t field, const double *in,
                     uint64_t n);
/* Read or overwrite lbm.f / lbm.g as k-major [k][9] doubles (n9 = 9*nx*ny). */
int cavlbm_get_pops(cavlbm_solver *s, int32_t which, double *out, uint64_t n9);
int cavlbm_set_pops(cavlbm_solver *s, int32_t which, const double *in,
                    uint64_t n9);

/* post::nusselt_*, total_mass, total_f_mass, max_uy, y_plus_hot,
 * check_finite and the progress-line maxima, computed on the device in FP64. */
int cavlbm_diagnostics(cavlbm_solver *s, cavlbm_diag *out);
/* post::max_uy_on_row (post.rs:137-150). */
int cavlbm_max_uy_on_row(cavlbm_solver *s, uint64_t iy, double *w_max,
                         double *x_norm);
/* Public scalars of CavitySolver (lz, lx, beta_eff, nu, alpha, ...). */
int cavlbm_derived_params(const cavlbm_solver *s, cavlbm_derived *out);
/* CavitySolver::time_step (solver.rs:67). */
int cavlbm_time_step(const cavlbm_solver *s, uint64_t *out);
int cavlbm_set_time_step(cavlbm_solver *s, uint64_t t);

/* Hot-wall temperature profile theta_hot[iy] (ny values) used by the
 * isothermal condition at ix = 0 (solver.rs:957-963) in place of T_HOT;
 * NULL restores the constant +0.5. */
int cavlbm_set_hot_wall(cavlbm_solver *s, const double *theta_hot, uint64_t ny);

/* CavitySolver::set_dynamic_sgs (solver.rs:215-220): switches the dynamic
 * mixed-scale model (compute_dynamic_sgs, solver.rs:562-828) on or off.  When
 * on, every step first computes nu_t and alpha_t per node from the post-stream
 * state and the collision uses them; once computed they stay in use after the
 * model is switched off (the reference keeps its Some(..) arrays).  Returns
 * CAVLBM_E_UNSUPPORTED for a cavity of a batch. */
int cavlbm_set_dynamic_sgs(cavlbm_solver *s, int32_t enable, double smag_alpha,
                           double nu_sm_min, double eta_test);
/* The fields cs_dyn_mean, cs_dyn_max, kappa_dyn_mean, kappa_dyn_max
 * (solver.rs:95-102, set at :821-824): out[0..3]; zeros before the first
 * dynamic step. */
int cavlbm_dynamic_stats(cavlbm_solver *s, double out[4]);
/* The per-node arrays nu_t_pre (which = 0) or alpha_t_pre (which = 1) of the
 * last dynamic step (solver.rs:89-93), n = nx*ny values in [iy*nx + ix]
 * order; CAVLBM_E_PARAM before the model has run. */
int cavlbm_get_sgs_pre(cavlbm_solver *s, int32_t which, double *out, uint64_t n);
/* CavitySolver::set_jobs (solver.rs:373-383): accepted and ignored. */
int cavlbm_set_jobs(uint64_t jobs);

/* One crate-private phase: collide(step) (solver.rs:421), stream (:838),
 * apply_thermal_boundary_conditions (:951), update_macroscopic (:298),
 * enforce_wall_velocity (:992), apply_bounce_back (:920).  `step` is used by
 * CAVLBM_PHASE_COLLIDE only. */
int cavlbm_phase(cavlbm_solver *s, int32_t phase, uint64_t step);

/* Timing helpers used by the benchmark (no reference counterpart): run
 * nsteps on the handle's stream between HIP events and return the elapsed
 * milliseconds; cavlbm_time_step_kernel brackets each of `launches` fused
 * step kernels with its own event pair and returns the mean kernel time. */
int cavlbm_time_run(cavlbm_solver *s, uint64_t nsteps, double *ms);
int cavlbm_time_step_kernel(cavlbm_solver *s, uint64_t launches, double *mean_ms);
/* Temporal blocking of the fused run loop (no reference counterpart; a
 * performance setting of this implementation): the run loop advances up to
 * `steps` (1..4) steps per kernel launch.  The defaults are 4 for FP64
 * storage and 3 for FP32 and FP16-offset (whose populations pass through the
 * storage's rounding between the steps of a launch as between launches).  A
 * block of n >= 2 steps needs a lattice that fits that block's tiling:
 * (ny - 1) % 12 != 0 for FP64 n = 2, % 16 for FP64 n = 3, 4, % 8 for FP32 /
 * FP16-offset n = 2, % 16 for n = 3, 4.  It also needs no dynamic
 * model and a handle outside a batch.  Otherwise the loop takes the largest
 * block that fits, down to single steps.  Every choice gives the same
 * populations bit for bit.  cavlbm_steps_per_launch reports the steps per
 * launch of the run loop's steady state for this handle;
 * cavlbm_time_step_kernel times launches of that kind. */
int cavlbm_set_time_blocking(cavlbm_solver *s, int32_t steps);
int cavlbm_steps_per_launch(const cavlbm_solver *s, int32_t *out);
/* One launch per dynamic-SGS step (no reference counterpart; a performance
 * setting of this implementation).  With a test filter of 1..3 nodes
 * (eta_test < 3.5) a dynamic step from the C-form is one kernel in every
 * storage: streaming, the model's filter, similarity scales, Lilly
 * contractions and coefficient stage (FP64), and the collision, per tile
 * through LDS.  Otherwise, or with one_launch = 0, it is four (the f planes
 * and fields of the post-stream state, the filter, the contractions, the
 * collision with the coefficient stage).  Both give the same populations and
 * per-node values bit for bit.  The default is one launch.
 * cavlbm_dynamic_launches reports the launches per dynamic step for the
 * handle's current settings. */
int cavlbm_set_dynamic_fusion(cavlbm_solver *s, int32_t one_launch);
int cavlbm_dynamic_launches(const cavlbm_solver *s, int32_t *out);

/* ---- time averaging and snapshots (the run loop's helpers) ---------------
 * TimeAverager (post.rs:11-57) on the device: FP64 sums of ux, uy, T, nu_t and
 * ux^2 + uy^2 per node; means = sums * (1/count). */
int cavlbm_average_reset(cavlbm_solver *s);
int cavlbm_average_accumulate(cavlbm_solver *s);
int cavlbm_average_count(const cavlbm_solver *s, uint64_t *count);
int cavlbm_average_means(cavlbm_solver *s, double *ux, double *uy, double *t,
                         double *nu_t, double *uu, uint64_t n);

/* CAVSNAP1 files (snapshot.rs:10-95): magic, nx, ny (u64 LE), then rho, ux,
 * uy, T (f64 LE, k = iy*nx + ix).  Host-side I/O; errors are CAVLBM_E_IO with
 * the message in cavlbm_snapshot_last_error(). */
int cavlbm_snapshot_write(const char *path, uint64_t nx, uint64_t ny,
                          const double *rho, const double *ux,
                          const double *uy, const double *t);
/* Temporary file + rename, so the destination is never left half written. */
int cavlbm_snapshot_write_atomic(const char *path, uint64_t nx, uint64_t ny,
                                 const double *rho, const double *ux,
                                 const double *uy, const double *t);
/* With null field pointers only nx, ny are returned; otherwise each buffer
 * holds capacity >= nx*ny values. */
int cavlbm_snapshot_read(const char *path, uint64_t *nx, uint64_t *ny,
                         double *rho, double *ux, double *uy, double *t,
                         uint64_t capacity);
/* interpolate_field (snapshot.rs:100-131): bilinear, normalised coordinates. */
int cavlbm_interpolate_field(const double *src, uint64_t nx0, uint64_t ny0,
                             double *dst, uint64_t nx1, uint64_t ny1);
const char *cavlbm_snapshot_last_error(void);

/* ---- differentiation (FP64 storage only) -------------------------------
 * Objective over a window of K steps starting from the solver's current
 * state S_0 (left unchanged by these calls):
 *     J = (1/K) sum_{m=1..K} nusselt_hot(S_m)          (post.rs:68-78)
 * with the hot-wall profile theta_hot[iy] (ny values; NULL: constant 0.5) in
 * the isothermal condition (solver.rs:957-963) and the Smagorinsky constant
 * cs in collide_node (solver.rs:1120) during the window.  No counterpart in
 * the reference (which has no derivatives); for calibration / inverse
 * problems on the reference's model.
 * The device buffers of these calls (window states; for the adjoint the
 * checkpoints, two recomputed segments of checkpoint_every states -- the
 * recomputation of one overlaps the backward sweep of the other -- and the
 * adjoint planes, about (K / C + 2 C + 4) * 19 * nx * ny doubles) stay with
 * the solver until cavlbm_destroy, so a repeated call allocates nothing. */
int cavlbm_nu_window(cavlbm_solver *s, const double *theta_hot, uint64_t ny,
                     double cs, uint64_t window_steps, double *j_out);
/* J exactly as cavlbm_nu_window computes it (bitwise the same value), with
 * no device allocation: the window's buffers live in a static device array of
 * the library (20.5 MB on each GPU it runs on, held until the process ends).
 * No counterpart in the reference; for callers that cannot allocate device
 * memory (tight evaluation loops of an optimiser, harnesses that forbid
 * hipMalloc).  Limit: the lattice's row-padded plane (ny * nx rounded up to
 * a multiple of 64 per row) must hold at most 65536 nodes and ny <= 1024
 * (e.g. 128x512, 96x384), else CAVLBM_E_SIZE.  Every call on a GPU uses that
 * GPU's one array: calls that overlap on one GPU (from several threads or
 * solvers at once) race and must not be made. */
int cavlbm_nu_window_static(cavlbm_solver *s, const double *theta_hot, uint64_t ny,
                            double cs, uint64_t window_steps, double *j_out);

/* J and its gradient by the hand-written adjoint of the fused step, with
 * checkpoints every checkpoint_every steps and recomputation in between.
 * grad_theta has ny entries (rows 0 and ny-1 are exactly 0: the adiabatic
 * rows overwrite the corners, solver.rs:974-987). */
int cavlbm_adjoint_nu_window(cavlbm_solver *s, const double *theta_hot,
                             uint64_t ny, double cs, uint64_t window_steps,
                             uint64_t checkpoint_every, double *j_out,
                             double *grad_theta, double *grad_cs);

/* J and its directional derivative dJ[v] along v = (dtheta_hot[ny], dcs) by
 * forward-mode (tangent-linear) FP64 steps; J is bitwise the value of
 * cavlbm_nu_window. */
int cavlbm_tangent_nu_window(cavlbm_solver *s, const double *theta_hot,
                             uint64_t ny, double cs, const double *dtheta_hot,
                             double dcs, uint64_t window_steps, double *j_out,
                             double *dj_out);

/* ---- batched form -------------------------------------------------------
 * A batch advances many independent cavities (sizes, Ra, Cs may differ) with
 * one kernel launch per step, captured in HIP graphs after the buoyancy
 * spin-up.  No counterpart in the reference, which runs one cavity per
 * process (src/main.rs:42-443); each cavity behaves exactly like a
 * cavlbm_solver created from the same parameters (bitwise equal results). */
typedef struct cavlbm_batch cavlbm_batch;

typedef struct {
  int32_t tile_x, tile_y;   /* chosen output tile (nodes) */
  int32_t xcd_remap;        /* 1: tiles remapped so each XCD owns contiguous rows */
  int32_t candidates;       /* configurations timed */
  double best_ms_per_step;  /* of the chosen configuration */
  double first_ms_per_step; /* of the first candidate (64x4, remapped) */
  int32_t steps_per_launch; /* chosen steps per launch (1: single-step tiles) */
  int32_t reserved;
} cavlbm_tune_result;

/* n parameter sets (same storage, geq_order and device; nu_start_mult 1). */
int cavlbm_batch_create(const cavlbm_params *cavities, uint64_t n,
                        cavlbm_batch **out);
int cavlbm_batch_destroy(cavlbm_batch *b);
/* nsteps steps of every cavity. */
int cavlbm_batch_run(cavlbm_batch *b, uint64_t nsteps);
int cavlbm_batch_synchronize(cavlbm_batch *b);
/* Borrowed handle of cavity i for reads, diagnostics and setters (valid
 * until cavlbm_batch_destroy; do not destroy it; do not step it alone). */
int cavlbm_batch_solver(cavlbm_batch *b, uint64_t i, cavlbm_solver **out);
/* Times every single-step tile shape and every N-step block (N = 2..4, when
 * every cavity fits its tiling), each with and without the XCD remap, for
 * trial_steps steps and keeps the fastest; the cavities' state is restored
 * afterwards.  A new batch advances the storage's default steps per launch
 * (4 FP64, 3 FP32 and FP16-offset) when every cavity fits, single steps
 * otherwise.  The batch
 * must have been run at least one step (CAVLBM_E_PARAM otherwise); tune after
 * the spin-up so that the trials time the graph-replayed steady steps. */
int cavlbm_batch_autotune(cavlbm_batch *b, uint64_t trial_steps,
                          cavlbm_tune_result *out);
/* Total lattice nodes of the batch. */
int cavlbm_batch_nodes(const cavlbm_batch *b, uint64_t *out);
int cavlbm_batch_time_run(cavlbm_batch *b, uint64_t nsteps, double *ms);
const char *cavlbm_batch_last_error(const cavlbm_batch *b);

/* ===========================================================================
 * Octree (multi-level) cavity: the reference's src/octree CavitySolver
 * (solver.rs:118-1206) with its per-level phases (cavity.rs:167-656) and
 * interface transfers (interface.rs:164-229) on the device.  Levels: a coarse
 * base grid L0 over the whole cavity, two full-height blocks L1 (dx/2) on the
 * isothermal walls, and with three levels an L2 strip (dx/4) in each block.
 * Grids are numbered as in the reference's snapshots (solver.rs:1048-1066):
 * 0 = L0, 1 = hot L1, 2 = cold L1, 3 = hot L2, 4 = cold L2; per-grid arrays
 * are k = iy*nx + ix with the grid's own nx, ny (ghost layers included).
 * FP64 only; the dynamic SGS of the octree is not on this path.
 * ======================================================================== */
typedef struct cavlbm_octree cavlbm_octree;

/* The arguments of octree CavitySolver::new (solver.rs:340-353). */
typedef struct {
  uint64_t nx0, ny0, bl_width;
  double ra, pr, utarget;
  uint64_t seed;
  uint64_t jobs;           /* recorded, ignored on the GPU */
  int32_t odd_exact;       /* xi_odd = (1/2) tau-_fine / tau-_coarse */
  int32_t levels;          /* 2 or 3 */
  double nu_start_mult;
  int32_t has_nu_override; /* Option<f64> nu_override */
  int32_t geq_order;       /* 2: quadratic g equilibrium (default), 1: linear */
  double nu_override;
  int32_t device;
  int32_t reserved;
} cavlbm_octree_params;

/* Calibration and level constants (solver.rs:363-468). */
typedef struct {
  double nu_lu, alpha_lu, beta_eff_coarse, beta_eff_fine;
  double tau[3][3];        /* [L0, L1, L2][tau+, tau-, tau_t] */
  double xi_a[3], xi_b[3]; /* interfaces L0-L1 and L1-L2: even, odd, thermal */
  uint64_t spin_up_steps, n_nodes_total;
  int32_t levels, n_grids;
} cavlbm_octree_derived;

/* Diagnostics of run_octree's progress lines (cavity.rs:672-790,
 * solver.rs:949-1041). */
typedef struct {
  double nusselt_hot, nusselt_cold, nusselt_hot_link0;
  double nusselt_hot_fine;  /* hot L1 block, scale lx_coarse = nx0 - 1 */
  double max_u_all, max_rho_dev, max_rho_dev_ifaces, masse;
  int32_t guard;            /* check_guards: 0 none, 1 max_u > 0.10, 2 non-finite */
  int32_t reserved;
  uint64_t time_step;
} cavlbm_octree_diag;

/* Defaults of src/cli.rs:121-130 (64x256, bl_width 8, Ra 1e6, Pr 0.71,
 * utarget 0.02, seed 42, 2 levels) on device 0. */
int cavlbm_octree_params_default(cavlbm_octree_params *p);
/* CavitySolver::new + init_cavity (solver.rs:340-575); its asserts are
 * CAVLBM_E_PARAM. */
int cavlbm_octree_create(const cavlbm_octree_params *p, cavlbm_octree **out);
int cavlbm_octree_destroy(cavlbm_octree *o);
/* step_cavity (solver.rs:647-947); run = n of them. */
int cavlbm_octree_step(cavlbm_octree *o);
int cavlbm_octree_run(cavlbm_octree *o, uint64_t nsteps);
int cavlbm_octree_synchronize(cavlbm_octree *o);
int cavlbm_octree_time_run(cavlbm_octree *o, uint64_t nsteps, double *ms);
int cavlbm_octree_time_step(const cavlbm_octree *o, uint64_t *out);
int cavlbm_octree_derived_params(const cavlbm_octree *o, cavlbm_octree_derived *out);
int cavlbm_octree_grid_dims(const cavlbm_octree *o, int32_t grid, uint64_t *nx,
                            uint64_t *ny);
/* Per-grid state: fields CAVLBM_FIELD_{RHO,UX,UY,T,FY_NODE}, populations
 * [k][9]; the setters overwrite the state as the reference's public level
 * fields allow. */
int cavlbm_octree_get_field(cavlbm_octree *o, int32_t grid, int32_t field, double *out,
                            uint64_t n);
int cavlbm_octree_set_field(cavlbm_octree *o, int32_t grid, int32_t field, const double *in,
                            uint64_t n);
int cavlbm_octree_get_pops(cavlbm_octree *o, int32_t grid, int32_t which, double *out,
                           uint64_t n9);
int cavlbm_octree_set_pops(cavlbm_octree *o, int32_t grid, int32_t which, const double *in,
                           uint64_t n9);
/* set_from_snapshot (solver.rs:1095-1175): rho, ux, uy, T of every grid
 * concatenated in grid order (total = sum of nx*ny); f = feq, g = geq, wall
 * velocities zeroed on every level, time_step = 0. */
int cavlbm_octree_set_from_snapshot(cavlbm_octree *o, const double *rho, const double *ux,
                                    const double *uy, const double *t, uint64_t total);
int cavlbm_octree_diagnostics(cavlbm_octree *o, cavlbm_octree_diag *out);
/* CavitySolver::set_dynamic_sgs (solver.rs:475-480): the dynamic
 * mixed-scale model (src/octree/smag.rs) on every level before each of its
 * collisions, with per-node relaxation times, and the interface factors
 * recomputed every step from the levels' mean nu_t / alpha_t
 * (solver.rs:647-947).  The means are deterministic tree sums on the device
 * (the reference sums sequentially), and the model's mixed forms use the
 * device's pow, so the path agrees with a CPU evaluation to a tolerance.
 * cavlbm_octree_dynamic_stats: cs_dyn_mean, cs_dyn_max, kappa_dyn_mean,
 * kappa_dyn_max of L0's last model pass, nu_t_max over every level (zeros
 * before the model first runs).  cavlbm_octree_get_sgs_pre: a grid's
 * nu_t_pre (which 0) or alpha_t_pre (1), n = its node count.
 * cavlbm_octree_interface_factors: xi even / odd / thermal of interface A
 * (L0-L1), then interface B (L1-L2) of the hot and of the cold block (zeros
 * with two levels); the factors in use, which the model recomputes. */
int cavlbm_octree_set_dynamic_sgs(cavlbm_octree *o, int32_t enable, double smag_alpha,
                                  double nu_sm_min, double eta_test);
int cavlbm_octree_dynamic_stats(cavlbm_octree *o, double out[5]);
int cavlbm_octree_get_sgs_pre(cavlbm_octree *o, int32_t grid, int32_t which, double *out,
                              uint64_t n);
int cavlbm_octree_interface_factors(cavlbm_octree *o, double out[9]);
const char *cavlbm_octree_last_error(const cavlbm_octree *o);

/* OCTSNAP1 multi-level snapshots (src/octree/snapshot.rs:1-170), host-side
 * file I/O like the CAVSNAP1 functions (errors: CAVLBM_E_IO, message in
 * cavlbm_snapshot_last_error()).  dims holds nx, ny of each grid (2*n_grids
 * values); the field arrays are the grids' fields concatenated in grid
 * order.  write: encode + (atomic != 0) temporary file and rename
 * (snapshot.rs:57-79, :140-161). */
int cavlbm_octsnap_write(const char *path, uint32_t levels, uint64_t step, uint32_t n_grids,
                         const uint64_t *dims, const double *rho, const double *ux,
                         const double *uy, const double *t, int32_t atomic);
/* read_snapshot + decode (snapshot.rs:83-137, :164-170), exact length
 * checked.  With null field pointers only the header is returned (dims needs
 * room for 10 values); else each field buffer holds `capacity` >= the total
 * node count. */
int cavlbm_octsnap_read(const char *path, uint32_t *levels, uint64_t *step, uint32_t *n_grids,
                        uint64_t *dims, double *rho, double *ux, double *uy, double *t,
                        uint64_t capacity);

/* Message of the last failure on this handle (or of cavlbm_create when s is
 * NULL).  Valid until the next call on the same handle / thread. */
const char *cavlbm_last_error(const cavlbm_solver *s);
/* CAVLBM_ABI_VERSION of the loaded library. */
int cavlbm_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif

/*
 * cavity_oracle.h — CPU oracle for the uniform-grid thermal LBM cavity step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * uniform solver (bbrangeo/cfd_cavity, src/uniform/solver.rs, post.rs,
 * rng.rs).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker / the timed CPU baseline.  The
 * product (libcavlbm, HIP) never links or calls it.
 *
 * Layout mirrors the reference: populations are AoS [k][9] doubles with
 * k = iy*nx + ix (src/uniform/solver.rs:162), per-node fields are [k].
 *
 * Direction order (the crate's own table is not vendored and stays unpinned):
 *   i : 0      1      2      3      4      5      6      7      8
 *   c : (0,0)  (1,0)  (1,1)  (0,1)  (-1,1) (-1,0) (-1,-1) (0,-1) (1,-1)
 *   opp(i) = i+4 for i=1..4 (consistent with swap_opposite iterating 1..4,
 *   src/uniform/solver.rs:938-946).
 */
#ifndef CAVITY_ORACLE_H
#define CAVITY_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_solver orc_solver;

enum { ORC_RHO = 0, ORC_UX = 1, ORC_UY = 2, ORC_T = 3, ORC_NU_T = 4, ORC_FY = 5 };
enum { ORC_F = 0, ORC_G = 1 };

/* CavitySolver::new (src/uniform/solver.rs:108-210).  Returns NULL and writes
 * a message to err on the conditions the reference asserts on. */
orc_solver *orc_new(size_t nx, size_t ny, double ra, double pr, double cs,
                    double pr_t, double utarget, uint64_t seed, size_t jobs,
                    double alpha_mult, double nu_start_mult,
                    int has_nu_override, double nu_override, int geq_order,
                    char *err, size_t errlen);
void orc_free(orc_solver *s);
void orc_set_jobs(orc_solver *s, size_t jobs);

/* step() and its phases (src/uniform/solver.rs:279-293). */
void orc_step(orc_solver *s);
void orc_run(orc_solver *s, uint64_t nsteps);
void orc_collide(orc_solver *s, uint64_t step);
void orc_stream(orc_solver *s);
void orc_apply_bounce_back(orc_solver *s);
void orc_thermal_bc(orc_solver *s);
void orc_update_macroscopic(orc_solver *s);
void orc_enforce_wall_velocity(orc_solver *s);
int orc_init_from_fields(orc_solver *s, const double *rho, const double *ux,
                         const double *uy, const double *t, size_t n);

/* Optional hot-wall temperature profile theta_hot[iy] (ny values) in place of
 * the constant T_HOT; NULL restores the constant. */
void orc_set_hot_wall(orc_solver *s, const double *theta_hot);
void orc_set_cs(orc_solver *s, double cs);

/* Raw state access (numpy wraps these pointers). */
double *orc_field(orc_solver *s, int which);
double *orc_pops(orc_solver *s, int which);
size_t orc_nx(const orc_solver *s);
size_t orc_ny(const orc_solver *s);
uint64_t orc_time_step(const orc_solver *s);
void orc_set_time_step(orc_solver *s, uint64_t t);
/* lz, lx, nu, alpha, alpha_phys, beta_eff, tau, tau_t, spin_up, cs, pr_t,
 * nu_start_mult, t_ref */
void orc_derived(const orc_solver *s, double out[13]);
/* Per-step scalars of collide(step): beta, nu_eff, alpha_eff, tau_mol. */
void orc_step_scalars(const orc_solver *s, uint64_t step, double out[4]);

/* Diagnostics (src/uniform/post.rs, src/uniform/solver.rs:1012-1020). */
double orc_total_mass(const orc_solver *s);
double orc_total_f_mass(const orc_solver *s);
double orc_nusselt_hot(const orc_solver *s);
double orc_nusselt_hot_wall_link(const orc_solver *s);
double orc_nusselt_cold(const orc_solver *s);
double orc_nusselt_volume(const orc_solver *s);
double orc_max_uy(const orc_solver *s);
double orc_max_uy_on_row(const orc_solver *s, size_t iy, double *x_norm);
double orc_max_u(const orc_solver *s);
double orc_nu_t_max(const orc_solver *s);
double orc_wall_shear_hot(const orc_solver *s, size_t iy);
double orc_y_plus_hot(const orc_solver *s);
int orc_check_finite(const orc_solver *s);

/* Dynamic mixed-scale SGS (solver.rs:215-220, :562-828, :1195-1261). */
void orc_set_dynamic_sgs(orc_solver *s, int on, double smag_alpha, double nu_sm_min,
                         double eta_test);
void orc_compute_dynamic_sgs(orc_solver *s);
/* The model on given arrays (f k-major [k][9]); stats: C_nu mean, max,
 * C_kappa mean, max.  Used by the uniform solver and by each octree level. */
void orc_dynamic_model(size_t nx, size_t ny, const double *f9, const double *rho,
                       const double *ux, const double *uy, const double *temp, double tau,
                       double smag_alpha, double eta_test, double nu_mol, double alpha_mol,
                       double nu_sm_min, double *nu_t_out, double *alpha_t_out,
                       double stats[4]);
double orc_lilly_coefficient(double num, double den);
double orc_mixed_factor(double a, double b, double alpha);
/* cs_dyn_mean, cs_dyn_max, kappa_dyn_mean, kappa_dyn_max */
void orc_dynamic_stats(const orc_solver *s, double out[4]);
/* nu_t_pre (0) or alpha_t_pre (1); NULL before the first dynamic step */
double *orc_sgs_pre(orc_solver *s, int which);

/* Crate primitives as restated here (call sites listed in the .c file). */
double orc_feq(int i, double rho, double ux, double uy);
double orc_geq(int i, double t, double ux, double uy, int geq_order);
void orc_lattice(int cx[9], int cy[9], int opp[9], double w[9]);

/* One collide_node call on caller-owned node arrays (for closed-form tests). */
void orc_collide_node(double f[9], double g[9], double *nu_t_out,
                      double *fy_out, double rho, double ux, double uy,
                      double temp, double tau_mol, double beta, double t_ref,
                      double cs, double pr_t, double nu, double alpha,
                      int geq_order);

/* Transient Poiseuille channel built from collide_node's TRT+Guo split
 * (src/uniform/tests.rs:246-376, restated with the force along +y): periodic
 * in y, full-way bounce-back walls in x, uniform body force fy*rho.  Returns
 * max over nodes of the Guo-corrected u_y after nsteps. */
double orc_poiseuille_channel(size_t nx, size_t ny, double nu, double fy,
                              uint64_t nsteps);

/* xorshift64* (src/uniform/rng.rs:5-31). */
uint64_t orc_rng_next_u64(uint64_t *state);
double orc_rng_next_f64(uint64_t *state);

#ifdef __cplusplus
}
#endif
#endif

/*
 * octree_oracle.h — CPU oracle for the octree (multi-level) cavity solver.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's octree
 * path (bbrangeo/cfd_cavity src/octree/cavity.rs, grid.rs, interface.rs,
 * lbm.rs, solver.rs) used to check the HIP octree kernels.  Only tests/ and
 * the smoke check may load it; the product never links it.
 *
 * Levels are AoS [k][9] doubles with k = iy*nx + ix like the reference.  Grids
 * are numbered in the reference's snapshot order (solver.rs:1048-1066):
 * 0 = coarse L0, 1 = hot-wall L1 block, 2 = cold-wall L1 block, and with three
 * levels 3 = hot L2 strip, 4 = cold L2 strip.
 */
#ifndef OCTREE_ORACLE_H
#define OCTREE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_octree orc_octree;
typedef struct orc_oct2 orc_oct2;

enum { OCT_RHO = 0, OCT_UX = 1, OCT_UY = 2, OCT_T = 3, OCT_FY = 4 };

/* CavitySolver::new (src/octree/solver.rs:340-468): returns NULL and a
 * message on the conditions the reference asserts on. */
orc_octree *orc_oct_new(size_t nx0, size_t ny0, size_t bl_width, double ra, double pr,
                        double utarget, uint64_t seed, size_t jobs, int odd_exact,
                        int levels, double nu_start_mult, int has_nu_override,
                        double nu_override, int geq_order, char *err, size_t errlen);
void orc_oct_free(orc_octree *o);
void orc_oct_step(orc_octree *o);                /* step_cavity (solver.rs:647-947) */
void orc_oct_run(orc_octree *o, uint64_t n);
size_t orc_oct_ngrids(const orc_octree *o);
void orc_oct_grid_dims(const orc_octree *o, size_t gi, size_t *nx, size_t *ny);
double *orc_oct_field(orc_octree *o, size_t gi, int which);  /* OCT_* */
double *orc_oct_pops(orc_octree *o, size_t gi, int which);   /* 0 = f, 1 = g */
uint64_t orc_oct_time_step(const orc_octree *o);
/* nu_lu, alpha_lu, beta_eff_coarse, beta_eff_fine, spin_up_steps,
 * tau+ / tau- / tau_t of L0, L1, L2 (9), xi_even / xi_odd / xi_thermal of
 * interface A (3) and of interface B (3), n_nodes_total  -> 22 values */
void orc_oct_derived(const orc_octree *o, double out[22]);
/* Dynamic mixed-scale SGS (solver.rs:476-481, :647-947, smag.rs).  stats:
 * cs_dyn_mean, cs_dyn_max, kappa_dyn_mean, kappa_dyn_max (of L0, this step),
 * nu_t_max (all levels).  sgs_pre: the level's nu_t_pre (0) / alpha_t_pre (1)
 * in grid order.  xi: interface A (even, odd, thermal), then interface B of
 * the hot and the cold block (zeros with two levels). */
void orc_oct_set_dynamic_sgs(orc_octree *o, int on, double smag_alpha, double nu_sm_min,
                             double eta_test);
void orc_oct_dynamic_stats(const orc_octree *o, double out[5]);
double *orc_oct_sgs_pre(orc_octree *o, size_t gi, int which);
void orc_oct_xi(const orc_octree *o, double out[9]);
/* noise_stats (solver.rs:949-1000): max_rho_dev, max_rho_dev_ifaces, masse */
void orc_oct_noise_stats(const orc_octree *o, double out[3]);
double orc_oct_max_u_all(const orc_octree *o);
/* check_guards (solver.rs:1018-1041): 0 none, 1 max_u abort, 2 non-finite */
int orc_oct_check_guards(const orc_octree *o);
/* cavity.rs:672-740 on the coarse level / a fine block */
double orc_oct_nusselt_hot(const orc_octree *o);
double orc_oct_nusselt_cold(const orc_octree *o);
double orc_oct_nusselt_hot_link0(const orc_octree *o);
double orc_oct_nusselt_hot_fine(const orc_octree *o, size_t block, double lx_coarse);
/* set_from_snapshot (solver.rs:1095-1175): fields of every grid concatenated
 * in grid order; returns 0, or -1 when `total` does not match */
int orc_oct_set_from_snapshot(orc_octree *o, const double *rho, const double *ux,
                              const double *uy, const double *t, size_t total);

/* Interface tables of a grid (grid.rs:95-203): number of ghost / coincident
 * entries; entry j as (k_fine, 4 coarse indices, 4 weights) or (k_coarse,
 * 9 fine indices, 9 weights). */
size_t orc_oct_n_ghosts(const orc_octree *o, size_t gi);
size_t orc_oct_n_coincident(const orc_octree *o, size_t gi);
void orc_oct_ghost(const orc_octree *o, size_t gi, size_t j, size_t *k_fine, size_t kc[4],
                   double w[4]);
void orc_oct_coincident(const orc_octree *o, size_t gi, size_t j, size_t *k_coarse,
                        size_t kf[9], double w[9]);

/* Per-level phases on grid gi (cavity.rs:167-656), for pinning against the
 * uniform oracle.  `tau` is (tau+, tau-, tau_t). */
void orc_oct_collide_level(orc_octree *o, size_t gi, const double tau[3], double beta,
                           int interior_only);
/* collide_node_cavity (cavity.rs:68-161) on caller-owned node arrays; pre:
 * the per-node nu_t / alpha_t of the dynamic model set the relaxation times */
void orc_oct_collide_node(double f[9], double g[9], double *fy_out, double rho, double ux,
                          double uy, double temp, const double tau[3], double beta,
                          int geq_order, int pre, double nu_t, double alpha_t);
void orc_oct_stream_level(orc_octree *o, size_t gi);                     /* stream_cavity */
void orc_oct_stream_block_level(orc_octree *o, size_t gi, int wl, int wr, int wb, int wt);
void orc_oct_bc_level(orc_octree *o, size_t gi);                         /* uniform BCs */
void orc_oct_bc_block_level(orc_octree *o, size_t gi, int hot, int cold);
void orc_oct_macro_level(orc_octree *o, size_t gi);                      /* update_macroscopic_guo */
void orc_oct_walls_level(orc_octree *o, size_t gi);
void orc_oct_walls_block_level(orc_octree *o, size_t gi, int hot, int cold);

/* The two-level interface harness OctreeSolver (solver.rs:66-116) for the
 * interface tests: a coarse level, one fine block, the taus and xi. */
orc_oct2 *orc_oct2_new(size_t nx0, size_t ny0, size_t bx0, size_t by0, size_t bx1, size_t by1,
                       double nu, double alpha, int odd_exact);
void orc_oct2_free(orc_oct2 *s);
void orc_oct2_dims(const orc_oct2 *s, size_t *nx1, size_t *ny1);
double *orc_oct2_field(orc_oct2 *s, int level, int which);   /* level 0 coarse, 1 fine */
double *orc_oct2_pops(orc_oct2 *s, int level, int which);
void orc_oct2_xi(const orc_oct2 *s, double out[3]);
void orc_oct2_coarse_to_fine(orc_oct2 *s);
void orc_oct2_fine_to_coarse(orc_oct2 *s);
size_t orc_oct2_n_ghosts(const orc_oct2 *s);
size_t orc_oct2_n_coincident(const orc_oct2 *s);
size_t orc_oct2_ghost_k(const orc_oct2 *s, size_t j);
void orc_oct2_coincident(const orc_oct2 *s, size_t j, size_t *k_coarse, size_t kf[9],
                         double w[9]);

/* interface.rs primitives */
void orc_oct_relax_times(double nu, double alpha, double out[3]);
void orc_oct_rescaling_factors(const double tau_c[3], const double tau_f[3], size_t substeps,
                               int odd_exact, double out[3]);
void orc_oct_rescale_fneq(const double fneq[9], const double xi[3], double out[9]);
void orc_oct_rescale_fneq_inverse(const double fneq[9], const double xi[3], double out[9]);
void orc_oct_bilinear_stencil(double x_c, double y_c, size_t nx0, size_t ny0, size_t k[4],
                              double w[4]);

#ifdef __cplusplus
}
#endif
#endif

/*
 * cavity_oracle.c — CPU oracle for the uniform-grid thermal LBM cavity step.
 *
 * TEST INFRASTRUCTURE ONLY: a restatement of the reference's algorithm used to
 * check the HIP product and to time the reference's CPU path.  Nothing in the
 * product loads this file.  Build: oracle/Makefile (gcc -O3 -ffp-contract=off
 * -fopenmp), output oracle/libcavity_oracle.so.
 *
 * Each function cites the reference file:line it restates.  Floating-point
 * expressions keep the reference's operand order (no contraction), so that a
 * HIP kernel written with the same order is bitwise equal in FP64.
 *
 * Parity status (what pins each part; the tests live in tests/test_oracle_*.py):
 *   - feq: form fixed by the reference's own hand-written feq
 *     (src/uniform/tests.rs:314-316); moments checked in closed form.
 *   - geq: zeroth moment pinned (src/uniform/tests.rs:123-127); the velocity
 *     terms of the crate's geq are PARITY UNPINNED (crate not vendored);
 *     geq_order selects quadratic (default) or linear.
 *   - collide_node: Guo split pinned by the transient-Poiseuille closed form;
 *     moments of the source term and a one-step equilibrium collision pinned
 *     in closed form; full step pinned by an independent numpy textbook step.
 *   - xorshift64* and the initial noise: the state update pinned as a GF(2)
 *     matrix of full period 2^64 - 1 equal to the published shift triple,
 *     the multiplier recovered from outputs, the noise rebuilt by a matrix
 *     generator; the nu / alpha / beta spin-up ramps pinned by their closed
 *     forms through orc_step_scalars (tests/test_oracle_rng_ramp.py).
 *   - operand order inside boussinesq_force / compute_turbulent_viscosity /
 *     y_plus and the D2Q9 direction order: PARITY UNPINNED against the crate
 *     (values pinned, round-off order chosen here).
 */
#include "cavity_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define T_HOT 0.5
#define T_COLD (-0.5)
#define DELTA_T (T_HOT - T_COLD)
#define CS2 (1.0 / 3.0)
#define F64_EPSILON 2.220446049250313e-16

static const int CX[9] = {0, 1, 1, 0, -1, -1, -1, 0, 1};
static const int CY[9] = {0, 0, 1, 1, 1, 0, -1, -1, -1};
static const int OPP[9] = {0, 5, 6, 7, 8, 1, 2, 3, 4};
static const double W[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 36.0,
                            1.0 / 9.0,  1.0 / 36.0, 1.0 / 9.0,
                            1.0 / 36.0, 1.0 / 9.0,  1.0 / 36.0};

typedef struct {
  double f[9];
} node9;

struct orc_solver {
  size_t nx, ny, n;
  node9 *f, *g;     /* populations, AoS [k][9] */
  node9 *f_old, *g_old; /* stream() snapshot (the reference clones each step) */
  double *rho, *ux, *uy, *temp, *nu_t, *fy_node;
  double *theta_hot; /* NULL: constant T_HOT */
  double lz, lx, beta_eff, nu, alpha, alpha_phys, cs, pr_t, tau, tau_t, t_ref;
  double nu_start_mult;
  size_t jobs;
  uint64_t time_step, spin_up_steps;
  int geq_order;
  /* dynamic mixed-scale SGS (solver.rs:76-103) */
  int dynamic_smag;
  double smag_alpha, nu_sm_min, eta_test;
  double *nu_t_pre, *alpha_t_pre; /* NULL when off */
  double cs_dyn_mean, cs_dyn_max, kappa_dyn_mean, kappa_dyn_max;
};

/* ---- xorshift64* (src/uniform/rng.rs:11-30) ---------------------------- */
uint64_t orc_rng_next_u64(uint64_t *state) {
  uint64_t x = *state;
  x ^= x >> 12;
  x ^= x << 25;
  x ^= x >> 27;
  *state = x;
  return x * 0x2545F4914F6CDD1DULL;
}

double orc_rng_next_f64(uint64_t *state) {
  return (double)(orc_rng_next_u64(state) >> 11) / (double)(1ULL << 53);
}

/* ---- crate primitives ---------------------------------------------------- */
void orc_lattice(int cx[9], int cy[9], int opp[9], double w[9]) {
  for (int i = 0; i < 9; ++i) {
    cx[i] = CX[i];
    cy[i] = CY[i];
    opp[i] = OPP[i];
    w[i] = W[i];
  }
}

/* feq(i, rho, ux, uy): the textbook D2Q9 form that the reference writes out by
 * hand in src/uniform/tests.rs:312-316 (called at solver.rs:172, :253, :1091). */
double orc_feq(int i, double rho, double ux, double uy) {
  double cx = (double)CX[i], cy = (double)CY[i];
  double eu = cx * ux + cy * uy;
  double u2 = ux * ux + uy * uy;
  return W[i] * rho *
         (1.0 + eu / CS2 + eu * eu / (2.0 * CS2 * CS2) - u2 / (2.0 * CS2));
}

/* geq(i, T, ux, uy) (called at solver.rs:173, :254, :961, :967, :1092).
 * Velocity terms unpinned: quadratic (geq_order 2) or linear (geq_order 1). */
double orc_geq(int i, double t, double ux, double uy, int geq_order) {
  double cx = (double)CX[i], cy = (double)CY[i];
  double eu = cx * ux + cy * uy;
  if (geq_order == 1) return W[i] * t * (1.0 + eu / CS2);
  double u2 = ux * ux + uy * uy;
  return W[i] * t *
         (1.0 + eu / CS2 + eu * eu / (2.0 * CS2 * CS2) - u2 / (2.0 * CS2));
}

/* ---- CavitySolver::new (src/uniform/solver.rs:108-210) -------------------- */
static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) {
    fprintf(stderr, "cavity_oracle: out of memory\n");
    abort();
  }
  return p;
}

orc_solver *orc_new(size_t nx, size_t ny, double ra, double pr, double cs,
                    double pr_t, double utarget, uint64_t seed, size_t jobs,
                    double alpha_mult, double nu_start_mult,
                    int has_nu_override, double nu_override, int geq_order,
                    char *err, size_t errlen) {
  if (nx < 3 || ny < 3) {
    snprintf(err, errlen, "nx and ny must be >= 3 (got %zu x %zu)", nx, ny);
    return NULL;
  }
  double lz = (double)(ny - 1);
  double lx = (double)(nx - 1);
  double c_scale = 1.7;
  double nu_cal = pow(utarget * sqrt(pr) * lz / (c_scale * sqrt(ra)), 2.0 / 3.0);
  double nu;
  if (has_nu_override) {
    /* solver.rs:129-133 */
    if (!(nu_override > 0.0 && isfinite(nu_override))) {
      snprintf(err, errlen, "--nu must be > 0 and finite");
      return NULL;
    }
    if (!(0.5 + 3.0 * nu_override > 0.5005)) {
      snprintf(err, errlen, "--nu %g too low: tau+=%.5f close to 0.5",
               nu_override, 0.5 + 3.0 * nu_override);
      return NULL;
    }
    nu = nu_override;
  } else {
    nu = fmax(nu_cal, 0.002); /* solver.rs:135 */
  }
  double alpha_phys = nu / pr;
  double alpha = alpha_phys * alpha_mult;
  double beta_eff = ra * nu * alpha_phys / (lz * lz * lz); /* lz.powi(3) */
  double tau = 0.5 + 3.0 * nu;
  double tau_t = 0.5 + 3.0 * alpha;
  if (!has_nu_override && !(tau > 0.505)) { /* solver.rs:151-153 */
    snprintf(err, errlen, "tau=%g too close to 0.5", tau);
    return NULL;
  }
  if (!(tau_t > 0.505)) { /* solver.rs:154 */
    snprintf(err, errlen, "tau_t=%g too close to 0.5", tau_t);
    return NULL;
  }

  orc_solver *s = (orc_solver *)xcalloc(1, sizeof(orc_solver));
  s->nx = nx;
  s->ny = ny;
  s->n = nx * ny;
  s->f = (node9 *)xcalloc(s->n, sizeof(node9));
  s->g = (node9 *)xcalloc(s->n, sizeof(node9));
  s->f_old = (node9 *)xcalloc(s->n, sizeof(node9));
  s->g_old = (node9 *)xcalloc(s->n, sizeof(node9));
  s->rho = (double *)xcalloc(s->n, sizeof(double));
  s->ux = (double *)xcalloc(s->n, sizeof(double));
  s->uy = (double *)xcalloc(s->n, sizeof(double));
  s->temp = (double *)xcalloc(s->n, sizeof(double));
  s->nu_t = (double *)xcalloc(s->n, sizeof(double));
  s->fy_node = (double *)xcalloc(s->n, sizeof(double));
  s->lz = lz;
  s->lx = lx;
  s->beta_eff = beta_eff;
  s->nu = nu;
  s->alpha = alpha;
  s->alpha_phys = alpha_phys;
  s->cs = cs;
  s->pr_t = pr_t;
  s->tau = tau;
  s->tau_t = tau_t;
  s->t_ref = 0.0; /* ThermalD2Q9::new(.., 1.0, 0.0, -1.0); cf. octree/cavity.rs:85 */
  s->jobs = jobs ? jobs : 1;
  s->time_step = 0;
  {
    size_t su = ny / 10; /* solver.rs:194 */
    if (su < 200) su = 200;
    if (su > 2000) su = 2000;
    s->spin_up_steps = su;
  }
  s->nu_start_mult = fmax(nu_start_mult, 1.0);
  s->geq_order = geq_order == 1 ? 1 : 2;
  s->dynamic_smag = 0; /* solver.rs:199-208 */
  s->smag_alpha = 0.5;
  s->nu_sm_min = -0.1;
  s->eta_test = 2.0;

  /* init loop, iy outer / ix inner (solver.rs:160-176) */
  uint64_t rng = seed | 1ULL;
  for (size_t iy = 0; iy < ny; ++iy) {
    for (size_t ix = 0; ix < nx; ++ix) {
      size_t k = iy * nx + ix;
      double xf = (double)ix / (double)(nx - 1);
      double t_lin = T_HOT + (T_COLD - T_HOT) * xf;
      double pert = 1.0e-3 * (2.0 * orc_rng_next_f64(&rng) - 1.0);
      double t = t_lin + pert;
      s->temp[k] = t;
      s->rho[k] = 1.0;
      s->ux[k] = 0.0;
      s->uy[k] = 0.0;
      for (int i = 0; i < 9; ++i) {
        s->f[k].f[i] = orc_feq(i, 1.0, 0.0, 0.0);
        s->g[k].f[i] = orc_geq(i, t, 0.0, 0.0, s->geq_order);
      }
    }
  }
  return s;
}

void orc_free(orc_solver *s) {
  if (!s) return;
  free(s->f);
  free(s->g);
  free(s->f_old);
  free(s->g_old);
  free(s->rho);
  free(s->ux);
  free(s->uy);
  free(s->temp);
  free(s->nu_t);
  free(s->fy_node);
  free(s->theta_hot);
  free(s->nu_t_pre);
  free(s->alpha_t_pre);
  free(s);
}

void orc_set_jobs(orc_solver *s, size_t jobs) { s->jobs = jobs ? jobs : 1; }

void orc_set_hot_wall(orc_solver *s, const double *theta_hot) {
  free(s->theta_hot);
  s->theta_hot = NULL;
  if (theta_hot) {
    s->theta_hot = (double *)xcalloc(s->ny, sizeof(double));
    memcpy(s->theta_hot, theta_hot, s->ny * sizeof(double));
  }
}

/* The Smagorinsky constant of later collisions (the `cs` field read by
 * collide_node, solver.rs:1120); the window objective of the differentiation
 * form runs its steps with a chosen Cs. */
void orc_set_cs(orc_solver *s, double cs) { s->cs = cs; }

/* ---- init_from_fields (src/uniform/solver.rs:232-262) --------------------- */
int orc_init_from_fields(orc_solver *s, const double *rho, const double *ux,
                         const double *uy, const double *t, size_t n) {
  if (n != s->n) return -1; /* solver.rs:236-240 asserts */
  for (size_t k = 0; k < s->n; ++k) {
    double r = rho[k], un = ux[k], vn = uy[k], tn = t[k];
    s->rho[k] = r;
    s->ux[k] = un;
    s->uy[k] = vn;
    s->temp[k] = tn;
    for (int i = 0; i < 9; ++i) {
      s->f[k].f[i] = orc_feq(i, r, un, vn);
      s->g[k].f[i] = orc_geq(i, tn, un, vn, s->geq_order);
    }
  }
  orc_enforce_wall_velocity(s);
  s->time_step = 0;
  return 0;
}

/* ---- per-step scalars: collide() prologue + ramp_nu_alpha
 *      (src/uniform/solver.rs:390-410, :433-442) ---------------------------- */
typedef struct {
  double tau, beta, t_ref, cs, pr_t, nu, alpha, lambda;
} collide_ctx;

static collide_ctx make_ctx(const orc_solver *s, uint64_t step) {
  collide_ctx c;
  double beta = s->beta_eff;
  double nu_eff = s->nu, alpha_eff = s->alpha;
  if (step < s->spin_up_steps) {
    double x = (double)step / (double)s->spin_up_steps;
    beta = s->beta_eff * x;
    double nu_start = s->nu * s->nu_start_mult;
    nu_eff = nu_start * pow(s->nu / nu_start, x);
    double alpha_start = s->alpha * s->nu_start_mult;
    alpha_eff = alpha_start * pow(s->alpha / alpha_start, x);
  }
  c.tau = 0.5 + 3.0 * fmax(nu_eff, 0.002 / 3.0);
  c.beta = beta;
  c.t_ref = s->t_ref;
  c.cs = s->cs;
  c.pr_t = s->pr_t;
  c.nu = nu_eff;
  c.alpha = alpha_eff;
  c.lambda = 3.0 / 16.0;
  return c;
}

void orc_step_scalars(const orc_solver *s, uint64_t step, double out[4]) {
  collide_ctx c = make_ctx(s, step);
  out[0] = c.beta;
  out[1] = c.nu;
  out[2] = c.alpha;
  out[3] = c.tau;
}

/* ---- collide_node (src/uniform/solver.rs:1054-1185) ----------------------- */
static void collide_node_pre(double *f, double *g, double *nu_t_out, double *fy_out,
                             double rho, double ux, double uy, double temp,
                             const double *nu_t_pre, const double *alpha_t_pre,
                             const collide_ctx *c, int geq_order) {
  /* boussinesq_force(rho, beta, 1.0, T, t_ref) = rho*beta*(T - t_ref) */
  double fy = rho * c->beta * (temp - c->t_ref);
  *fy_out = fy;

  double feq_[9], geq_[9];
  for (int i = 0; i < 9; ++i) {
    feq_[i] = orc_feq(i, rho, ux, uy);
    geq_[i] = orc_geq(i, temp, ux, uy, geq_order);
  }
  double pi_xx = 0.0, pi_xy = 0.0, pi_yy = 0.0;
  for (int i = 0; i < 9; ++i) {
    double cx = (double)CX[i], cy = (double)CY[i];
    double fneq = f[i] - feq_[i];
    pi_xx += fneq * cx * cx;
    pi_xy += fneq * cx * cy;
    pi_yy += fneq * cy * cy;
  }
  double pi2 = pi_xx * pi_xx + 2.0 * pi_xy * pi_xy + pi_yy * pi_yy;
  double s_mag = sqrt(pi2) / (2.0 * CS2 * c->tau * fmax(rho, 1e-12));
  /* compute_turbulent_viscosity(s_mag, cs) = (cs*dx)^2*|S|, dx = 1, or the
   * precomputed dynamic value (solver.rs:1118-1121) */
  double nu_t = nu_t_pre ? *nu_t_pre : c->cs * c->cs * s_mag;
  *nu_t_out = nu_t;

  double tau_plus = 0.5 + 3.0 * (c->nu + nu_t);
  double tau_minus = 0.5 + c->lambda / (tau_plus - 0.5);
  double alpha_t = alpha_t_pre ? *alpha_t_pre : nu_t / c->pr_t; /* solver.rs:1136-1139 */
  double tau_t_plus = 0.5 + 3.0 * (c->alpha + alpha_t);
  double tau_t_minus = tau_t_plus;
  double a_f = 0.5 * (1.0 / tau_plus + 1.0 / tau_minus);
  double b_f = 0.5 * (1.0 / tau_plus - 1.0 / tau_minus);
  double a_g = 0.5 * (1.0 / tau_t_plus + 1.0 / tau_t_minus);
  double b_g = 0.5 * (1.0 / tau_t_plus - 1.0 / tau_t_minus);

  double d_f[9], d_g[9];
  for (int i = 0; i < 9; ++i) {
    d_f[i] = f[i] - feq_[i];
    d_g[i] = g[i] - geq_[i];
  }
  for (int i = 0; i < 9; ++i) {
    double cx = (double)CX[i], cy = (double)CY[i];
    int opp = OPP[i];
    double ei_u = cx * ux + cy * uy;
    double s_i = W[i] * ((cy - uy) / CS2 + ei_u * cy / (CS2 * CS2)) * fy;
    double s_minus = W[i] * (cy / CS2) * fy;
    double s_plus = s_i - s_minus;
    f[i] += -a_f * d_f[i] - b_f * d_f[opp] + (1.0 - 0.5 / tau_minus) * s_minus +
            (1.0 - 0.5 / tau_plus) * s_plus;
  }
  for (int i = 0; i < 9; ++i) {
    int opp = OPP[i];
    g[i] += -a_g * d_g[i] - b_g * d_g[opp];
  }
}

static void collide_node(double *f, double *g, double *nu_t_out, double *fy_out,
                         double rho, double ux, double uy, double temp,
                         const collide_ctx *c, int geq_order) {
  collide_node_pre(f, g, nu_t_out, fy_out, rho, ux, uy, temp, NULL, NULL, c, geq_order);
}

void orc_collide_node(double f[9], double g[9], double *nu_t_out,
                      double *fy_out, double rho, double ux, double uy,
                      double temp, double tau_mol, double beta, double t_ref,
                      double cs, double pr_t, double nu, double alpha,
                      int geq_order) {
  collide_ctx c = {tau_mol, beta, t_ref, cs, pr_t, nu, alpha, 3.0 / 16.0};
  collide_node(f, g, nu_t_out, fy_out, rho, ux, uy, temp, &c,
               geq_order == 1 ? 1 : 2);
}

/* ---- collide (src/uniform/solver.rs:421-540); rows in parallel when
 *      jobs > 1, bitwise equal to the serial loop (solver.rs:462-511) ------- */
void orc_collide(orc_solver *s, uint64_t step) {
  collide_ctx c = make_ctx(s, step);
  const long nx = (long)s->nx, ny = (long)s->ny;
  const int go = s->geq_order;
#pragma omp parallel for schedule(static) num_threads((int)s->jobs) if (s->jobs > 1)
  for (long iy = 0; iy < ny; ++iy) {
    for (long ix = 0; ix < nx; ++ix) {
      long k = iy * nx + ix;
      collide_node_pre(s->f[k].f, s->g[k].f, &s->nu_t[k], &s->fy_node[k], s->rho[k],
                       s->ux[k], s->uy[k], s->temp[k],
                       s->nu_t_pre ? &s->nu_t_pre[k] : NULL,
                       s->alpha_t_pre ? &s->alpha_t_pre[k] : NULL, &c, go);
    }
  }
}

/* ---- stream: pull with full-way bounce-back (src/uniform/solver.rs:838-913) */
void orc_stream(orc_solver *s) {
  const long nx = (long)s->nx, ny = (long)s->ny;
  memcpy(s->f_old, s->f, s->n * sizeof(node9));
  memcpy(s->g_old, s->g, s->n * sizeof(node9));
  const node9 *f_old = s->f_old, *g_old = s->g_old;
#pragma omp parallel for schedule(static) num_threads((int)s->jobs) if (s->jobs > 1)
  for (long iy = 0; iy < ny; ++iy) {
    for (long ix = 0; ix < nx; ++ix) {
      long k = iy * nx + ix;
      for (int i = 0; i < 9; ++i) {
        long sx = ix - CX[i], sy = iy - CY[i];
        if (sx >= 0 && sx < nx && sy >= 0 && sy < ny) {
          long src = sy * nx + sx;
          s->f[k].f[i] = f_old[src].f[i];
          s->g[k].f[i] = g_old[src].f[i];
        } else {
          s->f[k].f[i] = f_old[k].f[OPP[i]];
          s->g[k].f[i] = g_old[k].f[i];
        }
      }
    }
  }
}

/* ---- apply_bounce_back, test-only helper (src/uniform/solver.rs:920-946) -- */
static void swap_opposite(orc_solver *s, size_t k) {
  for (int i = 1; i < 5; ++i) {
    int j = OPP[i];
    double tmp = s->f[k].f[i];
    s->f[k].f[i] = s->f[k].f[j];
    s->f[k].f[j] = tmp;
  }
}

void orc_apply_bounce_back(orc_solver *s) {
  size_t nx = s->nx, ny = s->ny;
  for (size_t iy = 0; iy < ny; ++iy) {
    swap_opposite(s, iy * nx);
    swap_opposite(s, iy * nx + nx - 1);
  }
  for (size_t ix = 0; ix < nx; ++ix) {
    swap_opposite(s, ix);
    swap_opposite(s, (ny - 1) * nx + ix);
  }
}

/* ---- apply_thermal_boundary_conditions (src/uniform/solver.rs:951-988) ---- */
void orc_thermal_bc(orc_solver *s) {
  size_t nx = s->nx, ny = s->ny;
  for (size_t iy = 0; iy < ny; ++iy) {
    double th = s->theta_hot ? s->theta_hot[iy] : T_HOT;
    size_t k0 = iy * nx, k1 = iy * nx + (nx - 1);
    for (int i = 0; i < 9; ++i) {
      s->g[k0].f[i] = orc_geq(i, th, 0.0, 0.0, s->geq_order);
      s->g[k1].f[i] = orc_geq(i, T_COLD, 0.0, 0.0, s->geq_order);
    }
  }
  for (size_t ix = 0; ix < nx; ++ix) {
    s->g[ix] = s->g[nx + ix];
    s->g[(ny - 1) * nx + ix] = s->g[(ny - 2) * nx + ix];
  }
}

/* ---- update_macroscopic with the Guo velocity correction; the explicit
 *      per-node formulas of src/uniform/solver.rs:307-349 ------------------ */
void orc_update_macroscopic(orc_solver *s) {
  const long nx = (long)s->nx, ny = (long)s->ny;
#pragma omp parallel for schedule(static) num_threads((int)s->jobs) if (s->jobs > 1)
  for (long iy = 0; iy < ny; ++iy) {
    for (long ix = 0; ix < nx; ++ix) {
      long k = iy * nx + ix;
      double r = 0.0, un = 0.0, vn = 0.0, tn = 0.0;
      for (int j = 0; j < 9; ++j) {
        double fj = s->f[k].f[j], gj = s->g[k].f[j];
        r += fj;
        un += (double)CX[j] * fj;
        vn += (double)CY[j] * fj;
        tn += gj;
      }
      double fy = s->fy_node[k];
      if (r > F64_EPSILON) {
        un /= r;
        vn /= r;
        vn += 0.5 * fy / r;
      }
      s->rho[k] = r;
      s->ux[k] = un;
      s->uy[k] = vn;
      s->temp[k] = tn;
    }
  }
}

/* ---- enforce_wall_velocity (src/uniform/solver.rs:992-1009) --------------- */
void orc_enforce_wall_velocity(orc_solver *s) {
  size_t nx = s->nx, ny = s->ny;
  for (size_t iy = 0; iy < ny; ++iy) {
    size_t a = iy * nx, b = iy * nx + nx - 1;
    s->ux[a] = s->uy[a] = 0.0;
    s->ux[b] = s->uy[b] = 0.0;
  }
  for (size_t ix = 0; ix < nx; ++ix) {
    size_t a = ix, b = (ny - 1) * nx + ix;
    s->ux[a] = s->uy[a] = 0.0;
    s->ux[b] = s->uy[b] = 0.0;
  }
}


/* ---- dynamic mixed-scale SGS (src/uniform/solver.rs:562-828, helpers
 *      :1195-1261) ---------------------------------------------------------- */
#define C_DYN_MAX 0.04
#define TAU_TRT_FLOOR 0.505

void orc_set_dynamic_sgs(orc_solver *s, int on, double smag_alpha, double nu_sm_min,
                         double eta_test) { /* solver.rs:215-220 */
  s->dynamic_smag = on ? 1 : 0;
  s->smag_alpha = smag_alpha;
  s->nu_sm_min = nu_sm_min;
  s->eta_test = eta_test;
}

double orc_lilly_coefficient(double num, double den) { /* solver.rs:1251-1261 */
  if (!isfinite(den) || !isfinite(num) || den <= 0.0) return 0.0;
  double c = 0.5 * num / den;
  if (isfinite(c)) return c < 0.0 ? 0.0 : (c > C_DYN_MAX ? C_DYN_MAX : c);
  return 0.0;
}

double orc_mixed_factor(double a, double b, double alpha) { /* solver.rs:1195-1197 */
  return pow(a, alpha) * pow(b, 1.0 - alpha);
}

/* box_filter (solver.rs:1204-1224): window [i - half_left, i + width -
 * half_left) clamped to the domain, sum in (yy, xx) order, divided by count. */
static void box_filter(const double *field, size_t nx, size_t ny, size_t width, double *out) {
  size_t half_left = (width - 1) / 2;
  for (size_t iy = 0; iy < ny; ++iy) {
    size_t y0 = iy >= half_left ? iy - half_left : 0;
    size_t y1 = iy + width - half_left < ny ? iy + width - half_left : ny;
    for (size_t ix = 0; ix < nx; ++ix) {
      size_t x0 = ix >= half_left ? ix - half_left : 0;
      size_t x1 = ix + width - half_left < nx ? ix + width - half_left : nx;
      double sum = 0.0;
      for (size_t yy = y0; yy < y1; ++yy)
        for (size_t xx = x0; xx < x1; ++xx) sum += field[yy * nx + xx];
      double cnt = (double)((y1 - y0) * (x1 - x0));
      out[iy * nx + ix] = sum / cnt;
    }
  }
}

static double ddx(const double *f, size_t nx, size_t ix, size_t iy) { /* :1229-1234 */
  size_t xm = ix == 0 ? 0 : ix - 1;
  size_t xp = ix + 1 >= nx ? nx - 1 : ix + 1;
  return 0.5 * (f[iy * nx + xp] - f[iy * nx + xm]);
}

static double ddy(const double *f, size_t nx, size_t ny, size_t ix, size_t iy) { /* :1238-1242 */
  size_t ym = iy == 0 ? 0 : iy - 1;
  size_t yp = iy + 1 >= ny ? ny - 1 : iy + 1;
  return 0.5 * (f[yp * nx + ix] - f[ym * nx + ix]);
}

/* The model on given arrays: f k-major [k][9] (post-stream), the moments,
 * the molecular tau of the collision and the molecular nu / alpha of the
 * floors.  compute_dynamic_sgs (solver.rs:562-828) on the uniform solver;
 * compute_dynamic_sgs_level (src/octree/smag.rs:144-416, stated there as an
 * exact port of the uniform method) on one octree level.  stats: mean and
 * max of C_nu, mean and max of C_kappa (sequential sums). */
void orc_dynamic_model(size_t nx, size_t ny, const double *f9, const double *rho,
                       const double *ux, const double *uy, const double *temp, double tau,
                       double smag_alpha, double eta_test, double nu_mol, double alpha_mol,
                       double nu_sm_min, double *nu_t_out, double *alpha_t_out,
                       double stats[4]) {
  const size_t n = nx * ny;
  const double alpha_exp = smag_alpha;
  const double eta = fmax(eta_test, 1.0);
  /* (eta.round() as usize).max(1); every width from 2*max(nx, ny) + 1 on
   * averages the whole domain, and the cap keeps the cast defined */
  const double cap = 2.0 * (double)(nx > ny ? nx : ny) + 1.0;
  const double wr = round(eta);
  size_t width = (size_t)(wr < cap ? wr : cap);
  if (width < 1) width = 1;
  const double dx = 1.0, dtest = eta * dx, delta_t = DELTA_T;
  enum { UBX, UBY, TB, PXX, PXY, PYY, PTX, PTY, LXX, LXY, LYY, LTX, LTY,
         SHXX, SHXY, SHYY, SMH, SBXX, SBXY, SBYY, SMB, GTX, GTY, GBX, GBY,
         QSM, PHI, TMH, TMB, QBAR, PHIBAR, NUMV, DENV, NUMK, DENK, NUG, KAPG,
         NUMVB, DENVB, NUMKB, DENKB, NARR };
  double *a = (double *)xcalloc(n * NARR, sizeof(double));
#define A(i) (a + (size_t)(i) * n)
  box_filter(ux, nx, ny, width, A(UBX));
  box_filter(uy, nx, ny, width, A(UBY));
  box_filter(temp, nx, ny, width, A(TB));
  for (size_t k = 0; k < n; ++k) {
    A(PXX)[k] = ux[k] * ux[k];
    A(PXY)[k] = ux[k] * uy[k];
    A(PYY)[k] = uy[k] * uy[k];
    A(PTX)[k] = ux[k] * temp[k];
    A(PTY)[k] = uy[k] * temp[k];
  }
  box_filter(A(PXX), nx, ny, width, A(LXX));
  box_filter(A(PXY), nx, ny, width, A(LXY));
  box_filter(A(PYY), nx, ny, width, A(LYY));
  box_filter(A(PTX), nx, ny, width, A(LTX));
  box_filter(A(PTY), nx, ny, width, A(LTY));
  for (size_t iy = 0; iy < ny; ++iy)
    for (size_t ix = 0; ix < nx; ++ix) {
      size_t k = iy * nx + ix;
      double rho_k = rho[k], ux_k = ux[k], uy_k = uy[k], t_k = temp[k];
      double pi_xx = 0.0, pi_xy = 0.0, pi_yy = 0.0;
      for (int i = 0; i < 9; ++i) {
        double cx = (double)CX[i], cy = (double)CY[i];
        double fneq = f9[9 * k + (size_t)i] - orc_feq(i, rho_k, ux_k, uy_k);
        pi_xx += fneq * cx * cx;
        pi_xy += fneq * cx * cy;
        pi_yy += fneq * cy * cy;
      }
      double rho_m = fmax(rho_k, 1e-12);
      double denom = 2.0 * CS2 * tau * rho_m;
      double pi2 = pi_xx * pi_xx + 2.0 * pi_xy * pi_xy + pi_yy * pi_yy;
      A(SHXX)[k] = -pi_xx / denom;
      A(SHXY)[k] = -pi_xy / denom;
      A(SHYY)[k] = -pi_yy / denom;
      A(SMH)[k] = sqrt(pi2) / denom;
      double sxx = ddx(A(UBX), nx, ix, iy);
      double syy = ddy(A(UBY), nx, ny, ix, iy);
      double sxy = 0.5 * (ddx(A(UBY), nx, ix, iy) + ddy(A(UBX), nx, ny, ix, iy));
      A(SBXX)[k] = sxx;
      A(SBXY)[k] = sxy;
      A(SBYY)[k] = syy;
      A(SMB)[k] = sqrt(sxx * sxx + 2.0 * sxy * sxy + syy * syy);
      A(GTX)[k] = ddx(temp, nx, ix, iy);
      A(GTY)[k] = ddy(temp, nx, ny, ix, iy);
      A(GBX)[k] = ddx(A(TB), nx, ix, iy);
      A(GBY)[k] = ddy(A(TB), nx, ny, ix, iy);
      double dux = ux_k - A(UBX)[k], duy = uy_k - A(UBY)[k];
      double du2 = dux * dux + duy * duy;
      A(QSM)[k] = sqrt(0.5 * du2) / 2.0;
      double dt_local = t_k - A(TB)[k];
      A(PHI)[k] = 0.5 * fabs(dt_local) * sqrt(0.5 * du2);
      double th_x = A(GTX)[k] * A(SHXX)[k] + A(GTY)[k] * A(SHXY)[k];
      double th_y = A(GTX)[k] * A(SHXY)[k] + A(GTY)[k] * A(SHYY)[k];
      A(TMH)[k] = sqrt(2.0 * (th_x * th_x + th_y * th_y));
      double tb_x = A(GBX)[k] * A(SBXX)[k] + A(GBY)[k] * A(SBXY)[k];
      double tb_y = A(GBX)[k] * A(SBXY)[k] + A(GBY)[k] * A(SBYY)[k];
      A(TMB)[k] = sqrt(2.0 * (tb_x * tb_x + tb_y * tb_y));
    }
  box_filter(A(QSM), nx, ny, width, A(QBAR));
  box_filter(A(PHI), nx, ny, width, A(PHIBAR));
  for (size_t k = 0; k < n; ++k) {
    double nu_mix_t = pow(dtest * dtest * A(SMB)[k], alpha_exp) *
                      pow(dtest * A(QBAR)[k], 1.0 - alpha_exp);
    double nu_mix_g = pow(dx * dx * A(SMH)[k], alpha_exp) * pow(dx * A(QSM)[k], 1.0 - alpha_exp);
    double kap_mix_t = pow(dtest * dtest * dtest / delta_t * A(TMB)[k], alpha_exp) *
                       pow(dtest / delta_t * A(PHIBAR)[k], 1.0 - alpha_exp);
    double kap_mix_g = pow(dx * dx * dx / delta_t * A(TMH)[k], alpha_exp) *
                       pow(dx / delta_t * A(PHI)[k], 1.0 - alpha_exp);
    A(NUG)[k] = nu_mix_g;
    A(KAPG)[k] = kap_mix_g;
    double mxx = nu_mix_t * A(SBXX)[k] - nu_mix_g * A(SHXX)[k];
    double mxy = nu_mix_t * A(SBXY)[k] - nu_mix_g * A(SHXY)[k];
    double myy = nu_mix_t * A(SBYY)[k] - nu_mix_g * A(SHYY)[k];
    double mx = kap_mix_t * A(GBX)[k] - kap_mix_g * A(GTX)[k];
    double my = kap_mix_t * A(GBY)[k] - kap_mix_g * A(GTY)[k];
    double lxx = A(LXX)[k] - A(UBX)[k] * A(UBX)[k];
    double lxy = A(LXY)[k] - A(UBX)[k] * A(UBY)[k];
    double lyy = A(LYY)[k] - A(UBY)[k] * A(UBY)[k];
    double lkk3 = (lxx + lyy) / 3.0;
    double ldxx = lxx - lkk3, ldyy = lyy - lkk3;
    double ljx = A(LTX)[k] - A(UBX)[k] * A(TB)[k];
    double ljy = A(LTY)[k] - A(UBY)[k] * A(TB)[k];
    A(NUMV)[k] = ldxx * mxx + 2.0 * lxy * mxy + ldyy * myy;
    A(DENV)[k] = mxx * mxx + 2.0 * mxy * mxy + myy * myy;
    A(NUMK)[k] = ljx * mx + ljy * my;
    A(DENK)[k] = mx * mx + my * my;
  }
  box_filter(A(NUMV), nx, ny, width, A(NUMVB));
  box_filter(A(DENV), nx, ny, width, A(DENVB));
  box_filter(A(NUMK), nx, ny, width, A(NUMKB));
  box_filter(A(DENK), nx, ny, width, A(DENKB));
  double cnu_sum = 0.0, cnu_max = 0.0, ck_sum = 0.0, ck_max = 0.0;
  const double tau_floor_margin = (TAU_TRT_FLOOR - 0.5) / 3.0;
  for (size_t k = 0; k < n; ++k) {
    double cnu = orc_lilly_coefficient(A(NUMVB)[k], A(DENVB)[k]);
    double ck = orc_lilly_coefficient(A(NUMKB)[k], A(DENKB)[k]);
    cnu_sum += cnu;
    if (cnu > cnu_max) cnu_max = cnu;
    ck_sum += ck;
    if (ck > ck_max) ck_max = ck;
    double nt = cnu * A(NUG)[k];
    double nu_t_min = fmax(nu_sm_min * nu_mol, tau_floor_margin - nu_mol);
    nu_t_out[k] = nt < nu_t_min ? nu_t_min : nt;
    double at = ck * A(KAPG)[k];
    double alpha_t_min = fmax(nu_sm_min * alpha_mol, tau_floor_margin - alpha_mol);
    alpha_t_out[k] = at < alpha_t_min ? alpha_t_min : at;
  }
  stats[0] = cnu_sum / (double)n;
  stats[1] = cnu_max;
  stats[2] = ck_sum / (double)n;
  stats[3] = ck_max;
#undef A
  free(a);
}

void orc_compute_dynamic_sgs(orc_solver *s) {
  collide_ctx c = make_ctx(s, s->time_step);
  if (!s->nu_t_pre) s->nu_t_pre = (double *)xcalloc(s->n, sizeof(double));
  if (!s->alpha_t_pre) s->alpha_t_pre = (double *)xcalloc(s->n, sizeof(double));
  double st[4];
  orc_dynamic_model(s->nx, s->ny, &s->f[0].f[0], s->rho, s->ux, s->uy, s->temp, c.tau,
                    s->smag_alpha, s->eta_test, s->nu, s->alpha, s->nu_sm_min, s->nu_t_pre,
                    s->alpha_t_pre, st);
  s->cs_dyn_mean = st[0];
  s->cs_dyn_max = st[1];
  s->kappa_dyn_mean = st[2];
  s->kappa_dyn_max = st[3];
}

void orc_dynamic_stats(const orc_solver *s, double out[4]) {
  out[0] = s->cs_dyn_mean;
  out[1] = s->cs_dyn_max;
  out[2] = s->kappa_dyn_mean;
  out[3] = s->kappa_dyn_max;
}

double *orc_sgs_pre(orc_solver *s, int which) {
  return which == 0 ? s->nu_t_pre : s->alpha_t_pre;
}

/* ---- step (src/uniform/solver.rs:279-293) --------------------------------- */
void orc_step(orc_solver *s) {
  if (s->dynamic_smag) orc_compute_dynamic_sgs(s); /* solver.rs:284-286 */
  orc_collide(s, s->time_step);
  orc_stream(s);
  orc_thermal_bc(s);
  orc_update_macroscopic(s);
  orc_enforce_wall_velocity(s);
  s->time_step += 1;
}

void orc_run(orc_solver *s, uint64_t nsteps) {
  for (uint64_t i = 0; i < nsteps; ++i) orc_step(s);
}

/* ---- accessors ------------------------------------------------------------ */
double *orc_field(orc_solver *s, int which) {
  switch (which) {
    case ORC_RHO: return s->rho;
    case ORC_UX: return s->ux;
    case ORC_UY: return s->uy;
    case ORC_T: return s->temp;
    case ORC_NU_T: return s->nu_t;
    case ORC_FY: return s->fy_node;
    default: return NULL;
  }
}

double *orc_pops(orc_solver *s, int which) {
  return which == ORC_F ? (double *)s->f : which == ORC_G ? (double *)s->g : NULL;
}

size_t orc_nx(const orc_solver *s) { return s->nx; }
size_t orc_ny(const orc_solver *s) { return s->ny; }
uint64_t orc_time_step(const orc_solver *s) { return s->time_step; }
void orc_set_time_step(orc_solver *s, uint64_t t) { s->time_step = t; }

void orc_derived(const orc_solver *s, double out[13]) {
  out[0] = s->lz;
  out[1] = s->lx;
  out[2] = s->nu;
  out[3] = s->alpha;
  out[4] = s->alpha_phys;
  out[5] = s->beta_eff;
  out[6] = s->tau;
  out[7] = s->tau_t;
  out[8] = (double)s->spin_up_steps;
  out[9] = s->cs;
  out[10] = s->pr_t;
  out[11] = s->nu_start_mult;
  out[12] = s->t_ref;
}

/* ---- diagnostics ----------------------------------------------------------- */
double orc_total_mass(const orc_solver *s) { /* solver.rs:1012-1014 */
  double m = 0.0;
  for (size_t k = 0; k < s->n; ++k) m += s->rho[k];
  return m;
}

double orc_total_f_mass(const orc_solver *s) { /* solver.rs:1018-1020 */
  double m = 0.0;
  for (size_t k = 0; k < s->n; ++k)
    for (int i = 0; i < 9; ++i) m += s->f[k].f[i];
  return m;
}

double orc_nusselt_hot(const orc_solver *s) { /* post.rs:68-78 */
  double sum = 0.0;
  for (size_t iy = 0; iy < s->ny; ++iy) {
    double t1 = s->temp[iy * s->nx + 1], t2 = s->temp[iy * s->nx + 2];
    sum += fabs(t1 - t2);
  }
  double avg = sum / (double)s->ny;
  return avg * s->lz / DELTA_T;
}

double orc_nusselt_hot_wall_link(const orc_solver *s) { /* post.rs:86-96 */
  double sum = 0.0;
  for (size_t iy = 0; iy < s->ny; ++iy) {
    double t0 = s->temp[iy * s->nx], t1 = s->temp[iy * s->nx + 1];
    sum += fabs(t0 - t1);
  }
  double avg = sum / (double)s->ny;
  return avg * s->lz / DELTA_T;
}

double orc_nusselt_cold(const orc_solver *s) { /* post.rs:100-111 */
  double sum = 0.0;
  size_t nx = s->nx;
  for (size_t iy = 0; iy < s->ny; ++iy) {
    double t0 = s->temp[iy * nx + nx - 3], t1 = s->temp[iy * nx + nx - 2];
    sum += fabs(t1 - t0);
  }
  double avg = sum / (double)s->ny;
  return avg * s->lz / DELTA_T;
}

double orc_nusselt_volume(const orc_solver *s) { /* post.rs:116-132 */
  double sum = 0.0;
  size_t nx = s->nx, ny = s->ny;
  for (size_t iy = 1; iy < ny - 1; ++iy)
    for (size_t ix = 1; ix < nx - 1; ++ix) {
      size_t k = iy * nx + ix;
      sum += s->uy[k] * s->temp[k];
    }
  double conv = sum / (double)((nx - 2) * (ny - 2));
  return 1.0 + conv * s->lz / DELTA_T;
}

double orc_max_uy(const orc_solver *s) { /* post.rs:153-155 */
  double m = 0.0;
  for (size_t k = 0; k < s->n; ++k) m = fmax(m, fabs(s->uy[k]));
  return m;
}

double orc_max_uy_on_row(const orc_solver *s, size_t iy, double *x_norm) {
  /* post.rs:137-150 */
  double best = 0.0;
  size_t best_ix = 1;
  for (size_t ix = 1; ix < s->nx - 1; ++ix) {
    double v = fabs(s->uy[iy * s->nx + ix]);
    if (v > best) {
      best = v;
      best_ix = ix;
    }
  }
  if (x_norm) *x_norm = (double)best_ix / (double)(s->nx - 1);
  return best;
}

double orc_max_u(const orc_solver *s) { /* main.rs:192-196 */
  double m = 0.0;
  for (size_t k = 0; k < s->n; ++k) m = fmax(m, fmax(fabs(s->ux[k]), fabs(s->uy[k])));
  return m;
}

double orc_nu_t_max(const orc_solver *s) { /* main.rs:170-173 */
  double m = 0.0;
  for (size_t k = 0; k < s->n; ++k) m = fmax(m, fabs(s->nu_t[k]));
  return m;
}

double orc_wall_shear_hot(const orc_solver *s, size_t iy) { /* post.rs:182-187 */
  size_t k1 = iy * s->nx + 1, k0 = iy * s->nx;
  double nu_eff = s->nu + s->nu_t[k1];
  double du = fabs(s->uy[k1] - s->uy[k0]);
  return nu_eff * du;
}

double orc_y_plus_hot(const orc_solver *s) { /* post.rs:190-199 */
  double sum = 0.0;
  for (size_t iy = 1; iy < s->ny - 1; ++iy) {
    double tau_w = orc_wall_shear_hot(s, iy);
    double u_tau = sqrt(fmax(tau_w, 0.0));
    sum += 0.5 * u_tau / s->nu; /* turbulence::y_plus(y, u_tau, nu) = y*u_tau/nu */
  }
  return sum / (double)(s->ny - 2);
}

int orc_check_finite(const orc_solver *s) { /* post.rs:406-411 */
  for (size_t k = 0; k < s->n; ++k)
    if (!isfinite(s->temp[k])) return 0;
  for (size_t k = 0; k < s->n; ++k)
    if (!isfinite(s->ux[k])) return 0;
  for (size_t k = 0; k < s->n; ++k)
    if (!isfinite(s->uy[k])) return 0;
  for (size_t k = 0; k < s->n; ++k)
    if (!(isfinite(s->rho[k]) && s->rho[k] > 0.0)) return 0;
  return 1;
}

/* ---- Poiseuille channel from collide_node (src/uniform/tests.rs:246-376,
 *      axes swapped: periodic in y, bounce-back walls at x = -0.5, nx - 0.5,
 *      so H = nx; the body force is fy*rho through boussinesq_force with
 *      beta = fy, T = 1, t_ref = 0) ---------------------------------------- */
double orc_poiseuille_channel(size_t nx, size_t ny, double nu, double fy,
                              uint64_t nsteps) {
  size_t n = nx * ny;
  node9 *f = (node9 *)xcalloc(n, sizeof(node9));
  node9 *fp = (node9 *)xcalloc(n, sizeof(node9));
  node9 *g = (node9 *)xcalloc(n, sizeof(node9));
  for (size_t k = 0; k < n; ++k)
    for (int i = 0; i < 9; ++i) f[k].f[i] = W[i];
  collide_ctx c = {0.5 + 3.0 * nu, fy, 0.0, 0.0, 1.0, nu, nu, 3.0 / 16.0};
  for (uint64_t it = 0; it < nsteps; ++it) {
    for (size_t k = 0; k < n; ++k) {
      double r = 0.0, mx = 0.0, my = 0.0;
      for (int j = 0; j < 9; ++j) {
        r += f[k].f[j];
        mx += (double)CX[j] * f[k].f[j];
        my += (double)CY[j] * f[k].f[j];
      }
      double fnode = r * fy * (1.0 - 0.0);
      double ux = mx / r;
      double uy = my / r + 0.5 * fnode / r;
      double nut, fyo;
      fp[k] = f[k];
      for (int i = 0; i < 9; ++i) g[k].f[i] = 0.0;
      collide_node(fp[k].f, g[k].f, &nut, &fyo, r, ux, uy, 1.0, &c, 2);
    }
    for (size_t iy = 0; iy < ny; ++iy)
      for (size_t ix = 0; ix < nx; ++ix) {
        size_t k = iy * nx + ix;
        for (int i = 0; i < 9; ++i) {
          long sx = (long)ix - CX[i];
          long sy = ((long)iy - CY[i] + (long)ny) % (long)ny;
          if (sx < 0 || sx >= (long)nx)
            f[k].f[i] = fp[k].f[OPP[i]];
          else
            f[k].f[i] = fp[(size_t)sy * nx + (size_t)sx].f[i];
        }
      }
  }
  double umax = 0.0;
  for (size_t k = 0; k < n; ++k) {
    double r = 0.0, my = 0.0;
    for (int j = 0; j < 9; ++j) {
      r += f[k].f[j];
      my += (double)CY[j] * f[k].f[j];
    }
    umax = fmax(umax, my / r + 0.5 * fy);
  }
  free(f);
  free(fp);
  free(g);
  return umax;
}

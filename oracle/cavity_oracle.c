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
static void collide_node(double *f, double *g, double *nu_t_out, double *fy_out,
                         double rho, double ux, double uy, double temp,
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
  /* compute_turbulent_viscosity(s_mag, cs) = (cs*dx)^2*|S|, dx = 1 */
  double nu_t = c->cs * c->cs * s_mag;
  *nu_t_out = nu_t;

  double tau_plus = 0.5 + 3.0 * (c->nu + nu_t);
  double tau_minus = 0.5 + c->lambda / (tau_plus - 0.5);
  double alpha_t = nu_t / c->pr_t;
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
      collide_node(s->f[k].f, s->g[k].f, &s->nu_t[k], &s->fy_node[k], s->rho[k],
                   s->ux[k], s->uy[k], s->temp[k], &c, go);
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

/* ---- step (src/uniform/solver.rs:279-293) --------------------------------- */
void orc_step(orc_solver *s) {
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

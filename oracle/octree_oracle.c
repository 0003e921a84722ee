/*
 * octree_oracle.c — CPU oracle for the octree (multi-level) cavity solver.
 *
 * TEST INFRASTRUCTURE ONLY: a restatement of the reference's octree path used
 * to check the HIP octree kernels; nothing in the product loads it.  Built
 * into oracle/libcavity_oracle.so with -ffp-contract=off, so that expressions
 * evaluated in the reference's operand order are bitwise reproducible.
 *
 * Each function cites the reference file:line it restates.
 *
 * Parity status (tests in tests/test_oracle_octree.py):
 *   - per-level phases (collide_cavity, stream_cavity(_block), the thermal
 *     BCs, update_macroscopic_guo, the wall velocities): bitwise against the
 *     uniform oracle, which is pinned on its own — on the full coarse level,
 *     and on the interior of a block whose four sides are walls;
 *   - grid tables and interface transfers: the reference's V1 tests
 *     (src/octree/tests.rs:135-395), closed forms on linear equilibrium fields
 *     and an independent bilinear stencil;
 *   - the solver (init, step order, ramps, diagnostics): the reference's V2 and
 *     V6 tests (src/octree/tests.rs:396-575) and invariants; the order of the
 *     phases inside step_cavity beyond those is PARITY UNPINNED;
 *   - the dynamic mixed-scale SGS per level (src/octree/smag.rs, the model of
 *     the uniform oracle on each level's arrays, orc_dynamic_model) and its
 *     place in step_cavity (solver.rs:647-947: the model before every
 *     collision, the per-node relaxation times, the interface factors xi
 *     recomputed from the levels' mean nu_t / alpha_t): the reference's V5
 *     tests (src/octree/tests.rs:572-786), the level model against the
 *     independent numpy model on the coarse level, the per-level floors, and
 *     the per-node collision against the uniform collide_node.
 */
#include "octree_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cavity_oracle.h"

#define T_HOT 0.5
#define T_COLD (-0.5)
#define DELTA_T (T_HOT - T_COLD)
#define F64_EPSILON 2.220446049250313e-16
#define LAMBDA_TRT (3.0 / 16.0)
#define C_SCALE 1.7
#define MAX_U_ABORT 0.10
#define NU_TAU_RAMP 1000.0
#define CS2 (1.0 / 3.0)

static int CX[9], CY[9], OPP[9];
static double W[9];
static int tables_ready = 0;

static void tables(void) {
  if (!tables_ready) {
    orc_lattice(CX, CY, OPP, W);
    tables_ready = 1;
  }
}

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) {
    fprintf(stderr, "octree_oracle: out of memory\n");
    abort();
  }
  return p;
}

/* ---- LbmLevel (src/octree/lbm.rs:18-50) ---------------------------------- */
typedef struct {
  double f[9];
} node9;

typedef struct {
  size_t nx, ny;
  node9 *f, *g;
  double *rho, *ux, *uy, *temp;
} level;

static void level_init(level *l, size_t nx, size_t ny) {
  size_t n = nx * ny;
  l->nx = nx;
  l->ny = ny;
  l->f = (node9 *)xcalloc(n, sizeof(node9));
  l->g = (node9 *)xcalloc(n, sizeof(node9));
  l->rho = (double *)xcalloc(n, sizeof(double));
  l->ux = (double *)xcalloc(n, sizeof(double));
  l->uy = (double *)xcalloc(n, sizeof(double));
  l->temp = (double *)xcalloc(n, sizeof(double));
}

static void level_free(level *l) {
  free(l->f);
  free(l->g);
  free(l->rho);
  free(l->ux);
  free(l->uy);
  free(l->temp);
}

/* ---- RelaxTimes::from_nu_alpha (src/octree/lbm.rs:100-109) --------------- */
typedef struct {
  double tau_plus, tau_minus, tau_t;
} relax;

static relax relax_from(double nu_lu, double alpha_lu) {
  relax r;
  r.tau_plus = 0.5 + 3.0 * nu_lu;
  r.tau_minus = 0.5 + LAMBDA_TRT / (r.tau_plus - 0.5);
  r.tau_t = 0.5 + 3.0 * alpha_lu;
  return r;
}

void orc_oct_relax_times(double nu, double alpha, double out[3]) {
  relax r = relax_from(nu, alpha);
  out[0] = r.tau_plus;
  out[1] = r.tau_minus;
  out[2] = r.tau_t;
}

/* ---- RescalingFactors (src/octree/interface.rs:40-158) -------------------- */
typedef struct {
  double xi_even, xi_odd, xi_thermal;
} xi_t;

/* rescaling_factors_multilevel (interface.rs:141-158); for substeps = 2 it is
 * rescaling_factors (interface.rs:54-67): 1.0/2.0 = 0.5 exactly. */
static xi_t rescaling(relax c, relax f, size_t substeps, int odd_exact) {
  xi_t x;
  double dt_factor = 1.0 / (double)substeps;
  x.xi_even = dt_factor * f.tau_plus / c.tau_plus;
  x.xi_odd = odd_exact ? dt_factor * f.tau_minus / c.tau_minus : dt_factor;
  x.xi_thermal = dt_factor * f.tau_t / c.tau_t;
  return x;
}

void orc_oct_rescaling_factors(const double tc[3], const double tf[3], size_t substeps,
                               int odd_exact, double out[3]) {
  relax c = {tc[0], tc[1], tc[2]}, f = {tf[0], tf[1], tf[2]};
  xi_t x = rescaling(c, f, substeps, odd_exact);
  out[0] = x.xi_even;
  out[1] = x.xi_odd;
  out[2] = x.xi_thermal;
}

/* split_odd_even + rescale_fneq (interface.rs:72-97) */
static void rescale_fneq(const double fneq[9], const xi_t *xi, double out[9]) {
  tables();
  for (int i = 0; i < 9; ++i) {
    int opp = OPP[i];
    double even = 0.5 * (fneq[i] + fneq[opp]);
    double odd = 0.5 * (fneq[i] - fneq[opp]);
    out[i] = xi->xi_even * even + xi->xi_odd * odd;
  }
}

/* rescale_fneq_inverse (interface.rs:102-110) */
static void rescale_fneq_inverse(const double fneq[9], const xi_t *xi, double out[9]) {
  tables();
  for (int i = 0; i < 9; ++i) {
    int opp = OPP[i];
    double even = 0.5 * (fneq[i] + fneq[opp]);
    double odd = 0.5 * (fneq[i] - fneq[opp]);
    out[i] = even / xi->xi_even + odd / xi->xi_odd;
  }
}

void orc_oct_rescale_fneq(const double fneq[9], const double xi[3], double out[9]) {
  xi_t x = {xi[0], xi[1], xi[2]};
  rescale_fneq(fneq, &x, out);
}

void orc_oct_rescale_fneq_inverse(const double fneq[9], const double xi[3], double out[9]) {
  xi_t x = {xi[0], xi[1], xi[2]};
  rescale_fneq_inverse(fneq, &x, out);
}

/* ---- OctreeGrid (src/octree/grid.rs:95-229) ------------------------------ */
typedef struct {
  size_t k_fine;
  size_t kc[4];
  double w[4];
} ghost_entry;

typedef struct {
  size_t k_coarse;
  size_t kf[9];
  double w[9];
} coinc_entry;

typedef struct {
  size_t nx0, ny0, bx0, by0, bx1, by1, nx1, ny1;
  ghost_entry *ghosts;
  size_t n_ghosts;
  coinc_entry *coinc;
  size_t n_coinc;
} grid_t;

/* bilinear_stencil (grid.rs:216-229): base clamped to [0, n-2], so positions
 * outside the domain extrapolate linearly. */
static void bilinear_stencil(double x_c, double y_c, size_t nx0, size_t ny0, size_t k[4],
                             double w[4]) {
  long long ix = (long long)floor(x_c), iy = (long long)floor(y_c);
  long long mx = (long long)nx0 - 2, my = (long long)ny0 - 2;
  ix = ix < 0 ? 0 : (ix > mx ? mx : ix);
  iy = iy < 0 ? 0 : (iy > my ? my : iy);
  double fx = x_c - (double)ix;
  double fy = y_c - (double)iy;
  size_t ux = (size_t)ix, uy = (size_t)iy;
  k[0] = uy * nx0 + ux;
  w[0] = (1.0 - fx) * (1.0 - fy);
  k[1] = uy * nx0 + ux + 1;
  w[1] = fx * (1.0 - fy);
  k[2] = (uy + 1) * nx0 + ux;
  w[2] = (1.0 - fx) * fy;
  k[3] = (uy + 1) * nx0 + ux + 1;
  w[3] = fx * fy;
}

void orc_oct_bilinear_stencil(double x_c, double y_c, size_t nx0, size_t ny0, size_t k[4],
                              double w[4]) {
  bilinear_stencil(x_c, y_c, nx0, ny0, k, w);
}

/* OctreeGrid::build (grid.rs:121-203) */
static void grid_build(grid_t *g, size_t nx0, size_t ny0, size_t bx0, size_t by0, size_t bx1,
                       size_t by1, int wall_left, int wall_right, int wall_bottom,
                       int wall_top) {
  size_t fw = bx1 - bx0, fh = by1 - by0;
  memset(g, 0, sizeof(*g));
  g->nx0 = nx0;
  g->ny0 = ny0;
  g->bx0 = bx0;
  g->by0 = by0;
  g->bx1 = bx1;
  g->by1 = by1;
  g->nx1 = 2 * fw + 3;
  g->ny1 = 2 * fh + 3;
  const size_t nx1 = g->nx1, ny1 = g->ny1;
  g->ghosts = (ghost_entry *)xcalloc(2 * (nx1 + ny1), sizeof(ghost_entry));
  for (size_t b = 0; b < ny1; ++b)
    for (size_t a = 0; a < nx1; ++a)
      if (a == 0 || b == 0 || a == nx1 - 1 || b == ny1 - 1) {
        ghost_entry *e = &g->ghosts[g->n_ghosts++];
        double x_c = (double)bx0 - 0.5 + 0.5 * (double)a;
        double y_c = (double)by0 - 0.5 + 0.5 * (double)b;
        e->k_fine = b * nx1 + a;
        bilinear_stencil(x_c, y_c, nx0, ny0, e->kc, e->w);
      }
  g->coinc = (coinc_entry *)xcalloc(fw * fh, sizeof(coinc_entry));
  static const double wt[3] = {0.25, 0.5, 0.25};
  for (size_t cy = by0; cy < by1; ++cy)
    for (size_t cx = bx0; cx < bx1; ++cx) {
      if (cx == 0 || cx + 1 == nx0 || cy == 0 || cy + 1 == ny0) continue;
      if ((wall_left && cx == bx0) || (wall_right && cx + 1 == bx1) ||
          (wall_bottom && cy == by0) || (wall_top && cy + 1 == by1))
        continue;
      size_t a = 2 * (cx - bx0) + 1, b = 2 * (cy - by0) + 1;
      coinc_entry *e = &g->coinc[g->n_coinc++];
      e->k_coarse = cy * nx0 + cx;
      int idx = 0;
      for (int jb = 0; jb < 3; ++jb)
        for (int ja = 0; ja < 3; ++ja) {
          e->kf[idx] = (b + (size_t)jb - 1) * nx1 + (a + (size_t)ja - 1);
          e->w[idx] = wt[ja] * wt[jb];
          ++idx;
        }
    }
}

static void grid_free(grid_t *g) {
  free(g->ghosts);
  free(g->coinc);
}

/* ---- interface transfers (src/octree/interface.rs:164-229) --------------- */
/* refresh_moments (interface.rs:213-229): no epsilon guard, no Guo term */
static void refresh_moments(level *l, const ghost_entry *gh, size_t n) {
  for (size_t j = 0; j < n; ++j) {
    size_t k = gh[j].k_fine;
    double rho = 0.0, mx = 0.0, my = 0.0, t = 0.0;
    for (int i = 0; i < 9; ++i) {
      double fi = l->f[k].f[i];
      rho += fi;
      mx += fi * (double)CX[i];
      my += fi * (double)CY[i];
      t += l->g[k].f[i];
    }
    l->rho[k] = rho;
    l->ux[k] = mx / rho;
    l->uy[k] = my / rho;
    l->temp[k] = t;
  }
}

/* coarse_to_fine (interface.rs:164-195) */
static void coarse_to_fine(const grid_t *grid, const level *c, level *f, const xi_t *xi,
                           int geq_order) {
  tables();
  for (size_t j = 0; j < grid->n_ghosts; ++j) {
    const ghost_entry *gh = &grid->ghosts[j];
    size_t kf = gh->k_fine;
    double rho = 0.0, ux = 0.0, uy = 0.0, temp = 0.0;
    for (int s = 0; s < 4; ++s) {
      size_t kc = gh->kc[s];
      double w = gh->w[s];
      rho += w * c->rho[kc];
      ux += w * c->ux[kc];
      uy += w * c->uy[kc];
      temp += w * c->temp[kc];
    }
    double fneq_int[9] = {0}, gneq_int[9] = {0};
    for (int i = 0; i < 9; ++i)
      for (int s = 0; s < 4; ++s) {
        size_t kc = gh->kc[s];
        double w = gh->w[s];
        fneq_int[i] += w * (c->f[kc].f[i] - orc_feq(i, c->rho[kc], c->ux[kc], c->uy[kc]));
        gneq_int[i] +=
            w * (c->g[kc].f[i] - orc_geq(i, c->temp[kc], c->ux[kc], c->uy[kc], geq_order));
      }
    double fneq_fin[9];
    rescale_fneq(fneq_int, xi, fneq_fin);
    for (int i = 0; i < 9; ++i) {
      f->f[kf].f[i] = orc_feq(i, rho, ux, uy) + fneq_fin[i];
      f->g[kf].f[i] = orc_geq(i, temp, ux, uy, geq_order) + xi->xi_thermal * gneq_int[i];
    }
  }
  refresh_moments(f, grid->ghosts, grid->n_ghosts);
}

/* fine_to_coarse (interface.rs:200-210 with the 9-node stencil of grid.rs) */
static void fine_to_coarse(const grid_t *grid, const level *f, level *c, const xi_t *xi,
                           int geq_order) {
  tables();
  for (size_t j = 0; j < grid->n_coinc; ++j) {
    const coinc_entry *e = &grid->coinc[j];
    size_t kc = e->k_coarse;
    double rho = 0.0, ux = 0.0, uy = 0.0, temp = 0.0;
    for (int s = 0; s < 9; ++s) {
      size_t kf = e->kf[s];
      double w = e->w[s];
      rho += w * f->rho[kf];
      ux += w * f->ux[kf];
      uy += w * f->uy[kf];
      temp += w * f->temp[kf];
    }
    double fneq_avg[9] = {0}, gneq_avg[9] = {0};
    for (int i = 0; i < 9; ++i)
      for (int s = 0; s < 9; ++s) {
        size_t kf = e->kf[s];
        double w = e->w[s];
        fneq_avg[i] += w * (f->f[kf].f[i] - orc_feq(i, f->rho[kf], f->ux[kf], f->uy[kf]));
        gneq_avg[i] +=
            w * (f->g[kf].f[i] - orc_geq(i, f->temp[kf], f->ux[kf], f->uy[kf], geq_order));
      }
    double fneq_gross[9];
    rescale_fneq_inverse(fneq_avg, xi, fneq_gross);
    for (int i = 0; i < 9; ++i) {
      c->f[kc].f[i] = orc_feq(i, rho, ux, uy) + fneq_gross[i];
      c->g[kc].f[i] = orc_geq(i, temp, ux, uy, geq_order) + gneq_avg[i] / xi->xi_thermal;
    }
  }
}

/* refresh_moments_nodes (solver.rs:1188-1206): epsilon guard, no Guo term */
static void refresh_moments_nodes(level *l, const coinc_entry *e, size_t n) {
  for (size_t j = 0; j < n; ++j) {
    size_t k = e[j].k_coarse;
    double rho = 0.0, mx = 0.0, my = 0.0, t = 0.0;
    for (int i = 0; i < 9; ++i) {
      double fi = l->f[k].f[i];
      rho += fi;
      mx += (double)CX[i] * fi;
      my += (double)CY[i] * fi;
      t += l->g[k].f[i];
    }
    l->rho[k] = rho;
    l->temp[k] = t;
    if (rho > F64_EPSILON) {
      l->ux[k] = mx / rho;
      l->uy[k] = my / rho;
    } else {
      l->ux[k] = 0.0;
      l->uy[k] = 0.0;
    }
  }
}

/* ---- per-level phases (src/octree/cavity.rs) ----------------------------- */
/* collide_node_cavity (cavity.rs:68-161).  pre: the dynamic model's nu_t /
 * alpha_t of the node set the relaxation times (cavity.rs:104-113), else the
 * level's RelaxTimes. */
static void collide_node_cavity(double *f, double *g, double *fy_out, double rho, double ux,
                                double uy, double temp, const relax *tau, double beta_eff,
                                int geq_order, int pre, double nu_t, double alpha_t) {
  double fy = rho * beta_eff * (temp - 0.0); /* boussinesq_force(rho, beta, 1, T, 0) */
  *fy_out = fy;
  double f_eq[9], g_eq[9];
  for (int i = 0; i < 9; ++i) {
    f_eq[i] = orc_feq(i, rho, ux, uy);
    g_eq[i] = orc_geq(i, temp, ux, uy, geq_order);
  }
  double tau_plus = tau->tau_plus, tau_minus = tau->tau_minus, tau_t = tau->tau_t;
  if (pre) {
    tau_plus = 0.5 + 3.0 * (tau->tau_plus - 0.5) / 3.0 + 3.0 * nu_t;
    tau_minus = 0.5 + LAMBDA_TRT / (tau_plus - 0.5);
    tau_t = 0.5 + 3.0 * (tau->tau_t - 0.5) / 3.0 + 3.0 * alpha_t;
  }
  double a_f = 0.5 * (1.0 / tau_plus + 1.0 / tau_minus);
  double b_f = 0.5 * (1.0 / tau_plus - 1.0 / tau_minus);
  double a_g = 1.0 / tau_t;
  double b_g = 0.0;
  double d_f[9], d_g[9];
  for (int i = 0; i < 9; ++i) {
    d_f[i] = f[i] - f_eq[i];
    d_g[i] = g[i] - g_eq[i];
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

/* collide_cavity (cavity.rs:167-243): window [1, n-1)^2 when interior_only;
 * nu_t_pre / alpha_t_pre per node of the level, or NULL (model off). */
static void collide_cavity(level *l, const relax *tau, double beta_eff, double *fy_out,
                           int interior_only, int geq_order, const double *nu_t_pre,
                           const double *alpha_t_pre) {
  tables();
  size_t nx = l->nx;
  size_t ix0 = interior_only ? 1 : 0, ix1 = interior_only ? nx - 1 : nx;
  size_t iy0 = interior_only ? 1 : 0, iy1 = interior_only ? l->ny - 1 : l->ny;
  for (size_t iy = iy0; iy < iy1; ++iy)
    for (size_t ix = ix0; ix < ix1; ++ix) {
      size_t k = iy * nx + ix;
      const int pre = nu_t_pre != NULL && alpha_t_pre != NULL;
      collide_node_cavity(l->f[k].f, l->g[k].f, &fy_out[k], l->rho[k], l->ux[k], l->uy[k],
                          l->temp[k], tau, beta_eff, geq_order, pre, pre ? nu_t_pre[k] : 0.0,
                          pre ? alpha_t_pre[k] : 0.0);
    }
}

/* update_macroscopic_guo (cavity.rs:249-316), the single-thread form */
static void update_macroscopic_guo(level *l, const double *fy) {
  tables();
  for (size_t k = 0; k < l->nx * l->ny; ++k) {
    double rho = 0.0, mx = 0.0, my = 0.0, t = 0.0;
    for (int i = 0; i < 9; ++i) {
      double fi = l->f[k].f[i];
      rho += fi;
      mx += (double)CX[i] * fi;
      my += (double)CY[i] * fi;
      t += l->g[k].f[i];
    }
    l->rho[k] = rho;
    l->temp[k] = t;
    if (rho > F64_EPSILON) {
      l->ux[k] = mx / rho;
      l->uy[k] = my / rho + 0.5 * fy[k] / rho;
    } else {
      l->ux[k] = 0.0;
      l->uy[k] = 0.0;
    }
  }
}

/* stream (lbm.rs:143-171) = stream_cavity (cavity.rs:326-360) */
static void stream_level(level *l) {
  tables();
  size_t nx = l->nx, ny = l->ny, n = nx * ny;
  node9 *f_old = (node9 *)xcalloc(n, sizeof(node9)), *g_old = (node9 *)xcalloc(n, sizeof(node9));
  memcpy(f_old, l->f, n * sizeof(node9));
  memcpy(g_old, l->g, n * sizeof(node9));
  for (size_t iy = 0; iy < ny; ++iy)
    for (size_t ix = 0; ix < nx; ++ix) {
      size_t k = iy * nx + ix;
      for (int i = 0; i < 9; ++i) {
        long long sx = (long long)ix - CX[i], sy = (long long)iy - CY[i];
        if (sx >= 0 && sx < (long long)nx && sy >= 0 && sy < (long long)ny) {
          size_t src = (size_t)sy * nx + (size_t)sx;
          l->f[k].f[i] = f_old[src].f[i];
          l->g[k].f[i] = g_old[src].f[i];
        } else {
          l->f[k].f[i] = f_old[k].f[OPP[i]];
          l->g[k].f[i] = g_old[k].f[i];
        }
      }
    }
  free(f_old);
  free(g_old);
}

/* stream_cavity_block (cavity.rs:381-495): ghost layers not streamed; a
 * source in a wall-side ghost layer bounces back. */
static void stream_block(level *l, int wl, int wr, int wb, int wt) {
  tables();
  size_t nx = l->nx, ny = l->ny, n = nx * ny;
  node9 *f_old = (node9 *)xcalloc(n, sizeof(node9)), *g_old = (node9 *)xcalloc(n, sizeof(node9));
  memcpy(f_old, l->f, n * sizeof(node9));
  memcpy(g_old, l->g, n * sizeof(node9));
  for (size_t iy = 1; iy + 1 < ny; ++iy)
    for (size_t ix = 1; ix + 1 < nx; ++ix) {
      size_t k = iy * nx + ix;
      for (int i = 0; i < 9; ++i) {
        long long sx = (long long)ix - CX[i], sy = (long long)iy - CY[i];
        int wall_src = (sx == 0 && wl) || (sx == (long long)nx - 1 && wr) ||
                       (sy == 0 && wb) || (sy == (long long)ny - 1 && wt);
        if (wall_src) {
          l->f[k].f[i] = f_old[k].f[OPP[i]];
          l->g[k].f[i] = g_old[k].f[i];
        } else {
          size_t src = (size_t)sy * nx + (size_t)sx;
          l->f[k].f[i] = f_old[src].f[i];
          l->g[k].f[i] = g_old[src].f[i];
        }
      }
    }
  free(f_old);
  free(g_old);
}

/* apply_thermal_boundary_conditions (cavity.rs:500-533) */
static void thermal_bc(level *l, int geq_order) {
  size_t nx = l->nx, ny = l->ny;
  for (size_t iy = 0; iy < ny; ++iy) {
    size_t k0 = iy * nx, k1 = iy * nx + (nx - 1);
    for (int i = 0; i < 9; ++i) l->g[k0].f[i] = orc_geq(i, T_HOT, 0.0, 0.0, geq_order);
    for (int i = 0; i < 9; ++i) l->g[k1].f[i] = orc_geq(i, T_COLD, 0.0, 0.0, geq_order);
  }
  for (size_t ix = 0; ix < nx; ++ix) {
    l->g[ix] = l->g[nx + ix];
    l->g[(ny - 1) * nx + ix] = l->g[(ny - 2) * nx + ix];
  }
}

/* apply_thermal_boundary_conditions_block (cavity.rs:542-600) */
static void thermal_bc_block(level *l, int hot, int cold, int geq_order) {
  size_t nx = l->nx, ny = l->ny;
  for (size_t iy = 0; iy < ny; ++iy) {
    if (hot)
      for (int i = 0; i < 9; ++i) l->g[iy * nx + 1].f[i] = orc_geq(i, T_HOT, 0.0, 0.0, geq_order);
    if (cold)
      for (int i = 0; i < 9; ++i)
        l->g[iy * nx + nx - 2].f[i] = orc_geq(i, T_COLD, 0.0, 0.0, geq_order);
  }
  for (size_t ix = 1; ix + 1 < nx; ++ix) {
    l->g[nx + ix] = l->g[2 * nx + ix];
    l->g[(ny - 2) * nx + ix] = l->g[(ny - 3) * nx + ix];
  }
  for (size_t iy = 0; iy < ny; ++iy) {
    l->g[iy * nx] = l->g[iy * nx + 1];
    l->g[iy * nx + nx - 1] = l->g[iy * nx + nx - 2];
  }
  for (size_t ix = 0; ix < nx; ++ix) {
    l->g[ix] = l->g[nx + ix];
    l->g[(ny - 1) * nx + ix] = l->g[(ny - 2) * nx + ix];
  }
}

/* enforce_wall_velocity (cavity.rs:604-620) */
static void walls_level(level *l) {
  size_t nx = l->nx, ny = l->ny;
  for (size_t iy = 0; iy < ny; ++iy) {
    l->ux[iy * nx] = l->uy[iy * nx] = 0.0;
    l->ux[iy * nx + nx - 1] = l->uy[iy * nx + nx - 1] = 0.0;
  }
  for (size_t ix = 0; ix < nx; ++ix) {
    l->ux[ix] = l->uy[ix] = 0.0;
    l->ux[(ny - 1) * nx + ix] = l->uy[(ny - 1) * nx + ix] = 0.0;
  }
}

/* enforce_wall_velocity_block (cavity.rs:626-650) */
static void walls_block(level *l, int hot, int cold) {
  size_t nx = l->nx, ny = l->ny;
  for (size_t iy = 0; iy < ny; ++iy) {
    if (hot) l->ux[iy * nx + 1] = l->uy[iy * nx + 1] = 0.0;
    if (cold) l->ux[iy * nx + nx - 2] = l->uy[iy * nx + nx - 2] = 0.0;
  }
  for (size_t ix = 0; ix < nx; ++ix) {
    l->ux[nx + ix] = l->uy[nx + ix] = 0.0;
    l->ux[(ny - 2) * nx + ix] = l->uy[(ny - 2) * nx + ix] = 0.0;
  }
}

/* ---- the solver (src/octree/solver.rs:118-1206) -------------------------- */
typedef struct {
  grid_t grid;
  level lvl;
  relax tau;
  xi_t xi_down;
  double *fy;
  int hot, cold;
  double *nu_t_pre, *alpha_t_pre; /* dynamic model of the level (solver.rs:172-175) */
} nested_t;

typedef struct {
  grid_t grid;
  level fine;
  double *fy;
  int hot, cold;
  nested_t *sub;
  double *nu_t_pre, *alpha_t_pre; /* dynamic model of the level (solver.rs:137-140) */
} block_t;

struct orc_octree {
  size_t nx0, ny0;
  level coarse;
  block_t blocks[2];
  relax tau_coarse, tau_fine;
  xi_t xi;
  int levels;
  double nu_lu, alpha_lu;
  int odd_exact;
  double nu_start_mult;
  size_t spin_up_steps;
  uint64_t time_step;
  double beta_eff_coarse, beta_eff_fine;
  double *fy_coarse;
  size_t jobs;
  uint64_t rng;
  int geq_order;
  /* dynamic mixed-scale SGS (solver.rs:229-248) */
  int dynamic_smag;
  double smag_alpha, nu_sm_min, eta_test;
  double *nu_t_pre_coarse, *alpha_t_pre_coarse;
  double cs_dyn_mean, cs_dyn_max, kappa_dyn_mean, kappa_dyn_max, nu_t_max;
};

static double rng_f64(uint64_t *state) { return orc_rng_next_f64(state); }

/* init_cavity (solver.rs:500-575) */
static void init_cavity(orc_octree *o) {
  const size_t nx0 = o->nx0, ny0 = o->ny0;
  const int go = o->geq_order;
  for (size_t iy = 0; iy < ny0; ++iy)
    for (size_t ix = 0; ix < nx0; ++ix) {
      size_t k = iy * nx0 + ix;
      double xf = (double)ix / (double)(nx0 - 1);
      double t_lin = T_HOT + (T_COLD - T_HOT) * xf;
      double pert = 1.0e-3 * (2.0 * rng_f64(&o->rng) - 1.0);
      double t = t_lin + pert;
      o->coarse.rho[k] = 1.0;
      o->coarse.ux[k] = 0.0;
      o->coarse.uy[k] = 0.0;
      o->coarse.temp[k] = t;
      for (int i = 0; i < 9; ++i) {
        o->coarse.f[k].f[i] = orc_feq(i, 1.0, 0.0, 0.0);
        o->coarse.g[k].f[i] = orc_geq(i, t, 0.0, 0.0, go);
      }
    }
  for (int bi = 0; bi < 2; ++bi) {
    block_t *b = &o->blocks[bi];
    size_t nx1 = b->grid.nx1;
    for (size_t bb = 0; bb < b->grid.ny1; ++bb)
      for (size_t a = 0; a < nx1; ++a) {
        size_t k = bb * nx1 + a;
        double x_c = (double)b->grid.bx0 - 0.5 + 0.5 * (double)a;
        double y_c = (double)b->grid.by0 - 0.5 + 0.5 * (double)bb;
        size_t kc[4];
        double w[4];
        bilinear_stencil(x_c, y_c, nx0, ny0, kc, w);
        double t = 0.0;
        for (int s = 0; s < 4; ++s) t += w[s] * o->coarse.temp[kc[s]];
        b->fine.rho[k] = 1.0;
        b->fine.ux[k] = 0.0;
        b->fine.uy[k] = 0.0;
        b->fine.temp[k] = t;
        for (int i = 0; i < 9; ++i) {
          b->fine.f[k].f[i] = orc_feq(i, 1.0, 0.0, 0.0);
          b->fine.g[k].f[i] = orc_geq(i, t, 0.0, 0.0, go);
        }
      }
    if (b->sub) {
      nested_t *s = b->sub;
      size_t nx2 = s->grid.nx1;
      for (size_t b2 = 0; b2 < s->grid.ny1; ++b2)
        for (size_t a2 = 0; a2 < nx2; ++a2) {
          size_t k = b2 * nx2 + a2;
          double x_c = (double)b->grid.bx0 - 0.5 +
                       0.5 * ((double)s->grid.bx0 - 0.5 + 0.5 * (double)a2);
          double y_c = (double)b->grid.by0 - 0.5 +
                       0.5 * ((double)s->grid.by0 - 0.5 + 0.5 * (double)b2);
          size_t kc[4];
          double w[4];
          bilinear_stencil(x_c, y_c, nx0, ny0, kc, w);
          double t = 0.0;
          for (int q = 0; q < 4; ++q) t += w[q] * o->coarse.temp[kc[q]];
          s->lvl.rho[k] = 1.0;
          s->lvl.ux[k] = 0.0;
          s->lvl.uy[k] = 0.0;
          s->lvl.temp[k] = t;
          for (int i = 0; i < 9; ++i) {
            s->lvl.f[k].f[i] = orc_feq(i, 1.0, 0.0, 0.0);
            s->lvl.g[k].f[i] = orc_geq(i, t, 0.0, 0.0, go);
          }
        }
    }
  }
}

orc_octree *orc_oct_new(size_t nx0, size_t ny0, size_t bl_width, double ra, double pr,
                        double utarget, uint64_t seed, size_t jobs, int odd_exact,
                        int levels, double nu_start_mult, int has_nu_override,
                        double nu_override, int geq_order, char *err, size_t errlen) {
  tables();
  /* solver.rs:356-361 */
  if (bl_width < 2) {
    snprintf(err, errlen, "largeur de BL fine >= 2 cellules grossieres requise");
    return NULL;
  }
  if (!(2 * bl_width < nx0)) {
    snprintf(err, errlen, "les 2 blocs fins doivent laisser un noyau grossier");
    return NULL;
  }
  if (levels != 2 && levels != 3) {
    snprintf(err, errlen, "levels doit valoir 2 (config V2) ou 3 (escalier tanh V6)");
    return NULL;
  }
  if (ny0 < 2) {
    snprintf(err, errlen, "ny0 must be >= 2");
    return NULL;
  }
  /* calibration (solver.rs:363-392) */
  double lz = (double)(ny0 - 1);
  double nu_cal = pow(utarget * sqrt(pr) * lz / (C_SCALE * sqrt(ra)), 2.0 / 3.0);
  double nu;
  if (has_nu_override) {
    if (!(nu_override > 0.0 && isfinite(nu_override))) {
      snprintf(err, errlen, "--nu doit être > 0 et fini");
      return NULL;
    }
    nu = nu_override;
  } else {
    nu = fmax(nu_cal, 0.002);
  }
  double alpha = nu / pr;
  double beta_eff_coarse = ra * nu * alpha / (lz * lz * lz);

  orc_octree *o = (orc_octree *)xcalloc(1, sizeof(orc_octree));
  o->nx0 = nx0;
  o->ny0 = ny0;
  o->geq_order = geq_order == 1 ? 1 : 2;
  o->tau_coarse = relax_from(nu, alpha);
  o->tau_fine = relax_from(2.0 * nu, 2.0 * alpha);
  relax tau_sub = relax_from(4.0 * nu, 4.0 * alpha);
  o->xi = rescaling(o->tau_coarse, o->tau_fine, 2, odd_exact);
  xi_t xi_sub = rescaling(o->tau_fine, tau_sub, 2, odd_exact);
  level_init(&o->coarse, nx0, ny0);
  /* blocks (solver.rs:411-462) */
  for (int bi = 0; bi < 2; ++bi) {
    block_t *b = &o->blocks[bi];
    int hot = bi == 0, cold = bi == 1;
    size_t bx0 = hot ? 0 : nx0 - 1 - bl_width, bx1 = hot ? bl_width : nx0 - 1;
    grid_build(&b->grid, nx0, ny0, bx0, 0, bx1, ny0 - 1, 0, 0, 0, 0);
    level_init(&b->fine, b->grid.nx1, b->grid.ny1);
    b->fy = (double *)xcalloc(b->grid.nx1 * b->grid.ny1, sizeof(double));
    b->hot = hot;
    b->cold = cold;
    b->sub = NULL;
    b->nu_t_pre = (double *)xcalloc(b->grid.nx1 * b->grid.ny1, sizeof(double));
    b->alpha_t_pre = (double *)xcalloc(b->grid.nx1 * b->grid.ny1, sizeof(double));
    if (levels >= 3) {
      nested_t *s = (nested_t *)xcalloc(1, sizeof(nested_t));
      size_t sbx0 = hot ? 1 : b->grid.nx1 - 4, sbx1 = hot ? 3 : b->grid.nx1 - 2;
      grid_build(&s->grid, b->grid.nx1, b->grid.ny1, sbx0, 1, sbx1, b->grid.ny1 - 2, 1, 0, 1, 0);
      level_init(&s->lvl, s->grid.nx1, s->grid.ny1);
      s->tau = tau_sub;
      s->xi_down = xi_sub;
      s->fy = (double *)xcalloc(s->grid.nx1 * s->grid.ny1, sizeof(double));
      s->hot = hot;
      s->cold = cold;
      s->nu_t_pre = (double *)xcalloc(s->grid.nx1 * s->grid.ny1, sizeof(double));
      s->alpha_t_pre = (double *)xcalloc(s->grid.nx1 * s->grid.ny1, sizeof(double));
      b->sub = s;
    }
  }
  o->levels = levels;
  o->nu_lu = nu;
  o->alpha_lu = alpha;
  o->odd_exact = odd_exact;
  o->nu_start_mult = fmax(nu_start_mult, 1.0);
  o->spin_up_steps = ny0 / 10 > 2000 ? ny0 / 10 : 2000;
  o->time_step = 0;
  o->beta_eff_coarse = beta_eff_coarse;
  o->beta_eff_fine = 0.5 * beta_eff_coarse;
  o->fy_coarse = (double *)xcalloc(nx0 * ny0, sizeof(double));
  o->jobs = jobs ? jobs : 1;
  o->rng = seed | 1ULL;
  /* solver.rs:455-467 */
  o->dynamic_smag = 0;
  o->smag_alpha = 0.5;
  o->nu_sm_min = -0.1;
  o->eta_test = 2.0;
  o->nu_t_pre_coarse = (double *)xcalloc(nx0 * ny0, sizeof(double));
  o->alpha_t_pre_coarse = (double *)xcalloc(nx0 * ny0, sizeof(double));
  init_cavity(o);
  return o;
}

void orc_oct_free(orc_octree *o) {
  if (!o) return;
  level_free(&o->coarse);
  for (int bi = 0; bi < 2; ++bi) {
    block_t *b = &o->blocks[bi];
    grid_free(&b->grid);
    level_free(&b->fine);
    free(b->fy);
    free(b->nu_t_pre);
    free(b->alpha_t_pre);
    if (b->sub) {
      grid_free(&b->sub->grid);
      level_free(&b->sub->lvl);
      free(b->sub->fy);
      free(b->sub->nu_t_pre);
      free(b->sub->alpha_t_pre);
      free(b->sub);
    }
  }
  free(o->fy_coarse);
  free(o->nu_t_pre_coarse);
  free(o->alpha_t_pre_coarse);
  free(o);
}

/* ramp_nu_alpha (solver.rs:600-606) and ramp_beta_factor (solver.rs:614-619) */
static void ramp_nu_alpha(const orc_octree *o, uint64_t step, double *nu, double *alpha) {
  if (o->nu_start_mult == 1.0) {
    *nu = o->nu_lu;
    *alpha = o->alpha_lu;
    return;
  }
  double m = 1.0 + (o->nu_start_mult - 1.0) * exp(-(double)step / NU_TAU_RAMP);
  *nu = o->nu_lu * m;
  *alpha = o->alpha_lu * m;
}

static double ramp_beta_factor(const orc_octree *o, uint64_t step) {
  if (o->nu_start_mult == 1.0 || step >= o->spin_up_steps) return 1.0;
  return (double)step / (double)o->spin_up_steps;
}

/* mean_field (solver.rs:1176-1178): sequential sum over the buffer, then / n */
static double mean_field(const double *v, size_t n) {
  double sum = 0.0;
  for (size_t k = 0; k < n; ++k) sum += v[k];
  return sum / (double)n;
}

/* compute_dynamic_sgs_level (smag.rs:144-416) on a level: the uniform model
 * on the level's arrays with its molecular tau, nu and alpha. */
static void level_model(const orc_octree *o, const level *l, double tau_mol, double nu,
                        double alpha, double *nu_t, double *alpha_t, double st[4]) {
  orc_dynamic_model(l->nx, l->ny, &l->f[0].f[0], l->rho, l->ux, l->uy, l->temp, tau_mol,
                    o->smag_alpha, o->eta_test, nu, alpha, o->nu_sm_min, nu_t, alpha_t, st);
}

static double max_abs(const double *v, size_t n, double m) {
  for (size_t k = 0; k < n; ++k) m = fmax(m, fabs(v[k]));
  return m;
}

/* step_cavity (solver.rs:647-947) */
void orc_oct_step(orc_octree *o) {
  tables();
  const int go = o->geq_order;
  const int dyn = o->dynamic_smag;
  double nu_eff, alpha_eff;
  ramp_nu_alpha(o, o->time_step, &nu_eff, &alpha_eff);
  relax tau_coarse = relax_from(nu_eff, alpha_eff);
  relax tau_fine = relax_from(2.0 * nu_eff, 2.0 * alpha_eff);
  relax tau_sub_eff = relax_from(4.0 * nu_eff, 4.0 * alpha_eff);

  /* the model on L0 before its collision, then xi A from the effective
   * relaxation times: the ramped molecular ones plus the mean nu_t / alpha_t
   * of L0 (this step) and of the blocks (their last sub-step)
   * (solver.rs:660-723) */
  const double *nu_l0 = NULL, *al_l0 = NULL;
  if (dyn) {
    const size_t n0 = o->nx0 * o->ny0;
    double st[4];
    level_model(o, &o->coarse, 0.5 + 3.0 * nu_eff, nu_eff, alpha_eff, o->nu_t_pre_coarse,
                o->alpha_t_pre_coarse, st);
    o->cs_dyn_mean = st[0];
    o->cs_dyn_max = st[1];
    o->kappa_dyn_mean = st[2];
    o->kappa_dyn_max = st[3];
    nu_l0 = o->nu_t_pre_coarse;
    al_l0 = o->alpha_t_pre_coarse;
    const double nt0 = mean_field(o->nu_t_pre_coarse, n0);
    const double at0 = mean_field(o->alpha_t_pre_coarse, n0);
    double nt1 = 0.0, at1 = 0.0;
    for (int bi = 0; bi < 2; ++bi) {
      const block_t *b = &o->blocks[bi];
      const size_t n1 = b->grid.nx1 * b->grid.ny1;
      nt1 += mean_field(b->nu_t_pre, n1);
      at1 += mean_field(b->alpha_t_pre, n1);
    }
    const double nb = 2.0;
    relax tce = relax_from(nu_eff + nt0, alpha_eff + at0);
    relax tfe = relax_from(2.0 * nu_eff + nt1 / nb, 2.0 * alpha_eff + at1 / nb);
    o->xi = rescaling(tce, tfe, 2, o->odd_exact);
  }

  double beta_c = o->beta_eff_coarse * ramp_beta_factor(o, o->time_step);
  collide_cavity(&o->coarse, &tau_coarse, beta_c, o->fy_coarse, 0, go, nu_l0, al_l0);
  stream_level(&o->coarse);
  thermal_bc(&o->coarse, go);
  update_macroscopic_guo(&o->coarse, o->fy_coarse);
  walls_level(&o->coarse);

  xi_t xi = o->xi;
  for (int bi = 0; bi < 2; ++bi) coarse_to_fine(&o->blocks[bi].grid, &o->coarse, &o->blocks[bi].fine, &xi, go);

  double beta_ramp = ramp_beta_factor(o, o->time_step);
  double beta_f = o->beta_eff_fine * beta_ramp;
  double beta_sub = 0.25 * o->beta_eff_coarse * beta_ramp;
  const double tau_mol_l1 = 0.5 + 3.0 * 2.0 * nu_eff;
  const double tau_mol_l2 = 0.5 + 3.0 * 4.0 * nu_eff;
  double st[4];
  for (int bi = 0; bi < 2; ++bi) {
    block_t *b = &o->blocks[bi];
    int wl = b->hot, wr = b->cold;
    /* the model on L1 before sub-step 1 (solver.rs:767-786) */
    if (dyn)
      level_model(o, &b->fine, tau_mol_l1, 2.0 * nu_eff, 2.0 * alpha_eff, b->nu_t_pre,
                  b->alpha_t_pre, st);
    collide_cavity(&b->fine, &tau_fine, beta_f, b->fy, 1, go, dyn ? b->nu_t_pre : NULL,
                   dyn ? b->alpha_t_pre : NULL);
    stream_block(&b->fine, wl, wr, 1, 1);
    thermal_bc_block(&b->fine, wl, wr, go);
    update_macroscopic_guo(&b->fine, b->fy);
    walls_block(&b->fine, wl, wr);
    if (b->sub) {
      nested_t *s = b->sub;
      const size_t n1 = b->grid.nx1 * b->grid.ny1, n2 = s->grid.nx1 * s->grid.ny1;
      /* xi B from L1 (this sub-step) and L2 (its last sub-step) (solver.rs:807-830) */
      if (dyn) {
        const double nt1b = mean_field(b->nu_t_pre, n1), at1b = mean_field(b->alpha_t_pre, n1);
        const double nt2 = mean_field(s->nu_t_pre, n2), at2 = mean_field(s->alpha_t_pre, n2);
        relax l1e = relax_from(2.0 * nu_eff + nt1b, 2.0 * alpha_eff + at1b);
        relax l2e = relax_from(4.0 * nu_eff + nt2, 4.0 * alpha_eff + at2);
        s->xi_down = rescaling(l1e, l2e, 2, o->odd_exact);
      }
      xi_t xi_down = s->xi_down;
      coarse_to_fine(&s->grid, &b->fine, &s->lvl, &xi_down, go);
      for (int sp = 0; sp < 2; ++sp) {
        /* the model on L2 before each of its sub-steps (solver.rs:834-850) */
        if (dyn)
          level_model(o, &s->lvl, tau_mol_l2, 4.0 * nu_eff, 4.0 * alpha_eff, s->nu_t_pre,
                      s->alpha_t_pre, st);
        collide_cavity(&s->lvl, &tau_sub_eff, beta_sub, s->fy, 1, go, dyn ? s->nu_t_pre : NULL,
                       dyn ? s->alpha_t_pre : NULL);
        stream_block(&s->lvl, wl, wr, 1, 1);
        thermal_bc_block(&s->lvl, wl, wr, go);
        update_macroscopic_guo(&s->lvl, s->fy);
        walls_block(&s->lvl, wl, wr);
      }
      fine_to_coarse(&s->grid, &s->lvl, &b->fine, &xi_down, go);
      refresh_moments_nodes(&b->fine, s->grid.coinc, s->grid.n_coinc);
    }
    /* the model on L1 before sub-step 2; the reference passes the unramped
     * molecular nu / alpha here (solver.rs:879-892) */
    if (dyn)
      level_model(o, &b->fine, tau_mol_l1, 2.0 * o->nu_lu, 2.0 * o->alpha_lu, b->nu_t_pre,
                  b->alpha_t_pre, st);
    collide_cavity(&b->fine, &tau_fine, beta_f, b->fy, 1, go, dyn ? b->nu_t_pre : NULL,
                   dyn ? b->alpha_t_pre : NULL);
    stream_block(&b->fine, wl, wr, 1, 1);
    thermal_bc_block(&b->fine, wl, wr, go);
    update_macroscopic_guo(&b->fine, b->fy);
    walls_block(&b->fine, wl, wr);
    fine_to_coarse(&b->grid, &b->fine, &o->coarse, &xi, go);
  }
  /* nu_t max over the levels (solver.rs:918-931) */
  if (dyn) {
    double m = max_abs(o->nu_t_pre_coarse, o->nx0 * o->ny0, 0.0);
    for (int bi = 0; bi < 2; ++bi) {
      const block_t *b = &o->blocks[bi];
      m = max_abs(b->nu_t_pre, b->grid.nx1 * b->grid.ny1, m);
      if (b->sub) m = max_abs(b->sub->nu_t_pre, b->sub->grid.nx1 * b->sub->grid.ny1, m);
    }
    o->nu_t_max = m;
  }
  for (int bi = 0; bi < 2; ++bi)
    refresh_moments_nodes(&o->coarse, o->blocks[bi].grid.coinc, o->blocks[bi].grid.n_coinc);
  o->time_step += 1;
}

/* set_dynamic_sgs (solver.rs:476-481) */
void orc_oct_set_dynamic_sgs(orc_octree *o, int on, double smag_alpha, double nu_sm_min,
                             double eta_test) {
  o->dynamic_smag = on ? 1 : 0;
  o->smag_alpha = smag_alpha;
  o->nu_sm_min = nu_sm_min;
  o->eta_test = eta_test;
}

void orc_oct_dynamic_stats(const orc_octree *o, double out[5]) {
  out[0] = o->cs_dyn_mean;
  out[1] = o->cs_dyn_max;
  out[2] = o->kappa_dyn_mean;
  out[3] = o->kappa_dyn_max;
  out[4] = o->nu_t_max;
}

double *orc_oct_sgs_pre(orc_octree *o, size_t gi, int which) {
  if (gi == 0) return which == 0 ? o->nu_t_pre_coarse : o->alpha_t_pre_coarse;
  if (gi <= 2) return which == 0 ? o->blocks[gi - 1].nu_t_pre : o->blocks[gi - 1].alpha_t_pre;
  if (gi <= 4 && o->levels == 3) {
    nested_t *s = o->blocks[gi - 3].sub;
    return which == 0 ? s->nu_t_pre : s->alpha_t_pre;
  }
  return NULL;
}

void orc_oct_xi(const orc_octree *o, double out[9]) {
  out[0] = o->xi.xi_even;
  out[1] = o->xi.xi_odd;
  out[2] = o->xi.xi_thermal;
  for (int bi = 0; bi < 2; ++bi) {
    const nested_t *s = o->blocks[bi].sub;
    out[3 + 3 * bi] = s ? s->xi_down.xi_even : 0.0;
    out[4 + 3 * bi] = s ? s->xi_down.xi_odd : 0.0;
    out[5 + 3 * bi] = s ? s->xi_down.xi_thermal : 0.0;
  }
}

void orc_oct_run(orc_octree *o, uint64_t n) {
  for (uint64_t s = 0; s < n; ++s) orc_oct_step(o);
}

/* ---- grid access in snapshot order (solver.rs:1048-1066) ----------------- */
size_t orc_oct_ngrids(const orc_octree *o) { return o->levels == 2 ? 3 : 5; }

static level *grid_level(orc_octree *o, size_t gi) {
  if (gi == 0) return &o->coarse;
  if (gi <= 2) return &o->blocks[gi - 1].fine;
  if (gi <= 4 && o->levels == 3) return &o->blocks[gi - 3].sub->lvl;
  return NULL;
}

static const grid_t *grid_of(const orc_octree *o, size_t gi) {
  if (gi >= 1 && gi <= 2) return &o->blocks[gi - 1].grid;
  if (gi >= 3 && gi <= 4 && o->levels == 3) return &o->blocks[gi - 3].sub->grid;
  return NULL;
}

static double *grid_fy(orc_octree *o, size_t gi) {
  if (gi == 0) return o->fy_coarse;
  if (gi <= 2) return o->blocks[gi - 1].fy;
  return o->blocks[gi - 3].sub->fy;
}

void orc_oct_grid_dims(const orc_octree *o, size_t gi, size_t *nx, size_t *ny) {
  level *l = grid_level((orc_octree *)o, gi);
  *nx = l ? l->nx : 0;
  *ny = l ? l->ny : 0;
}

double *orc_oct_field(orc_octree *o, size_t gi, int which) {
  level *l = grid_level(o, gi);
  if (!l) return NULL;
  switch (which) {
    case OCT_RHO: return l->rho;
    case OCT_UX: return l->ux;
    case OCT_UY: return l->uy;
    case OCT_T: return l->temp;
    case OCT_FY: return grid_fy(o, gi);
  }
  return NULL;
}

double *orc_oct_pops(orc_octree *o, size_t gi, int which) {
  level *l = grid_level(o, gi);
  if (!l) return NULL;
  return which == 0 ? l->f[0].f : l->g[0].f;
}

uint64_t orc_oct_time_step(const orc_octree *o) { return o->time_step; }

void orc_oct_derived(const orc_octree *o, double out[22]) {
  double nu_eff = o->nu_lu, alpha_eff = o->alpha_lu;
  relax t2 = relax_from(4.0 * nu_eff, 4.0 * alpha_eff);
  xi_t xb = rescaling(o->tau_fine, t2, 2, o->odd_exact);
  out[0] = o->nu_lu;
  out[1] = o->alpha_lu;
  out[2] = o->beta_eff_coarse;
  out[3] = o->beta_eff_fine;
  out[4] = (double)o->spin_up_steps;
  const relax r[3] = {o->tau_coarse, o->tau_fine, t2};
  for (int l = 0; l < 3; ++l) {
    out[5 + 3 * l] = r[l].tau_plus;
    out[6 + 3 * l] = r[l].tau_minus;
    out[7 + 3 * l] = r[l].tau_t;
  }
  out[14] = o->xi.xi_even;
  out[15] = o->xi.xi_odd;
  out[16] = o->xi.xi_thermal;
  out[17] = xb.xi_even;
  out[18] = xb.xi_odd;
  out[19] = xb.xi_thermal;
  size_t total = o->nx0 * o->ny0;
  for (int bi = 0; bi < 2; ++bi) {
    total += o->blocks[bi].grid.nx1 * o->blocks[bi].grid.ny1;
    if (o->blocks[bi].sub) total += o->blocks[bi].sub->grid.nx1 * o->blocks[bi].sub->grid.ny1;
  }
  out[20] = (double)total;
  out[21] = (double)o->levels;
}

/* ---- diagnostics (cavity.rs:672-790, solver.rs:949-1041) ----------------- */
void orc_oct_noise_stats(const orc_octree *o, double out[3]) {
  size_t n0 = o->nx0 * o->ny0;
  double s = 0.0;
  for (size_t k = 0; k < n0; ++k) s += o->coarse.rho[k];
  double rho_bar = s / (double)n0;
  double max_dev = 0.0;
  for (size_t k = 0; k < n0; ++k) max_dev = fmax(max_dev, fabs(o->coarse.rho[k] - rho_bar));
  for (int bi = 0; bi < 2; ++bi) {
    const block_t *b = &o->blocks[bi];
    for (size_t k = 0; k < b->fine.nx * b->fine.ny; ++k)
      max_dev = fmax(max_dev, fabs(b->fine.rho[k] - rho_bar));
    if (b->sub)
      for (size_t k = 0; k < b->sub->lvl.nx * b->sub->lvl.ny; ++k)
        max_dev = fmax(max_dev, fabs(b->sub->lvl.rho[k] - rho_bar));
  }
  double max_if = 0.0;
  for (int bi = 0; bi < 2; ++bi) {
    const block_t *b = &o->blocks[bi];
    for (size_t j = 0; j < b->grid.n_ghosts; ++j)
      max_if = fmax(max_if, fabs(b->fine.rho[b->grid.ghosts[j].k_fine] - rho_bar));
    for (size_t j = 0; j < b->grid.n_coinc; ++j)
      max_if = fmax(max_if, fabs(o->coarse.rho[b->grid.coinc[j].k_coarse] - rho_bar));
    if (b->sub) {
      const nested_t *sb = b->sub;
      for (size_t j = 0; j < sb->grid.n_ghosts; ++j)
        max_if = fmax(max_if, fabs(sb->lvl.rho[sb->grid.ghosts[j].k_fine] - rho_bar));
      for (size_t j = 0; j < sb->grid.n_coinc; ++j)
        max_if = fmax(max_if, fabs(b->fine.rho[sb->grid.coinc[j].k_coarse] - rho_bar));
    }
  }
  out[0] = max_dev;
  out[1] = max_if;
  out[2] = rho_bar;
}

static double max_u_level(const level *l) {
  double m = 0.0;
  for (size_t k = 0; k < l->nx * l->ny; ++k) m = fmax(m, fmax(fabs(l->ux[k]), fabs(l->uy[k])));
  return m;
}

double orc_oct_max_u_all(const orc_octree *o) {
  double m = max_u_level(&o->coarse);
  for (int bi = 0; bi < 2; ++bi) {
    m = fmax(m, max_u_level(&o->blocks[bi].fine));
    if (o->blocks[bi].sub) m = fmax(m, max_u_level(&o->blocks[bi].sub->lvl));
  }
  return m;
}

static int level_finite(const level *l) {
  size_t n = l->nx * l->ny;
  for (size_t k = 0; k < n; ++k)
    if (!isfinite(l->rho[k]) || !isfinite(l->ux[k]) || !isfinite(l->uy[k]) ||
        !isfinite(l->temp[k]))
      return 0;
  for (size_t k = 0; k < n; ++k)
    for (int i = 0; i < 9; ++i)
      if (!isfinite(l->f[k].f[i]) || !isfinite(l->g[k].f[i])) return 0;
  return 1;
}

int orc_oct_check_guards(const orc_octree *o) {
  if (orc_oct_max_u_all(o) > MAX_U_ABORT) return 1;
  if (!level_finite(&o->coarse)) return 2;
  for (int bi = 0; bi < 2; ++bi) {
    if (!level_finite(&o->blocks[bi].fine)) return 2;
    if (o->blocks[bi].sub && !level_finite(&o->blocks[bi].sub->lvl)) return 2;
  }
  return 0;
}

double orc_oct_nusselt_hot(const orc_octree *o) {
  size_t nx = o->coarse.nx, ny = o->coarse.ny;
  double sum = 0.0;
  for (size_t iy = 0; iy < ny; ++iy)
    sum += fabs(o->coarse.temp[iy * nx + 1] - o->coarse.temp[iy * nx + 2]);
  double avg = sum / (double)ny;
  return avg * (double)(nx - 1) / DELTA_T;
}

double orc_oct_nusselt_hot_link0(const orc_octree *o) {
  size_t nx = o->coarse.nx, ny = o->coarse.ny;
  double sum = 0.0;
  for (size_t iy = 0; iy < ny; ++iy)
    sum += fabs(o->coarse.temp[iy * nx + 1] - o->coarse.temp[iy * nx]);
  double avg = sum / (double)ny;
  return avg * (double)(nx - 1) / DELTA_T;
}

double orc_oct_nusselt_cold(const orc_octree *o) {
  size_t nx = o->coarse.nx, ny = o->coarse.ny;
  double sum = 0.0;
  for (size_t iy = 0; iy < ny; ++iy)
    sum += fabs(o->coarse.temp[iy * nx + nx - 2] - o->coarse.temp[iy * nx + nx - 3]);
  double avg = sum / (double)ny;
  return avg * (double)(nx - 1) / DELTA_T;
}

double orc_oct_nusselt_hot_fine(const orc_octree *o, size_t block, double lx_coarse) {
  const level *f = &o->blocks[block].fine;
  size_t nx = f->nx, ny = f->ny;
  double sum = 0.0;
  for (size_t b = 1; b + 1 < ny; ++b)
    sum += fabs(f->temp[b * nx + 1] - f->temp[b * nx + 2]) * 2.0;
  double avg = sum / (double)(ny - 2);
  return avg * lx_coarse / DELTA_T;
}

/* set_from_snapshot (solver.rs:1095-1175), without the dimension checks that
 * the caller performs */
int orc_oct_set_from_snapshot(orc_octree *o, const double *rho, const double *ux,
                              const double *uy, const double *t, size_t total) {
  size_t need = 0, ng = orc_oct_ngrids(o);
  for (size_t gi = 0; gi < ng; ++gi) need += grid_level(o, gi)->nx * grid_level(o, gi)->ny;
  if (total != need) return -1;
  size_t off = 0;
  for (size_t gi = 0; gi < ng; ++gi) {
    level *l = grid_level(o, gi);
    for (size_t k = 0; k < l->nx * l->ny; ++k) {
      double r = rho[off + k], un = ux[off + k], vn = uy[off + k], tn = t[off + k];
      l->rho[k] = r;
      l->ux[k] = un;
      l->uy[k] = vn;
      l->temp[k] = tn;
      for (int i = 0; i < 9; ++i) {
        l->f[k].f[i] = orc_feq(i, r, un, vn);
        l->g[k].f[i] = orc_geq(i, tn, un, vn, o->geq_order);
      }
    }
    off += l->nx * l->ny;
  }
  walls_level(&o->coarse);
  for (int bi = 0; bi < 2; ++bi) {
    walls_block(&o->blocks[bi].fine, o->blocks[bi].hot, o->blocks[bi].cold);
    if (o->blocks[bi].sub) walls_block(&o->blocks[bi].sub->lvl, o->blocks[bi].hot, o->blocks[bi].cold);
  }
  o->time_step = 0;
  return 0;
}

size_t orc_oct_n_ghosts(const orc_octree *o, size_t gi) {
  const grid_t *g = grid_of(o, gi);
  return g ? g->n_ghosts : 0;
}

size_t orc_oct_n_coincident(const orc_octree *o, size_t gi) {
  const grid_t *g = grid_of(o, gi);
  return g ? g->n_coinc : 0;
}

void orc_oct_ghost(const orc_octree *o, size_t gi, size_t j, size_t *k_fine, size_t kc[4],
                   double w[4]) {
  const ghost_entry *e = &grid_of(o, gi)->ghosts[j];
  *k_fine = e->k_fine;
  memcpy(kc, e->kc, sizeof(e->kc));
  memcpy(w, e->w, sizeof(e->w));
}

void orc_oct_coincident(const orc_octree *o, size_t gi, size_t j, size_t *k_coarse,
                        size_t kf[9], double w[9]) {
  const coinc_entry *e = &grid_of(o, gi)->coinc[j];
  *k_coarse = e->k_coarse;
  memcpy(kf, e->kf, sizeof(e->kf));
  memcpy(w, e->w, sizeof(e->w));
}

/* ---- per-level phase entry points ---------------------------------------- */
void orc_oct_collide_level(orc_octree *o, size_t gi, const double tau[3], double beta,
                           int interior_only) {
  relax r = {tau[0], tau[1], tau[2]};
  collide_cavity(grid_level(o, gi), &r, beta, grid_fy(o, gi), interior_only, o->geq_order,
                 NULL, NULL);
}

void orc_oct_collide_node(double f[9], double g[9], double *fy_out, double rho, double ux,
                          double uy, double temp, const double tau[3], double beta,
                          int geq_order, int pre, double nu_t, double alpha_t) {
  tables();
  relax r = {tau[0], tau[1], tau[2]};
  collide_node_cavity(f, g, fy_out, rho, ux, uy, temp, &r, beta, geq_order == 1 ? 1 : 2, pre,
                      nu_t, alpha_t);
}
void orc_oct_stream_level(orc_octree *o, size_t gi) { stream_level(grid_level(o, gi)); }
void orc_oct_stream_block_level(orc_octree *o, size_t gi, int wl, int wr, int wb, int wt) {
  stream_block(grid_level(o, gi), wl, wr, wb, wt);
}
void orc_oct_bc_level(orc_octree *o, size_t gi) { thermal_bc(grid_level(o, gi), o->geq_order); }
void orc_oct_bc_block_level(orc_octree *o, size_t gi, int hot, int cold) {
  thermal_bc_block(grid_level(o, gi), hot, cold, o->geq_order);
}
void orc_oct_macro_level(orc_octree *o, size_t gi) {
  update_macroscopic_guo(grid_level(o, gi), grid_fy(o, gi));
}
void orc_oct_walls_level(orc_octree *o, size_t gi) { walls_level(grid_level(o, gi)); }
void orc_oct_walls_block_level(orc_octree *o, size_t gi, int hot, int cold) {
  walls_block(grid_level(o, gi), hot, cold);
}

/* ---- OctreeSolver, the two-level interface harness (solver.rs:66-116) ---- */
struct orc_oct2 {
  grid_t grid;
  level coarse, fine;
  relax tau_coarse, tau_fine;
  xi_t xi;
};

orc_oct2 *orc_oct2_new(size_t nx0, size_t ny0, size_t bx0, size_t by0, size_t bx1, size_t by1,
                       double nu, double alpha, int odd_exact) {
  tables();
  if (!(bx1 <= nx0 && by1 <= ny0 && bx0 < bx1 && by0 < by1)) return NULL; /* grid.rs:133-134 */
  orc_oct2 *s = (orc_oct2 *)xcalloc(1, sizeof(orc_oct2));
  grid_build(&s->grid, nx0, ny0, bx0, by0, bx1, by1, 0, 0, 0, 0);
  s->tau_coarse = relax_from(nu, alpha);
  s->tau_fine = relax_from(2.0 * nu, 2.0 * alpha);
  s->xi = rescaling(s->tau_coarse, s->tau_fine, 2, odd_exact);
  level_init(&s->coarse, nx0, ny0);
  level_init(&s->fine, s->grid.nx1, s->grid.ny1);
  return s;
}

void orc_oct2_free(orc_oct2 *s) {
  if (!s) return;
  grid_free(&s->grid);
  level_free(&s->coarse);
  level_free(&s->fine);
  free(s);
}

void orc_oct2_dims(const orc_oct2 *s, size_t *nx1, size_t *ny1) {
  *nx1 = s->grid.nx1;
  *ny1 = s->grid.ny1;
}

double *orc_oct2_field(orc_oct2 *s, int lv, int which) {
  level *l = lv == 0 ? &s->coarse : &s->fine;
  switch (which) {
    case OCT_RHO: return l->rho;
    case OCT_UX: return l->ux;
    case OCT_UY: return l->uy;
    case OCT_T: return l->temp;
  }
  return NULL;
}

double *orc_oct2_pops(orc_oct2 *s, int lv, int which) {
  level *l = lv == 0 ? &s->coarse : &s->fine;
  return which == 0 ? l->f[0].f : l->g[0].f;
}

void orc_oct2_xi(const orc_oct2 *s, double out[3]) {
  out[0] = s->xi.xi_even;
  out[1] = s->xi.xi_odd;
  out[2] = s->xi.xi_thermal;
}

void orc_oct2_coarse_to_fine(orc_oct2 *s) { coarse_to_fine(&s->grid, &s->coarse, &s->fine, &s->xi, 2); }
void orc_oct2_fine_to_coarse(orc_oct2 *s) { fine_to_coarse(&s->grid, &s->fine, &s->coarse, &s->xi, 2); }
size_t orc_oct2_n_ghosts(const orc_oct2 *s) { return s->grid.n_ghosts; }
size_t orc_oct2_n_coincident(const orc_oct2 *s) { return s->grid.n_coinc; }
size_t orc_oct2_ghost_k(const orc_oct2 *s, size_t j) { return s->grid.ghosts[j].k_fine; }
void orc_oct2_coincident(const orc_oct2 *s, size_t j, size_t *k_coarse, size_t kf[9],
                         double w[9]) {
  const coinc_entry *e = &s->grid.coinc[j];
  *k_coarse = e->k_coarse;
  memcpy(kf, e->kf, sizeof(e->kf));
  memcpy(w, e->w, sizeof(e->w));
}

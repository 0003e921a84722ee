// cavlbm_octree.cpp — the octree (multi-level) cavity solver behind the C-ABI
// (include/cavlbm.h, cavlbm_octree_*).  Reference: bbrangeo/cfd_cavity
// src/octree/solver.rs (CavitySolver, step_cavity), grid.rs (interface
// tables), interface.rs (rescaling factors), cavity.rs (diagnostics).
//
// All levels live on the device in M-form between phases; one coarse step is
// a fixed sequence of per-level collide / stream launches and interface
// transfers (lbm_octree.hip), captured once into a hipGraph when its
// per-step scalars are constant (no viscosity ramp).  Diagnostics copy the
// level fields back and reduce them on the host in the reference's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cavlbm.h"
#include "lbm_launch.hpp"
#include "lbm_octree.hpp"

using namespace cavlbm;

namespace {

thread_local std::string g_oct_create_error;

constexpr double kCScale = 1.7;        // calibration constant (cavity.rs:49)
constexpr double kMaxUAbort = 0.10;    // cavity.rs:41
constexpr double kNuTauRamp = 1000.0;  // solver.rs:58
constexpr double kLambdaTrt = 3.0 / 16.0;

struct Relax {
  double tp, tm, tt;
};

// RelaxTimes::from_nu_alpha (lbm.rs:100-109)
Relax relax_from(double nu, double alpha) {
  Relax r;
  r.tp = 0.5 + 3.0 * nu;
  r.tm = 0.5 + kLambdaTrt / (r.tp - 0.5);
  r.tt = 0.5 + 3.0 * alpha;
  return r;
}

// rescaling_factors_multilevel (interface.rs:141-158)
OctXi rescaling(const Relax& c, const Relax& f, int substeps, bool odd_exact) {
  const double dt = 1.0 / (double)substeps;
  OctXi x;
  x.even = dt * f.tp / c.tp;
  x.odd = odd_exact ? dt * f.tm / c.tm : dt;
  x.thermal = dt * f.tt / c.tt;
  x.inv_even = 1.0 / x.even;
  x.inv_odd = 1.0 / x.odd;
  x.inv_thermal = 1.0 / x.thermal;
  return x;
}

// bilinear_stencil (grid.rs:216-229)
void bilinear(double x_c, double y_c, size_t nx0, size_t ny0, int k[4], double w[4]) {
  long long ix = (long long)std::floor(x_c), iy = (long long)std::floor(y_c);
  ix = std::min<long long>(std::max<long long>(ix, 0), (long long)nx0 - 2);
  iy = std::min<long long>(std::max<long long>(iy, 0), (long long)ny0 - 2);
  const double fx = x_c - (double)ix, fy = y_c - (double)iy;
  k[0] = (int)(iy * (long long)nx0 + ix);
  w[0] = (1.0 - fx) * (1.0 - fy);
  k[1] = (int)(iy * (long long)nx0 + ix + 1);
  w[1] = fx * (1.0 - fy);
  k[2] = (int)((iy + 1) * (long long)nx0 + ix);
  w[2] = (1.0 - fx) * fy;
  k[3] = (int)((iy + 1) * (long long)nx0 + ix + 1);
  w[3] = fx * fy;
}

// OctreeGrid::build (grid.rs:121-203): the coarse side is nx0 x ny0, the fine
// block [bx0, bx1) x [by0, by1) in coarse cells with one ghost layer.
struct HostGrid {
  size_t nx0 = 0, ny0 = 0, bx0 = 0, by0 = 0, bx1 = 0, by1 = 0, nx1 = 0, ny1 = 0;
  std::vector<int> gk, gkc, ck, ckf;
  std::vector<double> gw, cw;
};

HostGrid build_grid(size_t nx0, size_t ny0, size_t bx0, size_t by0, size_t bx1, size_t by1,
                    bool wl, bool wr, bool wb, bool wt) {
  HostGrid g;
  g.nx0 = nx0;
  g.ny0 = ny0;
  g.bx0 = bx0;
  g.by0 = by0;
  g.bx1 = bx1;
  g.by1 = by1;
  g.nx1 = 2 * (bx1 - bx0) + 3;
  g.ny1 = 2 * (by1 - by0) + 3;
  for (size_t b = 0; b < g.ny1; ++b)
    for (size_t a = 0; a < g.nx1; ++a)
      if (a == 0 || b == 0 || a == g.nx1 - 1 || b == g.ny1 - 1) {
        int k[4];
        double w[4];
        bilinear((double)bx0 - 0.5 + 0.5 * (double)a, (double)by0 - 0.5 + 0.5 * (double)b, nx0,
                 ny0, k, w);
        g.gk.push_back((int)(b * g.nx1 + a));
        g.gkc.insert(g.gkc.end(), k, k + 4);
        g.gw.insert(g.gw.end(), w, w + 4);
      }
  static const double wt3[3] = {0.25, 0.5, 0.25};
  for (size_t cy = by0; cy < by1; ++cy)
    for (size_t cx = bx0; cx < bx1; ++cx) {
      if (cx == 0 || cx + 1 == nx0 || cy == 0 || cy + 1 == ny0) continue;
      if ((wl && cx == bx0) || (wr && cx + 1 == bx1) || (wb && cy == by0) ||
          (wt && cy + 1 == by1))
        continue;
      const size_t a = 2 * (cx - bx0) + 1, b = 2 * (cy - by0) + 1;
      g.ck.push_back((int)(cy * nx0 + cx));
      for (int jb = 0; jb < 3; ++jb)
        for (int ja = 0; ja < 3; ++ja) {
          g.ckf.push_back((int)((b + jb - 1) * g.nx1 + (a + ja - 1)));
          g.cw.push_back(wt3[ja] * wt3[jb]);
        }
    }
  return g;
}

}  // namespace

struct cavlbm_octree {
  cavlbm_octree_params p{};
  int levels = 2, ngrids = 3, geq = 2;
  size_t nx0 = 0, ny0 = 0;
  OctLevel lv[5]{};
  HostGrid hg[5];  // hg[gi], gi >= 1: tables whose fine side is grid gi
  OctGhostTab gt[5]{};
  OctCoincTab ct[5]{};
  std::vector<void*> dev;  // every device allocation
  Relax tau_coarse{}, tau_fine{}, tau_sub{};
  OctXi xi_a{}, xi_b{};
  double nu_lu = 0, alpha_lu = 0, nu_start_mult = 1, beta_eff_coarse = 0, beta_eff_fine = 0;
  uint64_t spin_up_steps = 0, time_step = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // The two M-form buffers of each level (the tiled sub-step kernel reads one
  // and writes the other); bit gi of a key: level gi's m is mbuf[gi][1].
  double* mbuf[5][2] = {};
  int tile_ty[5][3] = {};  // output rows of the tiled kernel, by level and sub-steps (0: none)
  bool tiled = true;       // use the tiled kernel where a level fits it
  // coarse-step graphs by the key at their start, and the key they end on
  hipGraphExec_t graph[32] = {};
  int graph_next[32] = {};
  bool graph_failed = false;
  // Dynamic mixed-scale SGS (solver.rs:229-248): per grid kDynPlanes FP64
  // planes (the model's scratch, C_nu 16, C_kappa 17, nu_t_pre 18, alpha_t_pre
  // 19), allocated and zeroed on the first enable; the interface factors on
  // the device ([0] A, [1] / [2] B of the hot / cold block), which the
  // transfers use from then on (the reference keeps the last recomputed
  // factors when the model is switched off).
  bool dyn_on = false;
  double dyn_alpha = 0.5, dyn_min = -0.1, dyn_eta = 2.0;
  double* dyn[5] = {};
  double* dyn_part[5] = {};  // per level: the coefficient pass's per-workgroup sums
  OctXi* xi_dev = nullptr;
  std::string err;
};

namespace {

int ofail(cavlbm_octree* o, int code, const std::string& m) {
  if (o)
    o->err = m;
  else
    g_oct_create_error = m;
  return code;
}

#define OHIP_TRY(o, expr)                                                              \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return ofail((o), CAVLBM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

void free_octree(cavlbm_octree* o) {
  for (auto& g : o->graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (void* p : o->dev)
    if (p) (void)hipFree(p);
  if (o->ev0) (void)hipEventDestroy(o->ev0);
  if (o->ev1) (void)hipEventDestroy(o->ev1);
  if (o->st) (void)hipStreamDestroy(o->st);
}

template <class T>
T* dalloc(cavlbm_octree* o, size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
  o->dev.push_back(p);
  return static_cast<T*>(p);
}

// Stream/BC rules of grid gi: the coarse level has domain walls; a block or
// strip bounces back from its wall-side ghost layers (stream_cavity_block
// with wall_left = hot, wall_right = cold, bottom and top walls), with the
// isothermal column of its wall (solver.rs:785-789, :846-850).
OctStream stream_rules(const cavlbm_octree* o, int gi) {
  OctStream s{};
  if (gi == 0) return s;
  const bool hot = gi == 1 || gi == 3;
  s.block = 1;
  s.wl = hot;
  s.wr = !hot;
  s.wb = 1;
  s.wt = 1;
  s.hot = hot;
  s.cold = !hot;
  (void)o;
  return s;
}

// ramp_nu_alpha (solver.rs:600-606) and ramp_beta_factor (solver.rs:614-619)
void ramp(const cavlbm_octree* o, uint64_t step, double& nu, double& alpha) {
  if (o->nu_start_mult == 1.0) {
    nu = o->nu_lu;
    alpha = o->alpha_lu;
    return;
  }
  const double m = 1.0 + (o->nu_start_mult - 1.0) * std::exp(-(double)step / kNuTauRamp);
  nu = o->nu_lu * m;
  alpha = o->alpha_lu * m;
}

double ramp_beta(const cavlbm_octree* o, uint64_t step) {
  if (o->nu_start_mult == 1.0 || step >= o->spin_up_steps) return 1.0;
  return (double)step / (double)o->spin_up_steps;
}

OctCollide collide_params(double beta, double nu, double alpha, const Relax& r, bool interior) {
  OctCollide p{};
  p.ctx.beta = beta;
  p.ctx.nu = nu;
  p.ctx.alpha = alpha;
  p.ctx.tau = r.tp;  // enters |S| only, multiplied by Cs = 0
  p.ctx.t_ref = 0.0;
  p.ctx.cs = 0.0;
  p.ctx.pr_t = 1.0;
  p.ctx.inv_pr_t = 1.0;
  // the level's RelaxTimes as relaxation factors, in collide_node's order
  // with its nu_t = alpha_t = +0 (Cs = 0)
  p.rf = relax_factors<double>(0.5 + 3.0 * (nu + 0.0), 0.5 + 3.0 * (alpha + 0.0));
  p.interior_only = interior ? 1 : 0;
  return p;
}

// Which buffer holds each level's M-form, as a key (bit gi set: mbuf[gi][1]).
int mkey(const cavlbm_octree* o) {
  int key = 0;
  for (int gi = 0; gi < o->ngrids; ++gi)
    if (o->lv[gi].m == o->mbuf[gi][1]) key |= 1 << gi;
  return key;
}

void set_mkey(cavlbm_octree* o, int key) {
  for (int gi = 0; gi < o->ngrids; ++gi) {
    const int b = (key >> gi) & 1;
    o->lv[gi].m = o->mbuf[gi][b];
    o->lv[gi].m2 = o->mbuf[gi][b ^ 1];
  }
}

// ns whole sub-steps of `count` levels in one tiled launch when every level
// fits the tiling (then their M-forms move to the other buffer); false:
// nothing enqueued.
bool tiled_substeps(cavlbm_octree* o, int g0, int count, int ns, const OctCollide& c,
                    hipError_t& e) {
  const int g1 = g0 + (count > 1 ? 1 : 0);
  const int ty = o->tile_ty[g0][ns];
  if (!o->tiled || ty == 0 || o->tile_ty[g1][ns] != ty) return false;
  OctLevel L[2] = {o->lv[g0], o->lv[g1]};
  OctStream S[2] = {stream_rules(o, g0), stream_rules(o, g1)};
  e = launch_oct_tile(L, S, count, ty, ns, c, o->geq, o->st);
  if (e == hipSuccess)
    for (int i = 0; i < count; ++i) std::swap(o->lv[g0 + i].m, o->lv[g0 + i].m2);
  return true;
}

// One sub-step of `count` levels of the same kind (the two blocks, or the two
// strips): collide_cavity, then stream + BCs + moments + walls.
hipError_t substep(cavlbm_octree* o, int g0, int count, const OctCollide& c) {
  hipError_t te;
  if (tiled_substeps(o, g0, count, 1, c, te)) return te;
  OctLevel L[2] = {o->lv[g0], o->lv[g0 + (count > 1 ? 1 : 0)]};
  OctStream S[2] = {stream_rules(o, g0), stream_rules(o, g0 + (count > 1 ? 1 : 0))};
  hipError_t e = launch_oct_collide(L, count, c, o->geq, o->st);
  if (e != hipSuccess) return e;
  return launch_oct_stream(L, S, count, o->geq, o->st);
}

// Two sub-steps with no transfer in between: collide, the fused stream of
// the first and collision of the second (C-form -> second C-form), and the
// stream of the second from that second C-form.
hipError_t double_substep(cavlbm_octree* o, int g0, int count, const OctCollide& c) {
  hipError_t te;
  if (tiled_substeps(o, g0, count, 2, c, te)) return te;
  OctLevel L[2] = {o->lv[g0], o->lv[g0 + (count > 1 ? 1 : 0)]};
  OctStream S[2] = {stream_rules(o, g0), stream_rules(o, g0 + (count > 1 ? 1 : 0))};
  hipError_t e = launch_oct_collide(L, count, c, o->geq, o->st);
  if (e == hipSuccess) e = launch_oct_stream_collide(L, S, count, c, o->geq, o->st);
  if (e != hipSuccess) return e;
  OctLevel L2[2] = {L[0], L[1]};
  for (auto& l : L2) std::swap(l.c, l.c2);
  return launch_oct_stream(L2, S, count, o->geq, o->st);
}

// compute_dynamic_sgs_level (smag.rs:144-416) on grid gi's current M-form:
// the uniform model's filter, contractions and coefficient passes with the
// level's molecular tau, nu and alpha; nu_t / alpha_t land in planes 18 / 19.
// The model passes' geometry of level gi (flat layout, kTileX x kTileY
// workgroups).
Geom model_geom(const cavlbm_octree* o, int gi) {
  const OctLevel& L = o->lv[gi];
  Geom geo{};
  geo.nx = L.nx;
  geo.ny = L.ny;
  geo.pitch = L.nx;
  geo.nbx = (L.nx + kTileX - 1) / kTileX;
  geo.nby = (L.ny + kTileY - 1) / kTileY;
  geo.plane = L.plane;
  return geo;
}

// smag.rs:170-175: eta = max(eta_test, 1), width = round(eta) nodes (from
// 2 max(nx, ny) + 1 on, every width averages the whole level).
int level_width(const cavlbm_octree* o, int gi) {
  const OctLevel& L = o->lv[gi];
  const double eta = std::fmax(o->dyn_eta, 1.0);
  const double cap = 2.0 * (double)std::max(L.nx, L.ny) + 1.0;
  const double wr = std::round(eta);
  const int w = (int)(wr < cap ? wr : cap);
  return w < 1 ? 1 : w;
}

void level_model_args(const cavlbm_octree* o, int gi, double tau_mol, double nu, double alpha,
                      Geom& geo, DynSgsArgs& a) {
  const OctLevel& L = o->lv[gi];
  geo = model_geom(o, gi);
  const size_t P = (size_t)L.plane;
  a = DynSgsArgs{};
  a.mpops = L.m;
  a.rho = L.m + kOctRho * P;
  a.ux = L.m + kOctUx * P;
  a.uy = L.m + kOctUy * P;
  a.temp = L.m + kOctT * P;
  a.tau = tau_mol;
  a.width = level_width(o, gi);
  a.half_left = (a.width - 1) / 2;
  a.dtest = std::fmax(o->dyn_eta, 1.0) * 1.0;
  a.alpha_exp = o->dyn_alpha;
  a.beta_exp = 1.0 - o->dyn_alpha;
  const double margin = (kTauTrtFloor - 0.5) / 3.0;
  a.nu_t_min = std::fmax(o->dyn_min * nu, margin - nu);
  a.alpha_t_min = std::fmax(o->dyn_min * alpha, margin - alpha);
  double* q = o->dyn[gi];
  a.ubx = q + 0 * P;
  a.uby = q + 1 * P;
  a.tb = q + 2 * P;
  a.lb = q + 3 * P;
  a.qsm = q + 8 * P;
  a.phi = q + 9 * P;
  a.terms = q + 10 * P;
  a.nu_mix_g = q + 14 * P;
  a.kap_mix_g = q + 15 * P;
  a.cnu = q + 16 * P;
  a.ck = q + 17 * P;
  a.nut_pre = q + 18 * P;
  a.at_pre = q + 19 * P;
  a.part = o->dyn_part[gi];
}

// The model on level gi, or on levels gi and gi + 1 (the hot and cold blocks
// or strips, same molecular nu / alpha) in the same launches.
hipError_t level_model(cavlbm_octree* o, int gi, double tau_mol, double nu, double alpha,
                       int count = 1) {
  Geom geo[2];
  DynSgsArgs a[2];
  for (int i = 0; i < count; ++i) level_model_args(o, gi + i, tau_mol, nu, alpha, geo[i], a[i]);
  return launch_dynamic_model(geo, a, count, o->st);
}

const double* nut_of(const cavlbm_octree* o, int gi) {
  return o->dyn[gi] + 18 * (size_t)o->lv[gi].plane;
}
const double* at_of(const cavlbm_octree* o, int gi) {
  return o->dyn[gi] + 19 * (size_t)o->lv[gi].plane;
}

// The per-node relaxation of levels g0 (and g0 + 1) in a collision pass.
void with_model(const cavlbm_octree* o, OctCollide& c, int g0, int count, const Relax& r) {
  for (int i = 0; i < count; ++i) {
    c.nut[i] = nut_of(o, g0 + i);
    c.at[i] = at_of(o, g0 + i);
  }
  c.tp = r.tp;
  c.tt = r.tt;
}

// The factors of one interface from the mean nu_t / alpha_t of its coarse
// grid gc and of the fine grids gf[0..nf) (solver.rs:690-722, :807-830).
OctXiUpdate xi_update(const cavlbm_octree* o, int gc, const int* gf, int nf, double nu_c,
                      double alpha_c, double nu_f, double alpha_f, OctXi* out) {
  OctXiUpdate u{};
  u.part_c = o->dyn_part[gc];
  u.nb_c = dyn_model_blocks(model_geom(o, gc), level_width(o, gc));
  u.n_c = o->lv[gc].plane;
  for (int i = 0; i < nf; ++i) {
    u.part_f[i] = o->dyn_part[gf[i]];
    u.nb_f[i] = dyn_model_blocks(model_geom(o, gf[i]), level_width(o, gf[i]));
    u.n_f[i] = o->lv[gf[i]].plane;
  }
  u.nf = nf;
  u.nu_c = nu_c;
  u.alpha_c = alpha_c;
  u.nu_f = nu_f;
  u.alpha_f = alpha_f;
  u.odd_exact = o->p.odd_exact != 0;
  u.out = out;
  return u;
}

// step_cavity (solver.rs:647-947): enqueues the launches of the coarse step
// `step`.  With the dynamic model the model runs on each level before each
// of its collisions, which use the per-node nu_t / alpha_t, and the interface
// factors are recomputed on the device (the tiled kernel then takes one
// sub-step per launch).  The reference walks the hot block and
// then the cold one; the two touch disjoint fine levels and disjoint coarse
// nodes after the coarse-to-fine injection, so both advance in the same
// launches.
hipError_t enqueue_step(cavlbm_octree* o, uint64_t step) {
  double nu_eff, alpha_eff;
  ramp(o, step, nu_eff, alpha_eff);
  const Relax tc = relax_from(nu_eff, alpha_eff);
  const Relax tf = relax_from(2.0 * nu_eff, 2.0 * alpha_eff);
  const Relax ts = relax_from(4.0 * nu_eff, 4.0 * alpha_eff);
  const double br = ramp_beta(o, step);
  const double beta_c = o->beta_eff_coarse * br;
  const double beta_f = o->beta_eff_fine * br;
  const double beta_s = 0.25 * o->beta_eff_coarse * br;
  OctCollide c0 = collide_params(beta_c, nu_eff, alpha_eff, tc, false);
  OctCollide c1 = collide_params(beta_f, 2.0 * nu_eff, 2.0 * alpha_eff, tf, true);
  OctCollide c2 = collide_params(beta_s, 4.0 * nu_eff, 4.0 * alpha_eff, ts, true);
  const bool dyn = o->dyn_on;
  // the molecular tau of each level's model (solver.rs:665, :761-763)
  const double tau_l0 = 0.5 + 3.0 * nu_eff;
  const double tau_l1 = 0.5 + 3.0 * 2.0 * nu_eff;
  const double tau_l2 = 0.5 + 3.0 * 4.0 * nu_eff;
  // interface factors: the device's once the model has been enabled
  const OctXi* xa[2] = {o->xi_dev, o->xi_dev};
  const OctXi* xb[2] = {o->xi_dev ? o->xi_dev + 1 : nullptr, o->xi_dev ? o->xi_dev + 2 : nullptr};
  const OctXi* const* xi_a = o->xi_dev ? xa : nullptr;
  const OctXi* const* xi_b = o->xi_dev ? xb : nullptr;
  // level descriptors are read right before each transfer and model pass: a
  // tiled sub-step moves its levels' M-forms to the other buffer
  auto coarse2 = [o](OctLevel* L) { L[0] = L[1] = o->lv[0]; };
  auto pair_of = [o](OctLevel* L, int g0) {
    L[0] = o->lv[g0];
    L[1] = o->lv[g0 + 1];
  };
  OctLevel lc[2], lf[2];
  hipError_t e;
#define OCT_Q(x)                   \
  do {                             \
    e = (x);                       \
    if (e != hipSuccess) return e; \
  } while (0)
  if (dyn) {
    // the model on L0, then interface A from L0 (this step) and the blocks
    // (their last sub-step) (solver.rs:663-723)
    OCT_Q(level_model(o, 0, tau_l0, nu_eff, alpha_eff));
    const int blocks[2] = {1, 2};
    const OctXiUpdate u = xi_update(o, 0, blocks, 2, nu_eff, alpha_eff, 2.0 * nu_eff,
                                    2.0 * alpha_eff, o->xi_dev);
    OCT_Q(launch_oct_xi_update(&u, 1, o->st));
    with_model(o, c0, 0, 1, tc);
    with_model(o, c1, 1, 2, tf);
    with_model(o, c2, 3, o->levels == 3 ? 2 : 0, ts);
  }
  OCT_Q(substep(o, 0, 1, c0));
  coarse2(lc);
  pair_of(lf, 1);
  OCT_Q(launch_oct_coarse_to_fine(lc, lf, &o->gt[1], 2, o->xi_a, xi_a, o->geq, o->st));
  if (dyn) {  // the blocks' model before sub-step 1 (solver.rs:767-786)
    OCT_Q(level_model(o, 1, tau_l1, 2.0 * nu_eff, 2.0 * alpha_eff, 2));
  }
  if (o->levels == 3) {
    OCT_Q(substep(o, 1, 2, c1));
    if (dyn) {  // interface B of each block from L1 (now) and L2 (last) (:807-830)
      const int h[1] = {3}, c[1] = {4};
      const OctXiUpdate u[2] = {
          xi_update(o, 1, h, 1, 2.0 * nu_eff, 2.0 * alpha_eff, 4.0 * nu_eff, 4.0 * alpha_eff,
                    o->xi_dev + 1),
          xi_update(o, 2, c, 1, 2.0 * nu_eff, 2.0 * alpha_eff, 4.0 * nu_eff, 4.0 * alpha_eff,
                    o->xi_dev + 2)};
      OCT_Q(launch_oct_xi_update(u, 2, o->st));
    }
    pair_of(lc, 1);
    pair_of(lf, 3);
    OCT_Q(launch_oct_coarse_to_fine(lc, lf, &o->gt[3], 2, o->xi_b, xi_b, o->geq, o->st));
    if (dyn) {
      for (int sp = 0; sp < 2; ++sp) {  // the strips' model before each sub-step (:834-850)
        OCT_Q(level_model(o, 3, tau_l2, 4.0 * nu_eff, 4.0 * alpha_eff, 2));
        OCT_Q(substep(o, 3, 2, c2));
      }
    } else {
      OCT_Q(double_substep(o, 3, 2, c2));
    }
    pair_of(lf, 3);
    pair_of(lc, 1);
    OCT_Q(launch_oct_fine_to_coarse(lf, lc, &o->ct[3], 2, o->xi_b, xi_b, o->geq, o->st));
    if (dyn) {  // before sub-step 2 the reference passes the unramped nu / alpha (:879-892)
      OCT_Q(level_model(o, 1, tau_l1, 2.0 * o->nu_lu, 2.0 * o->alpha_lu, 2));
    }
    OCT_Q(substep(o, 1, 2, c1));
  } else if (dyn) {
    OCT_Q(substep(o, 1, 2, c1));
    OCT_Q(level_model(o, 1, tau_l1, 2.0 * o->nu_lu, 2.0 * o->alpha_lu, 2));
    OCT_Q(substep(o, 1, 2, c1));
  } else {
    OCT_Q(double_substep(o, 1, 2, c1));
  }
  pair_of(lf, 1);
  coarse2(lc);
  OCT_Q(launch_oct_fine_to_coarse(lf, lc, &o->ct[1], 2, o->xi_a, xi_a, o->geq, o->st));
#undef OCT_Q
  return hipSuccess;
}

// Captures the coarse step from the current buffers; *next: the key its
// replay ends on (the buffers are left as they were).
hipGraphExec_t capture_step(cavlbm_octree* o, int* next) {
  hipGraph_t g = nullptr;
  hipGraphExec_t ex = nullptr;
  const int key = mkey(o);
  if (hipStreamBeginCapture(o->st, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
  const bool ok = enqueue_step(o, o->time_step) == hipSuccess;
  *next = mkey(o);
  set_mkey(o, key);
  if (hipStreamEndCapture(o->st, &g) != hipSuccess || !ok) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return nullptr;
  }
  if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) ex = nullptr;
  (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  return ex;
}

int run_steps(cavlbm_octree* o, uint64_t n) {
  for (uint64_t s = 0; s < n; ++s) {
    // constant per-step scalars without the viscosity ramp: replay one graph
    if (o->nu_start_mult == 1.0 && !o->graph_failed && graph_replay_enabled()) {
      const int key = mkey(o);
      if (!o->graph[key]) o->graph[key] = capture_step(o, &o->graph_next[key]);
      if (!o->graph[key]) {
        o->graph_failed = true;
      } else {
        OHIP_TRY(o, hipGraphLaunch(o->graph[key], o->st));
        set_mkey(o, o->graph_next[key]);
        o->time_step += 1;
        continue;
      }
    }
    OHIP_TRY(o, enqueue_step(o, o->time_step));
    o->time_step += 1;
  }
  return 0;
}

bool grid_ok(const cavlbm_octree* o, int32_t grid) { return grid >= 0 && grid < o->ngrids; }

int plane_of_field(int32_t field) {
  switch (field) {
    case CAVLBM_FIELD_RHO: return kOctRho;
    case CAVLBM_FIELD_UX: return kOctUx;
    case CAVLBM_FIELD_UY: return kOctUy;
    case CAVLBM_FIELD_T: return kOctT;
    case CAVLBM_FIELD_FY_NODE: return kOctFy;
  }
  return -1;
}

int copy_down(cavlbm_octree* o, const double* dev, double* host, size_t n) {
  OHIP_TRY(o, hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, o->st));
  OHIP_TRY(o, hipStreamSynchronize(o->st));
  return 0;
}

int copy_up(cavlbm_octree* o, const double* host, double* dev, size_t n) {
  OHIP_TRY(o, hipMemcpyAsync(dev, host, n * sizeof(double), hipMemcpyHostToDevice, o->st));
  OHIP_TRY(o, hipStreamSynchronize(o->st));
  return 0;
}

}  // namespace

extern "C" {

int cavlbm_octree_params_default(cavlbm_octree_params* p) {
  if (!p) return CAVLBM_E_PARAM;
  std::memset(p, 0, sizeof(*p));
  p->nx0 = 64;
  p->ny0 = 256;
  p->bl_width = 8;
  p->ra = 1.0e6;
  p->pr = 0.71;
  p->utarget = 0.02;
  p->seed = 42;
  p->jobs = 1;
  p->odd_exact = 0;
  p->levels = 2;
  p->nu_start_mult = 1.0;
  p->geq_order = 2;
  return 0;
}

int cavlbm_octree_create(const cavlbm_octree_params* pin, cavlbm_octree** out) {
  if (!pin || !out) return ofail(nullptr, CAVLBM_E_PARAM, "null argument");
  *out = nullptr;
  const cavlbm_octree_params p = *pin;
  // solver.rs:356-361
  if (p.bl_width < 2)
    return ofail(nullptr, CAVLBM_E_PARAM, "largeur de BL fine >= 2 cellules grossieres requise");
  if (!(2 * p.bl_width < p.nx0))
    return ofail(nullptr, CAVLBM_E_PARAM, "les 2 blocs fins doivent laisser un noyau grossier");
  if (p.levels != 2 && p.levels != 3)
    return ofail(nullptr, CAVLBM_E_PARAM,
                 "levels doit valoir 2 (config V2) ou 3 (escalier tanh V6)");
  if (p.ny0 < 3 || p.nx0 > (1u << 16) || p.ny0 > (1u << 16))
    return ofail(nullptr, CAVLBM_E_PARAM, "ny0 must be >= 3 and nx0, ny0 <= 65536");
  if (p.geq_order != 1 && p.geq_order != 2)
    return ofail(nullptr, CAVLBM_E_PARAM, "geq_order must be 1 or 2");
  // calibration (solver.rs:363-392)
  const double lz = (double)(p.ny0 - 1);
  const double nu_cal =
      std::pow(p.utarget * std::sqrt(p.pr) * lz / (kCScale * std::sqrt(p.ra)), 2.0 / 3.0);
  double nu;
  if (p.has_nu_override) {
    if (!(p.nu_override > 0.0 && std::isfinite(p.nu_override)))
      return ofail(nullptr, CAVLBM_E_PARAM, "--nu doit être > 0 et fini");
    nu = p.nu_override;
  } else {
    nu = std::fmax(nu_cal, 0.002);
  }
  const double alpha = nu / p.pr;

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return ofail(nullptr, CAVLBM_E_NODEVICE, "no HIP device visible");
  if (p.device < 0 || p.device >= ndev) return ofail(nullptr, CAVLBM_E_PARAM, "bad device ordinal");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, p.device) != hipSuccess)
    return ofail(nullptr, CAVLBM_E_NODEVICE, "hipGetDeviceProperties failed");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return ofail(nullptr, CAVLBM_E_NODEVICE,
                 std::string("libcavlbm is built for gfx950, device is ") + prop.gcnArchName);
  if (hipSetDevice(p.device) != hipSuccess) return ofail(nullptr, CAVLBM_E_HIP, "hipSetDevice failed");

  auto* o = new cavlbm_octree();
  o->p = p;
  o->levels = p.levels;
  o->ngrids = p.levels == 2 ? 3 : 5;
  o->geq = p.geq_order;
  o->nx0 = p.nx0;
  o->ny0 = p.ny0;
  o->nu_lu = nu;
  o->alpha_lu = alpha;
  o->beta_eff_coarse = p.ra * nu * alpha / (lz * lz * lz);
  o->beta_eff_fine = 0.5 * o->beta_eff_coarse;
  o->tau_coarse = relax_from(nu, alpha);
  o->tau_fine = relax_from(2.0 * nu, 2.0 * alpha);
  o->tau_sub = relax_from(4.0 * nu, 4.0 * alpha);
  o->xi_a = rescaling(o->tau_coarse, o->tau_fine, 2, p.odd_exact != 0);
  o->xi_b = rescaling(o->tau_fine, o->tau_sub, 2, p.odd_exact != 0);
  o->nu_start_mult = std::fmax(p.nu_start_mult, 1.0);
  o->spin_up_steps = std::max<uint64_t>(p.ny0 / 10, 2000);
  o->time_step = 0;

  auto fail_free = [&](const std::string& m) {
    free_octree(o);
    delete o;
    return ofail(nullptr, CAVLBM_E_HIP, m);
  };
  if (hipStreamCreateWithFlags(&o->st, hipStreamNonBlocking) != hipSuccess)
    return fail_free("hipStreamCreate failed");
  if (hipEventCreate(&o->ev0) != hipSuccess || hipEventCreate(&o->ev1) != hipSuccess)
    return fail_free("hipEventCreate failed");

  // grids (solver.rs:411-462): blocks [0, bl) and [nx0-1-bl, nx0-1) over the
  // full height [0, ny0-1); strips [1, 3) / [nx1-4, nx1-2) x [1, ny1-2) of a
  // block, with the wall column and the bottom row kept out of restriction.
  const size_t bl = p.bl_width;
  o->hg[1] = build_grid(p.nx0, p.ny0, 0, 0, bl, p.ny0 - 1, false, false, false, false);
  o->hg[2] = build_grid(p.nx0, p.ny0, p.nx0 - 1 - bl, 0, p.nx0 - 1, p.ny0 - 1, false, false,
                        false, false);
  if (p.levels == 3) {
    for (int b = 1; b <= 2; ++b) {
      const HostGrid& pg = o->hg[b];
      const bool hot = b == 1;
      const size_t sbx0 = hot ? 1 : pg.nx1 - 4, sbx1 = hot ? 3 : pg.nx1 - 2;
      o->hg[b + 2] = build_grid(pg.nx1, pg.ny1, sbx0, 1, sbx1, pg.ny1 - 2, true, false, true, false);
    }
  }
  size_t dims[5][2] = {{p.nx0, p.ny0}};
  for (int gi = 1; gi < o->ngrids; ++gi) {
    dims[gi][0] = o->hg[gi].nx1;
    dims[gi][1] = o->hg[gi].ny1;
  }
  for (int gi = 0; gi < o->ngrids; ++gi) {
    OctLevel& L = o->lv[gi];
    L.nx = (int)dims[gi][0];
    L.ny = (int)dims[gi][1];
    L.plane = (long long)L.nx * L.ny;
    L.m = dalloc<double>(o, (size_t)kOctM * L.plane);
    L.m2 = dalloc<double>(o, (size_t)kOctM * L.plane);
    L.c = dalloc<double>(o, (size_t)18 * L.plane);
    L.c2 = dalloc<double>(o, (size_t)18 * L.plane);
    if (!L.m || !L.m2 || !L.c || !L.c2) return fail_free("device allocation failed");
    if (hipMemsetAsync(L.m, 0, (size_t)kOctM * L.plane * sizeof(double), o->st) != hipSuccess ||
        hipMemsetAsync(L.m2, 0, (size_t)kOctM * L.plane * sizeof(double), o->st) != hipSuccess)
      return fail_free("device memset failed");
    o->mbuf[gi][0] = L.m;
    o->mbuf[gi][1] = L.m2;
    // tiled sub-steps: the tallest tile of at most 16 rows that fits (more
    // workgroups for these latency-bound levels, bounded recomputation)
    for (int ns = 1; ns <= 2; ++ns)
      for (int ty = 16; ty >= 3; --ty)
        if (oct_tile_fits(L.nx, L.ny, ty, ns)) {
          o->tile_ty[gi][ns] = ty;
          break;
        }
  }
  for (int gi = 1; gi < o->ngrids; ++gi) {
    const HostGrid& h = o->hg[gi];
    int* gk = dalloc<int>(o, h.gk.size());
    int* gkc = dalloc<int>(o, h.gkc.size());
    double* gw = dalloc<double>(o, h.gw.size());
    int* ck = dalloc<int>(o, h.ck.size());
    int* ckf = dalloc<int>(o, h.ckf.size());
    double* cw = dalloc<double>(o, h.cw.size());
    if (!gk || !gkc || !gw || !ck || !ckf || !cw) return fail_free("device allocation failed");
    bool ok = true;
    ok &= hipMemcpyAsync(gk, h.gk.data(), h.gk.size() * 4, hipMemcpyHostToDevice, o->st) == hipSuccess;
    ok &= hipMemcpyAsync(gkc, h.gkc.data(), h.gkc.size() * 4, hipMemcpyHostToDevice, o->st) == hipSuccess;
    ok &= hipMemcpyAsync(gw, h.gw.data(), h.gw.size() * 8, hipMemcpyHostToDevice, o->st) == hipSuccess;
    if (!h.ck.empty()) {
      ok &= hipMemcpyAsync(ck, h.ck.data(), h.ck.size() * 4, hipMemcpyHostToDevice, o->st) == hipSuccess;
      ok &= hipMemcpyAsync(ckf, h.ckf.data(), h.ckf.size() * 4, hipMemcpyHostToDevice, o->st) == hipSuccess;
      ok &= hipMemcpyAsync(cw, h.cw.data(), h.cw.size() * 8, hipMemcpyHostToDevice, o->st) == hipSuccess;
    }
    if (!ok) return fail_free("table upload failed");
    o->gt[gi] = OctGhostTab{(int)h.gk.size(), gk, gkc, gw};
    o->ct[gi] = OctCoincTab{(int)h.ck.size(), ck, ckf, cw};
  }

  // init_cavity (solver.rs:500-575): coarse T linear in x plus the xorshift64*
  // noise in (iy outer, ix inner) order; fine levels take its bilinear
  // interpolation at their node positions; rho = 1, u = 0, equilibria.
  std::vector<std::vector<double>> temp(o->ngrids);
  temp[0].resize(p.nx0 * p.ny0);
  uint64_t rng = p.seed | 1ULL;
  for (size_t iy = 0; iy < p.ny0; ++iy)
    for (size_t ix = 0; ix < p.nx0; ++ix) {
      uint64_t x = rng;
      x ^= x >> 12;
      x ^= x << 25;
      x ^= x >> 27;
      rng = x;
      const double u01 = (double)((x * 0x2545F4914F6CDD1DULL) >> 11) / (double)(1ULL << 53);
      const double xf = (double)ix / (double)(p.nx0 - 1);
      const double t_lin = kTHot + (kTCold - kTHot) * xf;
      const double pert = 1.0e-3 * (2.0 * u01 - 1.0);
      temp[0][iy * p.nx0 + ix] = t_lin + pert;
    }
  auto interp = [&](double x_c, double y_c) {
    int k[4];
    double w[4];
    bilinear(x_c, y_c, p.nx0, p.ny0, k, w);
    double t = 0.0;
    for (int q = 0; q < 4; ++q) t += w[q] * temp[0][k[q]];
    return t;
  };
  for (int b = 1; b <= 2; ++b) {
    const HostGrid& h = o->hg[b];
    temp[b].resize(h.nx1 * h.ny1);
    for (size_t bb = 0; bb < h.ny1; ++bb)
      for (size_t a = 0; a < h.nx1; ++a)
        temp[b][bb * h.nx1 + a] =
            interp((double)h.bx0 - 0.5 + 0.5 * (double)a, (double)h.by0 - 0.5 + 0.5 * (double)bb);
    if (p.levels == 3) {
      const HostGrid& s = o->hg[b + 2];
      temp[b + 2].resize(s.nx1 * s.ny1);
      for (size_t b2 = 0; b2 < s.ny1; ++b2)
        for (size_t a2 = 0; a2 < s.nx1; ++a2)
          temp[b + 2][b2 * s.nx1 + a2] =
              interp((double)h.bx0 - 0.5 + 0.5 * ((double)s.bx0 - 0.5 + 0.5 * (double)a2),
                     (double)h.by0 - 0.5 + 0.5 * ((double)s.by0 - 0.5 + 0.5 * (double)b2));
    }
  }
  for (int gi = 0; gi < o->ngrids; ++gi) {
    const OctLevel& L = o->lv[gi];
    std::vector<double> one((size_t)L.plane, 1.0);
    if (copy_up(o, one.data(), L.m + kOctRho * L.plane, one.size()) ||
        copy_up(o, temp[gi].data(), L.m + kOctT * L.plane, temp[gi].size()))
      return fail_free(o->err);
    if (launch_oct_init_eq(L, o->geq, o->st) != hipSuccess) return fail_free("init kernel failed");
  }
  if (hipStreamSynchronize(o->st) != hipSuccess) return fail_free("init failed");
  *out = o;
  return 0;
}

int cavlbm_octree_destroy(cavlbm_octree* o) {
  if (!o) return 0;
  (void)hipSetDevice(o->p.device);
  if (o->st) (void)hipStreamSynchronize(o->st);
  free_octree(o);
  delete o;
  return 0;
}

int cavlbm_octree_step(cavlbm_octree* o) { return cavlbm_octree_run(o, 1); }

int cavlbm_octree_run(cavlbm_octree* o, uint64_t nsteps) {
  if (!o) return CAVLBM_E_PARAM;
  OHIP_TRY(o, hipSetDevice(o->p.device));
  return run_steps(o, nsteps);
}

int cavlbm_octree_synchronize(cavlbm_octree* o) {
  if (!o) return CAVLBM_E_PARAM;
  OHIP_TRY(o, hipStreamSynchronize(o->st));
  return 0;
}

int cavlbm_octree_time_run(cavlbm_octree* o, uint64_t nsteps, double* ms) {
  if (!o || !ms) return CAVLBM_E_PARAM;
  OHIP_TRY(o, hipSetDevice(o->p.device));
  OHIP_TRY(o, hipEventRecord(o->ev0, o->st));
  int rc = run_steps(o, nsteps);
  if (rc) return rc;
  OHIP_TRY(o, hipEventRecord(o->ev1, o->st));
  OHIP_TRY(o, hipEventSynchronize(o->ev1));
  float t = 0.f;
  OHIP_TRY(o, hipEventElapsedTime(&t, o->ev0, o->ev1));
  *ms = t;
  return 0;
}

int cavlbm_octree_time_step(const cavlbm_octree* o, uint64_t* out) {
  if (!o || !out) return CAVLBM_E_PARAM;
  *out = o->time_step;
  return 0;
}

int cavlbm_octree_derived_params(const cavlbm_octree* o, cavlbm_octree_derived* d) {
  if (!o || !d) return CAVLBM_E_PARAM;
  std::memset(d, 0, sizeof(*d));
  d->nu_lu = o->nu_lu;
  d->alpha_lu = o->alpha_lu;
  d->beta_eff_coarse = o->beta_eff_coarse;
  d->beta_eff_fine = o->beta_eff_fine;
  const Relax r[3] = {o->tau_coarse, o->tau_fine, o->tau_sub};
  for (int l = 0; l < 3; ++l) {
    d->tau[l][0] = r[l].tp;
    d->tau[l][1] = r[l].tm;
    d->tau[l][2] = r[l].tt;
  }
  d->xi_a[0] = o->xi_a.even;
  d->xi_a[1] = o->xi_a.odd;
  d->xi_a[2] = o->xi_a.thermal;
  d->xi_b[0] = o->xi_b.even;
  d->xi_b[1] = o->xi_b.odd;
  d->xi_b[2] = o->xi_b.thermal;
  d->spin_up_steps = o->spin_up_steps;
  uint64_t total = 0;
  for (int gi = 0; gi < o->ngrids; ++gi) total += (uint64_t)o->lv[gi].plane;
  d->n_nodes_total = total;
  d->levels = o->levels;
  d->n_grids = o->ngrids;
  return 0;
}

int cavlbm_octree_grid_dims(const cavlbm_octree* o, int32_t grid, uint64_t* nx, uint64_t* ny) {
  if (!o || !nx || !ny || !grid_ok(o, grid)) return CAVLBM_E_PARAM;
  *nx = (uint64_t)o->lv[grid].nx;
  *ny = (uint64_t)o->lv[grid].ny;
  return 0;
}

int cavlbm_octree_get_field(cavlbm_octree* o, int32_t grid, int32_t field, double* out,
                            uint64_t n) {
  if (!o || !out || !grid_ok(o, grid)) return CAVLBM_E_PARAM;
  const int pl = plane_of_field(field);
  if (pl < 0) return ofail(o, CAVLBM_E_PARAM, "field not held by the octree levels");
  const OctLevel& L = o->lv[grid];
  if (n != (uint64_t)L.plane) return ofail(o, CAVLBM_E_SIZE, "nx*ny values expected");
  OHIP_TRY(o, hipSetDevice(o->p.device));
  return copy_down(o, L.m + (size_t)pl * L.plane, out, n);
}

int cavlbm_octree_set_field(cavlbm_octree* o, int32_t grid, int32_t field, const double* in,
                            uint64_t n) {
  if (!o || !in || !grid_ok(o, grid)) return CAVLBM_E_PARAM;
  const int pl = plane_of_field(field);
  if (pl < 0) return ofail(o, CAVLBM_E_PARAM, "field not held by the octree levels");
  const OctLevel& L = o->lv[grid];
  if (n != (uint64_t)L.plane) return ofail(o, CAVLBM_E_SIZE, "nx*ny values expected");
  OHIP_TRY(o, hipSetDevice(o->p.device));
  return copy_up(o, in, L.m + (size_t)pl * L.plane, n);
}

int cavlbm_octree_get_pops(cavlbm_octree* o, int32_t grid, int32_t which, double* out,
                           uint64_t n9) {
  if (!o || !out || !grid_ok(o, grid) || (which != 0 && which != 1)) return CAVLBM_E_PARAM;
  const OctLevel& L = o->lv[grid];
  const size_t n = (size_t)L.plane;
  if (n9 != 9 * n) return ofail(o, CAVLBM_E_SIZE, "9*nx*ny values expected");
  OHIP_TRY(o, hipSetDevice(o->p.device));
  std::vector<double> soa(9 * n);
  int rc = copy_down(o, L.m + (size_t)which * 9 * n, soa.data(), soa.size());
  if (rc) return rc;
  for (size_t k = 0; k < n; ++k)
    for (int i = 0; i < 9; ++i) out[k * 9 + i] = soa[(size_t)i * n + k];
  return 0;
}

int cavlbm_octree_set_pops(cavlbm_octree* o, int32_t grid, int32_t which, const double* in,
                           uint64_t n9) {
  if (!o || !in || !grid_ok(o, grid) || (which != 0 && which != 1)) return CAVLBM_E_PARAM;
  const OctLevel& L = o->lv[grid];
  const size_t n = (size_t)L.plane;
  if (n9 != 9 * n) return ofail(o, CAVLBM_E_SIZE, "9*nx*ny values expected");
  OHIP_TRY(o, hipSetDevice(o->p.device));
  std::vector<double> soa(9 * n);
  for (size_t k = 0; k < n; ++k)
    for (int i = 0; i < 9; ++i) soa[(size_t)i * n + k] = in[k * 9 + i];
  return copy_up(o, soa.data(), L.m + (size_t)which * 9 * n, soa.size());
}

int cavlbm_octree_set_from_snapshot(cavlbm_octree* o, const double* rho, const double* ux,
                                    const double* uy, const double* t, uint64_t total) {
  if (!o || !rho || !ux || !uy || !t) return CAVLBM_E_PARAM;
  uint64_t need = 0;
  for (int gi = 0; gi < o->ngrids; ++gi) need += (uint64_t)o->lv[gi].plane;
  if (total != need) return ofail(o, CAVLBM_E_SIZE, "fields of every grid expected");
  OHIP_TRY(o, hipSetDevice(o->p.device));
  size_t off = 0;
  for (int gi = 0; gi < o->ngrids; ++gi) {
    const OctLevel& L = o->lv[gi];
    const size_t n = (size_t)L.plane;
    int rc = copy_up(o, rho + off, L.m + kOctRho * n, n);
    if (!rc) rc = copy_up(o, ux + off, L.m + kOctUx * n, n);
    if (!rc) rc = copy_up(o, uy + off, L.m + kOctUy * n, n);
    if (!rc) rc = copy_up(o, t + off, L.m + kOctT * n, n);
    if (rc) return rc;
    OHIP_TRY(o, launch_oct_init_eq(L, o->geq, o->st));
    OHIP_TRY(o, launch_oct_walls(L, stream_rules(o, gi), o->st));
    off += n;
  }
  OHIP_TRY(o, hipStreamSynchronize(o->st));
  o->time_step = 0;
  return 0;
}

int cavlbm_octree_diagnostics(cavlbm_octree* o, cavlbm_octree_diag* d) {
  if (!o || !d) return CAVLBM_E_PARAM;
  OHIP_TRY(o, hipSetDevice(o->p.device));
  std::memset(d, 0, sizeof(*d));
  std::vector<std::vector<double>> m(o->ngrids);
  for (int gi = 0; gi < o->ngrids; ++gi) {
    const OctLevel& L = o->lv[gi];
    m[gi].resize((size_t)kOctM * L.plane);
    int rc = copy_down(o, L.m, m[gi].data(), m[gi].size());
    if (rc) return rc;
  }
  auto fld = [&](int gi, int pl) { return m[gi].data() + (size_t)pl * o->lv[gi].plane; };
  const size_t nx = o->nx0, ny = o->ny0, n0 = nx * ny;
  const double* tc = fld(0, kOctT);
  // cavity.rs:672-715 (Nusselt on the coarse level, scale Lx = nx0 - 1)
  double sh = 0.0, sc = 0.0, sl = 0.0;
  for (size_t iy = 0; iy < ny; ++iy) {
    sh += std::fabs(tc[iy * nx + 1] - tc[iy * nx + 2]);
    sl += std::fabs(tc[iy * nx + 1] - tc[iy * nx]);
    sc += std::fabs(tc[iy * nx + nx - 2] - tc[iy * nx + nx - 3]);
  }
  const double dT = kTHot - kTCold;
  d->nusselt_hot = sh / (double)ny * (double)(nx - 1) / dT;
  d->nusselt_cold = sc / (double)ny * (double)(nx - 1) / dT;
  d->nusselt_hot_link0 = sl / (double)ny * (double)(nx - 1) / dT;
  {  // cavity.rs:728-740 on the hot block
    const OctLevel& F = o->lv[1];
    const double* tf = fld(1, kOctT);
    double s = 0.0;
    for (int b = 1; b + 1 < F.ny; ++b)
      s += std::fabs(tf[b * F.nx + 1] - tf[b * F.nx + 2]) * 2.0;
    d->nusselt_hot_fine = s / (double)(F.ny - 2) * (double)(nx - 1) / dT;
  }
  // noise_stats (solver.rs:949-1000)
  const double* rc = fld(0, kOctRho);
  double rs = 0.0;
  for (size_t k = 0; k < n0; ++k) rs += rc[k];
  const double rho_bar = rs / (double)n0;
  double dev = 0.0;
  for (int gi = 0; gi < o->ngrids; ++gi) {
    const double* r = fld(gi, kOctRho);
    for (long long k = 0; k < o->lv[gi].plane; ++k) dev = std::fmax(dev, std::fabs(r[k] - rho_bar));
  }
  double devi = 0.0;
  for (int b = 1; b <= 2; ++b) {
    const double* rf = fld(b, kOctRho);
    for (int k : o->hg[b].gk) devi = std::fmax(devi, std::fabs(rf[k] - rho_bar));
    for (int k : o->hg[b].ck) devi = std::fmax(devi, std::fabs(rc[k] - rho_bar));
    if (o->levels == 3) {
      const double* rs2 = fld(b + 2, kOctRho);
      for (int k : o->hg[b + 2].gk) devi = std::fmax(devi, std::fabs(rs2[k] - rho_bar));
      for (int k : o->hg[b + 2].ck) devi = std::fmax(devi, std::fabs(rf[k] - rho_bar));
    }
  }
  d->max_rho_dev = dev;
  d->max_rho_dev_ifaces = devi;
  d->masse = rho_bar;
  // max_u_all and check_guards (solver.rs:1003-1041)
  double mu = 0.0;
  bool finite = true;
  // levels in the reference's order: coarse, then each block followed by its strip
  int order[5] = {0, 1, 3, 2, 4};
  for (int j = 0; j < o->ngrids; ++j) {
    const int gi = o->levels == 3 ? order[j] : j;
    const double* u = fld(gi, kOctUx);
    const double* v = fld(gi, kOctUy);
    for (long long k = 0; k < o->lv[gi].plane; ++k)
      mu = std::fmax(mu, std::fmax(std::fabs(u[k]), std::fabs(v[k])));
    for (int pl = 0; pl < kOctFy && finite; ++pl) {
      const double* a = fld(gi, pl);
      for (long long k = 0; k < o->lv[gi].plane; ++k)
        if (!std::isfinite(a[k])) {
          finite = false;
          break;
        }
    }
  }
  d->max_u_all = mu;
  d->guard = mu > kMaxUAbort ? 1 : (finite ? 0 : 2);
  d->time_step = o->time_step;
  return 0;
}

int cavlbm_octree_set_dynamic_sgs(cavlbm_octree* o, int32_t enable, double smag_alpha,
                                  double nu_sm_min, double eta_test) {
  if (!o) return CAVLBM_E_PARAM;
  OHIP_TRY(o, hipSetDevice(o->p.device));
  if (enable && !o->xi_dev) {
    // the buffers of every level start at zero (solver.rs:414-462), the
    // device factors at the molecular ones
    for (int gi = 0; gi < o->ngrids; ++gi) {
      const size_t n = (size_t)kDynPlanes * (size_t)o->lv[gi].plane;
      o->dyn[gi] = dalloc<double>(o, n);
      // entries of either form of the model (tiled, or the three passes)
      const Geom mg = model_geom(o, gi);
      const size_t np = 2 * (size_t)std::max(nblocks_of(mg), dyn_model_blocks(mg, 1));
      o->dyn_part[gi] = dalloc<double>(o, np);
      if (!o->dyn[gi] || !o->dyn_part[gi])
        return ofail(o, CAVLBM_E_HIP, "hipMalloc of the dynamic SGS planes");
      OHIP_TRY(o, hipMemsetAsync(o->dyn[gi], 0, n * sizeof(double), o->st));
      OHIP_TRY(o, hipMemsetAsync(o->dyn_part[gi], 0, np * sizeof(double), o->st));
    }
    OctXi* x = dalloc<OctXi>(o, 3);
    if (!x) return ofail(o, CAVLBM_E_HIP, "hipMalloc of the interface factors");
    const OctXi init[3] = {o->xi_a, o->xi_b, o->xi_b};
    OHIP_TRY(o, hipMemcpyAsync(x, init, sizeof(init), hipMemcpyHostToDevice, o->st));
    OHIP_TRY(o, hipStreamSynchronize(o->st));
    o->xi_dev = x;
  }
  // graphs captured with the previous settings
  for (auto& g : o->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  o->graph_failed = false;
  o->dyn_on = enable != 0;
  o->dyn_alpha = smag_alpha;
  o->dyn_min = nu_sm_min;
  o->dyn_eta = eta_test;
  return 0;
}

int cavlbm_octree_dynamic_stats(cavlbm_octree* o, double out[5]) {
  if (!o || !out) return CAVLBM_E_PARAM;
  for (int i = 0; i < 5; ++i) out[i] = 0.0;
  if (!o->xi_dev) return 0;
  OHIP_TRY(o, hipSetDevice(o->p.device));
  // C_nu / C_kappa of L0's last model pass: sequential mean and max from +0
  // (smag.rs:397-415)
  const size_t n0 = (size_t)o->lv[0].plane;
  std::vector<double> c(2 * n0);
  int rc = copy_down(o, o->dyn[0] + 16 * n0, c.data(), 2 * n0);
  if (rc) return rc;
  double sn = 0.0, mn = 0.0, sk = 0.0, mk = 0.0;
  for (size_t k = 0; k < n0; ++k) {
    sn += c[k];
    if (c[k] > mn) mn = c[k];
    sk += c[n0 + k];
    if (c[n0 + k] > mk) mk = c[n0 + k];
  }
  out[0] = sn / (double)n0;
  out[1] = mn;
  out[2] = sk / (double)n0;
  out[3] = mk;
  // nu_t max over every level (solver.rs:918-931)
  double m = 0.0;
  for (int gi = 0; gi < o->ngrids; ++gi) {
    const size_t n = (size_t)o->lv[gi].plane;
    std::vector<double> v(n);
    rc = copy_down(o, o->dyn[gi] + 18 * n, v.data(), n);
    if (rc) return rc;
    for (double x : v) m = std::fmax(m, std::fabs(x));
  }
  out[4] = m;
  return 0;
}

int cavlbm_octree_get_sgs_pre(cavlbm_octree* o, int32_t grid, int32_t which, double* out,
                              uint64_t n) {
  if (!o || !out) return CAVLBM_E_PARAM;
  if (!grid_ok(o, grid) || (which != 0 && which != 1))
    return ofail(o, CAVLBM_E_PARAM, "grid or which out of range");
  const size_t P = (size_t)o->lv[grid].plane;
  if (n != P) return ofail(o, CAVLBM_E_SIZE, "buffer size differs from the grid's node count");
  if (!o->xi_dev) {
    std::fill(out, out + P, 0.0);
    return 0;
  }
  OHIP_TRY(o, hipSetDevice(o->p.device));
  return copy_down(o, o->dyn[grid] + (18 + which) * P, out, P);
}

int cavlbm_octree_interface_factors(cavlbm_octree* o, double out[9]) {
  if (!o || !out) return CAVLBM_E_PARAM;
  OctXi x[3] = {o->xi_a, o->xi_b, o->xi_b};
  if (o->xi_dev) {
    OHIP_TRY(o, hipSetDevice(o->p.device));
    OHIP_TRY(o, hipMemcpyAsync(x, o->xi_dev, sizeof(x), hipMemcpyDeviceToHost, o->st));
    OHIP_TRY(o, hipStreamSynchronize(o->st));
  }
  for (int i = 0; i < 3; ++i) {
    const bool none = i > 0 && o->levels == 2;
    out[3 * i] = none ? 0.0 : x[i].even;
    out[3 * i + 1] = none ? 0.0 : x[i].odd;
    out[3 * i + 2] = none ? 0.0 : x[i].thermal;
  }
  return 0;
}

const char* cavlbm_octree_last_error(const cavlbm_octree* o) {
  return o ? o->err.c_str() : g_oct_create_error.c_str();
}

}  // extern "C"

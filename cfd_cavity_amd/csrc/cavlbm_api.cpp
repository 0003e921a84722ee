// cavlbm_api.cpp — the C-ABI (include/cavlbm.h) over the HIP kernels.
//
// State representation.  The reference keeps the post-stream populations and
// the moments after each step (src/uniform/solver.rs:279-293).  On the device
// the solver holds that state in one or both of two forms:
//   M-form: post-stream f, g (FP64, 18 planes) and the fields rho, ux, uy, T
//           exactly as the reference stores them;
//   C-form: the post-collision populations of the last collision (storage
//           type, ping-pong buffers).  Streaming, the thermal BCs, the moments
//           with the Guo correction and the wall velocity are recomputed from
//           it by the next fused step kernel, or by a materialize pass when
//           the caller reads the state.
// fy_node and nu_t belong to both forms (the last collision's values).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cavlbm.h"
#include "lbm_launch.hpp"

using namespace cavlbm;

namespace {

thread_local std::string g_create_error;

// Steps per captured graph: a multiple of every steps-per-launch 1..4 with an
// even launch count, so a replay ends on the buffers it started from.
constexpr int kGraphSteps = 48;

struct WinSlot {
  void* p;
  size_t bytes;
};

}  // namespace

struct cavlbm_solver {
  cavlbm_params p{};
  cavlbm_derived d{};
  Geom geo{};
  int storage = 0, geq = 2;
  size_t st_bytes = 8, r_bytes = 8;
  hipStream_t st = nullptr;
  void* cpop[2] = {nullptr, nullptr};
  int cur = 0;
  double* mpops = nullptr;
  double* mtmp = nullptr;
  double* fld[4] = {nullptr, nullptr, nullptr, nullptr};  // rho, ux, uy, T
  double* scratch_field = nullptr;                        // FP64 conversion buffer
  void* fy = nullptr;
  void* fy_alt = nullptr;  // second fy plane of the N-step kernels (pointers swap)
  void* nut = nullptr;
  double* theta = nullptr;  // ny values, always present (0.5 when no profile)
  double* red_part = nullptr;
  double* red_out = nullptr;
  bool c_valid = false, m_valid = false;
  uint64_t time_step = 0;
  // by starting buffer parity cur | (fy parity << 1)
  hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
  int fy_par = 0;  // swaps of fy / fy_alt so far, mod 2
  // dynamic-SGS steps, by starting buffer parity cur | (fy parity << 1)
  hipGraphExec_t dgraph[4] = {nullptr, nullptr, nullptr, nullptr};
  // the same for steps with the dynamic model's kept values (model off)
  hipGraphExec_t pgraph[4] = {nullptr, nullptr, nullptr, nullptr};
  bool graph_failed = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double* avg = nullptr;  // TimeAverager sums (5 planes), allocated on first use
  uint64_t avg_n = 0;
  bool in_batch = false;  // owned by a cavlbm_batch (stepped by the batched kernel)
  int tblock = 1;         // most steps per launch of the fused run loop (1..4)
  Geom geon[kMaxBlockSteps + 1] = {};   // tile grid of the N-step kernel, by N
  bool fitsn[kMaxBlockSteps + 1] = {};  // the lattice and storage have an N-step form
  // Dynamic mixed-scale SGS (solver.rs:76-103).  dyn_on: compute nu_t/alpha_t
  // before every collision; dyn_pre: the per-node values exist and are used by
  // the collision (they persist, as the reference's Option does, after the
  // model is switched off).
  bool dyn_on = false, dyn_pre = false;
  double dyn_alpha = 0.5, dyn_min = -0.1, dyn_eta = 2.0;
  double* dyn = nullptr;  // kDynPlanes FP64 planes, allocated on first enable
  bool dyn_fuse = true;   // one launch per dynamic step where it applies
  // Window objective / adjoint / tangent buffers (states, checkpoints, the
  // recomputed segment, adjoint planes), kept between calls until destroy.
  std::vector<WinSlot> win;
  hipStream_t st2 = nullptr;  // adjoint: recomputation beside the backward sweep
  hipEvent_t wev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  std::string err;
};

namespace {

int fail(cavlbm_solver* s, int code, const std::string& msg) {
  if (s)
    s->err = msg;
  else
    g_create_error = msg;
  return code;
}

#define HIP_TRY(s, expr)                                                              \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail((s), CAVLBM_E_HIP,                                                  \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                 \
  } while (0)

// Per-step scalars of collide(step): beta ramp, ramp_nu_alpha, tau_mol
// (src/uniform/solver.rs:390-410, :433-442).
StepCtx make_ctx(const cavlbm_solver* s, uint64_t step) {
  StepCtx c;
  const cavlbm_derived& d = s->d;
  double beta = d.beta_eff, nu_eff = d.nu, alpha_eff = d.alpha;
  if (step < d.spin_up_steps) {
    const double x = (double)step / (double)d.spin_up_steps;
    beta = d.beta_eff * x;
    const double nu_start = d.nu * d.nu_start_mult;
    nu_eff = nu_start * std::pow(d.nu / nu_start, x);
    const double alpha_start = d.alpha * d.nu_start_mult;
    alpha_eff = alpha_start * std::pow(d.alpha / alpha_start, x);
  }
  c.beta = beta;
  c.nu = nu_eff;
  c.alpha = alpha_eff;
  c.tau = 0.5 + 3.0 * std::fmax(nu_eff, 0.002 / 3.0);
  c.t_ref = d.t_ref;
  c.cs = d.cs;
  c.pr_t = d.pr_t;
  c.inv_pr_t = 1.0 / d.pr_t;
  return c;
}

StepBufs bufs(cavlbm_solver* s) {
  StepBufs b{};
  b.src = s->cpop[s->cur];
  b.dst = s->cpop[s->cur ^ 1];
  b.fy = s->fy;
  b.nut = s->nut;
  b.mpops = s->mpops;
  b.rho = s->fld[0];
  b.ux = s->fld[1];
  b.uy = s->fld[2];
  b.temp = s->fld[3];
  b.nut_pre = nullptr;
  b.at_pre = nullptr;
  b.dc = DynCoef{};
  if (s->dyn) {
    const size_t P = (size_t)s->geo.plane;
    b.nut_pre = s->dyn + 18 * P;
    b.at_pre = s->dyn + 19 * P;
  }
  return b;
}

hipError_t launch_b(cavlbm_solver* s, int mode, const StepCtx& ctx, const StepBufs& b,
                    int write_aux) {
  const Walls w{s->theta};
  if (s->storage == ST_FP64)
    return launch_step_fp64(s->geq, mode, s->geo, ctx, w, b, write_aux, s->st);
  return launch_step_lowp(s->storage, s->geq, mode, s->geo, ctx, w, b, write_aux, s->st);
}

hipError_t launch(cavlbm_solver* s, int mode, const StepCtx& ctx, int write_aux) {
  const Walls w{s->theta};
  const StepBufs b = bufs(s);
  if (s->storage == ST_FP64)
    return launch_step_fp64(s->geq, mode, s->geo, ctx, w, b, write_aux, s->st);
  return launch_step_lowp(s->storage, s->geq, mode, s->geo, ctx, w, b, write_aux, s->st);
}

int materialize(cavlbm_solver* s) {
  if (s->m_valid) return 0;
  if (!s->c_valid) return fail(s, CAVLBM_E_PARAM, "internal: no valid state");
  HIP_TRY(s, launch(s, MODE_MATERIALIZE, make_ctx(s, s->time_step), 0));
  s->m_valid = true;
  return 0;
}

// Steps per launch for a block of at most `left` steps (lbm_stepn.hpp): the
// largest n <= min(tblock, left) whose N-step tiling fits the lattice, for
// storage equal to the arithmetic type, outside a batch and without the
// dynamic model; 1 (the single-step kernel) otherwise.  `steady` ignores
// whether the C-form is current (what the run loop settles on).
int block_steps(const cavlbm_solver* s, uint64_t left, bool steady = false) {
  if ((!s->c_valid && !steady) || s->in_batch || s->dyn_on || s->dyn_pre) return 1;
  const int top = (int)std::min<uint64_t>((uint64_t)s->tblock, left);
  for (int n = top; n >= 2; --n)
    if (s->fitsn[n]) return n;
  return 1;
}

// collide(time_step + j) with stream in between, j = 0..n-1, from cpop[cur]
// to cpop[cur ^ 1] in one launch; `steady` uses the constant post-spin-up
// scalars (graph capture).  The caller advances cur once and swaps the fy
// planes.
hipError_t launch_block(cavlbm_solver* s, int n, bool steady, int write_aux) {
  StepCtxN cn{};
  for (int j = 0; j < n; ++j)
    cn.c[j] = make_ctx(s, steady ? s->d.spin_up_steps : s->time_step + (uint64_t)j);
  const Walls w{s->theta};
  const void* src = s->cpop[s->cur];
  void* dst = s->cpop[s->cur ^ 1];
  if (s->storage == ST_FP64)
    return launch_stepn_fp64(s->geq, n, s->geon[n], cn, w, src, dst, s->fy, s->fy_alt, s->nut,
                             write_aux, s->st);
  return launch_stepn_lowp(s->storage, s->geq, n, s->geon[n], cn, w, src, dst, s->fy, s->fy_alt,
                           s->nut, write_aux, s->st);
}

void drop_graphs(cavlbm_solver* s) {
  for (auto& g : s->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  for (auto& g : s->dgraph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  for (auto& g : s->pgraph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

void swap_fy(cavlbm_solver* s) {
  std::swap(s->fy, s->fy_alt);
  s->fy_par ^= 1;
}

// Scratch layout of the dynamic model, planes of s->dyn: ubx, uby, tb, 5 box
// products, qsm, phi, 4 Lilly terms, nu_mix, kappa_mix, C_nu, C_kappa,
// nu_t_pre, alpha_t_pre (indices 0..19).
DynSgsArgs dyn_args(const cavlbm_solver* s, const StepCtx& ctx) {
  const size_t P = (size_t)s->geo.plane;
  DynSgsArgs a{};
  a.mpops = s->mpops;
  a.rho = s->fld[0];
  a.ux = s->fld[1];
  a.uy = s->fld[2];
  a.temp = s->fld[3];
  a.tau = ctx.tau;
  // compute_dynamic_sgs (solver.rs:573-582): eta = max(eta_test, 1), width =
  // round(eta) nodes.  Every width from 2*max(nx, ny) + 1 on averages the whole
  // domain, so larger ones are capped there.
  const double eta = std::fmax(s->dyn_eta, 1.0);
  const double cap = 2.0 * (double)std::max(s->p.nx, s->p.ny) + 1.0;
  const double wr = std::round(eta);
  a.width = (int)(wr < cap ? wr : cap);
  if (a.width < 1) a.width = 1;
  a.half_left = (a.width - 1) / 2;
  a.dtest = eta * 1.0;
  a.alpha_exp = s->dyn_alpha;
  a.beta_exp = 1.0 - s->dyn_alpha;
  const double margin = (kTauTrtFloor - 0.5) / 3.0;
  a.nu_t_min = std::fmax(s->dyn_min * s->d.nu, margin - s->d.nu);
  a.alpha_t_min = std::fmax(s->dyn_min * s->d.alpha, margin - s->d.alpha);
  double* q = s->dyn;
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
  return a;
}

// The launches of one dynamic step from the C-form.  A test filter of 1..3
// nodes (dyn_launches 1): one launch (k_dyn_step for the storage), which
// reads fy and writes fy_alt; the caller swaps them (fy_swap).  Otherwise
// (4): the post-stream f planes and fields the model reads
// (MODE_MATERIALIZE_F, skipped when the M-form is current), the model's
// filter and terms, and the collision, which pulls the C-form again, runs the
// coefficient stage per node and stores its planes with the auxiliary
// outputs (MODE_FUSED_DYN).
int dyn_launches(const cavlbm_solver* s, int width) {
  return s->dyn_fuse && dyn_step_fits(width) ? 1 : 4;
}

hipError_t dyn_fused_launches(cavlbm_solver* s, const StepCtx& ctx, const DynSgsArgs& args,
                              int write_aux, bool& fy_swap) {
  fy_swap = false;
  if (dyn_launches(s, args.width) == 1) {
    fy_swap = true;
    return launch_dyn_step(s->storage, s->geq, s->geo, ctx, Walls{s->theta}, s->cpop[s->cur],
                           s->cpop[s->cur ^ 1], s->fy, s->fy_alt, s->nut, args, write_aux,
                           s->st);
  }
  hipError_t e = hipSuccess;
  if (!s->m_valid) e = launch(s, MODE_MATERIALIZE_F, ctx, 0);
  if (e == hipSuccess) e = launch_dynamic_sgs(&s->geo, &args, 1, false, s->st);
  if (e != hipSuccess) return e;
  StepBufs b = bufs(s);
  b.dc = dyn_coef_of(args);
  return launch_b(s, MODE_FUSED_DYN, ctx, b, write_aux);
}

// A step with the per-node SGS values (solver.rs:279-293 with
// compute_dynamic_sgs first): the model reads the post-stream state and the
// collision takes nu_t / alpha_t from it.  From the C-form the fused launches
// above, or, with the model off and its last values kept, one fused step
// whose collision reads them (MODE_FUSED_PRE); from the M-form (after init or
// setters) the model's three passes and the collision from the M-form
// (MODE_FIRST_PRE).
int dyn_step(cavlbm_solver* s, int write_aux) {
  const StepCtx ctx = make_ctx(s, s->time_step);
  if (!s->dyn_on && s->c_valid) {
    HIP_TRY(s, launch(s, MODE_FUSED_PRE, ctx, write_aux));
    s->cur ^= 1;
    s->m_valid = false;
    s->time_step += 1;
    return 0;
  }
  if (s->dyn_on && s->c_valid) {
    bool fy_swap;
    HIP_TRY(s, dyn_fused_launches(s, ctx, dyn_args(s, ctx), write_aux, fy_swap));
    if (fy_swap) swap_fy(s);
    s->dyn_pre = true;
    s->cur ^= 1;
    s->m_valid = false;
    s->time_step += 1;
    return 0;
  }
  int rc = materialize(s);
  if (rc) return rc;
  if (s->dyn_on) {
    const DynSgsArgs da = dyn_args(s, ctx);
    HIP_TRY(s, launch_dynamic_sgs(&s->geo, &da, 1, true, s->st));
    s->dyn_pre = true;
  }
  HIP_TRY(s, launch(s, MODE_FIRST_PRE, ctx, write_aux));
  s->cur ^= 1;
  s->c_valid = true;
  s->m_valid = false;
  s->time_step += 1;
  return 0;
}

// One step: collide(time_step) on the current state, stream/BC/moments/walls
// deferred to the next kernel.  Advances the C-form buffers.
int one_step(cavlbm_solver* s, int write_aux) {
  if (s->dyn_on || s->dyn_pre) return dyn_step(s, write_aux);
  const StepCtx ctx = make_ctx(s, s->time_step);
  HIP_TRY(s, launch(s, s->c_valid ? MODE_FUSED : MODE_FIRST, ctx, write_aux));
  s->cur ^= 1;
  s->c_valid = true;
  s->m_valid = false;
  s->time_step += 1;
  return 0;
}

// n steps in one launch: collide(time_step + j), stream, ... (n >= 2).
int block_step(cavlbm_solver* s, int n, int write_aux) {
  HIP_TRY(s, launch_block(s, n, false, write_aux));
  s->cur ^= 1;
  swap_fy(s);
  s->m_valid = false;
  s->time_step += (uint64_t)n;
  return 0;
}

// Captures kGraphSteps fused steps from the current buffers (cur, fy): n
// steps per launch (the run loop's steady choice), kGraphSteps / n launches,
// an even count, so the replay ends on the buffers it started from.
// With the dynamic model off and its values kept, the steps are single
// MODE_FUSED_PRE launches.
hipGraphExec_t capture(cavlbm_solver* s) {
  const StepCtx ctx = make_ctx(s, s->d.spin_up_steps);  // constant after spin-up
  const int mode = s->dyn_pre ? MODE_FUSED_PRE : MODE_FUSED;
  const int saved = s->cur;
  void* const fy_saved = s->fy;
  void* const fy_alt_saved = s->fy_alt;
  const int fy_par_saved = s->fy_par;
  const int n = block_steps(s, kGraphSteps);
  if (kGraphSteps % n != 0 || (kGraphSteps / n) % 2 != 0) return nullptr;
  hipGraph_t g = nullptr;
  hipGraphExec_t ex = nullptr;
  if (hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
  bool ok = true;
  for (int i = 0; i < kGraphSteps && ok; i += n) {
    if (n > 1) {
      ok = launch_block(s, n, true, 0) == hipSuccess;
      swap_fy(s);
    } else {
      ok = launch(s, mode, ctx, 0) == hipSuccess;
    }
    s->cur ^= 1;
  }
  s->cur = saved;
  s->fy = fy_saved;
  s->fy_alt = fy_alt_saved;
  s->fy_par = fy_par_saved;
  if (hipStreamEndCapture(s->st, &g) != hipSuccess || !ok) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return nullptr;
  }
  if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) ex = nullptr;
  (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  return ex;
}

// Captures kGraphSteps dynamic-SGS steps (dyn_fused_launches each) from the
// current buffers (cur, fy); kGraphSteps is even, so the replay ends on them.
hipGraphExec_t capture_dyn(cavlbm_solver* s) {
  const StepCtx ctx = make_ctx(s, s->d.spin_up_steps);
  const DynSgsArgs args = dyn_args(s, ctx);
  const int saved = s->cur;
  void* const fy_saved = s->fy;
  void* const fy_alt_saved = s->fy_alt;
  const int fy_par_saved = s->fy_par;
  hipGraph_t g = nullptr;
  hipGraphExec_t ex = nullptr;
  if (hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
  bool ok = true;
  const bool m_saved = s->m_valid;
  for (int i = 0; i < kGraphSteps && ok; ++i) {
    s->m_valid = false;  // every captured step materializes its f planes
    bool fy_swap;
    ok = dyn_fused_launches(s, ctx, args, 0, fy_swap) == hipSuccess;
    if (fy_swap) swap_fy(s);
    s->cur ^= 1;
  }
  s->cur = saved;
  s->fy = fy_saved;
  s->fy_alt = fy_alt_saved;
  s->fy_par = fy_par_saved;
  s->m_valid = m_saved;
  if (hipStreamEndCapture(s->st, &g) != hipSuccess || !ok) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return nullptr;
  }
  if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) ex = nullptr;
  (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  return ex;
}

int run_steps(cavlbm_solver* s, uint64_t n) {
  uint64_t left = n;
  while (left > 0) {
    // dynamic-SGS steps after the spin-up: graphs of kGraphSteps steps
    if (s->dyn_on && s->c_valid && !s->graph_failed && graph_replay_enabled() &&
        s->time_step >= s->d.spin_up_steps &&
        left > (uint64_t)kGraphSteps) {
      hipGraphExec_t& g = s->dgraph[s->cur | (s->fy_par << 1)];
      if (!g) g = capture_dyn(s);
      if (!g) {
        s->graph_failed = true;
        continue;
      }
      HIP_TRY(s, hipGraphLaunch(g, s->st));
      s->time_step += kGraphSteps;
      s->m_valid = false;
      s->dyn_pre = true;
      left -= kGraphSteps;
      continue;
    }
    // Steady part: whole graphs of kGraphSteps fused steps (constant per-step
    // scalars after spin-up), keeping the final step direct so that it writes
    // nu_t; with the dynamic model's kept values, graphs of their own.
    if (s->c_valid && !s->graph_failed && graph_replay_enabled() && !s->dyn_on &&
        s->time_step >= s->d.spin_up_steps && left > (uint64_t)kGraphSteps) {
      hipGraphExec_t& g = (s->dyn_pre ? s->pgraph : s->graph)[s->cur | (s->fy_par << 1)];
      if (!g) g = capture(s);
      if (!g) {
        s->graph_failed = true;
        continue;
      }
      HIP_TRY(s, hipGraphLaunch(g, s->st));
      s->time_step += kGraphSteps;
      s->m_valid = false;
      left -= kGraphSteps;
      continue;
    }
    const int n = block_steps(s, left);
    if (n > 1) {
      int rc = block_step(s, n, left == (uint64_t)n ? 1 : 0);
      if (rc) return rc;
      left -= (uint64_t)n;
      continue;
    }
    int rc = one_step(s, left == 1 ? 1 : 0);
    if (rc) return rc;
    --left;
  }
  return 0;
}

void free_all(cavlbm_solver* s) {
  drop_graphs(s);
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(s->cpop[0]);
  f(s->cpop[1]);
  f(s->mpops);
  f(s->mtmp);
  for (auto* p : s->fld) f(p);
  f(s->scratch_field);
  f(s->fy);
  f(s->fy_alt);
  f(s->nut);
  f(s->theta);
  f(s->red_part);
  f(s->red_out);
  f(s->avg);
  f(s->dyn);
  for (auto& w : s->win) f(w.p);
  s->win.clear();
  for (hipEvent_t e : s->wev)
    if (e) (void)hipEventDestroy(e);
  if (s->st2) (void)hipStreamDestroy(s->st2);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->st) (void)hipStreamDestroy(s->st);
}

// Host <-> device copies of one pitched per-node FP64 field.
int field_down(cavlbm_solver* s, const double* dev, double* host) {
  HIP_TRY(s, hipMemcpy2DAsync(host, s->p.nx * sizeof(double), dev,
                              (size_t)s->geo.pitch * sizeof(double), s->p.nx * sizeof(double),
                              s->p.ny, hipMemcpyDeviceToHost, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  return 0;
}

int field_up(cavlbm_solver* s, const double* host, double* dev) {
  HIP_TRY(s, hipMemcpy2DAsync(dev, (size_t)s->geo.pitch * sizeof(double), host,
                              s->p.nx * sizeof(double), s->p.nx * sizeof(double), s->p.ny,
                              hipMemcpyHostToDevice, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  return 0;
}

int fill_theta(cavlbm_solver* s, const double* theta) {
  std::vector<double> h(s->p.ny, 0.5);
  if (theta) std::copy(theta, theta + s->p.ny, h.begin());
  HIP_TRY(s, hipMemcpyAsync(s->theta, h.data(), h.size() * sizeof(double),
                            hipMemcpyHostToDevice, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  return 0;
}

// Before any host-side modification: make the M-form current, and mark the
// C-form stale afterwards.
int begin_modify(cavlbm_solver* s) {
  int rc = materialize(s);
  if (rc) return rc;
  s->c_valid = false;
  return 0;
}

}  // namespace

extern "C" {

int cavlbm_abi_version(void) { return CAVLBM_ABI_VERSION; }

int cavlbm_params_default(cavlbm_params* p) {
  if (!p) return CAVLBM_E_PARAM;
  std::memset(p, 0, sizeof(*p));
  p->nx = 64;
  p->ny = 256;
  p->ra = 1.0e6;
  p->pr = 0.71;
  p->cs = 0.10;
  p->pr_t = 0.9;
  p->utarget = 0.02;
  p->seed = 42;
  p->jobs = 1;
  p->alpha_mult = 1.0;
  p->nu_start_mult = 1.0;
  p->has_nu_override = 0;
  p->nu_override = 0.0;
  p->storage = CAVLBM_STORAGE_FP64;
  p->geq_order = 2;
  p->device = 0;
  return 0;
}

int cavlbm_create(const cavlbm_params* pin, cavlbm_solver** out) {
  if (!pin || !out) return fail(nullptr, CAVLBM_E_PARAM, "null argument");
  *out = nullptr;
  const cavlbm_params p = *pin;
  if (p.nx < 3 || p.ny < 3 || p.nx > (1u << 20) || p.ny > (1u << 20))
    return fail(nullptr, CAVLBM_E_PARAM, "nx and ny must be in [3, 2^20]");
  if (p.storage < 0 || p.storage > 2) return fail(nullptr, CAVLBM_E_PARAM, "bad storage");
  if (p.geq_order != 1 && p.geq_order != 2)
    return fail(nullptr, CAVLBM_E_PARAM, "geq_order must be 1 or 2");

  // Lattice-unit calibration (src/uniform/solver.rs:123-154).
  cavlbm_derived d{};
  d.lz = (double)(p.ny - 1);
  d.lx = (double)(p.nx - 1);
  const double c_scale = 1.7;
  const double nu_cal =
      std::pow(p.utarget * std::sqrt(p.pr) * d.lz / (c_scale * std::sqrt(p.ra)), 2.0 / 3.0);
  if (p.has_nu_override) {
    if (!(p.nu_override > 0.0 && std::isfinite(p.nu_override)))
      return fail(nullptr, CAVLBM_E_PARAM, "--nu must be > 0 and finite");
    if (!(0.5 + 3.0 * p.nu_override > 0.5005))
      return fail(nullptr, CAVLBM_E_PARAM, "--nu too low: tau+ close to 0.5");
    d.nu = p.nu_override;
  } else {
    d.nu = std::fmax(nu_cal, 0.002);
  }
  d.alpha_phys = d.nu / p.pr;
  d.alpha = d.alpha_phys * p.alpha_mult;
  d.beta_eff = p.ra * d.nu * d.alpha_phys / (d.lz * d.lz * d.lz);
  d.tau = 0.5 + 3.0 * d.nu;
  d.tau_t = 0.5 + 3.0 * d.alpha;
  if (!p.has_nu_override && !(d.tau > 0.505))
    return fail(nullptr, CAVLBM_E_PARAM,
                "tau=" + std::to_string(d.tau) + " too close to 0.5 (raise utarget)");
  if (!(d.tau_t > 0.505))
    return fail(nullptr, CAVLBM_E_PARAM,
                "tau_t=" + std::to_string(d.tau_t) + " too close to 0.5 (adjust alpha_mult)");
  d.cs = p.cs;
  d.pr_t = p.pr_t;
  d.nu_start_mult = std::fmax(p.nu_start_mult, 1.0);
  d.t_ref = 0.0;
  d.spin_up_steps = std::min<uint64_t>(std::max<uint64_t>(p.ny / 10, 200), 2000);

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(nullptr, CAVLBM_E_NODEVICE, "no HIP device visible");
  if (p.device < 0 || p.device >= ndev) return fail(nullptr, CAVLBM_E_PARAM, "bad device ordinal");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, p.device) != hipSuccess)
    return fail(nullptr, CAVLBM_E_NODEVICE, "hipGetDeviceProperties failed");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(nullptr, CAVLBM_E_NODEVICE,
                std::string("libcavlbm is built for gfx950, device is ") + prop.gcnArchName);
  if (hipSetDevice(p.device) != hipSuccess)
    return fail(nullptr, CAVLBM_E_HIP, "hipSetDevice failed");

  auto* s = new cavlbm_solver();
  s->p = p;
  s->d = d;
  s->storage = p.storage;
  s->geq = p.geq_order;
  s->st_bytes = p.storage == ST_FP64 ? 8 : (p.storage == ST_FP32 ? 4 : 2);
  s->r_bytes = p.storage == ST_FP64 ? 8 : 4;
  Geom& g = s->geo;
  g.nx = (int)p.nx;
  g.ny = (int)p.ny;
  g.pitch = (int)((p.nx + 63) / 64 * 64);
  g.nbx = (g.nx + kTileX - 1) / kTileX;
  g.nby = (g.ny + kTileY - 1) / kTileY;
  g.plane = (long long)g.pitch * g.ny;
  for (int n = 2; n <= kMaxBlockSteps; ++n) {
    s->geon[n] = g;
    s->fitsn[n] = stepn_grid((int)p.storage, n, s->geon[n]);
  }
  s->tblock = p.storage == ST_FP64
                  ? kDefaultBlockFp64
                  : (p.storage == ST_FP32 ? kDefaultBlockFp32 : kDefaultBlockFp16);
  if (18 * g.plane >= (1LL << 31)) {
    delete s;
    return fail(nullptr, CAVLBM_E_PARAM, "lattice too large: 18*pitch*ny must be < 2^31");
  }

  auto fail_free = [&](const std::string& m) {
    free_all(s);
    delete s;
    return fail(nullptr, CAVLBM_E_HIP, m);
  };
  const size_t P = (size_t)g.plane;
  if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess)
    return fail_free("hipStreamCreate failed");
  bool ok = true;
  // FP16 sets carry one FP32 temperature plane after the 18 population planes.
  const size_t set_bytes = 18 * P * s->st_bytes + (p.storage == ST_FP16_OFFSET ? P * 4 : 0);
  ok &= hipMalloc(&s->cpop[0], set_bytes) == hipSuccess;
  ok &= hipMalloc(&s->cpop[1], set_bytes) == hipSuccess;
  ok &= hipMalloc((void**)&s->mpops, 18 * P * sizeof(double)) == hipSuccess;
  for (auto& f : s->fld) ok &= hipMalloc((void**)&f, P * sizeof(double)) == hipSuccess;
  ok &= hipMalloc((void**)&s->scratch_field, P * sizeof(double)) == hipSuccess;
  ok &= hipMalloc(&s->fy, P * s->r_bytes) == hipSuccess;
  ok &= hipMalloc(&s->fy_alt, P * s->r_bytes) == hipSuccess;
  ok &= hipMalloc(&s->nut, P * s->r_bytes) == hipSuccess;
  ok &= hipMalloc((void**)&s->theta, p.ny * sizeof(double)) == hipSuccess;
  ok &= hipMalloc((void**)&s->red_part, (size_t)kDiagSlots * diag_blocks(g) * sizeof(double)) ==
        hipSuccess;
  ok &= hipMalloc((void**)&s->red_out, kDiagSlots * sizeof(double)) == hipSuccess;
  ok &= hipEventCreate(&s->ev0) == hipSuccess;
  ok &= hipEventCreate(&s->ev1) == hipSuccess;
  if (!ok) return fail_free("device allocation failed");
  ok &= hipMemsetAsync(s->fy, 0, P * s->r_bytes, s->st) == hipSuccess;
  ok &= hipMemsetAsync(s->fy_alt, 0, P * s->r_bytes, s->st) == hipSuccess;
  ok &= hipMemsetAsync(s->nut, 0, P * s->r_bytes, s->st) == hipSuccess;
  ok &= hipMemsetAsync(s->mpops, 0, 18 * P * sizeof(double), s->st) == hipSuccess;
  for (auto& f : s->fld) ok &= hipMemsetAsync(f, 0, P * sizeof(double), s->st) == hipSuccess;
  if (!ok) return fail_free("device memset failed");
  if (fill_theta(s, nullptr)) return fail_free(s->err);

  // Initial fields (src/uniform/solver.rs:160-176): T linear in x plus the
  // xorshift64* noise drawn in (iy outer, ix inner) order, rho = 1, u = 0;
  // the device then sets f = feq(1, 0, 0) and g = geq(T, 0, 0).
  std::vector<double> t(p.nx * p.ny), one(p.nx * p.ny, 1.0), zero(p.nx * p.ny, 0.0);
  uint64_t rng = p.seed | 1ULL;
  for (uint64_t iy = 0; iy < p.ny; ++iy)
    for (uint64_t ix = 0; ix < p.nx; ++ix) {
      uint64_t x = rng;
      x ^= x >> 12;
      x ^= x << 25;
      x ^= x >> 27;
      rng = x;
      const uint64_t r = x * 0x2545F4914F6CDD1DULL;
      const double u01 = (double)(r >> 11) / (double)(1ULL << 53);
      const double xf = (double)ix / (double)(p.nx - 1);
      const double t_lin = 0.5 + (-0.5 - 0.5) * xf;
      const double pert = 1.0e-3 * (2.0 * u01 - 1.0);
      t[iy * p.nx + ix] = t_lin + pert;
    }
  if (field_up(s, one.data(), s->fld[0]) || field_up(s, zero.data(), s->fld[1]) ||
      field_up(s, zero.data(), s->fld[2]) || field_up(s, t.data(), s->fld[3]))
    return fail_free(s->err);
  if (launch_init_eq(g, s->geq, s->mpops, s->fld[0], s->fld[1], s->fld[2], s->fld[3], s->st) !=
          hipSuccess ||
      hipStreamSynchronize(s->st) != hipSuccess)
    return fail_free("init kernel failed");
  s->m_valid = true;
  s->c_valid = false;
  s->time_step = 0;
  *out = s;
  return 0;
}

int cavlbm_destroy(cavlbm_solver* s) {
  if (!s) return 0;
  (void)hipSetDevice(s->p.device);
  if (s->st) (void)hipStreamSynchronize(s->st);
  if (s->st2) (void)hipStreamSynchronize(s->st2);
  free_all(s);
  delete s;
  return 0;
}

int cavlbm_step(cavlbm_solver* s) { return cavlbm_run(s, 1); }

int cavlbm_run(cavlbm_solver* s, uint64_t nsteps) {
  if (!s) return CAVLBM_E_PARAM;
  HIP_TRY(s, hipSetDevice(s->p.device));
  return run_steps(s, nsteps);
}

int cavlbm_synchronize(cavlbm_solver* s) {
  if (!s) return CAVLBM_E_PARAM;
  HIP_TRY(s, hipStreamSynchronize(s->st));
  return 0;
}

int cavlbm_init_from_fields(cavlbm_solver* s, const double* rho, const double* ux,
                            const double* uy, const double* t, uint64_t n) {
  if (!s) return CAVLBM_E_PARAM;
  if (!rho || !ux || !uy || !t) return fail(s, CAVLBM_E_PARAM, "null field");
  if (n != s->p.nx * s->p.ny)
    return fail(s, CAVLBM_E_SIZE, "init_from_fields: expected " +
                                      std::to_string(s->p.nx * s->p.ny) + " values, got " +
                                      std::to_string(n));
  HIP_TRY(s, hipSetDevice(s->p.device));
  if (field_up(s, rho, s->fld[0]) || field_up(s, ux, s->fld[1]) ||
      field_up(s, uy, s->fld[2]) || field_up(s, t, s->fld[3]))
    return CAVLBM_E_HIP;
  HIP_TRY(s, launch_init_eq(s->geo, s->geq, s->mpops, s->fld[0], s->fld[1], s->fld[2],
                            s->fld[3], s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  s->m_valid = true;
  s->c_valid = false;
  s->time_step = 0;
  return 0;
}

int cavlbm_get_field(cavlbm_solver* s, int32_t field, double* out, uint64_t n) {
  if (!s) return CAVLBM_E_PARAM;
  if (!out) return fail(s, CAVLBM_E_PARAM, "null output");
  if (n != s->p.nx * s->p.ny) return fail(s, CAVLBM_E_SIZE, "get_field: size mismatch");
  HIP_TRY(s, hipSetDevice(s->p.device));
  if (field >= CAVLBM_FIELD_RHO && field <= CAVLBM_FIELD_T) {
    int rc = materialize(s);
    if (rc) return rc;
    return field_down(s, s->fld[field], out);
  }
  if (field == CAVLBM_FIELD_NU_T || field == CAVLBM_FIELD_FY_NODE) {
    const void* src = field == CAVLBM_FIELD_NU_T ? s->nut : s->fy;
    HIP_TRY(s, launch_field_convert(s->geo, s->storage, src, s->scratch_field, 1, nullptr,
                                    nullptr, s->st));
    return field_down(s, s->scratch_field, out);
  }
  return fail(s, CAVLBM_E_PARAM, "unknown field");
}

int cavlbm_set_field(cavlbm_solver* s, int32_t field, const double* in, uint64_t n) {
  if (!s) return CAVLBM_E_PARAM;
  if (!in) return fail(s, CAVLBM_E_PARAM, "null input");
  if (n != s->p.nx * s->p.ny) return fail(s, CAVLBM_E_SIZE, "set_field: size mismatch");
  if (field < CAVLBM_FIELD_RHO || field > CAVLBM_FIELD_FY_NODE)
    return fail(s, CAVLBM_E_PARAM, "unknown field");
  HIP_TRY(s, hipSetDevice(s->p.device));
  int rc = begin_modify(s);
  if (rc) return rc;
  if (field <= CAVLBM_FIELD_T) return field_up(s, in, s->fld[field]);
  if (field_up(s, in, s->scratch_field)) return CAVLBM_E_HIP;
  void* dst = field == CAVLBM_FIELD_NU_T ? s->nut : s->fy;
  HIP_TRY(s, launch_field_convert(s->geo, s->storage, nullptr, nullptr, 0, dst,
                                  s->scratch_field, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  return 0;
}

int cavlbm_get_pops(cavlbm_solver* s, int32_t which, double* out, uint64_t n9) {
  if (!s) return CAVLBM_E_PARAM;
  if (!out) return fail(s, CAVLBM_E_PARAM, "null output");
  if (which != CAVLBM_POPS_F && which != CAVLBM_POPS_G)
    return fail(s, CAVLBM_E_PARAM, "which must be F or G");
  const uint64_t n = s->p.nx * s->p.ny;
  if (n9 != 9 * n) return fail(s, CAVLBM_E_SIZE, "get_pops: expected 9*nx*ny values");
  HIP_TRY(s, hipSetDevice(s->p.device));
  int rc = materialize(s);
  if (rc) return rc;
  std::vector<double> plane(n);
  for (int i = 0; i < 9; ++i) {
    rc = field_down(s, s->mpops + (size_t)(9 * which + i) * s->geo.plane, plane.data());
    if (rc) return rc;
    for (uint64_t k = 0; k < n; ++k) out[k * 9 + i] = plane[k];
  }
  return 0;
}

int cavlbm_set_pops(cavlbm_solver* s, int32_t which, const double* in, uint64_t n9) {
  if (!s) return CAVLBM_E_PARAM;
  if (!in) return fail(s, CAVLBM_E_PARAM, "null input");
  if (which != CAVLBM_POPS_F && which != CAVLBM_POPS_G)
    return fail(s, CAVLBM_E_PARAM, "which must be F or G");
  const uint64_t n = s->p.nx * s->p.ny;
  if (n9 != 9 * n) return fail(s, CAVLBM_E_SIZE, "set_pops: expected 9*nx*ny values");
  HIP_TRY(s, hipSetDevice(s->p.device));
  int rc = begin_modify(s);
  if (rc) return rc;
  std::vector<double> plane(n);
  for (int i = 0; i < 9; ++i) {
    for (uint64_t k = 0; k < n; ++k) plane[k] = in[k * 9 + i];
    rc = field_up(s, plane.data(), s->mpops + (size_t)(9 * which + i) * s->geo.plane);
    if (rc) return rc;
  }
  return 0;
}

int cavlbm_diagnostics(cavlbm_solver* s, cavlbm_diag* out) {
  if (!s) return CAVLBM_E_PARAM;
  if (!out) return fail(s, CAVLBM_E_PARAM, "null output");
  HIP_TRY(s, hipSetDevice(s->p.device));
  int rc = materialize(s);
  if (rc) return rc;
  HIP_TRY(s, launch_diagnostics(s->geo, s->storage, s->mpops, s->fld[0], s->fld[1], s->fld[2],
                                s->fld[3], s->nut, s->d.nu, s->red_part, s->red_out, s->st));
  double r[kDiagSlots];
  HIP_TRY(s, hipMemcpyAsync(r, s->red_out, sizeof(r), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  const double ny = (double)s->p.ny, lz = s->d.lz, dT = 1.0;
  out->nusselt_hot = r[D_NU_HOT] / ny * lz / dT;
  out->nusselt_cold = r[D_NU_COLD] / ny * lz / dT;
  out->nusselt_hot_wall_link = r[D_NU_LINK] / ny * lz / dT;
  const double conv = r[D_NU_VOL] / (double)((s->p.nx - 2) * (s->p.ny - 2));
  out->nusselt_volume = 1.0 + conv * lz / dT;
  out->total_mass = r[D_MASS];
  out->total_f_mass = r[D_FMASS];
  out->max_uy = r[D_MAX_UY];
  out->max_u = r[D_MAX_U];
  out->nu_t_max = r[D_NUT_MAX];
  out->y_plus_hot = r[D_YPLUS] / (double)(s->p.ny - 2);
  out->all_finite = r[D_NONFINITE] == 0.0 ? 1 : 0;
  out->reserved = 0;
  out->time_step = s->time_step;
  return 0;
}

int cavlbm_max_uy_on_row(cavlbm_solver* s, uint64_t iy, double* w_max, double* x_norm) {
  if (!s) return CAVLBM_E_PARAM;
  if (iy >= s->p.ny) return fail(s, CAVLBM_E_PARAM, "row out of range");
  HIP_TRY(s, hipSetDevice(s->p.device));
  int rc = materialize(s);
  if (rc) return rc;
  std::vector<double> row(s->p.nx);
  HIP_TRY(s, hipMemcpyAsync(row.data(), s->fld[2] + (size_t)iy * s->geo.pitch,
                            s->p.nx * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  double best = 0.0;
  uint64_t best_ix = 1;
  for (uint64_t ix = 1; ix + 1 < s->p.nx; ++ix) {
    const double v = std::fabs(row[ix]);
    if (v > best) {
      best = v;
      best_ix = ix;
    }
  }
  if (w_max) *w_max = best;
  if (x_norm) *x_norm = (double)best_ix / (double)(s->p.nx - 1);
  return 0;
}

int cavlbm_derived_params(const cavlbm_solver* s, cavlbm_derived* out) {
  if (!s || !out) return CAVLBM_E_PARAM;
  *out = s->d;
  return 0;
}

int cavlbm_time_step(const cavlbm_solver* s, uint64_t* out) {
  if (!s || !out) return CAVLBM_E_PARAM;
  *out = s->time_step;
  return 0;
}

int cavlbm_set_time_step(cavlbm_solver* s, uint64_t t) {
  if (!s) return CAVLBM_E_PARAM;
  s->time_step = t;
  return 0;
}

int cavlbm_set_hot_wall(cavlbm_solver* s, const double* theta_hot, uint64_t ny) {
  if (!s) return CAVLBM_E_PARAM;
  if (theta_hot && ny != s->p.ny) return fail(s, CAVLBM_E_SIZE, "theta_hot must have ny values");
  HIP_TRY(s, hipSetDevice(s->p.device));
  int rc = begin_modify(s);
  if (rc) return rc;
  return fill_theta(s, theta_hot);
}

int cavlbm_set_dynamic_sgs(cavlbm_solver* s, int32_t enable, double smag_alpha,
                           double nu_sm_min, double eta_test) {
  if (!s) return CAVLBM_E_PARAM;
  if (enable && s->in_batch)
    return fail(s, CAVLBM_E_UNSUPPORTED,
                "dynamic SGS is not available for cavities of a batch");
  HIP_TRY(s, hipSetDevice(s->p.device));
  if (enable && !s->dyn) {
    const size_t bytes = (size_t)kDynPlanes * (size_t)s->geo.plane * sizeof(double);
    HIP_TRY(s, hipMalloc((void**)&s->dyn, bytes));
    HIP_TRY(s, hipMemsetAsync(s->dyn, 0, bytes, s->st));
    HIP_TRY(s, hipStreamSynchronize(s->st));
  }
  for (auto& g : s->dgraph)  // captured with the previous model parameters
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  s->dyn_on = enable != 0;
  s->dyn_alpha = smag_alpha;
  s->dyn_min = nu_sm_min;
  s->dyn_eta = eta_test;
  return 0;
}

int cavlbm_dynamic_stats(cavlbm_solver* s, double out[4]) {
  if (!s || !out) return CAVLBM_E_PARAM;
  out[0] = out[1] = out[2] = out[3] = 0.0;
  if (!s->dyn) return 0;
  HIP_TRY(s, hipSetDevice(s->p.device));
  const size_t P = (size_t)s->geo.plane, n = s->p.nx * s->p.ny;
  std::vector<double> cnu(n), ck(n);
  int rc = field_down(s, s->dyn + 16 * P, cnu.data());
  if (!rc) rc = field_down(s, s->dyn + 17 * P, ck.data());
  if (rc) return rc;
  // sequential sums and maxima in node order (solver.rs:784-799, :821-824)
  double cs = 0.0, cm = 0.0, ks = 0.0, km = 0.0;
  for (size_t k = 0; k < n; ++k) {
    cs += cnu[k];
    if (cnu[k] > cm) cm = cnu[k];
    ks += ck[k];
    if (ck[k] > km) km = ck[k];
  }
  out[0] = cs / (double)n;
  out[1] = cm;
  out[2] = ks / (double)n;
  out[3] = km;
  return 0;
}

int cavlbm_get_sgs_pre(cavlbm_solver* s, int32_t which, double* out, uint64_t n) {
  if (!s || !out || (which != 0 && which != 1)) return CAVLBM_E_PARAM;
  if (n != s->p.nx * s->p.ny) return fail(s, CAVLBM_E_SIZE, "nx*ny values expected");
  if (!s->dyn_pre) return fail(s, CAVLBM_E_PARAM, "the dynamic SGS model has not run");
  HIP_TRY(s, hipSetDevice(s->p.device));
  const size_t P = (size_t)s->geo.plane;
  return field_down(s, s->dyn + (18 + which) * P, out);
}

int cavlbm_set_jobs(uint64_t) { return 0; }

int cavlbm_phase(cavlbm_solver* s, int32_t phase, uint64_t step) {
  if (!s) return CAVLBM_E_PARAM;
  HIP_TRY(s, hipSetDevice(s->p.device));
  int rc = begin_modify(s);
  if (rc) return rc;
  const Walls w{s->theta};
  switch (phase) {
    case CAVLBM_PHASE_COLLIDE:
      if (s->dyn_pre)
        return fail(s, CAVLBM_E_UNSUPPORTED,
                    "collide phase alone with dynamic SGS values is not provided");
      HIP_TRY(s, launch(s, MODE_COLLIDE_M, make_ctx(s, step), 1));
      break;
    case CAVLBM_PHASE_STREAM:
      if (!s->mtmp) HIP_TRY(s, hipMalloc((void**)&s->mtmp, 18 * (size_t)s->geo.plane * 8));
      HIP_TRY(s, launch_phase_stream(s->geo, s->mpops, s->mtmp, s->st));
      std::swap(s->mpops, s->mtmp);
      break;
    case CAVLBM_PHASE_THERMAL_BC:
      HIP_TRY(s, launch_phase_thermal_bc(s->geo, s->geq, w, s->mpops, s->st));
      break;
    case CAVLBM_PHASE_UPDATE_MACROSCOPIC:
      HIP_TRY(s, launch_phase_macro(s->geo, s->storage, s->mpops, s->fy, s->fld[0], s->fld[1],
                                    s->fld[2], s->fld[3], s->st));
      break;
    case CAVLBM_PHASE_ENFORCE_WALL_VELOCITY:
      HIP_TRY(s, launch_phase_walls(s->geo, s->fld[1], s->fld[2], s->st));
      break;
    case CAVLBM_PHASE_BOUNCE_BACK:
      HIP_TRY(s, launch_phase_bounce_back(s->geo, s->mpops, s->st));
      break;
    default:
      return fail(s, CAVLBM_E_PARAM, "unknown phase");
  }
  HIP_TRY(s, hipStreamSynchronize(s->st));
  return 0;
}

int cavlbm_time_run(cavlbm_solver* s, uint64_t nsteps, double* ms) {
  if (!s || !ms) return CAVLBM_E_PARAM;
  HIP_TRY(s, hipSetDevice(s->p.device));
  HIP_TRY(s, hipEventRecord(s->ev0, s->st));
  int rc = run_steps(s, nsteps);
  if (rc) return rc;
  HIP_TRY(s, hipEventRecord(s->ev1, s->st));
  HIP_TRY(s, hipEventSynchronize(s->ev1));
  float t = 0.f;
  HIP_TRY(s, hipEventElapsedTime(&t, s->ev0, s->ev1));
  *ms = t;
  return 0;
}

int cavlbm_time_step_kernel(cavlbm_solver* s, uint64_t launches, double* mean_ms) {
  if (!s || !mean_ms || launches == 0) return CAVLBM_E_PARAM;
  HIP_TRY(s, hipSetDevice(s->p.device));
  if (!s->c_valid) {
    int rc = run_steps(s, 1);
    if (rc) return rc;
  }
  double total = 0.0;
  for (uint64_t i = 0; i < launches; ++i) {
    HIP_TRY(s, hipEventRecord(s->ev0, s->st));
    const int aux = i + 1 == launches ? 1 : 0;
    const int n = block_steps(s, kGraphSteps);
    int rc = n > 1 ? block_step(s, n, aux) : one_step(s, aux);
    if (rc) return rc;
    HIP_TRY(s, hipEventRecord(s->ev1, s->st));
    HIP_TRY(s, hipEventSynchronize(s->ev1));
    float t = 0.f;
    HIP_TRY(s, hipEventElapsedTime(&t, s->ev0, s->ev1));
    total += t;
  }
  *mean_ms = total / (double)launches;
  return 0;
}

int cavlbm_set_time_blocking(cavlbm_solver* s, int32_t steps) {
  if (!s) return CAVLBM_E_PARAM;
  if (steps < 1 || steps > kMaxBlockSteps)
    return fail(s, CAVLBM_E_PARAM, "time blocking must be 1 to 4 steps per launch");
  if (s->tblock != steps) {
    HIP_TRY(s, hipSetDevice(s->p.device));
    drop_graphs(s);
    s->graph_failed = false;
    s->tblock = steps;
  }
  return 0;
}

int cavlbm_steps_per_launch(const cavlbm_solver* s, int32_t* out) {
  if (!s || !out) return CAVLBM_E_PARAM;
  *out = block_steps(s, kGraphSteps, true);
  return 0;
}

int cavlbm_set_dynamic_fusion(cavlbm_solver* s, int32_t one_launch) {
  if (!s) return CAVLBM_E_PARAM;
  if (s->dyn_fuse != (one_launch != 0)) {
    HIP_TRY(s, hipSetDevice(s->p.device));
    drop_graphs(s);
    s->graph_failed = false;
    s->dyn_fuse = one_launch != 0;
  }
  return 0;
}

int cavlbm_dynamic_launches(const cavlbm_solver* s, int32_t* out) {
  if (!s || !out) return CAVLBM_E_PARAM;
  const StepCtx ctx = make_ctx(s, s->time_step);
  *out = dyn_launches(s, dyn_args(s, ctx).width);
  return 0;
}

const char* cavlbm_last_error(const cavlbm_solver* s) {
  return s ? s->err.c_str() : g_create_error.c_str();
}

int cavlbm_average_reset(cavlbm_solver* s) {
  if (!s) return CAVLBM_E_PARAM;
  HIP_TRY(s, hipSetDevice(s->p.device));
  const size_t bytes = 5 * (size_t)s->geo.plane * sizeof(double);
  if (!s->avg) HIP_TRY(s, hipMalloc((void**)&s->avg, bytes));
  HIP_TRY(s, hipMemsetAsync(s->avg, 0, bytes, s->st));
  s->avg_n = 0;
  return 0;
}

int cavlbm_average_accumulate(cavlbm_solver* s) {
  if (!s) return CAVLBM_E_PARAM;
  if (!s->avg) {
    int rc = cavlbm_average_reset(s);
    if (rc) return rc;
  }
  HIP_TRY(s, hipSetDevice(s->p.device));
  int rc = materialize(s);
  if (rc) return rc;
  HIP_TRY(s, launch_avg_accumulate(s->geo, s->storage, s->fld[1], s->fld[2], s->fld[3], s->nut,
                                   s->avg, s->st));
  s->avg_n += 1;
  return 0;
}

int cavlbm_average_count(const cavlbm_solver* s, uint64_t* count) {
  if (!s || !count) return CAVLBM_E_PARAM;
  *count = s->avg_n;
  return 0;
}

int cavlbm_average_means(cavlbm_solver* s, double* ux, double* uy, double* t, double* nu_t,
                         double* uu, uint64_t n) {
  if (!s) return CAVLBM_E_PARAM;
  if (n != s->p.nx * s->p.ny) return fail(s, CAVLBM_E_SIZE, "average_means: size mismatch");
  if (!s->avg || s->avg_n == 0) return fail(s, CAVLBM_E_PARAM, "no samples accumulated");
  HIP_TRY(s, hipSetDevice(s->p.device));
  // TimeAverager::means (post.rs:52-56): v * (1/n)
  const double inv = 1.0 / (double)s->avg_n;
  double* outs[5] = {ux, uy, t, nu_t, uu};
  for (int q = 0; q < 5; ++q) {
    if (!outs[q]) continue;
    int rc = field_down(s, s->avg + (size_t)q * s->geo.plane, outs[q]);
    if (rc) return rc;
    for (uint64_t k = 0; k < n; ++k) outs[q][k] *= inv;
  }
  return 0;
}

}  // extern "C"

// ===========================================================================
// Batched form
// ===========================================================================
struct cavlbm_batch {
  std::vector<cavlbm_solver*> cav;
  int storage = 0, geq = 2, device = 0;
  hipStream_t st = nullptr;
  CavDesc* d_desc = nullptr;
  int* d_tile_cav = nullptr;   // single-step tile -> cavity table
  int total_tiles = 0;
  int* d_tile_cavn = nullptr;  // N-step tile -> cavity table (nblk >= 2)
  int total_tiles_n = 0;
  int shape = 0, xcd = 1;
  int nblk = 1;                // steps per launch of the batched run loop
  uint64_t max_spin = 0;
  uint64_t nodes = 0;
  // by starting buffer parity cur | (fy parity << 1)
  hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
  bool graph_failed = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
};

namespace {

int bfail(cavlbm_batch* b, int code, const std::string& m) {
  if (b)
    b->err = m;
  else
    g_create_error = m;
  return code;
}

#define BHIP_TRY(b, expr)                                                             \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return bfail((b), CAVLBM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

void batch_drop_graphs(cavlbm_batch* b) {
  for (auto& g : b->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// Whether every cavity of the batch fits the N-step tiling.
bool batch_fits(const cavlbm_batch* b, int n) {
  if (n < 2 || n > kMaxBlockSteps) return false;
  for (const cavlbm_solver* s : b->cav)
    if (!s->fitsn[n]) return false;
  return true;
}

// (Re)builds the cavity descriptors and the tile tables for the current
// single-step tile shape and steps per launch; buffer pointers come from each
// solver's two C-form sets and two fy planes.
int batch_layout(cavlbm_batch* b) {
  const TileShape sh = kBatchShapes[b->shape];
  std::vector<CavDesc> desc(b->cav.size());
  std::vector<int> tile_cav, tile_cavn;
  int begin = 0, begin_n = 0;
  for (size_t c = 0; c < b->cav.size(); ++c) {
    cavlbm_solver* s = b->cav[c];
    CavDesc& d = desc[c];
    d.geo = s->geo;
    d.geo.nbx = (s->geo.nx + sh.tx - 1) / sh.tx;
    d.geo.nby = (s->geo.ny + sh.ty - 1) / sh.ty;
    d.geon = b->nblk >= 2 ? s->geon[b->nblk] : s->geo;
    d.ctx = make_ctx(s, s->d.spin_up_steps);
    d.beta_eff = s->d.beta_eff;
    d.spin_up = (int)s->d.spin_up_steps;
    d.tile_begin = begin;
    d.tile_begin_n = begin_n;
    d.walls = Walls{s->theta};
    for (int par = 0; par < 2; ++par) {
      const int saved = s->cur;
      s->cur = par;
      d.bufs[par] = bufs(s);
      s->cur = saved;
    }
    d.fyp[s->fy_par] = s->fy;
    d.fyp[s->fy_par ^ 1] = s->fy_alt;
    const int nt = d.geo.nbx * d.geo.nby;
    tile_cav.insert(tile_cav.end(), nt, (int)c);
    begin += nt;
    if (b->nblk >= 2) {
      const int ntn = d.geon.nbx * d.geon.nby;
      tile_cavn.insert(tile_cavn.end(), ntn, (int)c);
      begin_n += ntn;
    }
  }
  b->total_tiles = begin;
  b->total_tiles_n = begin_n;
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_tile_cav) (void)hipFree(b->d_tile_cav);
  if (b->d_tile_cavn) (void)hipFree(b->d_tile_cavn);
  b->d_desc = nullptr;
  b->d_tile_cav = nullptr;
  b->d_tile_cavn = nullptr;
  BHIP_TRY(b, hipMalloc((void**)&b->d_desc, desc.size() * sizeof(CavDesc)));
  BHIP_TRY(b, hipMalloc((void**)&b->d_tile_cav, tile_cav.size() * sizeof(int)));
  BHIP_TRY(b, hipMemcpyAsync(b->d_desc, desc.data(), desc.size() * sizeof(CavDesc),
                             hipMemcpyHostToDevice, b->st));
  BHIP_TRY(b, hipMemcpyAsync(b->d_tile_cav, tile_cav.data(), tile_cav.size() * sizeof(int),
                             hipMemcpyHostToDevice, b->st));
  if (!tile_cavn.empty()) {
    BHIP_TRY(b, hipMalloc((void**)&b->d_tile_cavn, tile_cavn.size() * sizeof(int)));
    BHIP_TRY(b, hipMemcpyAsync(b->d_tile_cavn, tile_cavn.data(), tile_cavn.size() * sizeof(int),
                               hipMemcpyHostToDevice, b->st));
  }
  BHIP_TRY(b, hipStreamSynchronize(b->st));
  batch_drop_graphs(b);
  b->graph_failed = false;
  return 0;
}

// One step of every cavity (single-step tiles).
hipError_t batch_launch(cavlbm_batch* b, int parity, int fpar, uint64_t step, int write_aux) {
  const int stepi = step > (uint64_t)INT32_MAX ? INT32_MAX : (int)step;
  if (b->storage == ST_FP64)
    return launch_step_batch_fp64(b->geq, b->shape, b->d_desc, b->d_tile_cav, b->total_tiles,
                                  parity, fpar, stepi, write_aux, b->xcd, b->st);
  return launch_step_batch_lowp(b->storage, b->geq, b->shape, b->d_desc, b->d_tile_cav,
                                b->total_tiles, parity, fpar, stepi, write_aux, b->xcd, b->st);
}

// nblk steps of every cavity (N-step tiles); reads fy plane fpar, writes fpar ^ 1.
hipError_t batch_launch_n(cavlbm_batch* b, int parity, int fpar, uint64_t step, int write_aux) {
  const int stepi = step > (uint64_t)INT32_MAX - kMaxBlockSteps ? INT32_MAX - kMaxBlockSteps
                                                                : (int)step;
  if (b->storage == ST_FP64)
    return launch_stepn_batch_fp64(b->geq, b->nblk, b->d_desc, b->d_tile_cavn, b->total_tiles_n,
                                   parity, fpar, stepi, write_aux, b->xcd, b->st);
  return launch_stepn_batch_lowp(b->storage, b->geq, b->nblk, b->d_desc, b->d_tile_cavn,
                                 b->total_tiles_n, parity, fpar, stepi, write_aux, b->xcd, b->st);
}

// Every cavity in C-form, with a common buffer parity and step count.  A
// cavity still in M-form (never stepped, or modified through a setter) takes
// its next step alone (MODE_FIRST); *consumed reports whether that happened.
int batch_prepare(cavlbm_batch* b, int write_aux, uint64_t* consumed) {
  *consumed = 0;
  for (cavlbm_solver* s : b->cav) BHIP_TRY(b, hipStreamSynchronize(s->st));
  bool any = false;
  for (cavlbm_solver* s : b->cav) any |= !s->c_valid;
  if (any) {
    // keep the cavities in lockstep: the ones already in C-form take the
    // same step through the fused kernel
    for (cavlbm_solver* s : b->cav)
      if (one_step(s, write_aux)) return bfail(b, CAVLBM_E_HIP, s->err);
    *consumed = 1;
  }
  for (cavlbm_solver* s : b->cav) BHIP_TRY(b, hipStreamSynchronize(s->st));
  const cavlbm_solver* s0 = b->cav[0];
  for (cavlbm_solver* s : b->cav)
    if (s->cur != s0->cur || s->fy_par != s0->fy_par || s->time_step != s0->time_step)
      return bfail(b, CAVLBM_E_PARAM, "batch cavities out of lockstep (stepped individually?)");
  return 0;
}

bool batch_ready(const cavlbm_batch* b) {
  for (const cavlbm_solver* s : b->cav)
    if (!s->c_valid) return false;
  return true;
}

// k steps taken by `launches` launches (one buffer flip each); an N-step
// launch also exchanges the fy planes.
void batch_advance(cavlbm_batch* b, uint64_t k, uint64_t launches, bool block) {
  for (cavlbm_solver* s : b->cav) {
    s->time_step += k;
    if (launches & 1) {
      s->cur ^= 1;
      if (block) swap_fy(s);
    }
    s->m_valid = false;
  }
}

// kGraphSteps steps from buffers (par, fpar): kGraphSteps / nblk N-step
// launches, or single steps; an even launch count either way.
hipGraphExec_t batch_capture(cavlbm_batch* b, int par, int fpar) {
  hipGraph_t g = nullptr;
  hipGraphExec_t ex = nullptr;
  if (hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
  bool ok = true;
  int p = par, q = fpar;
  const int n = b->nblk;
  for (int i = 0; i < kGraphSteps && ok; i += n) {
    if (n >= 2) {
      ok = batch_launch_n(b, p, q, b->max_spin, 0) == hipSuccess;
      q ^= 1;
    } else {
      ok = batch_launch(b, p, q, b->max_spin, 0) == hipSuccess;
    }
    p ^= 1;
  }
  if (hipStreamEndCapture(b->st, &g) != hipSuccess || !ok) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return nullptr;
  }
  if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) ex = nullptr;
  (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  return ex;
}

int batch_run_steps(cavlbm_batch* b, uint64_t n) {
  if (n == 0) return 0;
  uint64_t consumed = 0;
  int rc = batch_prepare(b, n == 1 ? 1 : 0, &consumed);
  if (rc) return rc;
  uint64_t left = n - consumed;
  cavlbm_solver* s0 = b->cav[0];
  while (left > 0) {
    if (!b->graph_failed && graph_replay_enabled() && s0->time_step >= b->max_spin &&
        left > (uint64_t)kGraphSteps) {
      hipGraphExec_t& g = b->graph[s0->cur | (s0->fy_par << 1)];
      if (!g) g = batch_capture(b, s0->cur, s0->fy_par);
      if (!g) {
        b->graph_failed = true;
        continue;
      }
      BHIP_TRY(b, hipGraphLaunch(g, b->st));
      batch_advance(b, kGraphSteps, 0, false);  // an even number of launches
      left -= kGraphSteps;
      continue;
    }
    const uint64_t n = (uint64_t)b->nblk;
    if (n >= 2 && left >= n) {
      BHIP_TRY(b, batch_launch_n(b, s0->cur, s0->fy_par, s0->time_step, left == n ? 1 : 0));
      batch_advance(b, n, 1, true);
      left -= n;
      continue;
    }
    BHIP_TRY(b, batch_launch(b, s0->cur, s0->fy_par, s0->time_step, left == 1 ? 1 : 0));
    batch_advance(b, 1, 1, false);
    --left;
  }
  return 0;
}

}  // namespace

extern "C" {

int cavlbm_batch_create(const cavlbm_params* cavities, uint64_t n, cavlbm_batch** out) {
  if (!cavities || !out || n == 0) return bfail(nullptr, CAVLBM_E_PARAM, "empty batch");
  *out = nullptr;
  const cavlbm_params& p0 = cavities[0];
  for (uint64_t i = 0; i < n; ++i) {
    const cavlbm_params& p = cavities[i];
    if (p.storage != p0.storage || p.geq_order != p0.geq_order || p.device != p0.device)
      return bfail(nullptr, CAVLBM_E_PARAM, "batch cavities must share storage, geq_order, device");
    if (p.nu_start_mult > 1.0)
      return bfail(nullptr, CAVLBM_E_UNSUPPORTED, "batch requires nu_start_mult = 1");
  }
  auto* b = new cavlbm_batch();
  b->storage = p0.storage;
  b->geq = p0.geq_order;
  b->device = p0.device;
  for (uint64_t i = 0; i < n; ++i) {
    cavlbm_solver* s = nullptr;
    int rc = cavlbm_create(&cavities[i], &s);
    if (rc) {
      cavlbm_batch_destroy(b);
      return rc;  // message already in cavlbm_last_error(NULL)
    }
    s->in_batch = true;
    b->cav.push_back(s);
    b->max_spin = std::max<uint64_t>(b->max_spin, s->d.spin_up_steps);
    b->nodes += s->p.nx * s->p.ny;
  }
  // steps per launch: the storage's default, or the largest that every
  // cavity's tiling fits
  const int want = b->storage == ST_FP64
                       ? kDefaultBlockFp64
                       : (b->storage == ST_FP32 ? kDefaultBlockFp32 : kDefaultBlockFp16);
  b->nblk = 1;
  for (int n = want; n >= 2; --n)
    if (batch_fits(b, n)) {
      b->nblk = n;
      break;
    }
  if (hipSetDevice(b->device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    cavlbm_batch_destroy(b);
    return bfail(nullptr, CAVLBM_E_HIP, "batch stream/event creation failed");
  }
  if (batch_layout(b)) {
    const std::string m = b->err;
    cavlbm_batch_destroy(b);
    return bfail(nullptr, CAVLBM_E_HIP, m);
  }
  *out = b;
  return 0;
}

int cavlbm_batch_destroy(cavlbm_batch* b) {
  if (!b) return 0;
  (void)hipSetDevice(b->device);
  if (b->st) (void)hipStreamSynchronize(b->st);
  batch_drop_graphs(b);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_tile_cav) (void)hipFree(b->d_tile_cav);
  if (b->d_tile_cavn) (void)hipFree(b->d_tile_cavn);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->st) (void)hipStreamDestroy(b->st);
  for (cavlbm_solver* s : b->cav) cavlbm_destroy(s);
  delete b;
  return 0;
}

int cavlbm_batch_run(cavlbm_batch* b, uint64_t nsteps) {
  if (!b) return CAVLBM_E_PARAM;
  BHIP_TRY(b, hipSetDevice(b->device));
  int rc = batch_run_steps(b, nsteps);
  if (rc) return rc;
  BHIP_TRY(b, hipStreamSynchronize(b->st));
  return 0;
}

int cavlbm_batch_synchronize(cavlbm_batch* b) {
  if (!b) return CAVLBM_E_PARAM;
  BHIP_TRY(b, hipStreamSynchronize(b->st));
  return 0;
}

int cavlbm_batch_solver(cavlbm_batch* b, uint64_t i, cavlbm_solver** out) {
  if (!b || !out) return CAVLBM_E_PARAM;
  if (i >= b->cav.size()) return bfail(b, CAVLBM_E_PARAM, "cavity index out of range");
  *out = b->cav[i];
  return 0;
}

int cavlbm_batch_nodes(const cavlbm_batch* b, uint64_t* out) {
  if (!b || !out) return CAVLBM_E_PARAM;
  *out = b->nodes;
  return 0;
}

int cavlbm_batch_time_run(cavlbm_batch* b, uint64_t nsteps, double* ms) {
  if (!b || !ms) return CAVLBM_E_PARAM;
  BHIP_TRY(b, hipSetDevice(b->device));
  BHIP_TRY(b, hipEventRecord(b->ev0, b->st));
  int rc = batch_run_steps(b, nsteps);
  if (rc) return rc;
  BHIP_TRY(b, hipEventRecord(b->ev1, b->st));
  BHIP_TRY(b, hipEventSynchronize(b->ev1));
  float t = 0.f;
  BHIP_TRY(b, hipEventElapsedTime(&t, b->ev0, b->ev1));
  *ms = t;
  return 0;
}

int cavlbm_batch_autotune(cavlbm_batch* b, uint64_t trial_steps, cavlbm_tune_result* out) {
  if (!b || !out || trial_steps == 0) return CAVLBM_E_PARAM;
  BHIP_TRY(b, hipSetDevice(b->device));
  if (!batch_ready(b))
    return bfail(b, CAVLBM_E_PARAM, "run the batch at least one step before autotuning");
  uint64_t consumed = 0;
  int rc = batch_prepare(b, 0, &consumed);
  if (rc) return rc;
  // Snapshot the state the trials will overwrite: the current C-form set,
  // fy_node and nu_t of every cavity, and the step counters.
  struct Snap {
    void* pops;
    void* fy;
    void* nut;
    size_t pops_bytes, field_bytes;
    int cur, fy_par;
    uint64_t t;
  };
  std::vector<Snap> snap;
  auto free_snap = [&]() {
    for (auto& z : snap) {
      if (z.pops) (void)hipFree(z.pops);
      if (z.fy) (void)hipFree(z.fy);
      if (z.nut) (void)hipFree(z.nut);
    }
  };
  for (cavlbm_solver* s : b->cav) {
    Snap z{};
    const size_t P = (size_t)s->geo.plane;
    z.pops_bytes = 18 * P * s->st_bytes + (s->storage == ST_FP16_OFFSET ? P * 4 : 0);
    z.field_bytes = P * s->r_bytes;
    z.cur = s->cur;
    z.fy_par = s->fy_par;
    z.t = s->time_step;
    bool ok = hipMalloc(&z.pops, z.pops_bytes) == hipSuccess &&
              hipMalloc(&z.fy, z.field_bytes) == hipSuccess &&
              hipMalloc(&z.nut, z.field_bytes) == hipSuccess;
    snap.push_back(z);
    if (!ok) {
      free_snap();
      return bfail(b, CAVLBM_E_HIP, "autotune snapshot allocation failed");
    }
    ok = hipMemcpyAsync(z.pops, s->cpop[s->cur], z.pops_bytes, hipMemcpyDeviceToDevice, b->st) ==
             hipSuccess &&
         hipMemcpyAsync(z.fy, s->fy, z.field_bytes, hipMemcpyDeviceToDevice, b->st) == hipSuccess &&
         hipMemcpyAsync(z.nut, s->nut, z.field_bytes, hipMemcpyDeviceToDevice, b->st) == hipSuccess;
    if (!ok) {
      free_snap();
      return bfail(b, CAVLBM_E_HIP, "autotune snapshot copy failed");
    }
  }
  auto restore = [&]() -> bool {
    bool ok = true;
    for (size_t c = 0; c < b->cav.size(); ++c) {
      cavlbm_solver* s = b->cav[c];
      const Snap& z = snap[c];
      s->cur = z.cur;
      if (s->fy_par != z.fy_par) swap_fy(s);
      s->time_step = z.t;
      s->c_valid = true;
      s->m_valid = false;
      ok &= hipMemcpyAsync(s->cpop[s->cur], z.pops, z.pops_bytes, hipMemcpyDeviceToDevice,
                           b->st) == hipSuccess;
      ok &= hipMemcpyAsync(s->fy, z.fy, z.field_bytes, hipMemcpyDeviceToDevice, b->st) ==
            hipSuccess;
      ok &= hipMemcpyAsync(s->nut, z.nut, z.field_bytes, hipMemcpyDeviceToDevice, b->st) ==
            hipSuccess;
    }
    ok &= hipStreamSynchronize(b->st) == hipSuccess;
    return ok;
  };
  // candidates: every single-step tile shape, then every N-step block the
  // batch fits, each with and without the XCD remap
  struct Cand {
    int shape, nblk, xcd;
  };
  std::vector<Cand> cands;
  for (int shape = 0; shape < kNumBatchShapes; ++shape)
    for (int xcd = 1; xcd >= 0; --xcd) cands.push_back({shape, 1, xcd});
  for (int n = 2; n <= kMaxBlockSteps; ++n)
    if (batch_fits(b, n))
      for (int xcd = 1; xcd >= 0; --xcd) cands.push_back({b->shape, n, xcd});
  double best = 1e300, first = 0.0;
  Cand best_c = cands[0];
  int count = 0;
  for (const Cand& cd : cands) {
    b->shape = cd.shape;
    b->nblk = cd.nblk;
    b->xcd = cd.xcd;
    if (batch_layout(b) || !restore()) {
      free_snap();
      return CAVLBM_E_HIP;
    }
    // one untimed pass (graph capture, caches), then the timed trial
    double ms = 0.0;
    if (batch_run_steps(b, trial_steps) || hipStreamSynchronize(b->st) != hipSuccess ||
        !restore() || cavlbm_batch_time_run(b, trial_steps, &ms)) {
      free_snap();
      return CAVLBM_E_HIP;
    }
    const double per = ms / (double)trial_steps;
    if (count == 0) first = per;
    if (per < best) {
      best = per;
      best_c = cd;
    }
    ++count;
  }
  b->shape = best_c.shape;
  b->nblk = best_c.nblk;
  b->xcd = best_c.xcd;
  const int lrc = batch_layout(b);
  const bool rok = restore();
  free_snap();
  if (lrc || !rok) return bfail(b, CAVLBM_E_HIP, "autotune restore failed");
  if (b->nblk >= 2) {
    stepn_tile_dims(b->storage, b->nblk, &out->tile_x, &out->tile_y);
  } else {
    out->tile_x = kBatchShapes[b->shape].tx;
    out->tile_y = kBatchShapes[b->shape].ty;
  }
  out->xcd_remap = b->xcd;
  out->candidates = count;
  out->best_ms_per_step = best;
  out->first_ms_per_step = first;
  out->steps_per_launch = b->nblk;
  return 0;
}

const char* cavlbm_batch_last_error(const cavlbm_batch* b) {
  return b ? b->err.c_str() : g_create_error.c_str();
}

}  // extern "C"

// ===========================================================================
// Window objective and its adjoint (FP64)
// ===========================================================================

// Device memory of the allocation-free window objective
// (cavlbm_nu_window_static): the window's two states (18 planes each), two
// force planes, the nu_t plane, the per-row Nusselt sums and the hot-wall
// profile, for lattices whose pitched plane holds at most kStaticPlane nodes
// and at most kStaticRows rows (20.5 MB per GPU, held until the process ends).
// Host code never uses the name `scratch` as an address: each device's copy
// is found with hipGetSymbolAddress (static_scratch below).
constexpr size_t kStaticPlane = 65536;  // e.g. 128x512, 96x384, 256x256
constexpr size_t kStaticRows = 1024;
constexpr size_t kStaticDoubles = 39 * kStaticPlane + 2 * kStaticRows;
static __device__ __align__(256) double scratch[kStaticDoubles];

namespace {

// Device work area of one window evaluation; freed on destruction.
// The device buffers of one window call, handed out in request order from
// the solver's window arena.  Every call of an entry point requests the same
// sequence, so a repeated call finds its buffers there and allocates nothing
// (a slot grows only when a call needs more than it holds).  alloc: contents
// undefined (the caller overwrites every node before reading); zeroed: an
// accumulator, cleared on the stream.
struct WindowWork {
  cavlbm_solver* s;
  size_t next = 0;
  template <class T>
  T* alloc(size_t n) {
    const size_t bytes = n * sizeof(T);
    if (next == s->win.size()) s->win.push_back({nullptr, 0});
    WinSlot& slot = s->win[next++];
    if (slot.bytes < bytes) {
      if (slot.p) (void)hipFree(slot.p);  // synchronizes: no launch still uses it
      slot = {nullptr, 0};
      if (hipMalloc(&slot.p, bytes) != hipSuccess) {
        slot.p = nullptr;
        return nullptr;
      }
      slot.bytes = bytes;
    }
    return static_cast<T*>(slot.p);
  }
  template <class T>
  T* zeroed(size_t n) {
    T* p = alloc<T>(n);
    if (p && hipMemsetAsync(p, 0, n * sizeof(T), s->st) != hipSuccess) return nullptr;
    return p;
  }
};

StepCtx window_ctx(const cavlbm_solver* s, uint64_t step, double cs) {
  StepCtx c = make_ctx(s, step);
  c.cs = cs;
  return c;
}

// One forward collision of the window into (dst, fy); src == nullptr: the
// first collision, from the solver's M-form state S_0.
// One window step: X_m from X_{m-1} (src) or, for m = 0, from the M-form.  A
// step from src also forms the Nusselt row terms of the state it pulls
// (MODE_WINDOW): accumulated into acc and / or their signs into sgn; fy_in is
// the force plane of X_{m-1} when it is not fy itself.
hipError_t window_step(cavlbm_solver* s, const double* src, double* dst, double* fy,
                       double* nut, const double* theta, const StepCtx& ctx,
                       const double* fy_in = nullptr, double* acc = nullptr,
                       double* sgn = nullptr, hipStream_t st = nullptr) {
  StepBufs b{};
  b.src = src;
  b.dst = dst;
  b.fy = fy;
  b.fy_in = fy_in ? fy_in : fy;
  b.nu_acc = acc;
  b.nu_sgn = sgn;
  b.nut = nut;
  b.mpops = s->mpops;
  b.rho = s->fld[0];
  b.ux = s->fld[1];
  b.uy = s->fld[2];
  b.temp = s->fld[3];
  return launch_step_fp64(s->geq, src ? MODE_WINDOW : MODE_FIRST, s->geo, ctx, Walls{theta}, b,
                          0, st ? st : s->st);
}

// The second stream and the events of the adjoint's overlapped recomputation,
// created on first use (default priority, like the solver stream).
int window_streams(cavlbm_solver* s) {
  if (!s->st2) HIP_TRY(s, hipStreamCreateWithFlags(&s->st2, hipStreamNonBlocking));
  for (hipEvent_t& e : s->wev)
    if (!e) HIP_TRY(s, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return 0;
}

// Waits for the second stream when leaving the adjoint, on every return path.
struct AuxStreamJoin {
  hipStream_t st;
  ~AuxStreamJoin() { (void)hipStreamSynchronize(st); }
};

// The checks every window entry point makes before touching the device.
int window_check(cavlbm_solver* s, const double* theta_hot, uint64_t ny, uint64_t K) {
  if (s->storage != ST_FP64)
    return fail(s, CAVLBM_E_UNSUPPORTED, "window objective and adjoint need FP64 storage");
  if (s->dyn_on || s->dyn_pre)
    return fail(s, CAVLBM_E_UNSUPPORTED,
                "window objective and adjoint are for the fixed-Cs Smagorinsky model");
  if (K == 0) return fail(s, CAVLBM_E_PARAM, "window_steps must be >= 1");
  if (theta_hot && ny != s->p.ny) return fail(s, CAVLBM_E_SIZE, "theta_hot must have ny values");
  return 0;
}

int window_setup(cavlbm_solver* s, const double* theta_hot, uint64_t ny, uint64_t K,
                 WindowWork& w, double*& theta) {
  int rc = window_check(s, theta_hot, ny, K);
  if (rc) return rc;
  HIP_TRY(s, hipSetDevice(s->p.device));
  rc = materialize(s);  // S_0 in M-form (the logical state is unchanged)
  if (rc) return rc;
  theta = w.alloc<double>(s->p.ny);
  if (!theta) return fail(s, CAVLBM_E_HIP, "window allocation failed");
  std::vector<double> h(s->p.ny, 0.5);
  if (theta_hot) std::copy(theta_hot, theta_hot + s->p.ny, h.begin());
  HIP_TRY(s, hipMemcpyAsync(theta, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                            s->st));
  return 0;
}

// The device buffers of a forward window: X_m in ping-pong, the force plane
// of X_m (the N-step launches read one plane and write the other), the nu_t
// plane the steps write, and the per-row Nusselt sums (zero on entry).
struct WindowBufs {
  double* X[2];
  double* fy[2];
  double* nut;
  double* acc;
};

// Forward window: J, and optionally the checkpoints (X_{qC}, fy_{qC}), whose
// slots come from w (required when ckpt is given).  hrows: ny host doubles
// receiving the row sums.
int window_forward(cavlbm_solver* s, const double* theta, double cs, uint64_t K, uint64_t C,
                   const WindowBufs& wb, WindowWork* w, std::vector<double*>* ckpt,
                   double* hrows, double* j_out) {
  const size_t P = (size_t)s->geo.plane;
  double* const* X = wb.X;
  double* const* fy = wb.fy;
  double* nut = wb.nut;
  double* acc = wb.acc;
  const uint64_t t0 = s->time_step;
  const Walls walls{theta};
  int cur = 0, fc = 0;
  for (uint64_t m = 0; m < K;) {
    // X_m = Collide(S_m): S_0 from the M-form, then steps from X_{m-1}, which
    // also accumulate Nu(S_m) pulled from X_{m-1}: n of them per launch (the
    // N-step window kernel), ending on every checkpoint step
    int n = 1;
    if (m > 0) {
      uint64_t room = K - m;
      if (ckpt) room = std::min<uint64_t>(room, (m + C - 1) / C * C - m + 1);
      for (int q = (int)std::min<uint64_t>((uint64_t)s->tblock, room); q >= 2; --q)
        if (s->fitsn[q]) {
          n = q;
          break;
        }
    }
    if (m == 0) {
      HIP_TRY(s, window_step(s, nullptr, X[cur], fy[fc], nut, theta, window_ctx(s, t0, cs)));
    } else if (n == 1) {
      HIP_TRY(s, window_step(s, X[cur], X[cur ^ 1], fy[fc], nut, theta,
                             window_ctx(s, t0 + m, cs), nullptr, acc));
      cur ^= 1;
    } else {
      StepCtxN cn{};
      for (int j = 0; j < n; ++j) cn.c[j] = window_ctx(s, t0 + m + (uint64_t)j, cs);
      HIP_TRY(s, launch_stepn_window_fp64(s->geq, n, s->geon[n], cn, walls, X[cur], X[cur ^ 1],
                                          fy[fc], fy[fc ^ 1], acc, s->st));
      cur ^= 1;
      fc ^= 1;
    }
    m += (uint64_t)n;
    if (ckpt && (m - 1) % C == 0) {  // X_{m-1}, the last state of this launch
      double* slot = w->alloc<double>(19 * P);
      if (!slot) return fail(s, CAVLBM_E_HIP, "checkpoint allocation failed");
      HIP_TRY(s, hipMemcpyAsync(slot, X[cur], 18 * P * sizeof(double), hipMemcpyDeviceToDevice,
                                s->st));
      HIP_TRY(s, hipMemcpyAsync(slot + 18 * P, fy[fc], P * sizeof(double),
                                hipMemcpyDeviceToDevice, s->st));
      ckpt->push_back(slot);
    }
  }
  // Nu(S_K) from X_{K-1}
  HIP_TRY(s, launch_nu_rows(s->geo, s->geq, X[cur], theta, acc, nullptr, s->st));
  HIP_TRY(s, hipMemcpyAsync(hrows, acc, s->p.ny * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  double sum = 0.0;
  for (uint64_t iy = 0; iy < s->p.ny; ++iy) sum += hrows[iy];
  *j_out = sum / (double)s->p.ny * s->d.lz / 1.0 / (double)K;
  return 0;
}

// The address of `scratch` on the solver's device (current), looked up once
// per device and kept.
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kScratchDevices] = {};

int static_scratch(cavlbm_solver* s, double** out) {
  const int d = s->p.device;
  if (d < 0 || d >= kScratchDevices) return fail(s, CAVLBM_E_PARAM, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[d]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(s, CAVLBM_E_HIP,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[d] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[d];
  return 0;
}

// The forward window on buffers from the solver's window arena.
int window_forward(cavlbm_solver* s, const double* theta, double cs, uint64_t K, uint64_t C,
                   WindowWork& w, std::vector<double*>* ckpt, double* j_out) {
  const size_t P = (size_t)s->geo.plane;
  WindowBufs wb{};
  wb.X[0] = w.alloc<double>(18 * P);
  wb.X[1] = w.alloc<double>(18 * P);
  wb.fy[0] = w.alloc<double>(P);
  wb.fy[1] = w.alloc<double>(P);
  wb.nut = w.alloc<double>(P);
  wb.acc = w.zeroed<double>(s->p.ny);
  if (!wb.X[0] || !wb.X[1] || !wb.fy[0] || !wb.fy[1] || !wb.nut || !wb.acc)
    return fail(s, CAVLBM_E_HIP, "window allocation failed");
  std::vector<double> h(s->p.ny);
  return window_forward(s, theta, cs, K, C, wb, &w, ckpt, h.data(), j_out);
}

}  // namespace

extern "C" {

int cavlbm_nu_window(cavlbm_solver* s, const double* theta_hot, uint64_t ny, double cs,
                     uint64_t window_steps, double* j_out) {
  if (!s || !j_out) return CAVLBM_E_PARAM;
  WindowWork w{s};
  double* theta = nullptr;
  int rc = window_setup(s, theta_hot, ny, window_steps, w, theta);
  if (rc) return rc;
  return window_forward(s, theta, cs, window_steps, 1, w, nullptr, j_out);
}

int cavlbm_nu_window_static(cavlbm_solver* s, const double* theta_hot, uint64_t ny, double cs,
                            uint64_t window_steps, double* j_out) {
  if (!s || !j_out) return CAVLBM_E_PARAM;
  int rc = window_check(s, theta_hot, ny, window_steps);
  if (rc) return rc;
  const size_t P = (size_t)s->geo.plane;
  const size_t rows = s->p.ny;
  if (P > kStaticPlane || rows > kStaticRows)
    return fail(s, CAVLBM_E_SIZE,
                "cavlbm_nu_window_static: the lattice's pitched plane (" + std::to_string(P) +
                    " nodes, " + std::to_string(rows) + " rows) exceeds the static scratch (" +
                    std::to_string(kStaticPlane) + " nodes, " + std::to_string(kStaticRows) +
                    " rows); use cavlbm_nu_window");
  HIP_TRY(s, hipSetDevice(s->p.device));
  double* base = nullptr;
  rc = static_scratch(s, &base);
  if (rc) return rc;
  rc = materialize(s);  // S_0 in M-form (the logical state is unchanged)
  if (rc) return rc;
  // P is a multiple of 64 (rows padded to 64 nodes), so every part stays
  // 256-byte aligned
  WindowBufs wb{};
  wb.X[0] = base;
  wb.X[1] = base + 18 * P;
  wb.fy[0] = base + 36 * P;
  wb.fy[1] = base + 37 * P;
  wb.nut = base + 38 * P;
  wb.acc = base + 39 * P;
  double* theta = wb.acc + (rows + 31) / 32 * 32;
  // the array keeps the previous call's contents: the row sums start at zero
  HIP_TRY(s, launch_fill_f64(wb.acc, 0.0, rows, s->st));
  if (theta_hot)
    HIP_TRY(s, hipMemcpyAsync(theta, theta_hot, rows * sizeof(double), hipMemcpyHostToDevice,
                              s->st));
  else
    HIP_TRY(s, launch_fill_f64(theta, 0.5, rows, s->st));
  double hrows[kStaticRows];
  return window_forward(s, theta, cs, window_steps, 1, wb, nullptr, nullptr, hrows, j_out);
}

int cavlbm_adjoint_nu_window(cavlbm_solver* s, const double* theta_hot, uint64_t ny, double cs,
                             uint64_t window_steps, uint64_t checkpoint_every, double* j_out,
                             double* grad_theta, double* grad_cs) {
  if (!s || !j_out || !grad_theta || !grad_cs) return CAVLBM_E_PARAM;
  const uint64_t K = window_steps;
  const uint64_t C = checkpoint_every ? checkpoint_every : 100;
  WindowWork w{s};
  double* theta = nullptr;
  int rc = window_setup(s, theta_hot, ny, K, w, theta);
  if (rc) return rc;
  std::vector<double*> ckpt;
  rc = window_forward(s, theta, cs, K, C, w, &ckpt, j_out);
  if (rc) return rc;

  const size_t P = (size_t)s->geo.plane;
  const uint64_t t0 = s->time_step;
  const uint64_t rows = s->p.ny;
  const uint64_t seg_len = std::min<uint64_t>(C, K);
  const uint64_t nseg = (K + C - 1) / C;
  // Two segment buffers (states and signs): the recomputation of segment q-1
  // runs on the second stream while the backward sweep of segment q reads the
  // other buffer on the solver stream (both are small launches that leave
  // most CUs idle on their own).
  std::vector<double*> seg[2] = {std::vector<double*>(seg_len), std::vector<double*>(seg_len)};
  double* sgn[2] = {nullptr, nullptr};
  for (int b = 0; b < (nseg > 1 ? 2 : 1); ++b) {
    for (auto& q : seg[b])
      if (!(q = w.alloc<double>(19 * P))) return fail(s, CAVLBM_E_HIP, "segment allocation failed");
    // sign of T(1,iy) - T(2,iy) of S_{m+1} for each m of a segment
    if (!(sgn[b] = w.alloc<double>(seg_len * rows)))
      return fail(s, CAVLBM_E_HIP, "segment allocation failed");
  }
  // mu_m of the backward step m, alternating (step m gathers mu_{m+1})
  double* mu[2] = {w.alloc<double>(19 * P), w.alloc<double>(19 * P)};
  double* nut = w.alloc<double>(P);
  double* th_acc = w.zeroed<double>(rows);
  double* cs_acc = w.zeroed<double>(P);
  double* cs_sum = w.alloc<double>(1);
  if (!mu[0] || !mu[1] || !nut || !th_acc || !cs_acc || !cs_sum)
    return fail(s, CAVLBM_E_HIP, "adjoint allocation failed");
  rc = window_streams(s);
  if (rc) return rc;
  hipStream_t st2 = s->st2;
  hipEvent_t* ev = s->wev;  // 0: forward done; 1 + b: segment b recomputed; 3 + b: b swept
  AuxStreamJoin join{st2};   // no work of this call outlives it, also on an error return
  const double obj_w = s->d.lz / ((double)rows * 1.0 * (double)K);

  // X_m, fy_m for m in [m_lo, m_hi) of segment q from its checkpoint into
  // buffer b, on the second stream (step m pulls X_{m-1} and stores the signs
  // of S_m, used at m - 1)
  auto recompute = [&](uint64_t q, int b) -> int {
    const uint64_t m_lo = q * C, m_hi = std::min<uint64_t>(m_lo + C, K);
    std::vector<double*>& sb = seg[b];
    HIP_TRY(s, hipMemcpyAsync(sb[0], ckpt[q], 19 * P * sizeof(double), hipMemcpyDeviceToDevice,
                              st2));
    for (uint64_t m = m_lo + 1; m < m_hi; ++m) {
      double* prev = sb[m - 1 - m_lo];
      double* next = sb[m - m_lo];
      HIP_TRY(s, window_step(s, prev, next, next + 18 * P, nut, theta, window_ctx(s, t0 + m, cs),
                             prev + 18 * P, nullptr, sgn[b] + (m - 1 - m_lo) * rows, st2));
    }
    HIP_TRY(s, launch_nu_rows(s->geo, s->geq, sb[m_hi - 1 - m_lo], theta, nullptr,
                              sgn[b] + (m_hi - 1 - m_lo) * rows, st2));
    HIP_TRY(s, hipEventRecord(ev[1 + b], st2));
    return 0;
  };

  AdjArgs a{};
  a.theta = theta;
  a.theta_acc = th_acc;
  a.cs_acc = cs_acc;
  a.obj_w = obj_w;
  int mc = 0;              // step m writes mu[mc]; mu[mc ^ 1] holds mu_{m+1}
  bool lam_zero = true;    // X_K is not used by J: its adjoint is zero
  HIP_TRY(s, hipEventRecord(ev[0], s->st));  // checkpoints and buffers ready
  HIP_TRY(s, hipStreamWaitEvent(st2, ev[0], 0));
  rc = recompute(nseg - 1, (int)((nseg - 1) & 1));
  if (rc) return rc;
  for (uint64_t q = nseg; q-- > 0;) {
    const int b = (int)(q & 1);
    HIP_TRY(s, hipStreamWaitEvent(s->st, ev[1 + b], 0));
    if (q > 0) {
      // buffer b ^ 1 was last read by the sweep of segment q + 1
      if (q + 1 < nseg) HIP_TRY(s, hipStreamWaitEvent(st2, ev[3 + (b ^ 1)], 0));
      rc = recompute(q - 1, b ^ 1);
      if (rc) return rc;
    }
    const uint64_t m_lo = q * C, m_hi = std::min<uint64_t>(m_lo + C, K);
    for (uint64_t m = m_hi; m-- > m_lo;) {
      double* Xm = seg[b][m - m_lo];
      a.sgn = sgn[b] + (m - m_lo) * rows;
      a.X = Xm;
      a.fyp = Xm + 18 * P;
      a.mu_next = lam_zero ? nullptr : mu[mc ^ 1];
      a.mu_next_fy = lam_zero ? nullptr : mu[mc ^ 1] + 18 * P;
      a.mu = mu[mc];
      a.mu_fy = mu[mc] + 18 * P;
      HIP_TRY(s, launch_adj_step(s->geo, s->geq, false, window_ctx(s, t0 + m + 1, cs), a, s->st));
      mc ^= 1;
      lam_zero = false;
    }
    HIP_TRY(s, hipEventRecord(ev[3 + b], s->st));
  }
  // the first collision X_0 = Collide(S_0): its Cs dependence
  a.X = nullptr;
  a.fyp = nullptr;
  a.mpops = s->mpops;
  a.mrho = s->fld[0];
  a.mux = s->fld[1];
  a.muy = s->fld[2];
  a.mtemp = s->fld[3];
  a.mu_next = mu[mc ^ 1];
  a.mu_next_fy = mu[mc ^ 1] + 18 * P;
  a.mu = nullptr;  // the first collision has no pulled values
  a.mu_fy = nullptr;
  a.sgn = nullptr;
  HIP_TRY(s, launch_adj_step(s->geo, s->geq, true, window_ctx(s, t0, cs), a, s->st));
  HIP_TRY(s, launch_sum_nodes(s->geo, cs_acc, cs_sum, s->st));
  std::vector<double> th(s->p.ny);
  double gcs = 0.0;
  HIP_TRY(s, hipMemcpyAsync(th.data(), th_acc, th.size() * sizeof(double), hipMemcpyDeviceToHost,
                            s->st));
  HIP_TRY(s, hipMemcpyAsync(&gcs, cs_sum, sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  // fold the per-row contributions of column 0 onto theta indices: the
  // corner rows use theta[1] and theta[ny-2]
  const uint64_t n = s->p.ny;
  for (uint64_t i = 0; i < n; ++i) grad_theta[i] = 0.0;
  for (uint64_t iy = 0; iy < n; ++iy) {
    const uint64_t iyc = iy == 0 ? 1 : (iy == n - 1 ? n - 2 : iy);
    grad_theta[iyc] += th[iy];
  }
  *grad_cs = gcs;
  return 0;
}

}  // extern "C"

extern "C" int cavlbm_tangent_nu_window(cavlbm_solver* s, const double* theta_hot, uint64_t ny,
                                        double cs, const double* dtheta_hot, double dcs,
                                        uint64_t window_steps, double* j_out, double* dj_out) {
  if (!s || !j_out || !dj_out) return CAVLBM_E_PARAM;
  const uint64_t K = window_steps;
  WindowWork w{s};
  double* theta = nullptr;
  int rc = window_setup(s, theta_hot, ny, K, w, theta);
  if (rc) return rc;
  if (dtheta_hot && ny != s->p.ny) return fail(s, CAVLBM_E_SIZE, "dtheta_hot must have ny values");
  const size_t P = (size_t)s->geo.plane;
  double* dtheta = w.zeroed<double>(s->p.ny);
  double* X[2] = {w.alloc<double>(18 * P), w.alloc<double>(18 * P)};
  double* Xd[2] = {w.alloc<double>(18 * P), w.alloc<double>(18 * P)};
  double* fy = w.alloc<double>(P);
  double* fyd = w.alloc<double>(P);
  double* acc = w.zeroed<double>(s->p.ny);
  double* dacc = w.zeroed<double>(s->p.ny);
  if (!dtheta || !X[0] || !X[1] || !Xd[0] || !Xd[1] || !fy || !fyd || !acc || !dacc)
    return fail(s, CAVLBM_E_HIP, "tangent allocation failed");
  if (dtheta_hot)
    HIP_TRY(s, hipMemcpyAsync(dtheta, dtheta_hot, s->p.ny * sizeof(double), hipMemcpyHostToDevice,
                              s->st));
  const uint64_t t0 = s->time_step;
  int cur = 0;
  for (uint64_t m = 0; m < K; ++m) {
    const bool first = m == 0;
    const int out = first ? cur : cur ^ 1;
    // (a step from X_{m-1} also accumulates Nu(S_m) and its tangent)
    HIP_TRY(s, launch_tangent_step(s->geo, s->geq, first, window_ctx(s, t0 + m, cs), dcs, theta,
                                   dtheta, X[cur], Xd[cur], fy, fyd, X[out], Xd[out], s->mpops,
                                   s->fld[0], s->fld[1], s->fld[2], s->fld[3],
                                   first ? nullptr : acc, first ? nullptr : dacc, s->st));
    cur = out;
  }
  // Nu(S_K) from X_{K-1}
  HIP_TRY(s, launch_nu_rows_dual(s->geo, X[cur], Xd[cur], theta, dtheta, acc, dacc, s->st));
  std::vector<double> h(s->p.ny), hd(s->p.ny);
  HIP_TRY(s, hipMemcpyAsync(h.data(), acc, h.size() * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIP_TRY(s, hipMemcpyAsync(hd.data(), dacc, hd.size() * sizeof(double), hipMemcpyDeviceToHost,
                            s->st));
  HIP_TRY(s, hipStreamSynchronize(s->st));
  double sum = 0.0, dsum = 0.0;
  for (double v : h) sum += v;
  for (double v : hd) dsum += v;
  *j_out = sum / (double)s->p.ny * s->d.lz / 1.0 / (double)K;
  *dj_out = dsum / (double)s->p.ny * s->d.lz / 1.0 / (double)K;
  return 0;
}

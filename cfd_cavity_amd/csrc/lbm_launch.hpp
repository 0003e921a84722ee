// lbm_launch.hpp — host-side launch entry points of the cavity kernels
// (internal to libcavlbm; the public boundary is include/cavlbm.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>

#include "lbm_device.hpp"
#include "lbm_stepn.hpp"

namespace cavlbm {

// Whether the steady run loops replay captured graphs (the default).  With
// CAVLBM_GRAPHS=0 in the environment every step block is launched directly:
// the same kernels and the same results, for tools that intercept graph
// launches and to compare the two launch paths.  Read once per process.
inline bool graph_replay_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("CAVLBM_GRAPHS");
    return !(e && e[0] == '0');
  }();
  return on;
}

// One step-kernel launch: storage S, g-equilibrium order geq (1 or 2), mode.
hipError_t launch_step_fp64(int geq, int mode, const Geom& geo, const StepCtx& ctx,
                            const Walls& walls, const StepBufs& b, int write_aux,
                            hipStream_t st);
hipError_t launch_step_lowp(int storage, int geq, int mode, const Geom& geo,
                            const StepCtx& ctx, const Walls& walls, const StepBufs& b,
                            int write_aux, hipStream_t st);

// N consecutive fused steps per launch (N = 2..kMaxBlockSteps, collisions
// with cn.c[0..N-1]), see lbm_stepn.hpp.  geo_n carries the tile grid of the
// storage type and N (stepn_grid); fy is read from fy_in and written to fy_out
// (distinct planes); nu_t of the last step is written when write_aux is set.
// hipErrorInvalidValue when the lattice does not fit the tiling or the storage
// has no N-step form.
hipError_t launch_stepn_fp64(int geq, int n, const Geom& geo_n, const StepCtxN& cn,
                             const Walls& walls, const void* src, void* dst, const void* fy_in,
                             void* fy_out, void* nut, int write_aux, hipStream_t st);
// The objective window's N steps (FP64): launch_stepn_fp64's steps, each of
// which also adds the hot-wall Nusselt term of the state it pulls to
// nu_acc[iy] (as MODE_WINDOW does for one step).
hipError_t launch_stepn_window_fp64(int geq, int n, const Geom& geo_n, const StepCtxN& cn,
                                    const Walls& walls, const void* src, void* dst,
                                    const void* fy_in, void* fy_out, double* nu_acc,
                                    hipStream_t st);
hipError_t launch_stepn_lowp(int storage, int geq, int n, const Geom& geo_n, const StepCtxN& cn,
                             const Walls& walls, const void* src, void* dst, const void* fy_in,
                             void* fy_out, void* nut, int write_aux, hipStream_t st);
// stepn_grid: whether the lattice fits the N-step tiling of the storage type
// (the condition of stepn_fits), and that tiling's grid (nbx, nby) written
// into geo_n; stepn_tile_dims: the output tile of that tiling (false when the
// storage has no N-step form).
bool stepn_grid(int storage, int n, Geom& geo_n);
bool stepn_tile_dims(int storage, int n, int* tx, int* ty);

// f = feq / g = geq from the M-form fields, then u = 0 on the walls.
hipError_t launch_init_eq(const Geom& geo, int geq, double* mpops, const double* rho,
                          double* ux, double* uy, const double* temp, hipStream_t st);

// Phase kernels on the M-form state (FP64), see lbm_aux.hip.
hipError_t launch_phase_stream(const Geom& geo, const double* src, double* dst,
                               hipStream_t st);
hipError_t launch_phase_thermal_bc(const Geom& geo, int geq, const Walls& walls,
                                   double* mpops, hipStream_t st);
hipError_t launch_phase_macro(const Geom& geo, int storage, const double* mpops,
                              const void* fy, double* rho, double* ux, double* uy,
                              double* temp, hipStream_t st);
hipError_t launch_phase_walls(const Geom& geo, double* ux, double* uy, hipStream_t st);
// p[0..n) = v (FP64), by a kernel
hipError_t launch_fill_f64(double* p, double v, size_t n, hipStream_t st);
hipError_t launch_phase_bounce_back(const Geom& geo, double* mpops, hipStream_t st);

// Diagnostics: FP64 reductions over the M-form fields.  `partials` holds
// kDiagSlots*nblocks doubles, `out` kDiagSlots doubles.
constexpr int kDiagSlots = 12;
enum DiagSlot {
  D_MASS = 0, D_FMASS, D_NU_HOT, D_NU_COLD, D_NU_LINK, D_NU_VOL, D_YPLUS,
  D_MAX_UY, D_MAX_U, D_NUT_MAX, D_NONFINITE, D_UNUSED
};
int diag_blocks(const Geom& geo);
hipError_t launch_diagnostics(const Geom& geo, int storage, const double* mpops,
                              const double* rho, const double* ux, const double* uy,
                              const double* temp, const void* nut, double nu,
                              double* partials, double* out, hipStream_t st);
// TimeAverager sums: 5 FP64 planes (ux, uy, T, nu_t, ux^2 + uy^2).
hipError_t launch_avg_accumulate(const Geom& geo, int storage, const double* ux, const double* uy,
                                 const double* temp, const void* nut, double* sums,
                                 hipStream_t st);
// Type conversion of an R-typed per-node field (fy_node / nu_t) to/from FP64.
hipError_t launch_field_convert(const Geom& geo, int storage, const void* src_r,
                                double* dst_d, int to_double, void* dst_r,
                                const double* src_d, hipStream_t st);

inline int nblocks_of(const Geom& g) { return g.nbx * g.nby; }

// row_tiles: geo with the tile grid of kTileX x kRowTileY tiles (one wave
// per workgroup), for the objective window's single steps, backward and
// tangent steps.
inline Geom row_tiles(const Geom& geo) {
  Geom g = geo;
  g.nby = (geo.ny + kRowTileY - 1) / kRowTileY;
  return g;
}

// ---------------------------------------------------------------------------
// Dynamic mixed-scale SGS (lbm_dynsgs.hip).  All arrays are FP64 pitched
// planes of the solver's geometry; the inputs are the M-form state.
constexpr double kTauTrtFloor = 0.505;     // TRT stability floor of tau+
constexpr double kDeltaT = kTHot - kTCold; // global temperature difference
struct DynSgsArgs {
  const double* mpops;  // post-stream f (planes 0..8 used)
  const double* rho;
  const double* ux;
  const double* uy;
  const double* temp;
  double tau;            // molecular tau of this step's collision
  int width, half_left;  // box filter: width nodes, half_left = (width-1)/2
  double dtest;          // test-filter width eta * dx
  double alpha_exp, beta_exp;       // alpha and 1 - alpha
  double nu_t_min, alpha_t_min;     // floors of nu_t / alpha_t
  double* ubx;    // 1 plane each
  double* uby;
  double* tb;
  double* lb;     // 5 planes: box of uu, uv, vv, uT, vT
  double* qsm;
  double* phi;
  double* terms;  // 4 planes: L:M, M:M, Lj Mj, Mj Mj
  double* nu_mix_g;
  double* kap_mix_g;
  double* cnu;    // out: C_nu per node
  double* ck;     // out: C_kappa per node
  double* nut_pre;  // out: nu_t per node
  double* at_pre;   // out: alpha_t per node
  // out (coefficient pass, when set): per workgroup of the pass, the sum of
  // its nodes' nu_t at part[b] and of alpha_t at part[nblocks + b]
  double* part;
};
constexpr int kDynPlanes = 20;  // planes of scratch + outputs above
// The model's passes: the filter and the terms, then (coeffs) the
// coefficient pass; without it the collision runs that stage itself
// (MODE_FUSED_DYN with the DynCoef of dyn_coef_of).  count (1 or 2)
// independent lattices share each launch.
hipError_t launch_dynamic_sgs(const Geom* geo, const DynSgsArgs* a, int count, bool coeffs,
                              hipStream_t st);
// One launch per dynamic step from the C-form of any storage (src -> dst):
// streaming, the model's passes (FP64), the coefficient stage and the
// collision, reading the force plane fy_in and writing fy_out (tiles read
// their halo's fy).  With write_aux the collision's nu_t and the coefficient
// planes are stored.  Test-filter widths 1..3.
inline bool dyn_step_fits(int width) { return width >= 1 && width <= 3; }
// The whole model (filter, terms and coefficient stage, writing C_nu, C_kappa,
// nu_t_pre, alpha_t_pre and, with part, the per-workgroup sums) on count (1
// or 2) lattices from their M-form: one tiled launch for widths 1..3, the
// three passes otherwise.  dyn_model_blocks: the workgroups, i.e. the entries
// of part per sum, the model writes for a lattice of that geometry.
hipError_t launch_dynamic_model(const Geom* geo, const DynSgsArgs* a, int count, hipStream_t st);
int dyn_model_blocks(const Geom& geo, int width);
hipError_t launch_dyn_step(int storage, int geq, const Geom& geo, const StepCtx& ctx,
                           const Walls& walls, const void* src, void* dst, const void* fy_in,
                           void* fy_out, void* nut, const DynSgsArgs& a, int write_aux,
                           hipStream_t st);
// launch_dyn_step's FP32 / FP16-offset forms (the object built with
// contraction, as the other FP32-arithmetic step kernels).
hipError_t launch_dyn_step_lowp(int storage, int geq, const Geom& geo, const StepCtx& ctx,
                                const Walls& walls, const void* src, void* dst,
                                const void* fy_in, void* fy_out, void* nut, const DynSgsArgs& a,
                                int write_aux, hipStream_t st);
__host__ __device__ inline DynCoef dyn_coef_of(const DynSgsArgs& a) {
  DynCoef d;
  d.terms = a.terms;
  d.nu_mix_g = a.nu_mix_g;
  d.kap_mix_g = a.kap_mix_g;
  d.cnu = a.cnu;
  d.ck = a.ck;
  d.nut_pre = a.nut_pre;
  d.at_pre = a.at_pre;
  d.width = a.width;
  d.half_left = a.half_left;
  d.nu_t_min = a.nu_t_min;
  d.alpha_t_min = a.alpha_t_min;
  return d;
}

// ---------------------------------------------------------------------------
// Batched step: many cavities advanced by one launch.  One descriptor per
// cavity (device memory); `tile_cav` maps every global tile to its cavity.
struct CavDesc {
  Geom geo;          // nbx / nby for the single-step tile shape in use
  Geom geon;         // nbx / nby for the N-step tiles of the batch's N
  StepCtx ctx;       // scalars after spin-up (beta = beta_eff)
  double beta_eff;   // beta during spin-up: beta_eff * (step / spin_up)
  int spin_up;
  int tile_begin;    // first global tile of this cavity (single-step table)
  int tile_begin_n;  // first global tile of this cavity (N-step table)
  Walls walls;
  StepBufs bufs[2];  // by buffer parity (src/dst exchanged); fy comes from fyp
  void* fyp[2];      // the two fy_node planes, by fy parity
};

// Tile shapes the batched kernel is instantiated for (autotuner candidates).
struct TileShape {
  int tx, ty;
};
constexpr int kNumBatchShapes = 4;
constexpr TileShape kBatchShapes[kNumBatchShapes] = {{64, 4}, {64, 2}, {64, 8}, {128, 4}};

// parity: the C-form buffer read; fpar: the fy plane read (and, for N-step
// launches, fpar ^ 1 the one written).
hipError_t launch_step_batch_fp64(int geq, int shape, const CavDesc* descs, const int* tile_cav,
                                  int total_tiles, int parity, int fpar, int step, int write_aux,
                                  int xcd_remap, hipStream_t st);
hipError_t launch_step_batch_lowp(int storage, int geq, int shape, const CavDesc* descs,
                                  const int* tile_cav, int total_tiles, int parity, int fpar,
                                  int step, int write_aux, int xcd_remap, hipStream_t st);
// N steps of every cavity per launch (lbm_stepn.hpp tiles, N-step table).
hipError_t launch_stepn_batch_fp64(int geq, int n, const CavDesc* descs, const int* tile_cav,
                                   int total_tiles, int parity, int fpar, int step, int write_aux,
                                   int xcd_remap, hipStream_t st);
hipError_t launch_stepn_batch_lowp(int storage, int geq, int n, const CavDesc* descs,
                                   const int* tile_cav, int total_tiles, int parity, int fpar,
                                   int step, int write_aux, int xcd_remap, hipStream_t st);

// ---------------------------------------------------------------------------
// Adjoint (reverse mode) of the FP64 fused step, see lbm_adjoint.hip.
struct AdjArgs {
  const double* X;       // X_m (18 planes); unused for the first collision
  const double* fyp;     // fy_m
  const double* mpops;   // S_0 populations (first collision only)
  const double* mrho;    // S_0 fields (first collision only)
  const double* mux;
  const double* muy;
  const double* mtemp;
  const double* mu_next;     // mu of the previous backward step (19 planes), null: zero
  const double* mu_next_fy;  // its fy plane (mu_next + 18 planes)
  const double* sgn;     // sign(T(1,iy) - T(2,iy)) of S_{m+1}, null: no objective term
  double obj_w;          // Lz / (ny dT K)
  const double* theta;   // hot-wall profile (ny)
  double* mu;            // out: adjoint of the 18 values each node pulled
  double* mu_fy;         // out: adjoint of fy_m
  double* theta_acc;     // += dJ/dtheta contribution, by node row of column 0
  double* cs_acc;        // += dJ/dCs contribution, per node
};
hipError_t launch_nu_rows(const Geom& geo, int geq, const double* X, const double* theta,
                          double* acc, double* sgn, hipStream_t st);
hipError_t launch_adj_step(const Geom& geo, int geq, bool first, const StepCtx& ctx,
                           const AdjArgs& a, hipStream_t st);
hipError_t launch_sum_nodes(const Geom& geo, const double* v, double* out, hipStream_t st);

// Tangent-linear (forward mode) step and Nusselt rows, see lbm_tangent.hip.
hipError_t launch_tangent_step(const Geom& geo, int geq, bool first, const StepCtx& ctx,
                               double dcs, const double* theta, const double* dtheta,
                               const double* X, const double* Xd, double* fy, double* fyd,
                               double* Y, double* Yd, const double* mpops, const double* mrho,
                               const double* mux, const double* muy, const double* mtemp,
                               double* acc, double* dacc, hipStream_t st);
hipError_t launch_nu_rows_dual(const Geom& geo, const double* X, const double* Xd,
                               const double* theta, const double* dtheta, double* acc,
                               double* dacc, hipStream_t st);

}  // namespace cavlbm

// lbm_octree.hpp — device data and launchers of the octree (multi-level)
// cavity solver (internal to libcavlbm; the public boundary is
// include/cavlbm.h).  Reference: bbrangeo/cfd_cavity src/octree.
#pragma once

#include <hip/hip_runtime.h>

#include "lbm_device.hpp"

namespace cavlbm {

// One level in HBM, flat layout k = iy*nx + ix (levels are narrow strips or a
// small base grid, so no row padding).  M-form (post-stream state, as the
// reference stores it): 23 planes f_0..f_8, g_0..g_8, rho, ux, uy, T, fy.
// C-form: the 18 post-collision population planes.
constexpr int kOctM = 23;
constexpr int kOctRho = 18, kOctUx = 19, kOctUy = 20, kOctT = 21, kOctFy = 22;

struct OctLevel {
  int nx, ny;
  long long plane;  // nx * ny
  double* m;
  double* c;
  double* c2;  // second C-form, for two sub-steps fused across the stream
  double* m2;  // second M-form: the tiled sub-step kernel reads m, writes m2
};

// Relaxation and force of one collision pass: beta of the level and the
// relaxation factors of its RelaxTimes (computed on the host with
// relax_factors from tau+ = 0.5 + 3 nu, tau_t = 0.5 + 3 alpha), handed to
// collide_node_fixed.
struct OctCollide {
  StepCtx ctx;        // beta, t_ref (= 0) of the level
  RelaxF<double> rf;  // the level's relaxation factors
  int interior_only;  // window [1, n-2]^2 (fine levels) or the whole level
  // Dynamic SGS: per-node nu_t / alpha_t of each level of the launch (null:
  // the fixed factors rf), with the level's tau+ and tau_t; the node's
  // relaxation times are 0.5 + 3 (tau - 0.5) / 3 + 3 nu_t (cavity.rs:104-113).
  const double* nut[2];
  const double* at[2];
  double tp, tt;
};

// The relaxation factors of node k of launch item `item`.
__device__ __forceinline__ RelaxF<double> oct_relax(const OctCollide& p, int item, int k) {
  if (!p.nut[item]) return p.rf;
  const double tp = 0.5 + 3.0 * (p.tp - 0.5) / 3.0 + 3.0 * p.nut[item][k];
  const double tt = 0.5 + 3.0 * (p.tt - 0.5) / 3.0 + 3.0 * p.at[item][k];
  return relax_factors<double>(tp, tt);
}

// Streaming / boundary rules of a level.
struct OctStream {
  int block;               // 0: coarse level (domain walls), 1: fine block
  int wl, wr, wb, wt;      // block: wall-side ghost layers (bounce-back sources)
  int hot, cold;           // block: isothermal wall columns a = 1 / a = nx-2
};

// Interface tables on the device (grid.rs:95-203), node indices in the flat
// layouts of the two levels.
struct OctGhostTab {
  int n;
  const int* k_fine;   // n
  const int* kc;       // 4n
  const double* w;     // 4n
};
struct OctCoincTab {
  int n;
  const int* k_coarse;  // n
  const int* kf;        // 9n
  const double* w;      // 9n
};
struct OctXi {
  double even, odd, thermal;
  double inv_even, inv_odd, inv_thermal;  // RN(1/xi), for the restriction's quotients
};

// Device-side update of the interface factors from the levels' mean nu_t /
// alpha_t (dynamic SGS, solver.rs:690-722 and :807-830):
//   xi = rescaling(relax(m_c nu + <nu_t>_c, m_c alpha + <alpha_t>_c),
//                  relax(m_f nu + <nu_t>_f, m_f alpha + <alpha_t>_f))
// where <.>_c is the mean over the coarse level and <.>_f the average over
// the fine levels given (the two blocks for interface A: their means added
// in order, then / 2).  A level's mean is the sum of the coefficient pass's
// per-workgroup sums (part: nu_t at [0, nb), alpha_t at [nb, 2 nb)), added
// by a fixed tree, over the level's n nodes: deterministic, not the
// reference's sequential order.
struct OctXiUpdate {
  const double* part_c;
  long long nb_c, n_c;
  const double* part_f[2];
  long long nb_f[2], n_f[2];
  int nf;                 // fine levels averaged (1 or 2)
  double nu_c, alpha_c;   // molecular (ramped) nu / alpha of each side
  double nu_f, alpha_f;
  int odd_exact;
  OctXi* out;
};

// Launches over `count` (1 or 2) independent levels or interfaces at once:
// the hot and cold blocks (or strips) of a level take one launch, their
// workgroups split between the two.
hipError_t launch_oct_init_eq(const OctLevel& L, int geq, hipStream_t st);
hipError_t launch_oct_walls(const OctLevel& L, const OctStream& s, hipStream_t st);
hipError_t launch_oct_collide(const OctLevel* L, int count, const OctCollide& p, int geq,
                              hipStream_t st);
hipError_t launch_oct_stream(const OctLevel* L, const OctStream* s, int count, int geq,
                             hipStream_t st);
hipError_t launch_oct_stream_collide(const OctLevel* L, const OctStream* s, int count,
                                     const OctCollide& p, int geq, hipStream_t st);
// NS (1 or 2) whole sub-steps of `count` levels in one launch on full-width
// row tiles of `ty` output rows: collision from the M-form (m), the sub-steps'
// streams and collisions through LDS, the new M-form written to m2 (the
// caller then exchanges m and m2).  oct_tile_fits: whether a level of nx x ny
// nodes takes tiles of ty rows (threads, LDS, and a last tile of at least 3
// rows so that the boundary rows' copies reach inside it).
bool oct_tile_fits(int nx, int ny, int ty, int ns);
hipError_t launch_oct_tile(const OctLevel* L, const OctStream* s, int count, int ty, int ns,
                           const OctCollide& p, int geq, hipStream_t st);

// xi_dev: null (the factors xi) or per item the device factors to use.
hipError_t launch_oct_coarse_to_fine(const OctLevel* coarse, const OctLevel* fine,
                                     const OctGhostTab* t, int count, const OctXi& xi,
                                     const OctXi* const* xi_dev, int geq, hipStream_t st);
hipError_t launch_oct_fine_to_coarse(const OctLevel* fine, const OctLevel* coarse,
                                     const OctCoincTab* t, int count, const OctXi& xi,
                                     const OctXi* const* xi_dev, int geq, hipStream_t st);
// One or two interface-factor updates (one workgroup each).
hipError_t launch_oct_xi_update(const OctXiUpdate* u, int count, hipStream_t st);

}  // namespace cavlbm

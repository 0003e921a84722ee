// lbm_stepn.hpp — several consecutive fused steps per launch (temporal
// blocking).
//
// The single-step kernel reads and writes the whole C-form state once per
// step, so it is bound by HBM bandwidth with the collision's arithmetic using
// about a third of the time.  Here one workgroup advances a tile of kTX x kTY
// nodes by N steps.  Sub-step 1 runs on the tile grown by N-1 nodes on every
// side, pulling from HBM; sub-step s (2..N) runs on the tile grown by N-s
// nodes, pulling the populations of sub-step s-1 from LDS; sub-step N's
// populations are stored.  The state crosses HBM once per N steps; the halo is
// recomputed by the neighbouring tiles.
//
// Every node's arithmetic is the single-step kernel's (pull_node /
// node_moments / collide_node), and the values a sub-step pulls are the ones
// the previous step would have stored and the next one loaded (for
// FP16-offset storage the populations go through the storage's own store and
// load), so the result equals N launches of the single-step kernel bit for
// bit in every storage form.
//
// fy is read at the halo nodes, which other tiles write in the same launch,
// so it is double-buffered: the launch reads fy_in and writes fy_out.
#pragma once

#include "lbm_device.hpp"

namespace cavlbm {

constexpr int kMaxBlockSteps = 4;

// Per-step scalars of the N collisions of one launch.  t_ref, cs, pr_t and
// inv_pr_t are constants of the solver, the same in every entry.
struct StepCtxN {
  StepCtx c[kMaxBlockSteps];
};

// The collision scalars of sub-step s (0-based): beta, nu, alpha and tau of
// that step, the solver's constants from entry 0.  Reading the constants
// once keeps the kernel's scalar registers for one copy of them (four copies
// spill to vector lanes).
__device__ __forceinline__ StepCtx sub_ctx(const StepCtxN& cn, int s) {
  StepCtx c = cn.c[0];
  c.beta = cn.c[s].beta;
  c.nu = cn.c[s].nu;
  c.alpha = cn.c[s].alpha;
  c.tau = cn.c[s].tau;
  return c;
}

// Tile shapes per storage type and steps per launch: output tile kTX x kTY,
// first sub-step on (kTX + 2(N-1)) x (kTY + 2(N-1)) nodes, kNT threads.
// Measured at 256x1024 (shapes compared within one run):
//   FP64: N=2 32x12 (against 32x16, 32x10, 24x16, 64x4, 32x6, 16x14, 32x4),
//         N=3 16x16 (against 16x12, 16x20, 32x8, 32x10, 32x12, 64x4, and
//         16x18, 16x19, 16x21, whose regions fill whole waves better but
//         which measured 5-12 % slower),
//         N=4 16x16 (against 16x12, 32x8) -- the fastest, 16 % below N=2;
//   FP32: N=2 64x8 (against 32x12..32x28, 64x12), N=3 32x16 (against 32x12,
//         64x8, 16x16) -- the fastest, 15 % below N=2 -- and N=4 16x16
//         (against 32x16).
// The FP64 kernels take 110-122 VGPRs and 60-74 KB of LDS, so two workgroups
// of at most 8 waves share a CU; the FP32 ones about 60 VGPRs.
template <int S, int N> struct TileN;
template <> struct TileN<ST_FP64, 2> { static constexpr int kTX = 32, kTY = 12, kNT = 512; };
template <> struct TileN<ST_FP64, 3> { static constexpr int kTX = 16, kTY = 16, kNT = 448; };
template <> struct TileN<ST_FP64, 4> { static constexpr int kTX = 16, kTY = 16, kNT = 512; };
template <> struct TileN<ST_FP32, 2> { static constexpr int kTX = 64, kTY = 8, kNT = 704; };
template <> struct TileN<ST_FP32, 3> { static constexpr int kTX = 32, kTY = 16, kNT = 768; };
template <> struct TileN<ST_FP32, 4> { static constexpr int kTX = 16, kTY = 16, kNT = 512; };
// FP16-offset storage: FP32 arithmetic, the FP32 tiles.
template <int N> struct TileN<ST_FP16_OFFSET, N> : TileN<ST_FP32, N> {};

// Steps per launch the run loop uses by default, per storage type.
constexpr int kDefaultBlockFp64 = 4;
constexpr int kDefaultBlockFp32 = 3;
constexpr int kDefaultBlockFp16 = 3;

// The value a population takes through a store to and a load from the
// storage (identity for FP64 and FP32; FP16-offset rounds the deviation from
// the rest state, the node's own temperature t for g, with the random words
// of node k and the values it stores).
template <int S>
__device__ __forceinline__ void storage_round(typename StorageT<S>::R (&f)[9],
                                              typename StorageT<S>::R (&g)[9],
                                              typename StorageT<S>::R t, unsigned k) {
  using ST = StorageT<S>;
  if (!ST::kTPlane) return;
  const SrKey rk = ST::key(k, f, g);
  typename ST::T q;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    ST::store_f(&q, 0u, i, f[i], rk);
    f[i] = ST::load_f(&q, 0u, i);
    ST::store_g(&q, 0u, i, g[i], t, rk);
    g[i] = ST::load_g(&q, 0u, i, t);
  }
}

// The adiabatic top row copies the pulls of row ny-2, which reach row ny-3.
// For sub-steps s < N the grown tile's first row lies below the top row; for
// the last one that row is inside the previous region unless the top row is
// the first row of its output tile ((ny - 1) % kTY == 0).  The host uses
// fewer steps per launch for such heights.
template <int S, int N>
__host__ __device__ inline bool stepn_fits(int nx, int ny) {
  return nx >= 3 && ny >= 3 && (ny - 1) % TileN<S, N>::kTY != 0;
}

// Issue priority of a wave that starts sub-step s of N: N - s (at most 3),
// so the waves of a tile that is behind are issued before those of a tile
// that is ahead.  Two workgroups share a CU and the hardware favours the
// older one; without the hint one of them finishes its first tile about 10 us
// into a launch and the other about 15 us, and each CU ends with one slot
// idle while the other finishes.  Measured at 256x1024 FP64 N = 4: 8.23 ->
// 8.03 us per step; favouring the tile ahead instead: 8.25.
__device__ __forceinline__ void stepn_prio(int p) {
  // the argument is a constant once the sub-step loop is unrolled
  if (p >= 3) __builtin_amdgcn_s_setprio(3);
  else if (p == 2) __builtin_amdgcn_s_setprio(2);
  else if (p == 1) __builtin_amdgcn_s_setprio(1);
  else __builtin_amdgcn_s_setprio(0);
}

// One tile's N steps.  INTERIOR: every node of the first sub-step's region
// lies in [1, nx-2] x [1, ny-2], so every pull is inside the lattice and no
// boundary condition applies (the same arithmetic without the checks).
// PRIO: the sub-step issue priority of stepn_prio (off for the batched
// kernel, whose tiles of many cavities measured 8 % slower with it).
// NU (FP64, the objective window): every sub-step also adds the hot-wall
// Nusselt term |T(1,iy) - T(2,iy)| of the state it pulls (post.rs:68-78) to
// nu_acc[iy], for the rows of this tile's output region only, so each row of
// each step is added once and in step order: T(1, iy) is the node's pulled
// temperature, T(2, iy) the sum of the g that node (2, iy) pulls, from +0 in
// direction order (the isothermal value when column 2 is the cold wall).
template <int S, int GEQ, int N, bool INTERIOR, bool NU = false, bool PRIO = true>
__device__ __forceinline__ void stepn_tile(const Geom& geo, const StepCtxN& cn,
                                           const Walls& walls,
                                           const typename StorageT<S>::T* src,
                                           typename StorageT<S>::T* dst,
                                           const typename StorageT<S>::R* fy_in,
                                           typename StorageT<S>::R* fy_out,
                                           typename StorageT<S>::R* nut, int write_aux,
                                           typename StorageT<S>::R* sh, int bx, int by,
                                           double* nu_acc = nullptr) {
  using ST = StorageT<S>;
  using R = typename ST::R;
  constexpr int kTX = TileN<S, N>::kTX, kTY = TileN<S, N>::kTY;
  constexpr int kH = N - 1;                         // halo of sub-step 1
  constexpr int kE1X = kTX + 2 * kH;
  constexpr int kE1 = kE1X * (kTY + 2 * kH);        // sub-step 1 nodes
  static_assert(!NU || (S == ST_FP64 && !INTERIOR), "the window form is FP64, wall tiles");
  const int nx = geo.nx, ny = geo.ny;
  const int t = (int)threadIdx.x;
  // node (1, iy) of this tile's output rows: it adds the Nusselt term
  auto nu_node = [&](int ix, int iy) {
    return NU && ix == 1 && bx == 0 && iy >= by * kTY && iy < (by + 1) * kTY;
  };

  // ---- sub-step 1: one thread per node of the grown tile, pulls from HBM
  if (PRIO) stepn_prio(N - 1);
  R f[9], g[9], fyc = R(0.0), nutc = R(0.0), tc = R(0.0);
  bool have = false;
  if (t < kE1) {
    const int ix = bx * kTX - kH + t % kE1X, iy = by * kTY - kH + t / kE1X;
    if (INTERIOR || (ix >= 0 && ix < nx && iy >= 0 && iy < ny)) {
      have = true;
      R rho, ux, uy;
      pull_node<S, GEQ, INTERIOR>(geo, walls, src, fy_in[iy * geo.pitch + ix], ix, iy, f, g,
                                  rho, ux, uy, tc);
      if constexpr (NU) {
        if (nu_node(ix, iy)) {
          R t2 = R(0.0);
          if (nx == 3) {
#pragma unroll
            for (int i = 0; i < 9; ++i) t2 += R(kW[i]) * R(kTCold);
          } else {
            const int gy = iy == 0 ? 1 : (iy == ny - 1 ? ny - 2 : iy);
            const unsigned P32 = (unsigned)geo.plane;
#pragma unroll
            for (int i = 0; i < 9; ++i)
              t2 += R(src[(unsigned)(9 + i) * P32 +
                          (unsigned)((gy - kCy[i]) * geo.pitch + (2 - kCx[i]))]);
          }
          nu_acc[iy] += fabs((double)(tc - t2));
        }
      }
      collide_node<R, GEQ>(f, g, rho, ux, uy, tc, cn.c[0], R(cn.c[0].cs), fyc, nutc);
      if (N > 1) storage_round<S>(f, g, tc, (unsigned)(iy * geo.pitch + ix));
    }
  }

  // ---- sub-steps 2..N: pull the previous sub-step's populations from LDS
#pragma unroll
  for (int s = 2; s <= N; ++s) {
    const int hp = N - (s - 1), hs = N - s;        // halos of the two regions
    const int px = kTX + 2 * hp;                    // previous region's width
    const int px0 = bx * kTX - hp, py0 = by * kTY - hp;
    const int cw = kTX + 2 * hs, ch = kTY + 2 * hs;
    if (PRIO) stepn_prio(N - s);
    if (have) {
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        sh[i * kE1 + t] = f[i];
        sh[(10 + i) * kE1 + t] = g[i];
      }
      sh[9 * kE1 + t] = fyc;
    }
    __syncthreads();
    const int ix = bx * kTX - hs + t % cw, iy = by * kTY - hs + t / cw;
    const bool act = t < cw * ch && (INTERIOR || (ix >= 0 && ix < nx && iy >= 0 && iy < ny));
    const int e = (iy - py0) * px + (ix - px0);     // this node in the previous region
    R f2[9], g2[9], fyp = R(0.0);
    if (act) {
      // pull with full-way bounce-back (solver.rs:890-912)
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        const int sx = ix - kCx[i], sy = iy - kCy[i];
        const bool in = INTERIOR || (sx >= 0 && sx < nx && sy >= 0 && sy < ny);
        f2[i] = in ? sh[i * kE1 + e - kCy[i] * px - kCx[i]] : sh[kOpp[i] * kE1 + e];
      }
      fyp = sh[9 * kE1 + e];
      // g: pull, then the thermal boundary conditions (solver.rs:951-988)
      if (!INTERIOR && (ix == 0 || ix == nx - 1)) {
        isothermal_g<R, GEQ>(walls, ix, iy, ny, g2);
      } else {
        const int gy = INTERIOR ? iy : (iy == 0 ? 1 : (iy == ny - 1 ? ny - 2 : iy));
        const int eg = (gy - py0) * px + (ix - px0);
#pragma unroll
        for (int i = 0; i < 9; ++i) g2[i] = sh[(10 + i) * kE1 + eg - kCy[i] * px - kCx[i]];
      }
      const bool wall = !INTERIOR && (ix == 0 || ix == nx - 1 || iy == 0 || iy == ny - 1);
      R rho, ux, uy;
      storage_moments<S>(f2, g2, fyp, wall, rho, ux, uy, tc);
      if constexpr (NU) {
        if (nu_node(ix, iy)) {
          // node (2, iy) pulls from the previous region: its sources lie
          // within one row of this tile's output rows (stepn_fits)
          R t2 = R(0.0);
          if (nx == 3) {
#pragma unroll
            for (int i = 0; i < 9; ++i) t2 += R(kW[i]) * R(kTCold);
          } else {
            const int gy = iy == 0 ? 1 : (iy == ny - 1 ? ny - 2 : iy);
            const int e2 = (gy - py0) * px + (2 - px0);
#pragma unroll
            for (int i = 0; i < 9; ++i) t2 += sh[(10 + i) * kE1 + e2 - kCy[i] * px - kCx[i]];
          }
          nu_acc[iy] += fabs((double)(tc - t2));
        }
      }
      collide_node<R, GEQ>(f2, g2, rho, ux, uy, tc, sub_ctx(cn, s - 1), R(cn.c[0].cs), fyc,
                           nutc);
      if (s < N) storage_round<S>(f2, g2, tc, (unsigned)(iy * geo.pitch + ix));
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        f[i] = f2[i];
        g[i] = g2[i];
      }
    }
    have = act;
    if (s < N) __syncthreads();  // the next sub-step rewrites LDS
  }
  if (!have) return;

  // ---- store the output tile (the last sub-step's region)
  const int ix = bx * kTX + t % kTX, iy = by * kTY + t / kTX;
  const int k = iy * geo.pitch + ix;
  const unsigned P32 = (unsigned)geo.plane;
  fy_out[k] = fyc;
  if (write_aux) nut[k] = nutc;
  const SrKey rk = ST::key((unsigned)k, f, g);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    ST::store_f(dst, (unsigned)i * P32 + (unsigned)k, i, f[i], rk);
    ST::store_g(dst, (unsigned)(9 + i) * P32 + (unsigned)k, i, g[i], tc, rk);
  }
  if (ST::kTPlane) reinterpret_cast<float*>(dst + (unsigned)18 * P32)[k] = (float)tc;
}

template <int S, int GEQ, int N>
__global__ __launch_bounds__((TileN<S, N>::kNT)) void k_stepn(
    Geom geo, StepCtxN cn, Walls walls, const typename StorageT<S>::T* src,
    typename StorageT<S>::T* dst, const typename StorageT<S>::R* fy_in,
    typename StorageT<S>::R* fy_out, typename StorageT<S>::R* nut, int write_aux) {
  using ST = StorageT<S>;
  using R = typename ST::R;
  constexpr int kTX = TileN<S, N>::kTX, kTY = TileN<S, N>::kTY;
  constexpr int kH = N - 1;                         // halo of sub-step 1
  constexpr int kE1X = kTX + 2 * kH;
  constexpr int kE1 = kE1X * (kTY + 2 * kH);        // sub-step 1 nodes
  static_assert(N >= 2 && N <= kMaxBlockSteps, "steps per launch");
  static_assert(kE1 <= TileN<S, N>::kNT, "tile does not fit the workgroup");
  // After a sub-step's collision: planes 0..8 f, plane 9 fy, planes 10..18 g.
  // Thread t's values sit at index t.  One exchange of all 19 planes per
  // sub-step (two barriers) measured 8-15 % faster than exchanging f and g in
  // turn through 10 planes; FP64 tiles take 60-74 KB, two workgroups per CU.
  __shared__ R sh[19 * kE1];

  int bx, by;
  tile_of_block(geo, bx, by);
  // a workgroup-uniform choice of the path without boundary checks
  const bool interior = bx * kTX - kH >= 1 && (bx + 1) * kTX + kH - 1 <= geo.nx - 2 &&
                        by * kTY - kH >= 1 && (by + 1) * kTY + kH - 1 <= geo.ny - 2;
  if (interior)
    stepn_tile<S, GEQ, N, true>(geo, cn, walls, src, dst, fy_in, fy_out, nut, write_aux, sh, bx,
                                by);
  else
    stepn_tile<S, GEQ, N, false>(geo, cn, walls, src, dst, fy_in, fy_out, nut, write_aux, sh, bx,
                                 by);
}

// N steps of the objective window (FP64): the N-step kernel whose sub-steps
// also add the hot-wall Nusselt terms of the states they pull to nu_acc.
template <int GEQ, int N>
__global__ __launch_bounds__((TileN<ST_FP64, N>::kNT)) void k_stepn_window(
    Geom geo, StepCtxN cn, Walls walls, const double* src, double* dst, const double* fy_in,
    double* fy_out, double* nu_acc) {
  constexpr int kTX = TileN<ST_FP64, N>::kTX, kTY = TileN<ST_FP64, N>::kTY;
  constexpr int kH = N - 1;
  constexpr int kE1 = (kTX + 2 * kH) * (kTY + 2 * kH);
  __shared__ double sh[19 * kE1];
  int bx, by;
  tile_of_block(geo, bx, by);
  const bool interior = bx * kTX - kH >= 1 && (bx + 1) * kTX + kH - 1 <= geo.nx - 2 &&
                        by * kTY - kH >= 1 && (by + 1) * kTY + kH - 1 <= geo.ny - 2;
  if (interior)
    stepn_tile<ST_FP64, GEQ, N, true>(geo, cn, walls, src, dst, fy_in, fy_out, nullptr, 0, sh,
                                      bx, by);
  else
    stepn_tile<ST_FP64, GEQ, N, false, true>(geo, cn, walls, src, dst, fy_in, fy_out, nullptr, 0,
                                             sh, bx, by, nu_acc);
}

}  // namespace cavlbm

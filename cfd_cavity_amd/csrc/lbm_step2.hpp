// lbm_step2.hpp — two consecutive fused steps per launch (temporal blocking).
//
// The single-step kernel reads and writes the whole C-form state once per
// step, so it is bound by HBM bandwidth with the collision's arithmetic using
// about a third of the time.  Here one workgroup advances a tile of
// kT2X x kT2Y nodes by two steps: it runs the first step on the tile and a
// one-node halo (read from HBM), keeps those post-collision populations in
// LDS, and runs the second step on the tile from them.  The state crosses HBM
// once per two steps; the halo is recomputed by the neighbouring tiles.
//
// Every node's arithmetic is the single-step kernel's (pull_node /
// node_moments / collide_node), and the values the second step pulls are the
// ones the first step would have stored, so the result equals two launches of
// the single-step kernel bit for bit for storage types equal to the
// arithmetic type (FP64, FP32).
//
// fy is read at the halo nodes, which other tiles write in the same launch,
// so it is double-buffered: the launch reads fy_in and writes fy_out.
#pragma once

#include "lbm_device.hpp"

namespace cavlbm {

// Tile shapes, per storage type (output tile kTX x kTY, first-step nodes
// (kTX + 2) x (kTY + 2), kNT threads).  Measured at 256x1024 against 32x16,
// 32x10, 24x16, 64x4, 32x6, 16x14, 32x4 (FP64) and 32x12..32x28, 64x8, 64x12
// (FP32) tiles:
//   FP64: 32 x 12 (rows of 256 B), 8 waves, 110 VGPRs and 38 KB of LDS, so
//         two workgroups share a CU; taller tiles leave one workgroup per CU,
//         smaller ones recompute more halo.
//   FP32: 64 x 8 (rows of 256 B), 11 waves, about 60 VGPRs and 26 KB of LDS.
template <int S> struct Tile2;
template <> struct Tile2<ST_FP64> {
  static constexpr int kTX = 32, kTY = 12, kNT = 512;
};
template <> struct Tile2<ST_FP32> {
  static constexpr int kTX = 64, kTY = 8, kNT = 704;
};

// The adiabatic top row copies the pulls of row ny-2, which reach row ny-3;
// that row is inside the tile's halo unless the top row is the first row of
// its tile ((ny - 1) % kTY == 0).  The host uses the single-step kernel for
// such heights.
template <int S>
__host__ __device__ inline bool step2_fits(int nx, int ny) {
  return nx >= 3 && ny >= 3 && (ny - 1) % Tile2<S>::kTY != 0;
}

template <int S, int GEQ>
__global__ __launch_bounds__(Tile2<S>::kNT) void k_step2(Geom geo, StepCtx c1, StepCtx c2,
                                                    Walls walls,
                                                    const typename StorageT<S>::T* src,
                                                    typename StorageT<S>::T* dst,
                                                    const typename StorageT<S>::R* fy_in,
                                                    typename StorageT<S>::R* fy_out,
                                                    typename StorageT<S>::R* nut, int write_aux) {
  using ST = StorageT<S>;
  using R = typename ST::R;
  constexpr int kT2X = Tile2<S>::kTX, kT2Y = Tile2<S>::kTY;
  constexpr int kE2X = kT2X + 2;           // tile + one-node halo
  constexpr int kExt2 = kE2X * (kT2Y + 2);  // first-step nodes
  constexpr int kOut2 = kT2X * kT2Y;        // second-step nodes
  static_assert(kExt2 <= Tile2<S>::kNT && kOut2 <= Tile2<S>::kNT, "tile does not fit");
  // planes 0..8: f (then g) after the first collision; plane 9: its fy
  __shared__ R sh[10 * kExt2];

  const int nx = geo.nx, ny = geo.ny;
  int bx, by;
  tile_of_block(geo, bx, by);
  const int x0 = bx * kT2X - 1, y0 = by * kT2Y - 1;  // extended tile origin
  const int t = (int)threadIdx.x;

  // ---- first step on the tile and its halo: one thread per extended node
  R f[9], g[9], fy1 = R(0.0);
  bool have = false;
  if (t < kExt2) {
    const int ix = x0 + t % kE2X, iy = y0 + t / kE2X;
    if (ix >= 0 && ix < nx && iy >= 0 && iy < ny) {
      have = true;
      R rho, ux, uy, temp, nut1;
      pull_node<S, GEQ>(geo, walls, src, fy_in[iy * geo.pitch + ix], ix, iy, f, g, rho, ux, uy,
                        temp);
      collide_node<R, GEQ>(f, g, rho, ux, uy, temp, c1, R(c1.cs), fy1, nut1);
    }
  }
  if (have) {
#pragma unroll
    for (int i = 0; i < 9; ++i) sh[i * kExt2 + t] = f[i];
    sh[9 * kExt2 + t] = fy1;
  }
  __syncthreads();

  // ---- second step on the tile: one thread per output node
  const int ox = t % kT2X, oy = t / kT2X;
  const int ix = x0 + 1 + ox, iy = y0 + 1 + oy;
  const bool act = t < kOut2 && ix < nx && iy < ny;
  const int e = (oy + 1) * kE2X + ox + 1;  // this node in the extended tile
  R f2[9], g2[9], fyp = R(0.0);
  if (act) {
    // pull with full-way bounce-back (solver.rs:890-912)
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const int sx = ix - kCx[i], sy = iy - kCy[i];
      const bool in = sx >= 0 && sx < nx && sy >= 0 && sy < ny;
      f2[i] = in ? sh[i * kExt2 + e - kCy[i] * kE2X - kCx[i]] : sh[kOpp[i] * kExt2 + e];
    }
    fyp = sh[9 * kExt2 + e];
  }
  __syncthreads();
  if (have) {
#pragma unroll
    for (int i = 0; i < 9; ++i) sh[i * kExt2 + t] = g[i];
  }
  __syncthreads();
  if (!act) return;
  // g: pull, then the thermal boundary conditions (solver.rs:951-988)
  if (ix == 0 || ix == nx - 1) {
    isothermal_g<R, GEQ>(walls, ix, iy, ny, g2);
  } else {
    const int gy = iy == 0 ? 1 : (iy == ny - 1 ? ny - 2 : iy);
    const int eg = (gy - y0) * kE2X + (ix - x0);
#pragma unroll
    for (int i = 0; i < 9; ++i) g2[i] = sh[i * kExt2 + eg - kCy[i] * kE2X - kCx[i]];
  }
  const bool wall = ix == 0 || ix == nx - 1 || iy == 0 || iy == ny - 1;
  R rho, ux, uy, temp, fy2, nut2;
  node_moments<R>(f2, g2, fyp, wall, rho, ux, uy, temp);
  collide_node<R, GEQ>(f2, g2, rho, ux, uy, temp, c2, R(c2.cs), fy2, nut2);
  const int k = iy * geo.pitch + ix;
  const unsigned P32 = (unsigned)geo.plane;
  fy_out[k] = fy2;
  if (write_aux) nut[k] = nut2;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    ST::store_f(dst, (unsigned)i * P32 + (unsigned)k, i, f2[i]);
    ST::store_g(dst, (unsigned)(9 + i) * P32 + (unsigned)k, i, g2[i], temp);
  }
}

}  // namespace cavlbm

// lbm_device.hpp — per-node D2Q9 double-distribution physics and the step
// kernel template of the MI355X cavity solver.
//
// Physics follows the reference's uniform solver (bbrangeo/cfd_cavity,
// src/uniform/solver.rs): collide_node :1054-1185, stream :838-913, thermal
// boundary conditions :951-988, update_macroscopic :307-349,
// enforce_wall_velocity :992-1009.  Every floating-point expression keeps the
// reference's operand order; the FP64 instantiation is compiled with
// -ffp-contract=off so that it is bitwise equal to a CPU evaluation in the
// same order.
//
// HBM layout (structure of arrays): a population set holds 18 planes, f_0..f_8
// then g_0..g_8; plane p, node (ix, iy) lives at p*plane + iy*pitch + ix.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cavlbm {

// Direction table: c_i = (0,0) (1,0) (1,1) (0,1) (-1,1) (-1,0) (-1,-1) (0,-1) (1,-1).
constexpr int kCx[9] = {0, 1, 1, 0, -1, -1, -1, 0, 1};
constexpr int kCy[9] = {0, 0, 1, 1, 1, 0, -1, -1, -1};
constexpr int kOpp[9] = {0, 5, 6, 7, 8, 1, 2, 3, 4};
constexpr double kW[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 36.0,
                          1.0 / 9.0,  1.0 / 36.0, 1.0 / 9.0,
                          1.0 / 36.0, 1.0 / 9.0,  1.0 / 36.0};
constexpr double kTHot = 0.5;
constexpr double kTCold = -0.5;
constexpr double kF64Eps = 2.220446049250313e-16;

enum Storage { ST_FP64 = 0, ST_FP32 = 1, ST_FP16_OFFSET = 2 };
enum Mode {
  MODE_FUSED = 0,        // C-form -> stream+BC+moments+walls -> collide -> C-form
  MODE_FIRST = 1,        // M-form pops+fields -> collide -> C-form
  MODE_MATERIALIZE = 2,  // C-form -> stream+BC+moments+walls -> M-form
  MODE_COLLIDE_M = 3     // M-form pops+fields -> collide -> M-form pops (in place)
};

struct Geom {
  int nx, ny, pitch;
  int nbx, nby;       // tile grid: tiles of kTileX x kTileY nodes
  long long plane;    // elements per population plane (pitch*ny)
};

// Per-step scalars computed on the host exactly as collide() does
// (solver.rs:433-454); t_ref/cs/pr_t are constants of the solver.
struct StepCtx {
  double beta, nu, alpha, tau, t_ref, cs, pr_t;
};

struct Walls {
  const double* theta_hot;  // ny values or nullptr (constant kTHot)
};

struct StepBufs {
  const void* src;  // C-form population set (storage type)
  void* dst;        // C-form population set (storage type)
  void* fy;         // fy_node, compute type
  void* nut;        // nu_t, compute type
  double* mpops;    // M-form population set (FP64)
  double* rho;      // M-form fields (FP64)
  double* ux;
  double* uy;
  double* temp;
};

constexpr int kTileX = 64;
constexpr int kTileY = 4;
constexpr int kBlock = kTileX * kTileY;

// ---------------------------------------------------------------------------
// storage traits
template <int S> struct StorageT;
template <> struct StorageT<ST_FP64> {
  using T = double;
  using R = double;
  __device__ static __forceinline__ R load_f(const T* p, long long o, int) { return p[o]; }
  __device__ static __forceinline__ R load_g(const T* p, long long o) { return p[o]; }
  __device__ static __forceinline__ void store_f(T* p, long long o, int, R v) { p[o] = v; }
  __device__ static __forceinline__ void store_g(T* p, long long o, R v) { p[o] = v; }
};
template <> struct StorageT<ST_FP32> {
  using T = float;
  using R = float;
  __device__ static __forceinline__ R load_f(const T* p, long long o, int) { return p[o]; }
  __device__ static __forceinline__ R load_g(const T* p, long long o) { return p[o]; }
  __device__ static __forceinline__ void store_f(T* p, long long o, int, R v) { p[o] = v; }
  __device__ static __forceinline__ void store_g(T* p, long long o, R v) { p[o] = v; }
};
// FP16 storage of f - w_i (the rest-state offset keeps ~11 significant bits
// on the O(1e-2) deviations) and of g itself (its rest value w_i*0 is zero).
template <> struct StorageT<ST_FP16_OFFSET> {
  using T = __half;
  using R = float;
  __device__ static __forceinline__ R load_f(const T* p, long long o, int i) {
    return __half2float(p[o]) + (float)kW[i];
  }
  __device__ static __forceinline__ R load_g(const T* p, long long o) { return __half2float(p[o]); }
  __device__ static __forceinline__ void store_f(T* p, long long o, int i, R v) {
    p[o] = __float2half_rn(v - (float)kW[i]);
  }
  __device__ static __forceinline__ void store_g(T* p, long long o, R v) { p[o] = __float2half_rn(v); }
};

// ---------------------------------------------------------------------------
// equilibria (the crate's feq/geq; feq form as written in
// src/uniform/tests.rs:314-316)
template <class R>
__device__ __forceinline__ R feq(int i, R rho, R ux, R uy) {
  const R cs2 = R(1.0) / R(3.0);
  const R cx = R(kCx[i]), cy = R(kCy[i]);
  const R eu = cx * ux + cy * uy;
  const R u2 = ux * ux + uy * uy;
  return R(kW[i]) * rho *
         (R(1.0) + eu / cs2 + eu * eu / (R(2.0) * cs2 * cs2) - u2 / (R(2.0) * cs2));
}

template <class R, int GEQ>
__device__ __forceinline__ R geq(int i, R t, R ux, R uy) {
  const R cs2 = R(1.0) / R(3.0);
  const R cx = R(kCx[i]), cy = R(kCy[i]);
  const R eu = cx * ux + cy * uy;
  if (GEQ == 1) return R(kW[i]) * t * (R(1.0) + eu / cs2);
  const R u2 = ux * ux + uy * uy;
  return R(kW[i]) * t *
         (R(1.0) + eu / cs2 + eu * eu / (R(2.0) * cs2 * cs2) - u2 / (R(2.0) * cs2));
}

// ---------------------------------------------------------------------------
// collide_node (solver.rs:1054-1185): TRT Lambda = 3/16 with Smagorinsky nu_t
// from pi_neq and the Guo odd/even force split for f; BGK for g.
template <class R, int GEQ>
__device__ __forceinline__ void collide_node(R (&f)[9], R (&g)[9], R rho, R ux, R uy,
                                             R temp, const StepCtx& c, R& fy_out,
                                             R& nut_out) {
  const R cs2 = R(1.0) / R(3.0);
  const R fy = rho * R(c.beta) * (temp - R(c.t_ref));
  fy_out = fy;

  R d_f[9], d_g[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    d_f[i] = f[i] - feq<R>(i, rho, ux, uy);
    d_g[i] = g[i] - geq<R, GEQ>(i, temp, ux, uy);
  }
  R pi_xx = R(0.0), pi_xy = R(0.0), pi_yy = R(0.0);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const R cx = R(kCx[i]), cy = R(kCy[i]);
    const R fneq = d_f[i];
    pi_xx += fneq * cx * cx;
    pi_xy += fneq * cx * cy;
    pi_yy += fneq * cy * cy;
  }
  const R pi2 = pi_xx * pi_xx + R(2.0) * pi_xy * pi_xy + pi_yy * pi_yy;
  const R s_mag = sqrt(pi2) / (R(2.0) * cs2 * R(c.tau) * fmax(rho, R(1e-12)));
  const R nu_t = R(c.cs) * R(c.cs) * s_mag;
  nut_out = nu_t;

  const R tau_plus = R(0.5) + R(3.0) * (R(c.nu) + nu_t);
  const R tau_minus = R(0.5) + R(3.0 / 16.0) / (tau_plus - R(0.5));
  const R alpha_t = nu_t / R(c.pr_t);
  const R tau_t_plus = R(0.5) + R(3.0) * (R(c.alpha) + alpha_t);
  const R tau_t_minus = tau_t_plus;
  const R a_f = R(0.5) * (R(1.0) / tau_plus + R(1.0) / tau_minus);
  const R b_f = R(0.5) * (R(1.0) / tau_plus - R(1.0) / tau_minus);
  const R a_g = R(0.5) * (R(1.0) / tau_t_plus + R(1.0) / tau_t_minus);
  const R b_g = R(0.5) * (R(1.0) / tau_t_plus - R(1.0) / tau_t_minus);
  const R k_minus = R(1.0) - R(0.5) / tau_minus;
  const R k_plus = R(1.0) - R(0.5) / tau_plus;

  R fo[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const R cx = R(kCx[i]), cy = R(kCy[i]);
    const R ei_u = cx * ux + cy * uy;
    const R s_i = R(kW[i]) * ((cy - uy) / cs2 + ei_u * cy / (cs2 * cs2)) * fy;
    const R s_minus = R(kW[i]) * (cy / cs2) * fy;
    const R s_plus = s_i - s_minus;
    fo[i] = f[i] + (-a_f * d_f[i] - b_f * d_f[kOpp[i]] + k_minus * s_minus +
                    k_plus * s_plus);
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    f[i] = fo[i];
    g[i] = g[i] + (-a_g * d_g[i] - b_g * d_g[kOpp[i]]);
  }
}

// ---------------------------------------------------------------------------
// Tile index of this workgroup, remapped so that each XCD (workgroups are
// dispatched round-robin over the 8 XCDs) owns a contiguous band of rows and
// the pull stencil's halo rows stay in that XCD's L2.
__device__ __forceinline__ void tile_of_block(const Geom& geo, int& bx, int& by) {
  const int total = geo.nbx * geo.nby;
  const int lin = blockIdx.x;
  const int per = total / 8, rem = total % 8;
  const int xcd = lin % 8, idx = lin / 8;
  const int t = xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
  bx = t % geo.nbx;
  by = t / geo.nbx;
}

// ---------------------------------------------------------------------------
// The step kernel: one thread per lattice node.
template <int S, int GEQ, int MODE>
__global__ __launch_bounds__(kBlock) void k_step(Geom geo, StepCtx ctx, Walls walls,
                                                  StepBufs b, int write_aux) {
  using ST = StorageT<S>;
  using T = typename ST::T;
  using R = typename ST::R;
  int bx, by;
  tile_of_block(geo, bx, by);
  const int ix = bx * kTileX + (int)(threadIdx.x % kTileX);
  const int iy = by * kTileY + (int)(threadIdx.x / kTileX);
  const int nx = geo.nx, ny = geo.ny;
  if (ix >= nx || iy >= ny) return;
  const long long P = geo.plane;
  const int pitch = geo.pitch;
  const long long k = (long long)iy * pitch + ix;
  const bool wall = ix == 0 || ix == nx - 1 || iy == 0 || iy == ny - 1;

  R f[9], g[9];
  R rho, ux, uy, temp;

  if (MODE == MODE_FUSED || MODE == MODE_MATERIALIZE) {
    const T* src = static_cast<const T*>(b.src);
    // pull streaming with full-way bounce-back for f (solver.rs:890-912)
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const int sx = ix - kCx[i], sy = iy - kCy[i];
      const bool in = sx >= 0 && sx < nx && sy >= 0 && sy < ny;
      const long long o = in ? (long long)i * P + (long long)sy * pitch + sx
                             : (long long)kOpp[i] * P + k;
      f[i] = ST::load_f(src, o, in ? i : kOpp[i]);
    }
    // g: pull, then the thermal boundary conditions (solver.rs:951-988):
    // isothermal columns, then adiabatic rows copying rows 1 / ny-2.
    if (ix == 0 || ix == nx - 1) {
      R th;
      if (ix == 0) {
        const int iyc = iy == 0 ? 1 : (iy == ny - 1 ? ny - 2 : iy);
        th = walls.theta_hot ? R(walls.theta_hot[iyc]) : R(kTHot);
      } else {
        th = R(kTCold);
      }
#pragma unroll
      for (int i = 0; i < 9; ++i) g[i] = geq<R, GEQ>(i, th, R(0.0), R(0.0));
    } else {
      const int gy = iy == 0 ? 1 : (iy == ny - 1 ? ny - 2 : iy);
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        const long long o = (long long)(9 + i) * P + (long long)(gy - kCy[i]) * pitch + (ix - kCx[i]);
        g[i] = ST::load_g(src, o);
      }
    }
    // update_macroscopic with the Guo correction (solver.rs:319-345)
    R r = R(0.0), un = R(0.0), vn = R(0.0), tn = R(0.0);
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      r += f[j];
      un += R(kCx[j]) * f[j];
      vn += R(kCy[j]) * f[j];
      tn += g[j];
    }
    const R fyp = static_cast<const R*>(b.fy)[k];
    if (r > R(kF64Eps)) {
      un /= r;
      vn /= r;
      vn += R(0.5) * fyp / r;
    }
    // enforce_wall_velocity (solver.rs:992-1009)
    if (wall) {
      un = R(0.0);
      vn = R(0.0);
    }
    rho = r;
    ux = un;
    uy = vn;
    temp = tn;
    if (MODE == MODE_MATERIALIZE) {
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        b.mpops[(long long)i * P + k] = (double)f[i];
        b.mpops[(long long)(9 + i) * P + k] = (double)g[i];
      }
      b.rho[k] = (double)rho;
      b.ux[k] = (double)ux;
      b.uy[k] = (double)uy;
      b.temp[k] = (double)temp;
      return;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      f[i] = R(b.mpops[(long long)i * P + k]);
      g[i] = R(b.mpops[(long long)(9 + i) * P + k]);
    }
    rho = R(b.rho[k]);
    ux = R(b.ux[k]);
    uy = R(b.uy[k]);
    temp = R(b.temp[k]);
  }

  R fy_new, nut_new;
  collide_node<R, GEQ>(f, g, rho, ux, uy, temp, ctx, fy_new, nut_new);
  static_cast<R*>(b.fy)[k] = fy_new;
  if (write_aux) static_cast<R*>(b.nut)[k] = nut_new;

  if (MODE == MODE_COLLIDE_M) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      b.mpops[(long long)i * P + k] = (double)f[i];
      b.mpops[(long long)(9 + i) * P + k] = (double)g[i];
    }
  } else {
    T* dst = static_cast<T*>(b.dst);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      ST::store_f(dst, (long long)i * P + k, i, f[i]);
      ST::store_g(dst, (long long)(9 + i) * P + k, g[i]);
    }
  }
}

}  // namespace cavlbm

// lbm_octree.hip — per-level kernels and interface transfers of the octree
// (multi-level) cavity solver.  Reference: bbrangeo/cfd_cavity
// src/octree/cavity.rs:167-656 (per-level phases), interface.rs:164-229
// (transfers), solver.rs:1188-1206 (moment refresh of restricted nodes).
//
// A level sub-step is two launches: k_oct_collide (M-form -> C-form: TRT+Guo
// collision on the level's window, plain copy elsewhere) and k_oct_stream
// (C-form -> M-form: pull streaming with the level's wall rules, the thermal
// boundary conditions resolved per node, the moments with the Guo correction
// and the wall velocities).  The transfers work on M-forms.  Built with
// -ffp-contract=off: the results are bitwise those of a CPU evaluation in
// the reference's operand order.
#include "lbm_octree.hpp"

namespace cavlbm {

namespace {

// One wave per workgroup: the levels and interfaces are small, so the
// launches' few thousand lanes spread over more CUs (64x256 base grid: two
// levels 45.1 -> 43.9 us per coarse step, three 78.3 -> 75.0 against 256).
constexpr int kOctBlock = 64;

__device__ __forceinline__ int node_id(long long n) {
  const long long k = (long long)blockIdx.x * kOctBlock + threadIdx.x;
  return k < n ? (int)k : -1;
}

// Up to two independent items (levels or interfaces) per launch: workgroups
// [0, nblk0) serve item 0, the rest item 1.
struct OctGroup {
  OctLevel L[2];
  OctStream S[2];
  int nblk0;
};
struct OctXferGroup {
  OctLevel Lc[2], Lf[2];
  OctGhostTab gt[2];
  OctCoincTab ct[2];
  const OctXi* xd[2];  // per item: device factors, or null (the launch's xi)
  int nblk0;
};

__device__ __forceinline__ int group_item(int nblk0, int& local_block) {
  const int b = (int)blockIdx.x;
  const int w = b >= nblk0 ? 1 : 0;
  local_block = w ? b - nblk0 : b;
  return w;
}

__device__ __forceinline__ int local_id(int local_block, long long n) {
  const long long k = (long long)local_block * kOctBlock + threadIdx.x;
  return k < n ? (int)k : -1;
}

// Sums of update_macroscopic_guo / refresh_moments (cavity.rs:296-315,
// interface.rs:213-229): products by 0 and -1 are exact and the zero ones
// leave the +0-started accumulators unchanged.
struct Moments {
  double r, mx, my, t;
};
__device__ __forceinline__ Moments moments(const double (&f)[9], const double (&g)[9]) {
  Moments m;
  m.r = 0.0 + f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
  m.mx = 0.0 + f[1] + f[2] - f[4] - f[5] - f[6] + f[8];
  m.my = 0.0 + f[2] + f[3] + f[4] - f[6] - f[7] - f[8];
  m.t = 0.0 + g[0] + g[1] + g[2] + g[3] + g[4] + g[5] + g[6] + g[7] + g[8];
  return m;
}

template <int GEQ>
__global__ __launch_bounds__(kOctBlock) void k_oct_init_eq(OctLevel L) {
  const int k = node_id(L.plane);
  if (k < 0) return;
  const long long P = L.plane;
  double* m = L.m;
  const double r = m[kOctRho * P + k], u = m[kOctUx * P + k], v = m[kOctUy * P + k],
               t = m[kOctT * P + k];
  EqTerms<double> e;
  eq_terms<double>(u, v, e);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    m[i * P + k] = eq_quad<double>(i, r, e);
    m[(9 + i) * P + k] = eq_g<double, GEQ>(i, t, e);
  }
}

__device__ __forceinline__ bool is_wall_node(const OctLevel& L, const OctStream& s, int ix,
                                             int iy) {
  if (!s.block) return ix == 0 || ix == L.nx - 1 || iy == 0 || iy == L.ny - 1;
  return (s.hot && ix == 1) || (s.cold && ix == L.nx - 2) || iy == 1 || iy == L.ny - 2;
}

// enforce_wall_velocity(_block) (cavity.rs:604-650) alone, after an injection
// of fields (solver.rs:1160-1168).
__global__ __launch_bounds__(kOctBlock) void k_oct_walls(OctLevel L, OctStream s) {
  const int k = node_id(L.plane);
  if (k < 0) return;
  const int ix = k % L.nx, iy = k / L.nx;
  if (is_wall_node(L, s, ix, iy)) {
    L.m[kOctUx * L.plane + k] = 0.0;
    L.m[kOctUy * L.plane + k] = 0.0;
  }
}

// collide_cavity (cavity.rs:167-243): collide_node_cavity is the uniform
// collide_node's force, equilibria and update with the level's fixed
// relaxation factors (tau+ = 0.5 + 3 nu, tau- = 0.5 + (3/16)/(tau+ - 0.5),
// tau_t = 0.5 + 3 alpha, a_g = 1/tau_t, b_g = 0: the level's RelaxTimes).
// Nodes outside the window keep their populations (ghost layers carry the
// injected values).
template <int GEQ>
__global__ __launch_bounds__(kOctBlock) void k_oct_collide(OctGroup grp, OctCollide p) {
  int lb;
  const int item = group_item(grp.nblk0, lb);
  const OctLevel& L = grp.L[item];
  const int k = local_id(lb, L.plane);
  if (k < 0) return;
  const long long P = L.plane;
  const int ix = k % L.nx, iy = k / L.nx;
  const bool in = !p.interior_only ||
                  (ix >= 1 && ix <= L.nx - 2 && iy >= 1 && iy <= L.ny - 2);
  if (!in) {
#pragma unroll
    for (int i = 0; i < 18; ++i) L.c[i * P + k] = L.m[i * P + k];
    return;
  }
  double f[9], g[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    f[i] = L.m[i * P + k];
    g[i] = L.m[(9 + i) * P + k];
  }
  double fy;
  collide_node_fixed<GEQ>(f, g, L.m[kOctRho * P + k], L.m[kOctUx * P + k], L.m[kOctUy * P + k],
                          L.m[kOctT * P + k], p.ctx.beta, p.ctx.t_ref, oct_relax(p, item, k),
                          fy);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    L.c[i * P + k] = f[i];
    L.c[(9 + i) * P + k] = g[i];
  }
  L.m[kOctFy * P + k] = fy;
}

// Pull source rule of one direction for an interior node: the block form
// bounces back from wall-side ghost layers (cavity.rs:423-446), the coarse
// form from outside the domain (lbm.rs:143-171).
__device__ __forceinline__ bool bounce(const OctLevel& L, const OctStream& s, int sx, int sy) {
  if (!s.block) return !(sx >= 0 && sx < L.nx && sy >= 0 && sy < L.ny);
  return (sx == 0 && s.wl) || (sx == L.nx - 1 && s.wr) || (sy == 0 && s.wb) ||
         (sy == L.ny - 1 && s.wt);
}

// Post-collision values of a level in HBM, C[i*plane + y*nx + x].
struct CGlobal {
  const double* C;
  long long P;
  int nx;
  __device__ __forceinline__ double operator()(int i, int x, int y) const {
    return C[i * P + y * nx + x];
  }
};

// stream + BCs + moments + walls of one node from post-collision values
// src(i, x, y) (HBM or a tile in LDS); fyp is the node's fy of the same
// sub-step's collision.
template <int GEQ, class Src>
__device__ __forceinline__ void stream_node(const OctLevel& L, const OctStream& s, const Src& src,
                                            double fyp, int ix, int iy, double (&f)[9],
                                            double (&g)[9], double& rho, double& ux,
                                            double& uy, double& temp) {
  const int nx = L.nx, ny = L.ny;

  // f: pull; the block form leaves its ghost layers as they are
  const bool interior = ix >= 1 && ix <= nx - 2 && iy >= 1 && iy <= ny - 2;
  if (!s.block || interior) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const int sx = ix - kCx[i], sy = iy - kCy[i];
      f[i] = bounce(L, s, sx, sy) ? src(kOpp[i], ix, iy) : src(i, sx, sy);
    }
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) f[i] = src(i, ix, iy);
  }

  // g: the thermal boundary conditions resolved per node.  Every node ends
  // with the wall equilibrium or with the streamed g of one interior node:
  //  coarse (cavity.rs:500-533): columns 0 / nx-1 isothermal for every row,
  //    then rows 0 / ny-1 copy rows 1 / ny-2;
  //  block (cavity.rs:542-600): columns 1 / nx-2 isothermal, rows 1 / ny-2
  //    copy rows 2 / ny-3, then the ghost columns and rows copy inwards.
  int gx, gy;
  int wall = 0;  // +1 hot, -1 cold
  if (!s.block) {
    wall = ix == 0 ? 1 : (ix == nx - 1 ? -1 : 0);
    gx = ix;
    gy = iy == 0 ? 1 : (iy == ny - 1 ? ny - 2 : iy);
  } else {
    const int b1 = iy == 0 ? 1 : (iy == ny - 1 ? ny - 2 : iy);
    gx = ix == 0 ? 1 : (ix == nx - 1 ? nx - 2 : ix);
    gy = b1 == 1 ? 2 : (b1 == ny - 2 ? ny - 3 : b1);
    wall = (s.hot && gx == 1) ? 1 : ((s.cold && gx == nx - 2) ? -1 : 0);
  }
  if (wall != 0) {
    EqTerms<double> e0;
    eq_terms<double>(0.0, 0.0, e0);
    const double tw = wall > 0 ? kTHot : kTCold;
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] = eq_g<double, GEQ>(i, tw, e0);
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const int sx = gx - kCx[i], sy = gy - kCy[i];
      g[i] = bounce(L, s, sx, sy) ? src(9 + i, gx, gy) : src(9 + i, sx, sy);
    }
  }

  // update_macroscopic_guo (cavity.rs:296-315), then the wall velocities
  const Moments mo = moments(f, g);
  ux = 0.0;
  uy = 0.0;
  if (mo.r > kF64Eps) {
    ux = mo.mx / mo.r;
    uy = mo.my / mo.r + 0.5 * fyp / mo.r;
  }
  if (is_wall_node(L, s, ix, iy)) {
    ux = 0.0;
    uy = 0.0;
  }
  rho = mo.r;
  temp = mo.t;
}

template <int GEQ>
__global__ __launch_bounds__(kOctBlock) void k_oct_stream(OctGroup grp) {
  int lb;
  const int item = group_item(grp.nblk0, lb);
  const OctLevel& L = grp.L[item];
  const int k = local_id(lb, L.plane);
  if (k < 0) return;
  const long long P = L.plane;
  double f[9], g[9], rho, ux, uy, temp;
  stream_node<GEQ>(L, grp.S[item], CGlobal{L.c, P, L.nx}, L.m[kOctFy * P + k], k % L.nx,
                   k / L.nx, f, g, rho, ux, uy, temp);
  double* M = L.m;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    M[i * P + k] = f[i];
    M[(9 + i) * P + k] = g[i];
  }
  M[kOctRho * P + k] = rho;
  M[kOctUx * P + k] = ux;
  M[kOctUy * P + k] = uy;
  M[kOctT * P + k] = temp;
}

// Two consecutive sub-steps of a level with no transfer between them: the
// first sub-step's stream + BCs + moments + walls and the second sub-step's
// collision in one pass, C-form (c) -> C-form (c2), exactly the two separate
// launches' arithmetic.  The M-form is not needed in between; the following
// stream launch reads c2.
template <int GEQ>
__global__ __launch_bounds__(kOctBlock) void k_oct_stream_collide(OctGroup grp, OctCollide p) {
  int lb;
  const int item = group_item(grp.nblk0, lb);
  const OctLevel& L = grp.L[item];
  const int k = local_id(lb, L.plane);
  if (k < 0) return;
  const long long P = L.plane;
  const int ix = k % L.nx, iy = k / L.nx;
  double f[9], g[9], rho, ux, uy, temp;
  stream_node<GEQ>(L, grp.S[item], CGlobal{L.c, P, L.nx}, L.m[kOctFy * P + k], ix, iy, f, g, rho,
                   ux, uy, temp);
  const bool in = !p.interior_only ||
                  (ix >= 1 && ix <= L.nx - 2 && iy >= 1 && iy <= L.ny - 2);
  if (in) {
    double fy;
    collide_node_fixed<GEQ>(f, g, rho, ux, uy, temp, p.ctx.beta, p.ctx.t_ref,
                            oct_relax(p, item, k), fy);
    L.m[kOctFy * P + k] = fy;
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    L.c2[i * P + k] = f[i];
    L.c2[(9 + i) * P + k] = g[i];
  }
}

// ---------------------------------------------------------------------------
// Whole sub-steps of a level in one launch.  A workgroup owns output rows
// [y0, y0 + ty) of a level, all columns.  Sub-step 1 collides the rows grown
// by NS (read from the M-form m); each sub-step's stream runs on rows grown by
// NS - s, pulling from the previous collision's values in LDS, and, before the
// last, collides again; the last stream writes the new M-form to m2.  Every
// node is computed with the separate launches' functions (collide_node,
// stream_node), so the result is bitwise theirs.  The rows outside a tile
// are recomputed by its neighbours; the level's m is only read, so no tile
// sees another's output.
struct OctTileGroup {
  OctLevel L[2];
  OctStream S[2];
  int ty;      // output rows per tile
  int ntile0;  // tiles of item 0
};

constexpr int kOctTileMaxThreads = 1024;
constexpr int kOctTileLdsBytes = 65536;  // dynamic LDS of one tile, at most

// Post-collision values of a tile's region in LDS: planes 0..8 f, 9 fy,
// 10..18 g, rows from r0, all nx columns.
struct CTile {
  const double* sh;
  int q, nx, r0;
  __device__ __forceinline__ double operator()(int i, int x, int y) const {
    return sh[(i < 9 ? i : i + 1) * q + (y - r0) * nx + x];
  }
};

template <int GEQ, int NS>
__global__ __launch_bounds__(kOctTileMaxThreads) void k_oct_tile(OctTileGroup grp,
                                                                  OctCollide p) {
  extern __shared__ double sh[];
  int lt;
  const int item = group_item(grp.ntile0, lt);
  const OctLevel& L = grp.L[item];
  const OctStream& S = grp.S[item];
  const int nx = L.nx, ny = L.ny;
  const long long P = L.plane;
  const int y0 = lt * grp.ty, y1 = min(y0 + grp.ty, ny);
  const int q = nx * (grp.ty + 2 * NS);  // LDS plane stride
  const int t = (int)threadIdx.x;

  // sub-step 1's collision on rows [y0 - NS, y1 + NS), from the M-form
  int pr0 = max(y0 - NS, 0);
  const int pr1 = min(y1 + NS, ny);
  double f[9], g[9], fyc = 0.0;
  bool have = t < nx * (pr1 - pr0);
  if (have) {
    const int ix = t % nx, iy = pr0 + t / nx;
    const int k = iy * nx + ix;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      f[i] = L.m[i * P + k];
      g[i] = L.m[(9 + i) * P + k];
    }
    fyc = L.m[kOctFy * P + k];
    const bool in = !p.interior_only || (ix >= 1 && ix <= nx - 2 && iy >= 1 && iy <= ny - 2);
    if (in)
      collide_node_fixed<GEQ>(f, g, L.m[kOctRho * P + k], L.m[kOctUx * P + k],
                              L.m[kOctUy * P + k], L.m[kOctT * P + k], p.ctx.beta, p.ctx.t_ref,
                              oct_relax(p, item, k), fyc);
  }

#pragma unroll
  for (int s = 1; s <= NS; ++s) {
    if (have) {
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        sh[i * q + t] = f[i];
        sh[(10 + i) * q + t] = g[i];
      }
      sh[9 * q + t] = fyc;
    }
    __syncthreads();
    // this sub-step's stream on rows [y0 - (NS - s), y1 + (NS - s))
    const int r0 = max(y0 - (NS - s), 0), r1 = min(y1 + (NS - s), ny);
    const bool act = t < nx * (r1 - r0);
    const int ix = t % nx, iy = r0 + t / nx;
    double f2[9], g2[9], rho = 0.0, ux = 0.0, uy = 0.0, temp = 0.0, fyp = 0.0;
    if (act) {
      fyp = sh[9 * q + (iy - pr0) * nx + ix];
      stream_node<GEQ>(L, S, CTile{sh, q, nx, pr0}, fyp, ix, iy, f2, g2, rho, ux, uy, temp);
    }
    if (s == NS) {
      if (act) {
        const int k = iy * nx + ix;
        double* M = L.m2;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
          M[i * P + k] = f2[i];
          M[(9 + i) * P + k] = g2[i];
        }
        M[kOctRho * P + k] = rho;
        M[kOctUx * P + k] = ux;
        M[kOctUy * P + k] = uy;
        M[kOctT * P + k] = temp;
        M[kOctFy * P + k] = fyp;
      }
    } else {
      __syncthreads();  // the next sub-step rewrites LDS
      fyc = fyp;
      if (act) {
        const bool in =
            !p.interior_only || (ix >= 1 && ix <= nx - 2 && iy >= 1 && iy <= ny - 2);
        if (in)
          collide_node_fixed<GEQ>(f2, g2, rho, ux, uy, temp, p.ctx.beta, p.ctx.t_ref,
                                  oct_relax(p, item, iy * nx + ix), fyc);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
          f[i] = f2[i];
          g[i] = g2[i];
        }
      }
      have = act;
      pr0 = r0;
    }
  }
}

// The interface transfers, one workgroup lane per (entry, stencil node):
// each lane forms its node's weighted moments and weighted non-equilibrium
// parts, and the entry's first lane adds them in stencil order (the order of
// the reference's accumulation loops) and injects the result.  Splitting the
// stencil over lanes shortens each thread's dependent chain, which bounds
// these small launches.
constexpr int kXferVals = 22;  // w*rho, w*ux, w*uy, w*T, 9 f and 9 g parts

template <int NS>
struct XferShape {
  static constexpr int kEntries = kOctBlock / NS;  // entries per workgroup
};

template <int GEQ, int NS>
__device__ __forceinline__ void xfer_lane(const OctLevel& Ls, int ks, double w,
                                          double (&c)[kXferVals]) {
  const long long P = Ls.plane;
  const double* m = Ls.m;
  const double r = m[kOctRho * P + ks], u = m[kOctUx * P + ks], v = m[kOctUy * P + ks],
               t = m[kOctT * P + ks];
  c[0] = w * r;
  c[1] = w * u;
  c[2] = w * v;
  c[3] = w * t;
  EqTerms<double> e;
  eq_terms<double>(u, v, e);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    c[4 + i] = w * (m[i * P + ks] - eq_quad<double>(i, r, e));
    c[13 + i] = w * (m[(9 + i) * P + ks] - eq_g<double, GEQ>(i, t, e));
  }
}

// coarse_to_fine (interface.rs:164-195): bilinear moments and non-equilibrium
// of the coarse level, rescaled by xi, injected on the ghost nodes, whose
// moments are then recomputed from the injected populations.
template <int GEQ>
__global__ __launch_bounds__(kOctBlock) void k_oct_c2f(OctXferGroup grp, OctXi xi_arg) {
  constexpr int NS = 4, EPB = XferShape<NS>::kEntries;
  __shared__ double sh[EPB * NS][kXferVals];
  int lb;
  const int item = group_item(grp.nblk0, lb);
  const OctLevel& Lc = grp.Lc[item];
  const OctLevel& Lf = grp.Lf[item];
  const OctGhostTab& t = grp.gt[item];
  const OctXi xi = grp.xd[item] ? *grp.xd[item] : xi_arg;
  const int e = (int)threadIdx.x / NS, q = (int)threadIdx.x % NS;
  const int j = lb * EPB + e;
  const bool valid = e < EPB && j < t.n;
  if (valid) {
    double c[kXferVals];
    xfer_lane<GEQ, NS>(Lc, t.kc[4 * j + q], t.w[4 * j + q], c);
#pragma unroll
    for (int v = 0; v < kXferVals; ++v) sh[threadIdx.x][v] = c[v];
  }
  __syncthreads();
  // the entry's lanes add the 22 quantities over the stencil, each in
  // stencil order, into the first row of the entry's block of sh
  double part[(kXferVals + NS - 1) / NS];
#pragma unroll
  for (int u = 0; u < (kXferVals + NS - 1) / NS; ++u) {
    const int v = q + u * NS;
    double a = 0.0;
    if (valid && v < kXferVals)
      for (int r = 0; r < NS; ++r) a += sh[e * NS + r][v];
    part[u] = a;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < (kXferVals + NS - 1) / NS; ++u) {
    const int v = q + u * NS;
    if (valid && v < kXferVals) sh[e * NS][v] = part[u];
  }
  __syncthreads();
  if (!valid || q != 0) return;
  double s[kXferVals];
#pragma unroll
  for (int v = 0; v < kXferVals; ++v) s[v] = sh[e * NS][v];
  const long long Pf = Lf.plane;
  const double rho = s[0], ux = s[1], uy = s[2], temp = s[3];
  const double* fneq = s + 4;
  const double* gneq = s + 13;
  EqTerms<double> eq;
  eq_terms<double>(ux, uy, eq);
  double f[9], g[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int o = kOpp[i];
    const double even = 0.5 * (fneq[i] + fneq[o]);
    const double odd = 0.5 * (fneq[i] - fneq[o]);
    f[i] = eq_quad<double>(i, rho, eq) + (xi.even * even + xi.odd * odd);
    g[i] = eq_g<double, GEQ>(i, temp, eq) + xi.thermal * gneq[i];
  }
  const int kf = t.k_fine[j];
  double* fm = Lf.m;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    fm[i * Pf + kf] = f[i];
    fm[(9 + i) * Pf + kf] = g[i];
  }
  const Moments mo = moments(f, g);
  fm[kOctRho * Pf + kf] = mo.r;
  fm[kOctUx * Pf + kf] = mo.mx / mo.r;
  fm[kOctUy * Pf + kf] = mo.my / mo.r;
  fm[kOctT * Pf + kf] = mo.t;
}

// fine_to_coarse (interface.rs:200-210 with the 9-node restriction of
// grid.rs:178-196), then refresh_moments_nodes (solver.rs:1188-1206) of the
// restricted node.
template <int GEQ>
__global__ __launch_bounds__(kOctBlock) void k_oct_f2c(OctXferGroup grp, OctXi xi_arg) {
  constexpr int NS = 9, EPB = XferShape<NS>::kEntries;
  __shared__ double sh[EPB * NS][kXferVals];
  int lb;
  const int item = group_item(grp.nblk0, lb);
  const OctLevel& Lf = grp.Lf[item];
  const OctLevel& Lc = grp.Lc[item];
  const OctCoincTab& t = grp.ct[item];
  const OctXi xi = grp.xd[item] ? *grp.xd[item] : xi_arg;
  const int e = (int)threadIdx.x / NS, q = (int)threadIdx.x % NS;
  const int j = lb * EPB + e;
  const bool valid = e < EPB && j < t.n;
  if (valid) {
    double c[kXferVals];
    xfer_lane<GEQ, NS>(Lf, t.kf[9 * j + q], t.w[9 * j + q], c);
#pragma unroll
    for (int v = 0; v < kXferVals; ++v) sh[threadIdx.x][v] = c[v];
  }
  __syncthreads();
  // the entry's lanes add the 22 quantities over the stencil, each in
  // stencil order, into the first row of the entry's block of sh
  double part[(kXferVals + NS - 1) / NS];
#pragma unroll
  for (int u = 0; u < (kXferVals + NS - 1) / NS; ++u) {
    const int v = q + u * NS;
    double a = 0.0;
    if (valid && v < kXferVals)
      for (int r = 0; r < NS; ++r) a += sh[e * NS + r][v];
    part[u] = a;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < (kXferVals + NS - 1) / NS; ++u) {
    const int v = q + u * NS;
    if (valid && v < kXferVals) sh[e * NS][v] = part[u];
  }
  __syncthreads();
  if (!valid || q != 0) return;
  double s[kXferVals];
#pragma unroll
  for (int v = 0; v < kXferVals; ++v) s[v] = sh[e * NS][v];
  const long long Pc = Lc.plane;
  const double rho = s[0], ux = s[1], uy = s[2], temp = s[3];
  const double* fneq = s + 4;
  const double* gneq = s + 13;
  EqTerms<double> eq;
  eq_terms<double>(ux, uy, eq);
  double f[9], g[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int o = kOpp[i];
    const double even = 0.5 * (fneq[i] + fneq[o]);
    const double odd = 0.5 * (fneq[i] - fneq[o]);
    // a / xi as divr(a, xi, RN(1/xi)): the correctly rounded quotient
    // (lbm_device.hpp); a zero dividend only changes the sign of a zero
    // that is added to the nonzero equilibrium or to +0
    f[i] = eq_quad<double>(i, rho, eq) + (Arith<double>::divr(even, xi.even, xi.inv_even) +
                                          Arith<double>::divr(odd, xi.odd, xi.inv_odd));
    g[i] = eq_g<double, GEQ>(i, temp, eq) +
           Arith<double>::divr(gneq[i], xi.thermal, xi.inv_thermal);
  }
  const int kc = t.k_coarse[j];
  double* cm = Lc.m;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    cm[i * Pc + kc] = f[i];
    cm[(9 + i) * Pc + kc] = g[i];
  }
  const Moments mo = moments(f, g);
  cm[kOctRho * Pc + kc] = mo.r;
  cm[kOctT * Pc + kc] = mo.t;
  cm[kOctUx * Pc + kc] = mo.r > kF64Eps ? mo.mx / mo.r : 0.0;
  cm[kOctUy * Pc + kc] = mo.r > kF64Eps ? mo.my / mo.r : 0.0;
}

inline dim3 grid_for(long long n) { return dim3((unsigned)((n + kOctBlock - 1) / kOctBlock)); }

// Interface factors from the levels' mean nu_t / alpha_t: one workgroup per
// update, over the per-workgroup sums of the coefficient pass.  A level's
// total is a fixed-order sum: thread t adds entries t, t + kXiThreads, ...
// from +0 in that order, each wave adds its 64 partials by a butterfly of
// lane exchanges (a + b and b + a are the same IEEE sum, so every lane holds
// the same total), and the first wave adds the waves' totals the same way.
// The coarse level and the fine levels of an update (up to kXiSets, each with
// nu_t and alpha_t) are reduced together, behind one barrier.
constexpr int kXiThreads = 256;
constexpr int kXiWaves = kXiThreads / 64;
constexpr int kXiSets = 3;
constexpr int kXiBatch = 4;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sums of a[s][0..len[s]) and b[s][0..len[s]) for s < ns, divided by div[s].
__device__ void block_means(const double* const (&a)[kXiSets], const double* const (&b)[kXiSets],
                            const long long (&len)[kXiSets], const long long (&div)[kXiSets],
                            int ns, double* red, double (&ma)[kXiSets], double (&mb)[kXiSets]) {
  const int t = (int)threadIdx.x;
  double sa[kXiSets], sb[kXiSets];
  long long nmax = 0;
#pragma unroll
  for (int s = 0; s < kXiSets; ++s) {
    sa[s] = 0.0;
    sb[s] = 0.0;
    if (s < ns && len[s] > nmax) nmax = len[s];
  }
  for (long long k0 = t; k0 < nmax; k0 += (long long)kXiBatch * kXiThreads) {
    double va[kXiSets][kXiBatch], vb[kXiSets][kXiBatch];
#pragma unroll
    for (int s = 0; s < kXiSets; ++s)
#pragma unroll
      for (int j = 0; j < kXiBatch; ++j) {
        const long long k = k0 + (long long)j * kXiThreads;
        const bool in = s < ns && k < len[s];
        va[s][j] = in ? a[s][k] : 0.0;
        vb[s][j] = in ? b[s][k] : 0.0;
      }
#pragma unroll
    for (int s = 0; s < kXiSets; ++s)
#pragma unroll
      for (int j = 0; j < kXiBatch; ++j) {
        if (s < ns && k0 + (long long)j * kXiThreads < len[s]) {
          sa[s] += va[s][j];
          sb[s] += vb[s][j];
        }
      }
  }
  const int w = t / 64, lane = t % 64;
#pragma unroll
  for (int s = 0; s < kXiSets; ++s) {
    const double ta = wave_sum(sa[s]), tb = wave_sum(sb[s]);
    if (lane == 0) {
      red[(2 * s) * kXiWaves + w] = ta;
      red[(2 * s + 1) * kXiWaves + w] = tb;
    }
  }
  __syncthreads();
  double* tot = red + 2 * kXiSets * kXiWaves;
  if (w == 0) {
    for (int s = 0; s < ns; ++s) {
      const double ta = wave_sum(lane < kXiWaves ? red[(2 * s) * kXiWaves + lane] : 0.0);
      const double tb = wave_sum(lane < kXiWaves ? red[(2 * s + 1) * kXiWaves + lane] : 0.0);
      if (lane == 0) {
        tot[2 * s] = ta;
        tot[2 * s + 1] = tb;
      }
    }
  }
  __syncthreads();
  for (int s = 0; s < ns; ++s) {
    ma[s] = tot[2 * s] / (double)div[s];
    mb[s] = tot[2 * s + 1] / (double)div[s];
  }
}

__global__ __launch_bounds__(kXiThreads) void k_oct_xi_update(OctXiUpdate u0, OctXiUpdate u1) {
  __shared__ double red[2 * kXiSets * kXiWaves + 2 * kXiSets];
  const OctXiUpdate& u = blockIdx.x == 0 ? u0 : u1;
  const double* const a[kXiSets] = {u.part_c, u.part_f[0], u.part_f[1]};
  const double* const b[kXiSets] = {u.part_c + u.nb_c, u.part_f[0] + u.nb_f[0],
                                    u.part_f[1] + u.nb_f[1]};
  const long long len[kXiSets] = {u.nb_c, u.nb_f[0], u.nb_f[1]};
  const long long div[kXiSets] = {u.n_c, u.n_f[0], u.n_f[1]};
  double ma[kXiSets], mb[kXiSets];
  block_means(a, b, len, div, 1 + u.nf, red, ma, mb);
  const double nt_c = ma[0], at_c = mb[0];
  double nt_f = 0.0, at_f = 0.0;
  for (int i = 0; i < u.nf; ++i) {
    nt_f += ma[1 + i];
    at_f += mb[1 + i];
  }
  if (threadIdx.x != 0) return;
  const double nf = (double)u.nf;
  // RelaxTimes::from_nu_alpha (lbm.rs:100-109) of both sides
  const double tpc = 0.5 + 3.0 * (u.nu_c + nt_c);
  const double tmc = 0.5 + (3.0 / 16.0) / (tpc - 0.5);
  const double ttc = 0.5 + 3.0 * (u.alpha_c + at_c);
  const double tpf = 0.5 + 3.0 * (u.nu_f + nt_f / nf);
  const double tmf = 0.5 + (3.0 / 16.0) / (tpf - 0.5);
  const double ttf = 0.5 + 3.0 * (u.alpha_f + at_f / nf);
  // rescaling_factors_multilevel (interface.rs:141-158), two sub-steps
  const double dt = 1.0 / 2.0;
  OctXi x;
  x.even = dt * tpf / tpc;
  x.odd = u.odd_exact ? dt * tmf / tmc : dt;
  x.thermal = dt * ttf / ttc;
  x.inv_even = 1.0 / x.even;
  x.inv_odd = 1.0 / x.odd;
  x.inv_thermal = 1.0 / x.thermal;
  *u.out = x;
}

}  // namespace

hipError_t launch_oct_xi_update(const OctXiUpdate* u, int count, hipStream_t st) {
  if (count < 1 || count > 2) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_oct_xi_update, dim3(count), dim3(kXiThreads), 0, st, u[0],
                     u[count - 1]);
  return hipGetLastError();
}

hipError_t launch_oct_init_eq(const OctLevel& L, int geq, hipStream_t st) {
  if (geq == 1)
    hipLaunchKernelGGL(k_oct_init_eq<1>, grid_for(L.plane), dim3(kOctBlock), 0, st, L);
  else
    hipLaunchKernelGGL(k_oct_init_eq<2>, grid_for(L.plane), dim3(kOctBlock), 0, st, L);
  return hipGetLastError();
}

hipError_t launch_oct_walls(const OctLevel& L, const OctStream& s, hipStream_t st) {
  hipLaunchKernelGGL(k_oct_walls, grid_for(L.plane), dim3(kOctBlock), 0, st, L, s);
  return hipGetLastError();
}

namespace {

inline int nblk(long long n) { return (int)((n + kOctBlock - 1) / kOctBlock); }

}  // namespace

hipError_t launch_oct_collide(const OctLevel* L, int count, const OctCollide& p, int geq,
                              hipStream_t st) {
  OctGroup g{};
  int total = 0;
  for (int i = 0; i < count; ++i) {
    g.L[i] = L[i];
    if (i == 0) g.nblk0 = nblk(L[i].plane);
    total += nblk(L[i].plane);
  }
  if (count == 1) g.nblk0 = total;
  if (geq == 1)
    hipLaunchKernelGGL(k_oct_collide<1>, dim3(total), dim3(kOctBlock), 0, st, g, p);
  else
    hipLaunchKernelGGL(k_oct_collide<2>, dim3(total), dim3(kOctBlock), 0, st, g, p);
  return hipGetLastError();
}

hipError_t launch_oct_stream(const OctLevel* L, const OctStream* s, int count, int geq,
                             hipStream_t st) {
  OctGroup g{};
  int total = 0;
  for (int i = 0; i < count; ++i) {
    g.L[i] = L[i];
    g.S[i] = s[i];
    if (i == 0) g.nblk0 = nblk(L[i].plane);
    total += nblk(L[i].plane);
  }
  if (geq == 1)
    hipLaunchKernelGGL(k_oct_stream<1>, dim3(total), dim3(kOctBlock), 0, st, g);
  else
    hipLaunchKernelGGL(k_oct_stream<2>, dim3(total), dim3(kOctBlock), 0, st, g);
  return hipGetLastError();
}

hipError_t launch_oct_stream_collide(const OctLevel* L, const OctStream* s, int count,
                                     const OctCollide& p, int geq, hipStream_t st) {
  OctGroup g{};
  int total = 0;
  for (int i = 0; i < count; ++i) {
    g.L[i] = L[i];
    g.S[i] = s[i];
    if (i == 0) g.nblk0 = nblk(L[i].plane);
    total += nblk(L[i].plane);
  }
  if (geq == 1)
    hipLaunchKernelGGL(k_oct_stream_collide<1>, dim3(total), dim3(kOctBlock), 0, st, g, p);
  else
    hipLaunchKernelGGL(k_oct_stream_collide<2>, dim3(total), dim3(kOctBlock), 0, st, g, p);
  return hipGetLastError();
}

bool oct_tile_fits(int nx, int ny, int ty, int ns) {
  // levels are at least 6 x 6 (the block rules reach 3 rows and columns in)
  if (ns < 1 || ns > 2 || ty < 3 || nx < 6 || ny < 6) return false;
  const long long rows = (long long)ty + 2 * ns;
  if (nx * rows > kOctTileMaxThreads) return false;
  if (19 * nx * rows * (long long)sizeof(double) > kOctTileLdsBytes) return false;
  const int last = ny % ty;
  return last == 0 || last >= 3;
}

hipError_t launch_oct_tile(const OctLevel* L, const OctStream* s, int count, int ty, int ns,
                           const OctCollide& p, int geq, hipStream_t st) {
  OctTileGroup g{};
  int total = 0, threads = 0;
  for (int i = 0; i < count; ++i) {
    if (!oct_tile_fits(L[i].nx, L[i].ny, ty, ns)) return hipErrorInvalidValue;
    g.L[i] = L[i];
    g.S[i] = s[i];
    const int nt = (L[i].ny + ty - 1) / ty;
    if (i == 0) g.ntile0 = nt;
    total += nt;
    threads = std::max(threads, L[i].nx * (ty + 2 * ns));
  }
  if (count == 1) g.ntile0 = total;
  g.ty = ty;
  threads = (threads + 63) / 64 * 64;
  int lds_nx = 0;
  for (int i = 0; i < count; ++i) lds_nx = std::max(lds_nx, L[i].nx);
  const size_t lds = (size_t)19 * lds_nx * (ty + 2 * ns) * sizeof(double);
#define CAVLBM_OCT_TILE(G, N)                                                                \
  if (geq == G && ns == N)                                                                   \
    hipLaunchKernelGGL((k_oct_tile<G, N>), dim3(total), dim3(threads), lds, st, g, p);
  CAVLBM_OCT_TILE(1, 1)
  CAVLBM_OCT_TILE(1, 2)
  CAVLBM_OCT_TILE(2, 1)
  CAVLBM_OCT_TILE(2, 2)
#undef CAVLBM_OCT_TILE
  return hipGetLastError();
}

hipError_t launch_oct_coarse_to_fine(const OctLevel* coarse, const OctLevel* fine,
                                     const OctGhostTab* t, int count, const OctXi& xi,
                                     const OctXi* const* xi_dev, int geq, hipStream_t st) {
  OctXferGroup g{};
  int total = 0;
  for (int i = 0; i < count; ++i) {
    g.Lc[i] = coarse[i];
    g.Lf[i] = fine[i];
    g.gt[i] = t[i];
    g.xd[i] = xi_dev ? xi_dev[i] : nullptr;
    const int nb = (t[i].n + XferShape<4>::kEntries - 1) / XferShape<4>::kEntries;
    if (i == 0) g.nblk0 = nb;
    total += nb;
  }
  if (total == 0) return hipSuccess;
  if (geq == 1)
    hipLaunchKernelGGL(k_oct_c2f<1>, dim3(total), dim3(kOctBlock), 0, st, g, xi);
  else
    hipLaunchKernelGGL(k_oct_c2f<2>, dim3(total), dim3(kOctBlock), 0, st, g, xi);
  return hipGetLastError();
}

hipError_t launch_oct_fine_to_coarse(const OctLevel* fine, const OctLevel* coarse,
                                     const OctCoincTab* t, int count, const OctXi& xi,
                                     const OctXi* const* xi_dev, int geq, hipStream_t st) {
  OctXferGroup g{};
  int total = 0;
  for (int i = 0; i < count; ++i) {
    g.Lf[i] = fine[i];
    g.Lc[i] = coarse[i];
    g.ct[i] = t[i];
    g.xd[i] = xi_dev ? xi_dev[i] : nullptr;
    const int nb = (t[i].n + XferShape<9>::kEntries - 1) / XferShape<9>::kEntries;
    if (i == 0) g.nblk0 = nb;
    total += nb;
  }
  if (total == 0) return hipSuccess;
  if (geq == 1)
    hipLaunchKernelGGL(k_oct_f2c<1>, dim3(total), dim3(kOctBlock), 0, st, g, xi);
  else
    hipLaunchKernelGGL(k_oct_f2c<2>, dim3(total), dim3(kOctBlock), 0, st, g, xi);
  return hipGetLastError();
}

}  // namespace cavlbm

// lbm_tangent.hip — forward-mode (tangent-linear) FP64 step for the window
// objective J = (1/K) sum_m nusselt_hot(S_m): the same per-node templates as
// the production step (equilibria, collide_node), instantiated on a dual
// number (value, derivative).  It gives dJ[v] for a direction
// v = (dtheta_hot, dCs) and checks the hand-written adjoint through the
// identity <grad J, v> = dJ[v].
//
// At the two non-differentiable points the same one-sided choice as the
// adjoint is made: d sqrt(x) at x = 0 is 0, d|x| at x = 0 is 0.
#include "lbm_launch.hpp"

namespace cavlbm {

struct Dual {
  double v, d;
  __device__ Dual() : v(0.0), d(0.0) {}
  __device__ Dual(double x) : v(x), d(0.0) {}
  __device__ Dual(double x, double dx) : v(x), d(dx) {}
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) {
  return {a.v * b.v, a.d * b.v + a.v * b.d};
}
// Division, square root and the constant-divisor quotient through the step's
// FP64 helpers (one reciprocal, two corrected quotients): the value parts
// are the IEEE results the forward step computes, with shorter sequences
// than the compiler's division.
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double y = Arith<double>::rcp(b.v);
  const double q = quot_f64(a.v, b.v, y);
  return {q, quot_f64(a.d - q * b.d, b.v, y)};
}
__device__ __forceinline__ Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
__device__ __forceinline__ bool operator>(Dual a, Dual b) { return a.v > b.v; }
__device__ __forceinline__ Dual sqrt(Dual a) {
  const double s = Arith<double>::sqrtp(a.v);
  return {s, a.v > 0.0 ? Arith<double>::div(a.d, 2.0 * s) : 0.0};
}
__device__ __forceinline__ Dual fmax(Dual a, Dual b) { return a.v >= b.v ? a : b; }

template <> struct Arith<Dual> {
  __device__ static __forceinline__ Dual divr(Dual a, Dual b, Dual) { return a / b; }
  __device__ static __forceinline__ Dual divc(Dual a, Dual b, Dual y) {
    const double q = Arith<double>::divc(a.v, b.v, y.v);
    return {q, Arith<double>::divc(a.d - q * b.d, b.v, y.v)};
  }
  __device__ static __forceinline__ Dual div(Dual a, Dual b) { return a / b; }
  __device__ static __forceinline__ Dual rcp(Dual b) { return Dual(1.0) / b; }
  __device__ static __forceinline__ Dual sqrtp(Dual x) { return sqrt(x); }
};

namespace {

// Pre-collision g of node (ix, iy) after streaming and the thermal BCs.
__device__ __forceinline__ void pulled_g_dual(const Geom& geo, const double* X,
                                              const double* Xd, const double* theta,
                                              const double* dtheta, int ix, int iy,
                                              Dual (&g)[9]) {
  const int nx = geo.nx, ny = geo.ny, pitch = geo.pitch;
  const long long P = geo.plane;
  if (ix == 0 || ix == nx - 1) {
    const int iyc = iy == 0 ? 1 : (iy == ny - 1 ? ny - 2 : iy);
    const Dual th = ix == 0 ? Dual(theta[iyc], dtheta[iyc]) : Dual(kTCold);
    // geq(th, 0, 0) = w_i * th
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] = Dual(kW[i]) * th;
    return;
  }
  const int gy = iy == 0 ? 1 : (iy == ny - 1 ? ny - 2 : iy);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const long long o = (9 + i) * P + (long long)(gy - kCy[i]) * pitch + (ix - kCx[i]);
    g[i] = Dual(X[o], Xd[o]);
  }
}

// acc[iy] += |T1 - T2|, dacc[iy] += d|T1 - T2| for dlt = T1 - T2.
__device__ __forceinline__ void nu_row_add(Dual dlt, double* acc, double* dacc, int iy) {
  acc[iy] += fabs(dlt.v);
  dacc[iy] += dlt.v > 0.0 ? dlt.d : (dlt.v < 0.0 ? -dlt.d : 0.0);
}

// One tangent step; with acc (a step that pulls, m >= 1) it also accumulates
// the Nusselt row terms of the state S_m it pulls, as the window step does.
template <int GEQ, bool FIRST>
__global__ __launch_bounds__(kRowBlock) void k_tangent_step(Geom geo, StepCtx ctx, double dcs,
                                                          const double* theta,
                                                          const double* dtheta,
                                                          const double* X, const double* Xd,
                                                          double* fy, double* fyd, double* Y,
                                                          double* Yd, const double* mpops,
                                                          const double* mrho, const double* mux,
                                                          const double* muy,
                                                          const double* mtemp, double* acc,
                                                          double* dacc) {
  int bx, by;
  tile_of_block(geo, bx, by);
  const int ix = bx * kTileX + (int)(threadIdx.x % kTileX);
  const int iy = by * kRowTileY + (int)(threadIdx.x / kTileX);  // geo: the row_tiles grid
  const int nx = geo.nx, ny = geo.ny, pitch = geo.pitch;
  if (ix >= nx || iy >= ny) return;
  const long long P = geo.plane;
  const long long k = (long long)iy * pitch + ix;
  const bool wall = ix == 0 || ix == nx - 1 || iy == 0 || iy == ny - 1;
  using A = Arith<Dual>;

  Dual f[9], g[9], rho, ux, uy, temp;
  if (FIRST) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      f[i] = Dual(mpops[i * P + k]);
      g[i] = Dual(mpops[(9 + i) * P + k]);
    }
    rho = Dual(mrho[k]);
    ux = Dual(mux[k]);
    uy = Dual(muy[k]);
    temp = Dual(mtemp[k]);
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const int sx = ix - kCx[i], sy = iy - kCy[i];
      const bool in = sx >= 0 && sx < nx && sy >= 0 && sy < ny;
      const long long o = in ? i * P + (long long)sy * pitch + sx : kOpp[i] * P + k;
      f[i] = Dual(X[o], Xd[o]);
    }
    pulled_g_dual(geo, X, Xd, theta, dtheta, ix, iy, g);
    // update_macroscopic with the Guo correction, same sums as the step kernel
    const Dual r = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    Dual un = f[1] + f[2] - f[4] - f[5] - f[6] + f[8];
    Dual vn = f[2] + f[3] + f[4] - f[6] - f[7] - f[8];
    const Dual tn = g[0] + g[1] + g[2] + g[3] + g[4] + g[5] + g[6] + g[7] + g[8];
    const Dual fyp(fy[k], fyd[k]);
    if (r > Dual(kF64Eps)) {
      const Dual ir = A::rcp(r);
      un = A::divr(un, r, ir);
      vn = A::divr(vn, r, ir);
      vn += A::divr(Dual(0.5) * fyp, r, ir);
    }
    if (wall) {
      un = Dual(0.0);
      vn = Dual(0.0);
    }
    rho = r;
    ux = un;
    uy = vn;
    temp = tn;
    // Nusselt row term of the pulled state (same sums as k_nu_rows_dual):
    // node (1, iy) adds |T(1,iy) - T(2,iy)| and its tangent
    if (acc && ix == 1) {
      Dual g2[9];
      pulled_g_dual(geo, X, Xd, theta, dtheta, 2, iy, g2);
      Dual t2(0.0);
#pragma unroll
      for (int i = 0; i < 9; ++i) t2 += g2[i];
      nu_row_add(tn - t2, acc, dacc, iy);
    }
  }
  Dual fy_new, nut_new;
  collide_node<Dual, GEQ>(f, g, rho, ux, uy, temp, ctx, Dual(ctx.cs, dcs), fy_new, nut_new);
  fy[k] = fy_new.v;
  fyd[k] = fy_new.d;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    Y[i * P + k] = f[i].v;
    Yd[i * P + k] = f[i].d;
    Y[(9 + i) * P + k] = g[i].v;
    Yd[(9 + i) * P + k] = g[i].d;
  }
}

// Nusselt rows with their tangent: acc += |T1 - T2|, dacc += d|T1 - T2|.
__global__ void k_nu_rows_dual(Geom geo, const double* X, const double* Xd, const double* theta,
                               const double* dtheta, double* acc, double* dacc) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  if (iy >= geo.ny) return;
  Dual g1[9], g2[9];
  pulled_g_dual(geo, X, Xd, theta, dtheta, 1, iy, g1);
  pulled_g_dual(geo, X, Xd, theta, dtheta, 2, iy, g2);
  Dual t1(0.0), t2(0.0);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    t1 += g1[i];
    t2 += g2[i];
  }
  nu_row_add(t1 - t2, acc, dacc, iy);
}

}  // namespace

hipError_t launch_tangent_step(const Geom& geo, int geq, bool first, const StepCtx& ctx,
                               double dcs, const double* theta, const double* dtheta,
                               const double* X, const double* Xd, double* fy, double* fyd,
                               double* Y, double* Yd, const double* mpops, const double* mrho,
                               const double* mux, const double* muy, const double* mtemp,
                               double* acc, double* dacc, hipStream_t st) {
  const Geom gr = row_tiles(geo);
  const dim3 grid(nblocks_of(gr)), block(kRowBlock);
#define CAVLBM_T(G, F)                                                                      \
  hipLaunchKernelGGL((k_tangent_step<G, F>), grid, block, 0, st, gr, ctx, dcs, theta, dtheta, \
                     X, Xd, fy, fyd, Y, Yd, mpops, mrho, mux, muy, mtemp, acc, dacc)
  if (geq == 1) {
    if (first) CAVLBM_T(1, true); else CAVLBM_T(1, false);
  } else {
    if (first) CAVLBM_T(2, true); else CAVLBM_T(2, false);
  }
#undef CAVLBM_T
  return hipGetLastError();
}

hipError_t launch_nu_rows_dual(const Geom& geo, const double* X, const double* Xd,
                               const double* theta, const double* dtheta, double* acc,
                               double* dacc, hipStream_t st) {
  const int tb = 256;
  hipLaunchKernelGGL(k_nu_rows_dual, dim3((geo.ny + tb - 1) / tb), dim3(tb), 0, st, geo, X, Xd,
                     theta, dtheta, acc, dacc);
  return hipGetLastError();
}

}  // namespace cavlbm

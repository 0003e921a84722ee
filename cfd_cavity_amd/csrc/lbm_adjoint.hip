// lbm_adjoint.hip — reverse-mode (adjoint) kernels of the fused FP64 step, for
// gradients of the window-averaged hot-wall Nusselt number
//     J = (1/K) sum_{m=1..K} Nu_hot(S_m),
//     Nu_hot(S) = (Lz/(ny dT)) sum_iy |T(1,iy) - T(2,iy)|    (src/uniform/post.rs:68-78)
// with respect to the hot-wall temperature profile theta_hot[iy] (the
// isothermal condition of src/uniform/solver.rs:957-963) and the Smagorinsky
// constant Cs (solver.rs:1120).
//
// State convention (as in the forward kernels): X_m is the post-collision
// population set of collision m, fy_m its Guo force; S_{m+1} = Mat(X_m, fy_m)
// is stream + thermal BCs + moments + wall velocity, X_{m+1} = Collide(S_{m+1}).
//
// One launch per backward step m, k_adj_node (one thread per node k):
//   1. lambda(k), the adjoint of X_{m+1}(k) (18 populations + fy), gathered
//      from mu_{m+1}, the output of the previous backward step: the sum of the
//      mu of every destination that read X_{m+1}(k) (the pull partner k + c_i,
//      the bounce-back partner, and for g the adiabatic rows 0 / ny-1, which
//      copy the values pulled by rows 1 / ny-2) -- the transpose of the pull,
//      as the forward step pulls its own inputs;
//   2. recompute S_{m+1} at k from X_m and the collision at k, then propagate
//      lambda(k) and the objective's dJ/dT(k) back to the adjoints mu_m(k) of
//      the 18 values node k pulled, of fy_m(k), of theta_hot (column 0) and
//      of Cs.
// mu alternates between two buffers (neighbours read mu_{m+1} while mu_m is
// written).
// The adjoint of sqrt(pi:pi) at pi = 0 is taken as 0, and that of |x| at 0.
#include "lbm_launch.hpp"

namespace cavlbm {

namespace {

constexpr double kCs2d = 1.0 / 3.0;
constexpr double kLambda = 3.0 / 16.0;

// geo carries the row_tiles grid.
__device__ __forceinline__ bool node_of(const Geom& geo, int& ix, int& iy) {
  int bx, by;
  tile_of_block(geo, bx, by);
  ix = bx * kTileX + (int)(threadIdx.x % kTileX);
  iy = by * kRowTileY + (int)(threadIdx.x / kTileX);
  return ix < geo.nx && iy < geo.ny;
}

// g of node (ix, iy) after streaming and the thermal BCs, from population set X.
template <int GEQ>
__device__ __forceinline__ void pulled_g(const Geom& geo, const double* X, const double* theta,
                                         int ix, int iy, double (&g)[9]) {
  const int nx = geo.nx, ny = geo.ny, pitch = geo.pitch;
  const long long P = geo.plane;
  if (ix == 0 || ix == nx - 1) {
    const int iyc = iy == 0 ? 1 : (iy == ny - 1 ? ny - 2 : iy);
    const double th = ix == 0 ? theta[iyc] : kTCold;
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] = kW[i] * th;
    return;
  }
  const int gy = iy == 0 ? 1 : (iy == ny - 1 ? ny - 2 : iy);
#pragma unroll
  for (int i = 0; i < 9; ++i)
    g[i] = X[(9 + i) * P + (long long)(gy - kCy[i]) * pitch + (ix - kCx[i])];
}

// Nusselt rows: T(1,iy), T(2,iy) of S = Mat(X); accumulates |T1 - T2| into
// acc[iy] and stores sign(T1 - T2) into sgn[iy] (either may be null).
template <int GEQ>
__global__ void k_nu_rows(Geom geo, const double* X, const double* theta, double* acc,
                          double* sgn) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  if (iy >= geo.ny) return;
  double g1[9], g2[9];
  pulled_g<GEQ>(geo, X, theta, 1, iy, g1);
  pulled_g<GEQ>(geo, X, theta, 2, iy, g2);
  double t1 = 0.0, t2 = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    t1 += g1[i];
    t2 += g2[i];
  }
  const double d = t1 - t2;
  if (acc) acc[iy] += fabs(d);
  if (sgn) sgn[iy] = d > 0.0 ? 1.0 : (d < 0.0 ? -1.0 : 0.0);
}

// lambda(j), the adjoint of the populations X(j) of a pulled state: the
// transpose of the pull with bounce-back and of the adiabatic row copies,
// applied to mu (the adjoints of the values each node pulled).
__device__ __forceinline__ void gather_lambda(const Geom& geo, const double* mu, int jx, int jy,
                                              double (&F)[9], double (&G)[9]) {
  const int nx = geo.nx, ny = geo.ny, pitch = geo.pitch;
  const long long P = geo.plane;
  const long long j = (long long)jy * pitch + jx;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int kx = jx + kCx[i], ky = jy + kCy[i];
    const bool in = kx >= 0 && kx < nx && ky >= 0 && ky < ny;
    // f: read by k = j + c_i (direction i), or, when that falls outside, by
    // j itself through the bounce-back of direction opp(i)
    F[i] = in ? mu[i * P + (long long)ky * pitch + kx] : mu[kOpp[i] * P + j];
    double s = 0.0;
    if (in && kx >= 1 && kx <= nx - 2 && ky >= 1 && ky <= ny - 2) {
      s = mu[(9 + i) * P + (long long)ky * pitch + kx];
      if (ky == 1) s += mu[(9 + i) * P + kx];
      if (ky == ny - 2) s += mu[(9 + i) * P + (long long)(ny - 1) * pitch + kx];
    }
    G[i] = s;
  }
}

template <int GEQ, bool FIRST>
__global__ __launch_bounds__(kRowBlock) void k_adj_node(Geom geo, StepCtx c, AdjArgs a) {
  int ix, iy;
  if (!node_of(geo, ix, iy)) return;
  const int nx = geo.nx, ny = geo.ny, pitch = geo.pitch;
  const long long P = geo.plane;
  const long long k = (long long)iy * pitch + ix;
  const bool wall = ix == 0 || ix == nx - 1 || iy == 0 || iy == ny - 1;
  const double cs2 = kCs2d;
  // divisions through the step's FP64 helpers: IEEE quotients of these
  // positive, normal divisors, with shorter sequences than the compiler's
  using A = Arith<double>;

  // ---- forward: pre-collision state of node k -----------------------------
  double f[9], g[9], rho, ux, uy, temp, fyp = 0.0, r = 0.0, mx = 0.0, my = 0.0;
  bool divided = false;
  if (FIRST) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      f[i] = a.mpops[i * P + k];
      g[i] = a.mpops[(9 + i) * P + k];
    }
    rho = a.mrho[k];
    ux = a.mux[k];
    uy = a.muy[k];
    temp = a.mtemp[k];
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const int sx = ix - kCx[i], sy = iy - kCy[i];
      const bool in = sx >= 0 && sx < nx && sy >= 0 && sy < ny;
      f[i] = in ? a.X[i * P + (long long)sy * pitch + sx] : a.X[kOpp[i] * P + k];
    }
    pulled_g<GEQ>(geo, a.X, a.theta, ix, iy, g);
    temp = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      r += f[i];
      mx += kCx[i] * f[i];
      my += kCy[i] * f[i];
      temp += g[i];
    }
    fyp = a.fyp[k];
    ux = mx;
    uy = my;
    if (r > kF64Eps) {
      ux = A::div(mx, r);
      uy = A::div(my, r) + A::div(0.5 * fyp, r);
      divided = true;
    }
    if (wall) {
      ux = 0.0;
      uy = 0.0;
    }
    rho = r;
  }

  // ---- forward: collision intermediates -----------------------------------
  const double fy = rho * c.beta * (temp - c.t_ref);
  const double usq = ux * ux + uy * uy;
  double eu[9], Q[9], d[9], e[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    eu[i] = kCx[i] * ux + kCy[i] * uy;
    Q[i] = 1.0 + 3.0 * eu[i] + 4.5 * eu[i] * eu[i] - 1.5 * usq;
    d[i] = f[i] - kW[i] * rho * Q[i];
    const double gq = GEQ == 1 ? 1.0 + 3.0 * eu[i] : Q[i];
    e[i] = g[i] - kW[i] * temp * gq;
  }
  double pxx = 0.0, pxy = 0.0, pyy = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    pxx += kCx[i] * kCx[i] * d[i];
    pxy += kCx[i] * kCy[i] * d[i];
    pyy += kCy[i] * kCy[i] * d[i];
  }
  const double p2 = pxx * pxx + 2.0 * pxy * pxy + pyy * pyy;
  const double sq = sqrt(p2);
  const double rho_m = fmax(rho, 1e-12);
  const double den = 2.0 * cs2 * c.tau * rho_m;
  const double smag = A::div(sq, den);
  const double nu_t = c.cs * c.cs * smag;
  const double tp = 0.5 + 3.0 * (c.nu + nu_t);
  const double tm = 0.5 + A::div(kLambda, tp - 0.5);
  const double tt = 0.5 + 3.0 * (c.alpha + A::div(nu_t, c.pr_t));
  const double itp = A::rcp(tp), itm = A::rcp(tm), itt = A::rcp(tt);
  const double af = 0.5 * (itp + itm), bf = 0.5 * (itp - itm), ag = itt;
  const double km = 1.0 - 0.5 * itm, kp = 1.0 - 0.5 * itp;
  double B[9], sm[9], sp[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    B[i] = A::divc(kCy[i] - uy, cs2, 3.0) + A::divc(eu[i] * kCy[i], cs2 * cs2, 9.0);
    const double s = kW[i] * B[i] * fy;
    sm[i] = kW[i] * (kCy[i] / cs2) * fy;
    sp[i] = s - sm[i];
  }

  // ---- reverse: collision ---------------------------------------------------
  double F[9], G[9], fyb = 0.0;
  if (a.mu_next) {
    gather_lambda(geo, a.mu_next, ix, iy, F, G);
    fyb = a.mu_next_fy[k];
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i] = G[i] = 0.0;
  }
  double fb[9], gb[9], db[9], eb[9];
  double ab = 0.0, bb = 0.0, kmb = 0.0, kpb = 0.0, agb = 0.0, uxb = 0.0, uyb = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    gb[i] = G[i];
    eb[i] = -ag * G[i];
    agb -= e[i] * G[i];
    fb[i] = F[i];
    db[i] = -af * F[i] - bf * F[kOpp[i]];
    ab -= d[i] * F[i];
    bb -= d[kOpp[i]] * F[i];
    kmb += sm[i] * F[i];
    kpb += sp[i] * F[i];
    const double sb = kp * F[i];          // adjoint of s_i
    const double smb = km * F[i] - sb;    // adjoint of sm_i (sp = s - sm)
    fyb += kW[i] * B[i] * sb + kW[i] * (kCy[i] / cs2) * smb;
    const double Bb = kW[i] * fy * sb;
    uyb += Bb * (-1.0 / cs2 + (double)(kCy[i] * kCy[i]) / (cs2 * cs2));
    uxb += Bb * ((double)(kCx[i] * kCy[i]) / (cs2 * cs2));
  }
  double tmb = 0.5 * kmb * itm * itm - 0.5 * itm * itm * (ab - bb);
  double tpb = 0.5 * kpb * itp * itp - 0.5 * itp * itp * (ab + bb);
  const double ttb = -agb * itt * itt;
  double nutb = A::div(3.0 * ttb, c.pr_t);
  tpb += A::div(-kLambda, (tp - 0.5) * (tp - 0.5)) * tmb;
  nutb += 3.0 * tpb;
  const double smagb = c.cs * c.cs * nutb;
  const double csb = 2.0 * c.cs * smag * nutb;
  double rhob = 0.0;
  if (p2 > 0.0) {
    const double p2b = A::div(smagb * 0.5, sq * den);
    const double pxxb = 2.0 * pxx * p2b, pxyb = 4.0 * pxy * p2b, pyyb = 2.0 * pyy * p2b;
#pragma unroll
    for (int i = 0; i < 9; ++i)
      db[i] += kCx[i] * kCx[i] * pxxb + kCx[i] * kCy[i] * pxyb + kCy[i] * kCy[i] * pyyb;
  }
  if (rho > 1e-12) rhob -= A::div(smag, rho_m) * smagb;
  double tempb = 0.0, usqb = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    fb[i] += db[i];
    gb[i] += eb[i];
    const double feqb = -db[i], geqb = -eb[i];
    rhob += kW[i] * Q[i] * feqb;
    double Qb = kW[i] * rho * feqb;
    double eub;
    if (GEQ == 1) {
      tempb += kW[i] * (1.0 + 3.0 * eu[i]) * geqb;
      eub = 3.0 * kW[i] * temp * geqb;
    } else {
      tempb += kW[i] * Q[i] * geqb;
      Qb += kW[i] * temp * geqb;
      eub = 0.0;
    }
    eub += Qb * (3.0 + 9.0 * eu[i]);
    usqb += -1.5 * Qb;
    uxb += kCx[i] * eub;
    uyb += kCy[i] * eub;
  }
  uxb += 2.0 * ux * usqb;
  uyb += 2.0 * uy * usqb;
  rhob += c.beta * (temp - c.t_ref) * fyb;
  tempb += rho * c.beta * fyb;

  // Cs contribution of this node
  a.cs_acc[k] += csb;
  if (FIRST) return;

  // ---- reverse: wall velocity, moments, objective ---------------------------
  if (wall) {
    uxb = 0.0;
    uyb = 0.0;
  }
  if (a.sgn && (ix == 1 || ix == 2)) {
    const double s = a.sgn[iy] * a.obj_w;
    tempb += ix == 1 ? s : -s;
  }
  double rb = rhob, mxb = uxb, myb = uyb, fypb = 0.0;
  if (divided) {
    mxb = A::div(uxb, r);
    myb = A::div(uyb, r);
    fypb = A::div(0.5 * uyb, r);
    rb += A::div(-(mx * uxb + (my + 0.5 * fyp) * uyb), r * r);
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    fb[i] += rb + kCx[i] * mxb + kCy[i] * myb;
    gb[i] += tempb;
  }

  // ---- outputs ----------------------------------------------------------------
#pragma unroll
  for (int i = 0; i < 9; ++i) a.mu[i * P + k] = fb[i];
  if (ix == 0 || ix == nx - 1) {
    // g set by the isothermal condition: g_i = w_i theta (hot) or constant
#pragma unroll
    for (int i = 0; i < 9; ++i) a.mu[(9 + i) * P + k] = 0.0;
    if (ix == 0) {
      double tb = 0.0;
#pragma unroll
      for (int i = 0; i < 9; ++i) tb += kW[i] * gb[i];
      a.theta_acc[iy] += tb;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) a.mu[(9 + i) * P + k] = gb[i];
  }
  a.mu_fy[k] = fypb;
}

__global__ void k_sum_nodes(Geom geo, const double* v, double* partial) {
  __shared__ double lds[1024 / 64];
  double acc = 0.0;
  const long long n = (long long)geo.pitch * geo.ny;
  for (long long q = threadIdx.x; q < n; q += blockDim.x) {
    const int x = (int)(q % geo.pitch);
    if (x < geo.nx) acc += v[q];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (threadIdx.x % 64 == 0) lds[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) r += lds[w];
    partial[0] = r;
  }
}

}  // namespace

hipError_t launch_nu_rows(const Geom& geo, int geq, const double* X, const double* theta,
                          double* acc, double* sgn, hipStream_t st) {
  const int tb = 256, nb = (geo.ny + tb - 1) / tb;
  if (geq == 1)
    hipLaunchKernelGGL(k_nu_rows<1>, dim3(nb), dim3(tb), 0, st, geo, X, theta, acc, sgn);
  else
    hipLaunchKernelGGL(k_nu_rows<2>, dim3(nb), dim3(tb), 0, st, geo, X, theta, acc, sgn);
  return hipGetLastError();
}

hipError_t launch_adj_step(const Geom& geo, int geq, bool first, const StepCtx& ctx,
                           const AdjArgs& a, hipStream_t st) {
  const Geom gr = row_tiles(geo);
  const dim3 grid(nblocks_of(gr)), block(kRowBlock);
  if (geq == 1) {
    if (first)
      hipLaunchKernelGGL((k_adj_node<1, true>), grid, block, 0, st, gr, ctx, a);
    else
      hipLaunchKernelGGL((k_adj_node<1, false>), grid, block, 0, st, gr, ctx, a);
  } else {
    if (first)
      hipLaunchKernelGGL((k_adj_node<2, true>), grid, block, 0, st, gr, ctx, a);
    else
      hipLaunchKernelGGL((k_adj_node<2, false>), grid, block, 0, st, gr, ctx, a);
  }
  return hipGetLastError();
}

hipError_t launch_sum_nodes(const Geom& geo, const double* v, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_nodes, dim3(1), dim3(1024), 0, st, geo, v, out);
  return hipGetLastError();
}

}  // namespace cavlbm

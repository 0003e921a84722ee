// lbm_aux.hip — initialisation, phase-split and diagnostics kernels of the
// cavity solver.  All arithmetic here is FP64 with -ffp-contract=off.
//
// References (bbrangeo/cfd_cavity): init f = feq / g = geq at
// src/uniform/solver.rs:171-174 and :252-255; stream :890-912; thermal BCs
// :951-988; update_macroscopic :319-345; enforce_wall_velocity :992-1009;
// apply_bounce_back :920-946; diagnostics src/uniform/post.rs:68-199, :406-411,
// src/uniform/solver.rs:1012-1020, src/main.rs:170-196.
#include "lbm_launch.hpp"

namespace cavlbm {

namespace {

__device__ __forceinline__ bool node_of_thread(const Geom& geo, int& ix, int& iy) {
  int bx, by;
  tile_of_block(geo, bx, by);
  ix = bx * kTileX + (int)(threadIdx.x % kTileX);
  iy = by * kTileY + (int)(threadIdx.x / kTileX);
  return ix < geo.nx && iy < geo.ny;
}

// f = feq(rho, u), g = geq(T, u) from the M-form fields; u = 0 on the walls
// afterwards (init_from_fields calls enforce_wall_velocity last).
template <int GEQ>
__global__ __launch_bounds__(kBlock) void k_init_eq(Geom geo, double* mpops,
                                                     const double* rho, double* ux,
                                                     double* uy, const double* temp) {
  int ix, iy;
  if (!node_of_thread(geo, ix, iy)) return;
  const long long k = (long long)iy * geo.pitch + ix;
  const long long P = geo.plane;
  const double r = rho[k], u = ux[k], v = uy[k], t = temp[k];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    mpops[(long long)i * P + k] = feq<double>(i, r, u, v);
    mpops[(long long)(9 + i) * P + k] = geq<double, GEQ>(i, t, u, v);
  }
  if (ix == 0 || ix == geo.nx - 1 || iy == 0 || iy == geo.ny - 1) {
    ux[k] = 0.0;
    uy[k] = 0.0;
  }
}

__global__ __launch_bounds__(kBlock) void k_stream(Geom geo, const double* src, double* dst) {
  int ix, iy;
  if (!node_of_thread(geo, ix, iy)) return;
  const long long k = (long long)iy * geo.pitch + ix;
  const long long P = geo.plane;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int sx = ix - kCx[i], sy = iy - kCy[i];
    if (sx >= 0 && sx < geo.nx && sy >= 0 && sy < geo.ny) {
      const long long s = (long long)sy * geo.pitch + sx;
      dst[(long long)i * P + k] = src[(long long)i * P + s];
      dst[(long long)(9 + i) * P + k] = src[(long long)(9 + i) * P + s];
    } else {
      dst[(long long)i * P + k] = src[(long long)kOpp[i] * P + k];
      dst[(long long)(9 + i) * P + k] = src[(long long)(9 + i) * P + k];
    }
  }
}

// Isothermal columns ix = 0 (theta_hot[iy]) and ix = nx-1 (T_COLD).
template <int GEQ>
__global__ void k_bc_columns(Geom geo, const double* theta_hot, double* mpops) {
  const int iy = blockIdx.x * blockDim.x + threadIdx.x;
  if (iy >= geo.ny) return;
  const long long P = geo.plane;
  const long long k0 = (long long)iy * geo.pitch, k1 = k0 + geo.nx - 1;
  const double th = theta_hot ? theta_hot[iy] : kTHot;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    mpops[(long long)(9 + i) * P + k0] = geq<double, GEQ>(i, th, 0.0, 0.0);
    mpops[(long long)(9 + i) * P + k1] = geq<double, GEQ>(i, kTCold, 0.0, 0.0);
  }
}

// Adiabatic rows: g(ix, 0) <- g(ix, 1), g(ix, ny-1) <- g(ix, ny-2).
__global__ void k_bc_rows(Geom geo, double* mpops) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  if (ix >= geo.nx) return;
  const long long P = geo.plane;
  const long long p = geo.pitch;
  const long long kb = ix, kt = (long long)(geo.ny - 1) * p + ix;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    double* gp = mpops + (long long)(9 + i) * P;
    gp[kb] = gp[kb + p];
    gp[kt] = gp[kt - p];
  }
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_macro(Geom geo, const double* mpops, const R* fy,
                                                   double* rho, double* ux, double* uy,
                                                   double* temp) {
  int ix, iy;
  if (!node_of_thread(geo, ix, iy)) return;
  const long long k = (long long)iy * geo.pitch + ix;
  const long long P = geo.plane;
  double r = 0.0, un = 0.0, vn = 0.0, tn = 0.0;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const double fj = mpops[(long long)j * P + k], gj = mpops[(long long)(9 + j) * P + k];
    r += fj;
    un += (double)kCx[j] * fj;
    vn += (double)kCy[j] * fj;
    tn += gj;
  }
  const double fyk = (double)fy[k];
  if (r > kF64Eps) {
    un /= r;
    vn /= r;
    vn += 0.5 * fyk / r;
  }
  rho[k] = r;
  ux[k] = un;
  uy[k] = vn;
  temp[k] = tn;
}

__global__ __launch_bounds__(kBlock) void k_walls(Geom geo, double* ux, double* uy) {
  int ix, iy;
  if (!node_of_thread(geo, ix, iy)) return;
  if (ix == 0 || ix == geo.nx - 1 || iy == 0 || iy == geo.ny - 1) {
    const long long k = (long long)iy * geo.pitch + ix;
    ux[k] = 0.0;
    uy[k] = 0.0;
  }
}

// apply_bounce_back swaps f_i <-> f_opp(i) (i = 1..4) once per wall the node
// lies on: corners are swapped twice, i.e. left unchanged.
__global__ __launch_bounds__(kBlock) void k_bounce_back(Geom geo, double* mpops) {
  int ix, iy;
  if (!node_of_thread(geo, ix, iy)) return;
  const int nswap = (ix == 0 || ix == geo.nx - 1 ? 1 : 0) + (iy == 0 || iy == geo.ny - 1 ? 1 : 0);
  if (nswap != 1) return;
  const long long k = (long long)iy * geo.pitch + ix;
  const long long P = geo.plane;
#pragma unroll
  for (int i = 1; i < 5; ++i) {
    const long long a = (long long)i * P + k, b = (long long)kOpp[i] * P + k;
    const double t = mpops[a];
    mpops[a] = mpops[b];
    mpops[b] = t;
  }
}

// ---------------------------------------------------------------------------
// Diagnostics: per-node contributions, workgroup reduction in FP64 with a
// fixed shuffle/LDS tree (deterministic), then one workgroup over partials.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}
__device__ __forceinline__ bool is_max_slot(int s) {
  return s == D_MAX_UY || s == D_MAX_U || s == D_NUT_MAX;
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_diag_partial(Geom geo, const double* mpops,
                                                          const double* rho, const double* ux,
                                                          const double* uy, const double* temp,
                                                          const R* nut, double nu,
                                                          double* partials) {
  __shared__ double lds[kDiagSlots][kBlock / 64];
  double v[kDiagSlots];
#pragma unroll
  for (int s = 0; s < kDiagSlots; ++s) v[s] = 0.0;
  int ix, iy;
  if (node_of_thread(geo, ix, iy)) {
    const int nx = geo.nx, ny = geo.ny;
    const long long p = geo.pitch;
    const long long k = (long long)iy * p + ix;
    const long long P = geo.plane;
    const double r = rho[k], u = ux[k], w = uy[k], t = temp[k];
    v[D_MASS] = r;
    double fm = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) fm += mpops[(long long)i * P + k];
    v[D_FMASS] = fm;
    if (ix == 1) v[D_NU_HOT] = fabs(t - temp[k + 1]);
    if (ix == nx - 2) v[D_NU_COLD] = fabs(t - temp[k - 1]);
    if (ix == 0) v[D_NU_LINK] = fabs(t - temp[k + 1]);
    if (ix >= 1 && ix <= nx - 2 && iy >= 1 && iy <= ny - 2) v[D_NU_VOL] = w * t;
    if (ix == 1 && iy >= 1 && iy <= ny - 2) {
      const double tau_w = (nu + (double)nut[k]) * fabs(w - uy[k - 1]);
      const double u_tau = sqrt(fmax(tau_w, 0.0));
      v[D_YPLUS] = 0.5 * u_tau / nu;
    }
    v[D_MAX_UY] = fabs(w);
    v[D_MAX_U] = fmax(fabs(u), fabs(w));
    v[D_NUT_MAX] = fabs((double)nut[k]);
    const bool ok = isfinite(t) && isfinite(u) && isfinite(w) && isfinite(r) && r > 0.0;
    v[D_NONFINITE] = ok ? 0.0 : 1.0;
  }
  const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
#pragma unroll
  for (int s = 0; s < kDiagSlots; ++s) {
    const double x = is_max_slot(s) ? wave_max(v[s]) : wave_sum(v[s]);
    if (lane == 0) lds[s][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x < kDiagSlots) {
    const int s = threadIdx.x;
    double acc = lds[s][0];
    for (int q = 1; q < kBlock / 64; ++q)
      acc = is_max_slot(s) ? fmax(acc, lds[s][q]) : acc + lds[s][q];
    partials[(long long)s * gridDim.x + blockIdx.x] = acc;
  }
}

__global__ __launch_bounds__(1024) void k_diag_final(int nparts, const double* partials,
                                                      double* out) {
  __shared__ double lds[16];
  const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
  for (int s = 0; s < kDiagSlots; ++s) {
    const bool mx = is_max_slot(s);
    double acc = 0.0;
    for (int q = threadIdx.x; q < nparts; q += blockDim.x) {
      const double x = partials[(long long)s * nparts + q];
      acc = mx ? fmax(acc, x) : acc + x;
    }
    acc = mx ? wave_max(acc) : wave_sum(acc);
    if (lane == 0) lds[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double r = lds[0];
      for (int q = 1; q < (int)(blockDim.x / 64); ++q) r = mx ? fmax(r, lds[q]) : r + lds[q];
      out[s] = r;
    }
    __syncthreads();
  }
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_convert(Geom geo, const R* src_r, double* dst_d,
                                                     int to_double, R* dst_r,
                                                     const double* src_d) {
  int ix, iy;
  if (!node_of_thread(geo, ix, iy)) return;
  const long long k = (long long)iy * geo.pitch + ix;
  if (to_double)
    dst_d[k] = (double)src_r[k];
  else
    dst_r[k] = (R)src_d[k];
}

// TimeAverager::accumulate (src/uniform/post.rs:37-49): per-node FP64 sums of
// ux, uy, T, nu_t and ux*ux + uy*uy (5 planes).
template <class R>
__global__ __launch_bounds__(kBlock) void k_avg_acc(Geom geo, const double* ux, const double* uy,
                                                     const double* temp, const R* nut,
                                                     double* sums) {
  int ix, iy;
  if (!node_of_thread(geo, ix, iy)) return;
  const long long k = (long long)iy * geo.pitch + ix;
  const long long P = geo.plane;
  const double u = ux[k], v = uy[k];
  sums[k] += u;
  sums[P + k] += v;
  sums[2 * P + k] += temp[k];
  sums[3 * P + k] += (double)nut[k];
  sums[4 * P + k] += u * u + v * v;
}

// p[0..n) = v: clears or sets a small device buffer with a kernel, which
// (unlike a memset node) stays correct inside captured graphs.
__global__ __launch_bounds__(256) void k_fill_f64(double* p, double v, unsigned n) {
  const unsigned i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

int diag_blocks(const Geom& geo) { return nblocks_of(geo); }

hipError_t launch_fill_f64(double* p, double v, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (n > 0xffffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v,
                     (unsigned)n);
  return hipGetLastError();
}

hipError_t launch_avg_accumulate(const Geom& geo, int storage, const double* ux, const double* uy,
                                 const double* temp, const void* nut, double* sums,
                                 hipStream_t st) {
  if (storage == ST_FP64)
    hipLaunchKernelGGL(k_avg_acc<double>, dim3(nblocks_of(geo)), dim3(kBlock), 0, st, geo, ux, uy,
                       temp, static_cast<const double*>(nut), sums);
  else
    hipLaunchKernelGGL(k_avg_acc<float>, dim3(nblocks_of(geo)), dim3(kBlock), 0, st, geo, ux, uy,
                       temp, static_cast<const float*>(nut), sums);
  return hipGetLastError();
}

hipError_t launch_init_eq(const Geom& geo, int geq, double* mpops, const double* rho,
                          double* ux, double* uy, const double* temp, hipStream_t st) {
  if (geq == 1)
    hipLaunchKernelGGL(k_init_eq<1>, dim3(nblocks_of(geo)), dim3(kBlock), 0, st, geo, mpops,
                       rho, ux, uy, temp);
  else
    hipLaunchKernelGGL(k_init_eq<2>, dim3(nblocks_of(geo)), dim3(kBlock), 0, st, geo, mpops,
                       rho, ux, uy, temp);
  return hipGetLastError();
}

hipError_t launch_phase_stream(const Geom& geo, const double* src, double* dst, hipStream_t st) {
  hipLaunchKernelGGL(k_stream, dim3(nblocks_of(geo)), dim3(kBlock), 0, st, geo, src, dst);
  return hipGetLastError();
}

hipError_t launch_phase_thermal_bc(const Geom& geo, int geq, const Walls& walls, double* mpops,
                                   hipStream_t st) {
  const int tb = 256;
  if (geq == 1)
    hipLaunchKernelGGL(k_bc_columns<1>, dim3((geo.ny + tb - 1) / tb), dim3(tb), 0, st, geo,
                       walls.theta_hot, mpops);
  else
    hipLaunchKernelGGL(k_bc_columns<2>, dim3((geo.ny + tb - 1) / tb), dim3(tb), 0, st, geo,
                       walls.theta_hot, mpops);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bc_rows, dim3((geo.nx + tb - 1) / tb), dim3(tb), 0, st, geo, mpops);
  return hipGetLastError();
}

hipError_t launch_phase_macro(const Geom& geo, int storage, const double* mpops, const void* fy,
                              double* rho, double* ux, double* uy, double* temp,
                              hipStream_t st) {
  if (storage == ST_FP64)
    hipLaunchKernelGGL(k_macro<double>, dim3(nblocks_of(geo)), dim3(kBlock), 0, st, geo, mpops,
                       static_cast<const double*>(fy), rho, ux, uy, temp);
  else
    hipLaunchKernelGGL(k_macro<float>, dim3(nblocks_of(geo)), dim3(kBlock), 0, st, geo, mpops,
                       static_cast<const float*>(fy), rho, ux, uy, temp);
  return hipGetLastError();
}

hipError_t launch_phase_walls(const Geom& geo, double* ux, double* uy, hipStream_t st) {
  hipLaunchKernelGGL(k_walls, dim3(nblocks_of(geo)), dim3(kBlock), 0, st, geo, ux, uy);
  return hipGetLastError();
}

hipError_t launch_phase_bounce_back(const Geom& geo, double* mpops, hipStream_t st) {
  hipLaunchKernelGGL(k_bounce_back, dim3(nblocks_of(geo)), dim3(kBlock), 0, st, geo, mpops);
  return hipGetLastError();
}

hipError_t launch_diagnostics(const Geom& geo, int storage, const double* mpops,
                              const double* rho, const double* ux, const double* uy,
                              const double* temp, const void* nut, double nu,
                              double* partials, double* out, hipStream_t st) {
  const int nb = nblocks_of(geo);
  if (storage == ST_FP64)
    hipLaunchKernelGGL(k_diag_partial<double>, dim3(nb), dim3(kBlock), 0, st, geo, mpops, rho,
                       ux, uy, temp, static_cast<const double*>(nut), nu, partials);
  else
    hipLaunchKernelGGL(k_diag_partial<float>, dim3(nb), dim3(kBlock), 0, st, geo, mpops, rho,
                       ux, uy, temp, static_cast<const float*>(nut), nu, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(1024), 0, st, nb, partials, out);
  return hipGetLastError();
}

hipError_t launch_field_convert(const Geom& geo, int storage, const void* src_r, double* dst_d,
                                int to_double, void* dst_r, const double* src_d,
                                hipStream_t st) {
  if (storage == ST_FP64)
    hipLaunchKernelGGL(k_convert<double>, dim3(nblocks_of(geo)), dim3(kBlock), 0, st, geo,
                       static_cast<const double*>(src_r), dst_d, to_double,
                       static_cast<double*>(dst_r), src_d);
  else
    hipLaunchKernelGGL(k_convert<float>, dim3(nblocks_of(geo)), dim3(kBlock), 0, st, geo,
                       static_cast<const float*>(src_r), dst_d, to_double,
                       static_cast<float*>(dst_r), src_d);
  return hipGetLastError();
}

}  // namespace cavlbm

// lbm_batch.hip — one launch advances a whole batch of independent cavities
// of different sizes and parameters by one step.  Each workgroup finds its
// cavity through a tile -> cavity table and runs the same per-node code as
// the single-cavity step kernel, so a batched cavity is bitwise equal to the
// same cavity stepped alone.  Built with the single-cavity kernels' flags
// per precision (FP64: no contraction; FP32 arithmetic: contraction inside
// each source expression), see build_lib.py.  The N-step form runs the
// single-cavity N-step tile code (lbm_stepn.hpp) on each cavity's tiles.
#include "lbm_launch.hpp"

namespace cavlbm {

namespace {

__device__ __forceinline__ int xcd_remap_block(int lin, int total) {
  const int per = total / 8, rem = total % 8;
  const int xcd = lin % 8, idx = lin / 8;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

template <int S, int GEQ, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_step_batch(const CavDesc* __restrict__ descs,
                                                         const int* __restrict__ tile_cav,
                                                         int total_tiles, int parity, int fpar,
                                                         int step, int write_aux, int xcd_remap) {
  const int lin = xcd_remap ? xcd_remap_block((int)blockIdx.x, total_tiles) : (int)blockIdx.x;
  const int c = tile_cav[lin];
  const CavDesc& d = descs[c];
  const int t = lin - d.tile_begin;
  const int bx = t % d.geo.nbx, by = t / d.geo.nbx;
  const int ix = bx * TX + (int)(threadIdx.x % TX);
  const int iy = by * TY + (int)(threadIdx.x / TX);
  if (ix >= d.geo.nx || iy >= d.geo.ny) return;
  StepCtx ctx = d.ctx;
  // beta ramp of collide() (solver.rs:433-437), same double arithmetic as
  // the host computes for a single cavity
  if (step < d.spin_up) ctx.beta = d.beta_eff * ((double)step / (double)d.spin_up);
  StepBufs b = d.bufs[parity];
  b.fy = d.fyp[fpar];
  step_node<S, GEQ, MODE_FUSED>(d.geo, ctx, d.walls, b, write_aux, ix, iy);
}

// N steps of a batch: the cavity's N-step tiles, the single-cavity tile code.
template <int S, int GEQ, int N>
__global__ __launch_bounds__((TileN<S, N>::kNT)) void k_stepn_batch(
    const CavDesc* __restrict__ descs, const int* __restrict__ tile_cav, int total_tiles,
    int parity, int fpar, int step, int write_aux, int xcd_remap) {
  using ST = StorageT<S>;
  using R = typename ST::R;
  using T = typename ST::T;
  constexpr int kTX = TileN<S, N>::kTX, kTY = TileN<S, N>::kTY, kH = N - 1;
  constexpr int kE1 = (kTX + 2 * kH) * (kTY + 2 * kH);
  __shared__ R sh[19 * kE1];
  const int lin = xcd_remap ? xcd_remap_block((int)blockIdx.x, total_tiles) : (int)blockIdx.x;
  const int c = tile_cav[lin];
  const CavDesc& d = descs[c];
  const int t = lin - d.tile_begin_n;
  const int bx = t % d.geon.nbx, by = t / d.geon.nbx;
  // the N collisions' scalars: beta ramp of collide() per step (solver.rs:433-437)
  StepCtxN cn;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    cn.c[j] = d.ctx;
    const int sj = step + j;
    if (sj < d.spin_up) cn.c[j].beta = d.beta_eff * ((double)sj / (double)d.spin_up);
  }
  const StepBufs& b = d.bufs[parity];
  const T* src = static_cast<const T*>(b.src);
  T* dst = static_cast<T*>(b.dst);
  const R* fy_in = static_cast<const R*>(d.fyp[fpar]);
  R* fy_out = static_cast<R*>(d.fyp[fpar ^ 1]);
  R* nut = static_cast<R*>(b.nut);
  const bool interior = bx * kTX - kH >= 1 && (bx + 1) * kTX + kH - 1 <= d.geon.nx - 2 &&
                        by * kTY - kH >= 1 && (by + 1) * kTY + kH - 1 <= d.geon.ny - 2;
  // without the sub-step priority hint (stepn_prio): the batch's tiles
  // measured 0.098 against 0.106 ms per batch step with it
  if (interior)
    stepn_tile<S, GEQ, N, true, false, false>(d.geon, cn, d.walls, src, dst, fy_in, fy_out, nut,
                                              write_aux, sh, bx, by);
  else
    stepn_tile<S, GEQ, N, false, false, false>(d.geon, cn, d.walls, src, dst, fy_in, fy_out, nut,
                                               write_aux, sh, bx, by);
}

template <int S, int GEQ, int N>
hipError_t gon(const CavDesc* descs, const int* tile_cav, int total, int parity, int fpar,
               int step, int write_aux, int xcd, hipStream_t st) {
  hipLaunchKernelGGL((k_stepn_batch<S, GEQ, N>), dim3(total), dim3(TileN<S, N>::kNT), 0, st,
                     descs, tile_cav, total, parity, fpar, step, write_aux, xcd);
  return hipGetLastError();
}

template <int S, int GEQ>
hipError_t by_n(int n, const CavDesc* d, const int* tc, int total, int par, int fpar, int step,
                int wa, int xcd, hipStream_t st) {
  switch (n) {
    case 2: return gon<S, GEQ, 2>(d, tc, total, par, fpar, step, wa, xcd, st);
    case 3: return gon<S, GEQ, 3>(d, tc, total, par, fpar, step, wa, xcd, st);
    case 4: return gon<S, GEQ, 4>(d, tc, total, par, fpar, step, wa, xcd, st);
    default: return hipErrorInvalidValue;
  }
}

template <int S, int GEQ, int SH>
hipError_t go(const CavDesc* descs, const int* tile_cav, int total, int parity, int fpar, int step,
              int write_aux, int xcd, hipStream_t st) {
  constexpr int TX = kBatchShapes[SH].tx, TY = kBatchShapes[SH].ty;
  hipLaunchKernelGGL((k_step_batch<S, GEQ, TX, TY>), dim3(total), dim3(TX * TY), 0, st, descs,
                     tile_cav, total, parity, fpar, step, write_aux, xcd);
  return hipGetLastError();
}

template <int S, int GEQ>
hipError_t by_shape(int shape, const CavDesc* d, const int* tc, int total, int par, int fpar,
                    int step, int wa, int xcd, hipStream_t st) {
  switch (shape) {
    case 0: return go<S, GEQ, 0>(d, tc, total, par, fpar, step, wa, xcd, st);
    case 1: return go<S, GEQ, 1>(d, tc, total, par, fpar, step, wa, xcd, st);
    case 2: return go<S, GEQ, 2>(d, tc, total, par, fpar, step, wa, xcd, st);
    case 3: return go<S, GEQ, 3>(d, tc, total, par, fpar, step, wa, xcd, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

#if CAVLBM_BATCH_PRECISION == 64
hipError_t launch_step_batch_fp64(int geq, int shape, const CavDesc* d, const int* tc, int total,
                                  int par, int fpar, int step, int wa, int xcd, hipStream_t st) {
  return geq == 1 ? by_shape<ST_FP64, 1>(shape, d, tc, total, par, fpar, step, wa, xcd, st)
                  : by_shape<ST_FP64, 2>(shape, d, tc, total, par, fpar, step, wa, xcd, st);
}
hipError_t launch_stepn_batch_fp64(int geq, int n, const CavDesc* d, const int* tc, int total,
                                   int par, int fpar, int step, int wa, int xcd, hipStream_t st) {
  return geq == 1 ? by_n<ST_FP64, 1>(n, d, tc, total, par, fpar, step, wa, xcd, st)
                  : by_n<ST_FP64, 2>(n, d, tc, total, par, fpar, step, wa, xcd, st);
}
#else
hipError_t launch_step_batch_lowp(int storage, int geq, int shape, const CavDesc* d,
                                  const int* tc, int total, int par, int fpar, int step, int wa,
                                  int xcd, hipStream_t st) {
  if (storage == ST_FP32)
    return geq == 1 ? by_shape<ST_FP32, 1>(shape, d, tc, total, par, fpar, step, wa, xcd, st)
                    : by_shape<ST_FP32, 2>(shape, d, tc, total, par, fpar, step, wa, xcd, st);
  return geq == 1
             ? by_shape<ST_FP16_OFFSET, 1>(shape, d, tc, total, par, fpar, step, wa, xcd, st)
             : by_shape<ST_FP16_OFFSET, 2>(shape, d, tc, total, par, fpar, step, wa, xcd, st);
}
hipError_t launch_stepn_batch_lowp(int storage, int geq, int n, const CavDesc* d, const int* tc,
                                   int total, int par, int fpar, int step, int wa, int xcd,
                                   hipStream_t st) {
  if (storage == ST_FP32)
    return geq == 1 ? by_n<ST_FP32, 1>(n, d, tc, total, par, fpar, step, wa, xcd, st)
                    : by_n<ST_FP32, 2>(n, d, tc, total, par, fpar, step, wa, xcd, st);
  if (storage == ST_FP16_OFFSET)
    return geq == 1 ? by_n<ST_FP16_OFFSET, 1>(n, d, tc, total, par, fpar, step, wa, xcd, st)
                    : by_n<ST_FP16_OFFSET, 2>(n, d, tc, total, par, fpar, step, wa, xcd, st);
  return hipErrorInvalidValue;
}
#endif

}  // namespace cavlbm

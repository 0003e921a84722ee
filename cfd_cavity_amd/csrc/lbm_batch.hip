// lbm_batch.hip — one launch advances a whole batch of independent cavities
// of different sizes and parameters by one step.  Each workgroup finds its
// cavity through a tile -> cavity table and runs the same per-node code as
// the single-cavity step kernel, so a batched cavity is bitwise equal to the
// same cavity stepped alone.  Built with the single-cavity kernels' flags
// per precision (FP64: no contraction; FP32 arithmetic: contraction allowed),
// see build_lib.py.
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
                                                         int total_tiles, int parity, int step,
                                                         int write_aux, int xcd_remap) {
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
  step_node<S, GEQ, MODE_FUSED>(d.geo, ctx, d.walls, d.bufs[parity], write_aux, ix, iy);
}

template <int S, int GEQ, int SH>
hipError_t go(const CavDesc* descs, const int* tile_cav, int total, int parity, int step,
              int write_aux, int xcd, hipStream_t st) {
  constexpr int TX = kBatchShapes[SH].tx, TY = kBatchShapes[SH].ty;
  hipLaunchKernelGGL((k_step_batch<S, GEQ, TX, TY>), dim3(total), dim3(TX * TY), 0, st, descs,
                     tile_cav, total, parity, step, write_aux, xcd);
  return hipGetLastError();
}

template <int S, int GEQ>
hipError_t by_shape(int shape, const CavDesc* d, const int* tc, int total, int par, int step,
                    int wa, int xcd, hipStream_t st) {
  switch (shape) {
    case 0: return go<S, GEQ, 0>(d, tc, total, par, step, wa, xcd, st);
    case 1: return go<S, GEQ, 1>(d, tc, total, par, step, wa, xcd, st);
    case 2: return go<S, GEQ, 2>(d, tc, total, par, step, wa, xcd, st);
    case 3: return go<S, GEQ, 3>(d, tc, total, par, step, wa, xcd, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

#if CAVLBM_BATCH_PRECISION == 64
hipError_t launch_step_batch_fp64(int geq, int shape, const CavDesc* d, const int* tc, int total,
                                  int par, int step, int wa, int xcd, hipStream_t st) {
  return geq == 1 ? by_shape<ST_FP64, 1>(shape, d, tc, total, par, step, wa, xcd, st)
                  : by_shape<ST_FP64, 2>(shape, d, tc, total, par, step, wa, xcd, st);
}
#else
hipError_t launch_step_batch_lowp(int storage, int geq, int shape, const CavDesc* d,
                                  const int* tc, int total, int par, int step, int wa, int xcd,
                                  hipStream_t st) {
  if (storage == ST_FP32)
    return geq == 1 ? by_shape<ST_FP32, 1>(shape, d, tc, total, par, step, wa, xcd, st)
                    : by_shape<ST_FP32, 2>(shape, d, tc, total, par, step, wa, xcd, st);
  return geq == 1 ? by_shape<ST_FP16_OFFSET, 1>(shape, d, tc, total, par, step, wa, xcd, st)
                  : by_shape<ST_FP16_OFFSET, 2>(shape, d, tc, total, par, step, wa, xcd, st);
}
#endif

}  // namespace cavlbm

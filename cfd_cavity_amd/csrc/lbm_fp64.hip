// lbm_fp64.hip — FP64 instantiations of the step kernel.  Built with
// -ffp-contract=off: no FMA contraction, IEEE division and square root, so the
// populations are bitwise equal to a CPU evaluation in the reference's order.
#include "lbm_launch.hpp"

namespace cavlbm {

template <int GEQ, int MODE>
static hipError_t go(const Geom& geo, const StepCtx& ctx, const Walls& walls,
                     const StepBufs& b, int write_aux, hipStream_t st) {
  const Geom g = MODE == MODE_WINDOW ? row_tiles(geo) : geo;
  hipLaunchKernelGGL((k_step<ST_FP64, GEQ, MODE>), dim3(nblocks_of(g)),
                     dim3(MODE == MODE_WINDOW ? kRowBlock : kBlock), 0, st, g, ctx, walls, b,
                     write_aux);
  return hipGetLastError();
}

hipError_t launch_step_fp64(int geq, int mode, const Geom& geo, const StepCtx& ctx,
                            const Walls& walls, const StepBufs& b, int write_aux,
                            hipStream_t st) {
#define CAVLBM_CASE(G, M) \
  if (geq == G && mode == M) return go<G, M>(geo, ctx, walls, b, write_aux, st);
  CAVLBM_CASE(2, MODE_FUSED)
  CAVLBM_CASE(2, MODE_FIRST)
  CAVLBM_CASE(2, MODE_MATERIALIZE)
  CAVLBM_CASE(2, MODE_COLLIDE_M)
  CAVLBM_CASE(2, MODE_FIRST_PRE)
  CAVLBM_CASE(2, MODE_MATERIALIZE_F)
  CAVLBM_CASE(2, MODE_FUSED_DYN)
  CAVLBM_CASE(2, MODE_FUSED_PRE)
  CAVLBM_CASE(2, MODE_WINDOW)
  CAVLBM_CASE(1, MODE_FUSED)
  CAVLBM_CASE(1, MODE_FIRST)
  CAVLBM_CASE(1, MODE_MATERIALIZE)
  CAVLBM_CASE(1, MODE_COLLIDE_M)
  CAVLBM_CASE(1, MODE_FIRST_PRE)
  CAVLBM_CASE(1, MODE_MATERIALIZE_F)
  CAVLBM_CASE(1, MODE_FUSED_DYN)
  CAVLBM_CASE(1, MODE_FUSED_PRE)
  CAVLBM_CASE(1, MODE_WINDOW)
#undef CAVLBM_CASE
  return hipErrorInvalidValue;
}

bool stepn_tile_dims(int storage, int n, int* tx, int* ty) {
#define CAVLBM_TILE(S, N)        \
  if (storage == S && n == N) {  \
    *tx = TileN<S, N>::kTX;      \
    *ty = TileN<S, N>::kTY;      \
    return true;                 \
  }
  CAVLBM_TILE(ST_FP64, 2)
  CAVLBM_TILE(ST_FP64, 3)
  CAVLBM_TILE(ST_FP64, 4)
  CAVLBM_TILE(ST_FP32, 2)
  CAVLBM_TILE(ST_FP32, 3)
  CAVLBM_TILE(ST_FP32, 4)
  CAVLBM_TILE(ST_FP16_OFFSET, 2)
  CAVLBM_TILE(ST_FP16_OFFSET, 3)
  CAVLBM_TILE(ST_FP16_OFFSET, 4)
#undef CAVLBM_TILE
  *tx = *ty = 0;
  return false;
}

bool stepn_grid(int storage, int n, Geom& geo_n) {
  int tx = 0, ty = 0;
  if (!stepn_tile_dims(storage, n, &tx, &ty)) return false;
  geo_n.nbx = (geo_n.nx + tx - 1) / tx;
  geo_n.nby = (geo_n.ny + ty - 1) / ty;
  return geo_n.nx >= 3 && geo_n.ny >= 3 && (geo_n.ny - 1) % ty != 0;
}

template <int GEQ, int N>
static hipError_t gon(const Geom& geo_n, const StepCtxN& cn, const Walls& walls, const void* src,
                      void* dst, const void* fy_in, void* fy_out, void* nut, int write_aux,
                      hipStream_t st) {
  if (!stepn_fits<ST_FP64, N>(geo_n.nx, geo_n.ny)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_stepn<ST_FP64, GEQ, N>), dim3(geo_n.nbx * geo_n.nby),
                     dim3(TileN<ST_FP64, N>::kNT), 0, st, geo_n, cn, walls,
                     static_cast<const double*>(src), static_cast<double*>(dst),
                     static_cast<const double*>(fy_in), static_cast<double*>(fy_out),
                     static_cast<double*>(nut), write_aux);
  return hipGetLastError();
}

hipError_t launch_stepn_fp64(int geq, int n, const Geom& geo_n, const StepCtxN& cn,
                             const Walls& walls, const void* src, void* dst, const void* fy_in,
                             void* fy_out, void* nut, int write_aux, hipStream_t st) {
#define CAVLBM_CASE(G, N) \
  if (geq == G && n == N) return gon<G, N>(geo_n, cn, walls, src, dst, fy_in, fy_out, nut, write_aux, st);
  CAVLBM_CASE(2, 2)
  CAVLBM_CASE(2, 3)
  CAVLBM_CASE(2, 4)
  CAVLBM_CASE(1, 2)
  CAVLBM_CASE(1, 3)
  CAVLBM_CASE(1, 4)
#undef CAVLBM_CASE
  return hipErrorInvalidValue;
}

template <int GEQ, int N>
static hipError_t gow(const Geom& geo_n, const StepCtxN& cn, const Walls& walls, const void* src,
                      void* dst, const void* fy_in, void* fy_out, double* nu_acc,
                      hipStream_t st) {
  if (!stepn_fits<ST_FP64, N>(geo_n.nx, geo_n.ny)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_stepn_window<GEQ, N>), dim3(geo_n.nbx * geo_n.nby),
                     dim3(TileN<ST_FP64, N>::kNT), 0, st, geo_n, cn, walls,
                     static_cast<const double*>(src), static_cast<double*>(dst),
                     static_cast<const double*>(fy_in), static_cast<double*>(fy_out), nu_acc);
  return hipGetLastError();
}

hipError_t launch_stepn_window_fp64(int geq, int n, const Geom& geo_n, const StepCtxN& cn,
                                    const Walls& walls, const void* src, void* dst,
                                    const void* fy_in, void* fy_out, double* nu_acc,
                                    hipStream_t st) {
#define CAVLBM_CASE(G, N) \
  if (geq == G && n == N) return gow<G, N>(geo_n, cn, walls, src, dst, fy_in, fy_out, nu_acc, st);
  CAVLBM_CASE(2, 2)
  CAVLBM_CASE(2, 3)
  CAVLBM_CASE(2, 4)
  CAVLBM_CASE(1, 2)
  CAVLBM_CASE(1, 3)
  CAVLBM_CASE(1, 4)
#undef CAVLBM_CASE
  return hipErrorInvalidValue;
}

}  // namespace cavlbm

// lbm_fp64.hip — FP64 instantiations of the step kernel.  Built with
// -ffp-contract=off: no FMA contraction, IEEE division and square root, so the
// populations are bitwise equal to a CPU evaluation in the reference's order.
#include "lbm_launch.hpp"
#include "lbm_step2.hpp"

namespace cavlbm {

template <int GEQ, int MODE>
static hipError_t go(const Geom& geo, const StepCtx& ctx, const Walls& walls,
                     const StepBufs& b, int write_aux, hipStream_t st) {
  hipLaunchKernelGGL((k_step<ST_FP64, GEQ, MODE>), dim3(nblocks_of(geo)), dim3(kBlock), 0,
                     st, geo, ctx, walls, b, write_aux);
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
  CAVLBM_CASE(1, MODE_FUSED)
  CAVLBM_CASE(1, MODE_FIRST)
  CAVLBM_CASE(1, MODE_MATERIALIZE)
  CAVLBM_CASE(1, MODE_COLLIDE_M)
  CAVLBM_CASE(1, MODE_FIRST_PRE)
#undef CAVLBM_CASE
  return hipErrorInvalidValue;
}

template <int GEQ>
static hipError_t go2(const Geom& geo2, const StepCtx& c1, const StepCtx& c2, const Walls& walls,
                      const void* src, void* dst, const void* fy_in, void* fy_out, void* nut,
                      int write_aux, hipStream_t st) {
  hipLaunchKernelGGL((k_step2<ST_FP64, GEQ>), dim3(geo2.nbx * geo2.nby), dim3(Tile2<ST_FP64>::kNT), 0, st,
                     geo2, c1, c2, walls, static_cast<const double*>(src),
                     static_cast<double*>(dst), static_cast<const double*>(fy_in),
                     static_cast<double*>(fy_out), static_cast<double*>(nut), write_aux);
  return hipGetLastError();
}

bool step2_grid(int storage, Geom& geo2) {
  bool ok = false;
  int tx = 1, ty = 1;
  if (storage == ST_FP64) {
    ok = step2_fits<ST_FP64>(geo2.nx, geo2.ny);
    tx = Tile2<ST_FP64>::kTX;
    ty = Tile2<ST_FP64>::kTY;
  } else if (storage == ST_FP32) {
    ok = step2_fits<ST_FP32>(geo2.nx, geo2.ny);
    tx = Tile2<ST_FP32>::kTX;
    ty = Tile2<ST_FP32>::kTY;
  }
  geo2.nbx = (geo2.nx + tx - 1) / tx;
  geo2.nby = (geo2.ny + ty - 1) / ty;
  return ok;
}

hipError_t launch_step2_fp64(int geq, const Geom& geo2, const StepCtx& c1, const StepCtx& c2,
                             const Walls& walls, const void* src, void* dst, const void* fy_in,
                             void* fy_out, void* nut, int write_aux, hipStream_t st) {
  if (!step2_fits<ST_FP64>(geo2.nx, geo2.ny)) return hipErrorInvalidValue;
  if (geq == 2) return go2<2>(geo2, c1, c2, walls, src, dst, fy_in, fy_out, nut, write_aux, st);
  if (geq == 1) return go2<1>(geo2, c1, c2, walls, src, dst, fy_in, fy_out, nut, write_aux, st);
  return hipErrorInvalidValue;
}

}  // namespace cavlbm

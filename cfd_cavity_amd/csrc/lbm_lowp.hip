// lbm_lowp.hip — FP32-arithmetic instantiations of the step kernel for FP32
// and FP16-offset population storage.  Built with FMA contraction inside each
// source expression (-ffp-contract=on), so the FMA formation follows the source
// and the single-step and two-step kernels round alike; the precision forms
// are checked against FP64 with stated tolerances.
#include "lbm_launch.hpp"
#include "lbm_step2.hpp"

namespace cavlbm {

template <int S, int GEQ, int MODE>
static hipError_t go(const Geom& geo, const StepCtx& ctx, const Walls& walls,
                     const StepBufs& b, int write_aux, hipStream_t st) {
  hipLaunchKernelGGL((k_step<S, GEQ, MODE>), dim3(nblocks_of(geo)), dim3(kBlock), 0, st,
                     geo, ctx, walls, b, write_aux);
  return hipGetLastError();
}

template <int S>
static hipError_t dispatch(int geq, int mode, const Geom& geo, const StepCtx& ctx,
                           const Walls& walls, const StepBufs& b, int write_aux,
                           hipStream_t st) {
#define CAVLBM_CASE(G, M) \
  if (geq == G && mode == M) return go<S, G, M>(geo, ctx, walls, b, write_aux, st);
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

hipError_t launch_step_lowp(int storage, int geq, int mode, const Geom& geo,
                            const StepCtx& ctx, const Walls& walls, const StepBufs& b,
                            int write_aux, hipStream_t st) {
  if (storage == ST_FP32) return dispatch<ST_FP32>(geq, mode, geo, ctx, walls, b, write_aux, st);
  if (storage == ST_FP16_OFFSET)
    return dispatch<ST_FP16_OFFSET>(geq, mode, geo, ctx, walls, b, write_aux, st);
  return hipErrorInvalidValue;
}

template <int GEQ>
static hipError_t go2(const Geom& geo2, const StepCtx& c1, const StepCtx& c2, const Walls& walls,
                      const void* src, void* dst, const void* fy_in, void* fy_out, void* nut,
                      int write_aux, hipStream_t st) {
  hipLaunchKernelGGL((k_step2<ST_FP32, GEQ>), dim3(geo2.nbx * geo2.nby), dim3(Tile2<ST_FP32>::kNT), 0, st,
                     geo2, c1, c2, walls, static_cast<const float*>(src),
                     static_cast<float*>(dst), static_cast<const float*>(fy_in),
                     static_cast<float*>(fy_out), static_cast<float*>(nut), write_aux);
  return hipGetLastError();
}

// Two fused steps per launch for FP32 storage (FP16-offset storage rounds the
// populations between steps and keeps the single-step kernel).
hipError_t launch_step2_lowp(int storage, int geq, const Geom& geo2, const StepCtx& c1,
                             const StepCtx& c2, const Walls& walls, const void* src, void* dst,
                             const void* fy_in, void* fy_out, void* nut, int write_aux,
                             hipStream_t st) {
  if (storage != ST_FP32 || !step2_fits<ST_FP32>(geo2.nx, geo2.ny)) return hipErrorInvalidValue;
  if (geq == 2) return go2<2>(geo2, c1, c2, walls, src, dst, fy_in, fy_out, nut, write_aux, st);
  if (geq == 1) return go2<1>(geo2, c1, c2, walls, src, dst, fy_in, fy_out, nut, write_aux, st);
  return hipErrorInvalidValue;
}

}  // namespace cavlbm

// lbm_lowp.hip — FP32-arithmetic instantiations of the step kernel for FP32
// and FP16-offset population storage.  Built with FMA contraction allowed; the
// precision forms are checked against FP64 with stated tolerances.
#include "lbm_launch.hpp"

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

}  // namespace cavlbm

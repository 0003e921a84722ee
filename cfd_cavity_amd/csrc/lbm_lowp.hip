// lbm_lowp.hip — FP32-arithmetic instantiations of the step kernel for FP32
// and FP16-offset population storage.  Built with FMA contraction inside each
// source expression (-ffp-contract=on), so the FMA formation follows the source
// and the single-step and two-step kernels round alike; the precision forms
// are checked against FP64 with stated tolerances.
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
  CAVLBM_CASE(2, MODE_MATERIALIZE_F)
  CAVLBM_CASE(2, MODE_FUSED_DYN)
  CAVLBM_CASE(2, MODE_FUSED_PRE)
  CAVLBM_CASE(1, MODE_FUSED)
  CAVLBM_CASE(1, MODE_FIRST)
  CAVLBM_CASE(1, MODE_MATERIALIZE)
  CAVLBM_CASE(1, MODE_COLLIDE_M)
  CAVLBM_CASE(1, MODE_FIRST_PRE)
  CAVLBM_CASE(1, MODE_MATERIALIZE_F)
  CAVLBM_CASE(1, MODE_FUSED_DYN)
  CAVLBM_CASE(1, MODE_FUSED_PRE)
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

template <int S, int GEQ, int N>
static hipError_t gon(const Geom& geo_n, const StepCtxN& cn, const Walls& walls, const void* src,
                      void* dst, const void* fy_in, void* fy_out, void* nut, int write_aux,
                      hipStream_t st) {
  using T = typename StorageT<S>::T;
  using R = typename StorageT<S>::R;
  if (!stepn_fits<S, N>(geo_n.nx, geo_n.ny)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_stepn<S, GEQ, N>), dim3(geo_n.nbx * geo_n.nby), dim3(TileN<S, N>::kNT), 0,
                     st, geo_n, cn, walls, static_cast<const T*>(src), static_cast<T*>(dst),
                     static_cast<const R*>(fy_in), static_cast<R*>(fy_out), static_cast<R*>(nut),
                     write_aux);
  return hipGetLastError();
}

template <int S>
static hipError_t dispatch_n(int geq, int n, const Geom& geo_n, const StepCtxN& cn,
                             const Walls& walls, const void* src, void* dst, const void* fy_in,
                             void* fy_out, void* nut, int write_aux, hipStream_t st) {
#define CAVLBM_CASE(G, N)                                                                  \
  if (geq == G && n == N)                                                                  \
    return gon<S, G, N>(geo_n, cn, walls, src, dst, fy_in, fy_out, nut, write_aux, st);
  CAVLBM_CASE(2, 2)
  CAVLBM_CASE(2, 3)
  CAVLBM_CASE(2, 4)
  CAVLBM_CASE(1, 2)
  CAVLBM_CASE(1, 3)
  CAVLBM_CASE(1, 4)
#undef CAVLBM_CASE
  return hipErrorInvalidValue;
}

// Several fused steps per launch for FP32 and FP16-offset storage (the
// latter rounds the populations through its store and load between the
// sub-steps, as single steps do).
hipError_t launch_stepn_lowp(int storage, int geq, int n, const Geom& geo_n, const StepCtxN& cn,
                             const Walls& walls, const void* src, void* dst, const void* fy_in,
                             void* fy_out, void* nut, int write_aux, hipStream_t st) {
  if (storage == ST_FP32)
    return dispatch_n<ST_FP32>(geq, n, geo_n, cn, walls, src, dst, fy_in, fy_out, nut, write_aux,
                               st);
  if (storage == ST_FP16_OFFSET)
    return dispatch_n<ST_FP16_OFFSET>(geq, n, geo_n, cn, walls, src, dst, fy_in, fy_out, nut,
                                      write_aux, st);
  return hipErrorInvalidValue;
}

}  // namespace cavlbm

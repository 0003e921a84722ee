// lbm_fp64.hip — FP64 instantiations of the step kernel.  Built with
// -ffp-contract=off: no FMA contraction, IEEE division and square root, so the
// populations are bitwise equal to a CPU evaluation in the reference's order.
#include "lbm_launch.hpp"

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

}  // namespace cavlbm

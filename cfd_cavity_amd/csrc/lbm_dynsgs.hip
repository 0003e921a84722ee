// lbm_dynsgs.hip — dynamic mixed-scale subgrid model on the device
// (reference: src/uniform/solver.rs:562-828 compute_dynamic_sgs, helpers
// box_filter :1204-1224, ddx/ddy :1229-1242, lilly_coefficient :1251-1261).
//
// Three passes over the M-form state (post-stream populations and moments),
// each one thread per node, FP64, built with -ffp-contract=off so that every
// sum, product, division and square root is the IEEE result in the
// reference's operand order:
//   k_dyn_filter : test-filter (box of `width` nodes, clamped one-sided at the
//                  walls) of u, v, T and of the products uu, uv, vv, uT, vT;
//                  the node-local similarity scales qsm^1/2 and Phi_sm^1/2;
//   k_dyn_terms  : grid-level S~ from pi_neq, test-level S- from centred
//                  differences of the filtered fields, temperature gradients,
//                  |T~| / |T-|, the interpolated mixed forms at both levels,
//                  the Germano M and Leonard L tensors and the four Lilly
//                  contractions;
//   k_dyn_coeffs : box of the contractions, C = clamp(num/(2 den), 0, 0.04),
//                  nu_t = C_nu * nu_mix, alpha_t = C_kappa * kappa_mix, both
//                  floored (dyn_coef_node; the solver's dynamic step runs it
//                  inside the collision instead, MODE_FUSED_DYN).
// The box sums run in the reference's (y outer, x inner) order from +0.
//
// The mixed forms a^alpha b^(1-alpha) use pow; the reference's powf is the C
// library's pow, which is not correctly rounded, so this path agrees with a
// CPU evaluation to a few ulp per node rather than bitwise.  For alpha = 0.5
// (the default) the factors are correctly rounded square roots.
#include "lbm_launch.hpp"

namespace cavlbm {

namespace {

__device__ __forceinline__ int node_of_thread(const Geom& geo, int& ix, int& iy) {
  const int bx = (int)blockIdx.x % geo.nbx, by = (int)blockIdx.x / geo.nbx;
  ix = bx * kTileX + (int)(threadIdx.x % kTileX);
  iy = by * kTileY + (int)(threadIdx.x / kTileX);
  return ix < geo.nx && iy < geo.ny;
}

__device__ __forceinline__ DynWin window(const Geom& geo, const DynSgsArgs& a, int ix, int iy) {
  return dyn_window(geo, a.width, a.half_left, ix, iy);
}

// Centred differences with one-sided clamping at the walls (solver.rs:1229-1242).
__device__ __forceinline__ double ddx(const Geom& geo, const double* f, int ix, int iy) {
  const int xm = ix == 0 ? 0 : ix - 1;
  const int xp = ix + 1 >= geo.nx ? geo.nx - 1 : ix + 1;
  return 0.5 * (f[iy * geo.pitch + xp] - f[iy * geo.pitch + xm]);
}

__device__ __forceinline__ double ddy(const Geom& geo, const double* f, int ix, int iy) {
  const int ym = iy == 0 ? 0 : iy - 1;
  const int yp = iy + 1 >= geo.ny ? geo.ny - 1 : iy + 1;
  return 0.5 * (f[yp * geo.pitch + ix] - f[ym * geo.pitch + ix]);
}

// x^e for the mixed forms; every base here is >= +0.
__device__ __forceinline__ double mixpow(double x, double e) {
  if (e == 0.5) return sqrt(x == 0.0 ? 0.0 : x);
  return pow(x, e);
}

__global__ __launch_bounds__(kBlock) void k_dyn_filter(Geom geo, DynSgsArgs a) {
  int ix, iy;
  if (!node_of_thread(geo, ix, iy)) return;
  const int k = iy * geo.pitch + ix;
  const DynWin w = window(geo, a, ix, iy);
  double s_u = 0.0, s_v = 0.0, s_t = 0.0, s_uu = 0.0, s_uv = 0.0, s_vv = 0.0, s_ut = 0.0,
         s_vt = 0.0;
  for (int yy = w.y0; yy < w.y1; ++yy)
    for (int xx = w.x0; xx < w.x1; ++xx) {
      const int j = yy * geo.pitch + xx;
      const double u = a.ux[j], v = a.uy[j], t = a.temp[j];
      s_u += u;
      s_v += v;
      s_t += t;
      s_uu += u * u;
      s_uv += u * v;
      s_vv += v * v;
      s_ut += u * t;
      s_vt += v * t;
    }
  const double ubx = s_u / w.cnt, uby = s_v / w.cnt, tb = s_t / w.cnt;
  a.ubx[k] = ubx;
  a.uby[k] = uby;
  a.tb[k] = tb;
  a.lb[0 * geo.plane + k] = s_uu / w.cnt;
  a.lb[1 * geo.plane + k] = s_uv / w.cnt;
  a.lb[2 * geo.plane + k] = s_vv / w.cnt;
  a.lb[3 * geo.plane + k] = s_ut / w.cnt;
  a.lb[4 * geo.plane + k] = s_vt / w.cnt;
  // similarity scales (solver.rs:688-698)
  const double dux = a.ux[k] - ubx, duy = a.uy[k] - uby;
  const double du2 = dux * dux + duy * duy;
  a.qsm[k] = sqrt(0.5 * du2) / 2.0;
  const double dt_local = a.temp[k] - tb;
  a.phi[k] = 0.5 * fabs(dt_local) * sqrt(0.5 * du2);
}

__global__ __launch_bounds__(kBlock) void k_dyn_terms(Geom geo, DynSgsArgs a) {
  int ix, iy;
  if (!node_of_thread(geo, ix, iy)) return;
  const int k = iy * geo.pitch + ix;
  const long long P = geo.plane;
  const double rho_k = a.rho[k], ux_k = a.ux[k], uy_k = a.uy[k];

  // S~ and |S~| from pi_neq with collide_node's formulas (solver.rs:638-657).
  // With c in {-1,0,1} the products fneq*c*c are exact and the zero ones
  // leave the +0-started accumulators unchanged.
  EqTerms<double> e;
  eq_terms<double>(ux_k, uy_k, e);
  double d[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) d[i] = a.mpops[i * P + k] - eq_quad<double>(i, rho_k, e);
  const double pi_xx = 0.0 + d[1] + d[2] + d[4] + d[5] + d[6] + d[8];
  const double pi_xy = 0.0 + d[2] - d[4] + d[6] - d[8];
  const double pi_yy = 0.0 + d[2] + d[3] + d[4] + d[6] + d[7] + d[8];
  const double rho_m = fmax(rho_k, 1e-12);
  const double denom = 2.0 * kCs2 * a.tau * rho_m;
  const double pi2 = pi_xx * pi_xx + 2.0 * pi_xy * pi_xy + pi_yy * pi_yy;
  const double shxx = -pi_xx / denom, shxy = -pi_xy / denom, shyy = -pi_yy / denom;
  const double smag_hat = sqrt(pi2) / denom;

  // S- from centred differences of the test-filtered velocity (:662-667)
  const double sbxx = ddx(geo, a.ubx, ix, iy);
  const double sbyy = ddy(geo, a.uby, ix, iy);
  const double sbxy = 0.5 * (ddx(geo, a.uby, ix, iy) + ddy(geo, a.ubx, ix, iy));
  const double smag_bar = sqrt(sbxx * sbxx + 2.0 * sbxy * sbxy + sbyy * sbyy);

  // temperature gradients, resolved and test-filtered (:671-674)
  const double gtx = ddx(geo, a.temp, ix, iy), gty = ddy(geo, a.temp, ix, iy);
  const double gbx = ddx(geo, a.tb, ix, iy), gby = ddy(geo, a.tb, ix, iy);

  const double qsm_h = a.qsm[k], phi_h = a.phi[k];

  // |T~| and |T-| (:702-707)
  const double th_x = gtx * shxx + gty * shxy;
  const double th_y = gtx * shxy + gty * shyy;
  const double tmag_hat = sqrt(2.0 * (th_x * th_x + th_y * th_y));
  const double tb_x = gbx * sbxx + gby * sbxy;
  const double tb_y = gbx * sbxy + gby * sbyy;
  const double tmag_bar = sqrt(2.0 * (tb_x * tb_x + tb_y * tb_y));

  // test-filtered similarity scales (:713-714)
  const DynWin w = window(geo, a, ix, iy);
  const double qbar_h = dyn_box(geo, w, a.qsm);
  const double phibar_h = dyn_box(geo, w, a.phi);

  // interpolated mixed forms at both levels (:733-740)
  const double dt = a.dtest, dx = 1.0, del = kDeltaT;
  const double ae = a.alpha_exp, be = a.beta_exp;
  const double nu_mix_t = mixpow(dt * dt * smag_bar, ae) * mixpow(dt * qbar_h, be);
  const double nu_mix_g = mixpow(dx * dx * smag_hat, ae) * mixpow(dx * qsm_h, be);
  const double kap_mix_t =
      mixpow(dt * dt * dt / del * tmag_bar, ae) * mixpow(dt / del * phibar_h, be);
  const double kap_mix_g =
      mixpow(dx * dx * dx / del * tmag_hat, ae) * mixpow(dx / del * phi_h, be);

  // Germano M (:748-755), Leonard L at the test filter (:760-769)
  const double mxx = nu_mix_t * sbxx - nu_mix_g * shxx;
  const double mxy = nu_mix_t * sbxy - nu_mix_g * shxy;
  const double myy = nu_mix_t * sbyy - nu_mix_g * shyy;
  const double mx = kap_mix_t * gbx - kap_mix_g * gtx;
  const double my = kap_mix_t * gby - kap_mix_g * gty;
  const double ubx = a.ubx[k], uby = a.uby[k], tb = a.tb[k];
  const double lxx = a.lb[0 * P + k] - ubx * ubx;
  const double lxy = a.lb[1 * P + k] - ubx * uby;
  const double lyy = a.lb[2 * P + k] - uby * uby;
  const double lkk3 = (lxx + lyy) / 3.0;
  const double ldxx = lxx - lkk3;
  const double ldyy = lyy - lkk3;
  const double ljx = a.lb[3 * P + k] - ubx * tb;
  const double ljy = a.lb[4 * P + k] - uby * tb;

  // Lilly contractions (:772-775)
  a.terms[0 * P + k] = ldxx * mxx + 2.0 * lxy * mxy + ldyy * myy;
  a.terms[1 * P + k] = mxx * mxx + 2.0 * mxy * mxy + myy * myy;
  a.terms[2 * P + k] = ljx * mx + ljy * my;
  a.terms[3 * P + k] = mx * mx + my * my;
  a.nu_mix_g[k] = nu_mix_g;
  a.kap_mix_g[k] = kap_mix_g;
}

__global__ __launch_bounds__(kBlock) void k_dyn_coeffs(Geom geo, DynSgsArgs a) {
  int ix, iy;
  if (!node_of_thread(geo, ix, iy)) return;
  double nut, at;
  dyn_coef_node(geo, dyn_coef_of(a), ix, iy, true, nut, at);
}

}  // namespace

hipError_t launch_dynamic_sgs(const Geom& geo, const DynSgsArgs& a, bool coeffs, hipStream_t st) {
  const dim3 grid(nblocks_of(geo)), block(kBlock);
  hipLaunchKernelGGL(k_dyn_filter, grid, block, 0, st, geo, a);
  hipLaunchKernelGGL(k_dyn_terms, grid, block, 0, st, geo, a);
  if (coeffs) hipLaunchKernelGGL(k_dyn_coeffs, grid, block, 0, st, geo, a);
  return hipGetLastError();
}

}  // namespace cavlbm

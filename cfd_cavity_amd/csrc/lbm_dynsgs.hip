// lbm_dynsgs.hip — dynamic mixed-scale subgrid model on the device
// (reference: src/uniform/solver.rs:562-828 compute_dynamic_sgs, helpers
// box_filter :1204-1224, ddx/ddy :1229-1242, lilly_coefficient :1251-1261).
//
// The model is FP64 with no contraction (the pragma below, whatever the
// object's flag) so that every sum, product, division and square root is the
// IEEE result in the reference's operand order.  The model's per-node
// arithmetic is written once (dyn_similarity, dyn_grid_strain,
// dyn_terms_node, dyn_coef_vals) and reached from two forms:
//
// Three passes over the M-form state (post-stream populations and moments),
// one thread per node, planes in HBM:
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
//                  floored (dyn_coef_node; a step from the C-form runs it
//                  inside the collision instead, MODE_FUSED_DYN).
//
// One launch per dynamic step from the C-form (k_dyn_step, any storage,
// test-filter widths 1..3): each workgroup pulls the populations of its tile
// grown by the model's stencil (the filter windows, the centred differences
// and the box of the contractions), runs the filter, the similarity scales,
// the contractions and the coefficient box through LDS, then pulls its own
// nodes again (from L2) and collides and stores them.  Tiles read the force
// plane fy of their halo nodes, so the step reads one fy plane and writes the
// other.
//
// The box sums run in the reference's (y outer, x inner) order from +0.
// The mixed forms a^alpha b^(1-alpha) use pow; the reference's powf is the C
// library's pow, which is not correctly rounded, so this path agrees with a
// CPU evaluation to a few ulp per node rather than bitwise.  For alpha = 0.5
// (the default) the factors are correctly rounded square roots.
//
// Two objects are built from this file: CAVLBM_DYN_OBJECT 64 (no
// contraction) holds everything but the low-precision one-launch steps;
// CAVLBM_DYN_OBJECT 32 (-ffp-contract=on) holds k_dyn_step for FP32 and
// FP16-offset storage, whose pull, collision and coefficient values (the
// functions of lbm_device.hpp, included before the pragma) then round as in
// the single-step kernels of those storages, while the model defined here
// stays uncontracted.
#include "lbm_launch.hpp"

#ifndef CAVLBM_DYN_OBJECT
#define CAVLBM_DYN_OBJECT 64
#endif

#pragma clang fp contract(off)

namespace cavlbm {

namespace {

// ---------------------------------------------------------------------------
// Per-node pieces of the model.

// Square roots of the model: Arith::sqrtp, the generic sequence without its
// 2^256 scaling of arguments below 2^-767, so IEEE sqrt for +0 and every
// argument from 2^-767 up.  The arguments are +0 or far above that: sums of
// squares of differences of velocities and temperatures (nonzero ones above
// about 2^-110, the ulp of values that are themselves differences of O(1e-2)
// populations divided by a density near 1), and products of such
// differences, which stay above about 2^-450.  sqrt_pz maps -0 to +0, as
// pow(-0, 0.5) does.
__device__ __forceinline__ double sqrt_m(double x) { return Arith<double>::sqrtp(x); }
__device__ __forceinline__ double sqrt_pz(double x) {
  const double r = Arith<double>::sqrtp(x);
  return x == 0.0 ? 0.0 : r;
}

// x^e for the mixed forms; every base here is >= +0.
__device__ __forceinline__ double mixpow(double x, double e) {
  if (e == 0.5) return sqrt_pz(x);
  return pow(x, e);
}

// Box mean of an accessor f(x, y) over a window, (y outer, x inner) from +0.
template <class Acc>
__device__ __forceinline__ double box_mean(const DynWin& w, const Acc& f) {
  double s = 0.0;
  for (int yy = w.y0; yy < w.y1; ++yy)
    for (int xx = w.x0; xx < w.x1; ++xx) s += f(xx, yy);
  return w.mean(s);
}

// Box sum over the unclamped W x W window whose first node is (x0, y0), for
// accessors that read +0 at every position outside the lattice (the LDS
// planes of a tile).  The unrolled loop adds the clamped window's terms in
// the same (y outer, x inner) order with +0 terms between them; a sum
// started at +0 is never -0, so s + (+0) = s and the result is the clamped
// window's sum bit for bit.
template <int W, class Acc>
__device__ __forceinline__ double box_sum_w(const Acc& f, int x0, int y0) {
  double s = 0.0;
#pragma unroll
  for (int dy = 0; dy < W; ++dy)
#pragma unroll
    for (int dx = 0; dx < W; ++dx) s += f(x0 + dx, y0 + dy);
  return s;
}

// Centred differences with one-sided clamping at the walls (solver.rs:1229-1242).
template <class Acc>
__device__ __forceinline__ double ddx(const Acc& f, int nx, int ix, int iy) {
  const int xm = ix == 0 ? 0 : ix - 1;
  const int xp = ix + 1 >= nx ? nx - 1 : ix + 1;
  return 0.5 * (f(xp, iy) - f(xm, iy));
}

template <class Acc>
__device__ __forceinline__ double ddy(const Acc& f, int ny, int ix, int iy) {
  const int ym = iy == 0 ? 0 : iy - 1;
  const int yp = iy + 1 >= ny ? ny - 1 : iy + 1;
  return 0.5 * (f(ix, yp) - f(ix, ym));
}

// Similarity scales of one node from its resolved and test-filtered values
// (solver.rs:688-698).
__device__ __forceinline__ void dyn_similarity(double ux, double uy, double t, double ubx,
                                               double uby, double tb, double& qsm, double& phi) {
  const double dux = ux - ubx, duy = uy - uby;
  const double du2 = dux * dux + duy * duy;
  const double sq = sqrt_m(0.5 * du2);
  qsm = sq / 2.0;
  const double dt_local = t - tb;
  phi = 0.5 * fabs(dt_local) * sq;
}

// S~ and |S~| from pi_neq of the post-stream populations with collide_node's
// formulas (solver.rs:638-657).  With c in {-1,0,1} the products fneq*c*c
// are exact and the zero ones leave the +0-started accumulators unchanged.
struct DynStrain {
  double xx, xy, yy, mag;
};

__device__ __forceinline__ DynStrain dyn_grid_strain(const double (&fp)[9], double rho_k,
                                                     double ux_k, double uy_k, double tau) {
  EqTerms<double> e;
  eq_terms<double>(ux_k, uy_k, e);
  double d[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) d[i] = fp[i] - eq_quad<double>(i, rho_k, e);
  const double pi_xx = 0.0 + d[1] + d[2] + d[4] + d[5] + d[6] + d[8];
  const double pi_xy = 0.0 + d[2] - d[4] + d[6] - d[8];
  const double pi_yy = 0.0 + d[2] + d[3] + d[4] + d[6] + d[7] + d[8];
  const double rho_m = fmax(rho_k, 1e-12);
  const double denom = 2.0 * kCs2 * tau * rho_m;
  const double pi2 = pi_xx * pi_xx + 2.0 * pi_xy * pi_xy + pi_yy * pi_yy;
  DynStrain s;
  s.xx = -pi_xx / denom;
  s.xy = -pi_xy / denom;
  s.yy = -pi_yy / denom;
  s.mag = sqrt_m(pi2) / denom;
  return s;
}

// Inputs of the contractions at one node besides the grid-level strain.
struct DynLocal {
  double sbxx, sbyy, sbxy;      // S- (centred differences of the filtered u, v)
  double gtx, gty, gbx, gby;    // resolved and filtered temperature gradients
  double qsm_h, phi_h;          // similarity scales of the node
  double qbar_h, phibar_h;      // and their test-filtered values
  double ubx, uby, tb;          // test-filtered u, v, T
  double lb[5];                 // box of uu, uv, vv, uT, vT
};

// The four Lilly contractions and the grid-level mixed forms of one node
// (solver.rs:662-775).
struct DynTerms {
  double t[4];  // L:M, M:M, Lj Mj, Mj Mj
  double nu_mix_g, kap_mix_g;
};

__device__ __forceinline__ DynTerms dyn_terms_node(const DynStrain& sh, const DynLocal& q,
                                                   const DynSgsArgs& a) {
  const double shxx = sh.xx, shxy = sh.xy, shyy = sh.yy, smag_hat = sh.mag;
  const double sbxx = q.sbxx, sbyy = q.sbyy, sbxy = q.sbxy;
  const double smag_bar = sqrt_m(sbxx * sbxx + 2.0 * sbxy * sbxy + sbyy * sbyy);
  const double gtx = q.gtx, gty = q.gty, gbx = q.gbx, gby = q.gby;

  // |T~| and |T-| (:702-707)
  const double th_x = gtx * shxx + gty * shxy;
  const double th_y = gtx * shxy + gty * shyy;
  const double tmag_hat = sqrt_m(2.0 * (th_x * th_x + th_y * th_y));
  const double tb_x = gbx * sbxx + gby * sbxy;
  const double tb_y = gbx * sbxy + gby * sbyy;
  const double tmag_bar = sqrt_m(2.0 * (tb_x * tb_x + tb_y * tb_y));

  // interpolated mixed forms at both levels (:733-740)
  const double dt = a.dtest, dx = 1.0, del = kDeltaT;
  const double ae = a.alpha_exp, be = a.beta_exp;
  const double nu_mix_t = mixpow(dt * dt * smag_bar, ae) * mixpow(dt * q.qbar_h, be);
  const double nu_mix_g = mixpow(dx * dx * smag_hat, ae) * mixpow(dx * q.qsm_h, be);
  const double kap_mix_t =
      mixpow(dt * dt * dt / del * tmag_bar, ae) * mixpow(dt / del * q.phibar_h, be);
  const double kap_mix_g =
      mixpow(dx * dx * dx / del * tmag_hat, ae) * mixpow(dx / del * q.phi_h, be);

  // Germano M (:748-755), Leonard L at the test filter (:760-769)
  const double mxx = nu_mix_t * sbxx - nu_mix_g * shxx;
  const double mxy = nu_mix_t * sbxy - nu_mix_g * shxy;
  const double myy = nu_mix_t * sbyy - nu_mix_g * shyy;
  const double mx = kap_mix_t * gbx - kap_mix_g * gtx;
  const double my = kap_mix_t * gby - kap_mix_g * gty;
  const double ubx = q.ubx, uby = q.uby, tb = q.tb;
  const double lxx = q.lb[0] - ubx * ubx;
  const double lxy = q.lb[1] - ubx * uby;
  const double lyy = q.lb[2] - uby * uby;
  // (lxx + lyy) / 3: 3 RN(1/3) = 1 - 2^-54, so one correction is exact; the
  // dividend is never -0 (each term is a +0-started box mean minus a square)
  const double lkk3 = Arith<double>::divc(lxx + lyy, 3.0, 1.0 / 3.0);
  const double ldxx = lxx - lkk3;
  const double ldyy = lyy - lkk3;
  const double ljx = q.lb[3] - ubx * tb;
  const double ljy = q.lb[4] - uby * tb;

  // Lilly contractions (:772-775)
  DynTerms r;
  r.t[0] = ldxx * mxx + 2.0 * lxy * mxy + ldyy * myy;
  r.t[1] = mxx * mxx + 2.0 * mxy * mxy + myy * myy;
  r.t[2] = ljx * mx + ljy * my;
  r.t[3] = mx * mx + my * my;
  r.nu_mix_g = nu_mix_g;
  r.kap_mix_g = kap_mix_g;
  return r;
}

// The S-, gradient and box inputs of node (ix, iy) from accessors of the
// resolved fields (u, v, T), the filtered fields (ub, vb, tb) and the
// similarity scales (qs, ph); lb is filled by the caller.  WF = 0: the
// similarity scales' box means over the clamped window w; WF > 0: over the
// unclamped WF x WF window from (ix - HL, iy - HL), for accessors that read
// +0 outside the lattice (box_sum_w).
template <int WF, int HL, class AR, class AF, class AQ>
__device__ __forceinline__ void dyn_local(const Geom& geo, const DynWin& w, int ix, int iy,
                                          const AR& rt, const AF& fu, const AF& fv,
                                          const AF& ft, const AQ& qs, const AQ& ph,
                                          DynLocal& q) {
  const int nx = geo.nx, ny = geo.ny;
  // S- from centred differences of the test-filtered velocity (:662-667)
  q.sbxx = ddx(fu, nx, ix, iy);
  q.sbyy = ddy(fv, ny, ix, iy);
  q.sbxy = 0.5 * (ddx(fv, nx, ix, iy) + ddy(fu, ny, ix, iy));
  // temperature gradients, resolved and test-filtered (:671-674)
  q.gtx = ddx(rt, nx, ix, iy);
  q.gty = ddy(rt, ny, ix, iy);
  q.gbx = ddx(ft, nx, ix, iy);
  q.gby = ddy(ft, ny, ix, iy);
  q.qsm_h = qs(ix, iy);
  q.phi_h = ph(ix, iy);
  // test-filtered similarity scales (:713-714)
  if constexpr (WF > 0) {
    q.qbar_h = w.mean_w<WF>(box_sum_w<WF>(qs, ix - HL, iy - HL));
    q.phibar_h = w.mean_w<WF>(box_sum_w<WF>(ph, ix - HL, iy - HL));
  } else {
    q.qbar_h = box_mean(w, qs);
    q.phibar_h = box_mean(w, ph);
  }
  q.ubx = fu(ix, iy);
  q.uby = fv(ix, iy);
  q.tb = ft(ix, iy);
}

// ---------------------------------------------------------------------------
// Three-pass form.

struct Plane {
  const double* p;
  int pitch;
  __device__ __forceinline__ double operator()(int x, int y) const { return p[y * pitch + x]; }
};

// The three passes take one or two lattices per launch (the octree's hot and
// cold blocks): blocks [0, nb0) work on item 0, the others on item 1.
struct DynSgsPair {
  Geom geo[2];
  DynSgsArgs a[2];
  int nb0;
};

// This block's item, and the node of this thread in it.
__device__ __forceinline__ int pair_node(const DynSgsPair& p, int& ix, int& iy, bool& in,
                                         int& blk) {
  const int b = (int)blockIdx.x;
  const int it = b < p.nb0 ? 0 : 1;
  blk = it ? b - p.nb0 : b;
  const Geom& geo = p.geo[it];
  const int bx = blk % geo.nbx, by = blk / geo.nbx;
  ix = bx * kTileX + (int)(threadIdx.x % kTileX);
  iy = by * kTileY + (int)(threadIdx.x / kTileX);
  in = ix < geo.nx && iy < geo.ny;
  return it;
}

__global__ __launch_bounds__(kBlock) void k_dyn_filter(DynSgsPair p) {
  int ix, iy;
  bool in;
  int blk;
  const int it = pair_node(p, ix, iy, in, blk);
  if (!in) return;
  const Geom& geo = p.geo[it];
  const DynSgsArgs& a = p.a[it];
  const int k = iy * geo.pitch + ix;
  const DynWin w = dyn_window(geo, a.width, a.half_left, ix, iy);
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
  const double ubx = w.mean(s_u), uby = w.mean(s_v), tb = w.mean(s_t);
  a.ubx[k] = ubx;
  a.uby[k] = uby;
  a.tb[k] = tb;
  a.lb[0 * geo.plane + k] = w.mean(s_uu);
  a.lb[1 * geo.plane + k] = w.mean(s_uv);
  a.lb[2 * geo.plane + k] = w.mean(s_vv);
  a.lb[3 * geo.plane + k] = w.mean(s_ut);
  a.lb[4 * geo.plane + k] = w.mean(s_vt);
  dyn_similarity(a.ux[k], a.uy[k], a.temp[k], ubx, uby, tb, a.qsm[k], a.phi[k]);
}

__global__ __launch_bounds__(kBlock) void k_dyn_terms(DynSgsPair p) {
  int ix, iy;
  bool in;
  int blk;
  const int it = pair_node(p, ix, iy, in, blk);
  if (!in) return;
  const Geom& geo = p.geo[it];
  const DynSgsArgs& a = p.a[it];
  const int k = iy * geo.pitch + ix;
  const long long P = geo.plane;
  double fp[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) fp[i] = a.mpops[i * P + k];
  const DynStrain sh = dyn_grid_strain(fp, a.rho[k], a.ux[k], a.uy[k], a.tau);
  const int pitch = geo.pitch;
  const DynWin w = dyn_window(geo, a.width, a.half_left, ix, iy);
  DynLocal q;
  dyn_local<0, 0>(geo, w, ix, iy, Plane{a.temp, pitch}, Plane{a.ubx, pitch},
                  Plane{a.uby, pitch}, Plane{a.tb, pitch}, Plane{a.qsm, pitch},
                  Plane{a.phi, pitch}, q);
#pragma unroll
  for (int j = 0; j < 5; ++j) q.lb[j] = a.lb[j * P + k];
  const DynTerms r = dyn_terms_node(sh, q, a);
#pragma unroll
  for (int j = 0; j < 4; ++j) a.terms[j * P + k] = r.t[j];
  a.nu_mix_g[k] = r.nu_mix_g;
  a.kap_mix_g[k] = r.kap_mix_g;
}

__global__ __launch_bounds__(kBlock) void k_dyn_coeffs(DynSgsPair p) {
  int ix, iy;
  bool in;
  int blk;
  const int it = pair_node(p, ix, iy, in, blk);
  const Geom& geo = p.geo[it];
  const DynSgsArgs& a = p.a[it];
  double nut = 0.0, at = 0.0;
  if (in) dyn_coef_node(geo, dyn_coef_of(a), ix, iy, true, nut, at);
  if (!a.part) return;  // uniform over the workgroup
  // the workgroup's sums: each wave's by a butterfly of lane exchanges, then
  // the waves' in order
  constexpr int kWaves = kBlock / 64;
  __shared__ double red[2 * kWaves];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    nut += __shfl_xor(nut, off, 64);
    at += __shfl_xor(at, off, 64);
  }
  const int w = (int)threadIdx.x / 64;
  if (threadIdx.x % 64 == 0) {
    red[w] = nut;
    red[kWaves + w] = at;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sn = 0.0, sa = 0.0;
    for (int i = 0; i < kWaves; ++i) {
      sn += red[i];
      sa += red[kWaves + i];
    }
    a.part[blk] = sn;
    a.part[geo.nbx * geo.nby + blk] = sa;
  }
}

// ---------------------------------------------------------------------------
// One-launch form.

constexpr int cmax(int a, int b) { return a > b ? a : b; }

// Node sets of a tile of TX x TY output nodes for the test-filter width W,
// as extensions below / above the tile (the same in x and y; windows are
// [i - kHL, i + kHR]):
//   T  the contractions, boxed by the outputs' windows;
//   Q  the similarity scales, boxed by T's windows;
//   F  the filtered fields, differenced at T +- 1 and used by Q;
//   R  the pulled nodes, boxed by F's windows.
// LDS holds u, v, T on R, ub, vb, tb on F, qsm, Phi on Q and the grid-level
// strain on T; after the contractions are computed, they and the mixed forms
// (6 planes on T) overwrite the R and F planes.  The planes that are box-summed
// (u, v, T on R, qsm, Phi on Q, the four contractions on T) hold +0 at their
// positions outside the lattice, so every box sum runs over the unclamped
// W x W window (box_sum_w) with no bounds in its loops.
template <int W, int TX, int TY>
struct DynStepTile {
  static constexpr int kNT = TX * TY;
  static constexpr int kHL = (W - 1) / 2, kHR = W - 1 - kHL;
  static constexpr int kTlo = kHL, kThi = kHR;
  static constexpr int kQlo = 2 * kHL, kQhi = 2 * kHR;
  static constexpr int kFlo = cmax(kHL + 1, 2 * kHL), kFhi = cmax(kHR + 1, 2 * kHR);
  static constexpr int kRlo = kFlo + kHL, kRhi = kFhi + kHR;
  static constexpr int kRW = TX + kRlo + kRhi, kRH = TY + kRlo + kRhi, kR = kRW * kRH;
  static constexpr int kFW = TX + kFlo + kFhi, kFH = TY + kFlo + kFhi, kF = kFW * kFH;
  static constexpr int kQW = TX + kQlo + kQhi, kQH = TY + kQlo + kQhi, kQ = kQW * kQH;
  static constexpr int kTW = TX + kTlo + kThi, kTH = TY + kTlo + kThi, kT = kTW * kTH;
  static constexpr int kTPer = (kT + kNT - 1) / kNT;  // contraction nodes per thread
  static constexpr int kHalo = kR - TX * TY;           // nodes of R outside the tile
  // LDS planes (doubles): R at 0, F at kOffF, Q at kOffQ, strain at kOffS;
  // contractions and mixed forms at 0 once R and F are dead.
  static constexpr int kOffF = 3 * kR, kOffQ = kOffF + 3 * kF, kOffS = kOffQ + 2 * kQ;
  static constexpr int kLds = kOffS + 4 * kT;
  static_assert(6 * kT <= kOffQ, "contractions must fit over the R and F planes");
};

// Halo node j (0 <= j < kHalo) of the tile whose first output node is
// (ox, oy): the kRlo rows of R below the tile, the kRhi rows above it (full
// width), then, for each tile row, its kRlo nodes left and kRhi nodes right.
template <int W, int TX, int TY>
__device__ __forceinline__ void dyn_halo_node(int j, int ox, int oy, int& x, int& y) {
  using DT = DynStepTile<W, TX, TY>;
  const int rx0 = ox - DT::kRlo, ry0 = oy - DT::kRlo;
  if (j < DT::kRlo * DT::kRW) {
    y = ry0 + j / DT::kRW;
    x = rx0 + j % DT::kRW;
    return;
  }
  j -= DT::kRlo * DT::kRW;
  if (j < DT::kRhi * DT::kRW) {
    y = oy + TY + j / DT::kRW;
    x = rx0 + j % DT::kRW;
    return;
  }
  j -= DT::kRhi * DT::kRW;
  constexpr int kE = DT::kRlo + DT::kRhi;
  y = oy + j / kE;
  const int c = j % kE;
  x = c < DT::kRlo ? rx0 + c : ox + TX + (c - DT::kRlo);
}

// LDS plane accessor in global node coordinates: plane of width `w` whose
// local (0, 0) is the global node (x0, y0).
struct LdsPlane {
  const double* p;
  int x0, y0, w;
  __device__ __forceinline__ double operator()(int x, int y) const {
    return p[(y - y0) * w + (x - x0)];
  }
};

// Steps 2 and 3 of a tile, once u, v, T of every lattice node of R and the
// grid-level strain of every node of T are in LDS and the workgroup has
// synchronised: the test filter on F (box_filter, solver.rs:1204-1224) and
// the similarity scales on Q; then the contractions and mixed forms on T,
// held in registers until R and F are no longer read and left in LDS planes
// 0..5 over T.  Ends synchronised.  `own` runs after the thread's
// contractions and before the workgroup waits for the others': work that
// reads no LDS (the one-launch step pulls its own node there, so that waves
// without a second contraction node stream while the others compute).
template <int W, int TX, int TY, class Own>
__device__ __forceinline__ void dyn_tile_terms(const Geom& geo, const DynSgsArgs& a, double* sh,
                                               int ox, int oy, Own&& own) {
  using DT = DynStepTile<W, TX, TY>;
  const int nx = geo.nx, ny = geo.ny;
  const int t = (int)threadIdx.x;
  const int rx0 = ox - DT::kRlo, ry0 = oy - DT::kRlo;
  const int fx0 = ox - DT::kFlo, fy0 = oy - DT::kFlo;
  const int qx0 = ox - DT::kQlo, qy0 = oy - DT::kQlo;
  const int tx0 = ox - DT::kTlo, ty0 = oy - DT::kTlo;
  const LdsPlane ru{sh + 0 * DT::kR, rx0, ry0, DT::kRW};
  const LdsPlane rv{sh + 1 * DT::kR, rx0, ry0, DT::kRW};
  const LdsPlane rt{sh + 2 * DT::kR, rx0, ry0, DT::kRW};
  const LdsPlane fu{sh + DT::kOffF + 0 * DT::kF, fx0, fy0, DT::kFW};
  const LdsPlane fv{sh + DT::kOffF + 1 * DT::kF, fx0, fy0, DT::kFW};
  const LdsPlane ft{sh + DT::kOffF + 2 * DT::kF, fx0, fy0, DT::kFW};
  const LdsPlane qs{sh + DT::kOffQ + 0 * DT::kQ, qx0, qy0, DT::kQW};
  const LdsPlane ph{sh + DT::kOffQ + 1 * DT::kQ, qx0, qy0, DT::kQW};

  // 2. test filter on F and similarity scales on Q (+0 at the positions of
  //    Q outside the lattice, for the unclamped box sums)
  for (int idx = t; idx < DT::kF; idx += DT::kNT) {
    const int x = fx0 + idx % DT::kFW, y = fy0 + idx / DT::kFW;
    const int lx = x - qx0, ly = y - qy0;
    const bool in_q = lx >= 0 && lx < DT::kQW && ly >= 0 && ly < DT::kQH;
    if (x < 0 || x >= nx || y < 0 || y >= ny) {
      if (in_q) {
        sh[DT::kOffQ + 0 * DT::kQ + ly * DT::kQW + lx] = 0.0;
        sh[DT::kOffQ + 1 * DT::kQ + ly * DT::kQW + lx] = 0.0;
      }
      continue;
    }
    const DynWin w = dyn_window_w<W>(geo, DT::kHL, x, y);
    double s_u = 0.0, s_v = 0.0, s_t = 0.0;
#pragma unroll
    for (int dy = 0; dy < W; ++dy)
#pragma unroll
      for (int dx = 0; dx < W; ++dx) {
        const int xx = x - DT::kHL + dx, yy = y - DT::kHL + dy;
        s_u += ru(xx, yy);
        s_v += rv(xx, yy);
        s_t += rt(xx, yy);
      }
    const double ubx = w.mean_w<W>(s_u), uby = w.mean_w<W>(s_v), tb = w.mean_w<W>(s_t);
    sh[DT::kOffF + 0 * DT::kF + idx] = ubx;
    sh[DT::kOffF + 1 * DT::kF + idx] = uby;
    sh[DT::kOffF + 2 * DT::kF + idx] = tb;
    if (in_q) {
      double q, p;
      dyn_similarity(ru(x, y), rv(x, y), rt(x, y), ubx, uby, tb, q, p);
      sh[DT::kOffQ + 0 * DT::kQ + ly * DT::kQW + lx] = q;
      sh[DT::kOffQ + 1 * DT::kQ + ly * DT::kQW + lx] = p;
    }
  }
  __syncthreads();

  // 3. contractions and mixed forms on T, held in registers until R and F
  //    are no longer read
  DynTerms keep[DT::kTPer];
  bool have[DT::kTPer];
#pragma unroll
  for (int j = 0; j < DT::kTPer; ++j) {
    const int idx = t + j * DT::kNT;
    const int x = tx0 + idx % DT::kTW, y = ty0 + idx / DT::kTW;
    have[j] = idx < DT::kT && x >= 0 && x < nx && y >= 0 && y < ny;
    if (!have[j]) continue;
    const DynWin w = dyn_window_w<W>(geo, DT::kHL, x, y);
    DynLocal q;
    dyn_local<W, DT::kHL>(geo, w, x, y, rt, fu, fv, ft, qs, ph, q);
    double s_uu = 0.0, s_uv = 0.0, s_vv = 0.0, s_ut = 0.0, s_vt = 0.0;
#pragma unroll
    for (int dy = 0; dy < W; ++dy)
#pragma unroll
      for (int dx = 0; dx < W; ++dx) {
        const int xx = x - DT::kHL + dx, yy = y - DT::kHL + dy;
        const double u = ru(xx, yy), v = rv(xx, yy), tt = rt(xx, yy);
        s_uu += u * u;
        s_uv += u * v;
        s_vv += v * v;
        s_ut += u * tt;
        s_vt += v * tt;
      }
    q.lb[0] = w.mean_w<W>(s_uu);
    q.lb[1] = w.mean_w<W>(s_uv);
    q.lb[2] = w.mean_w<W>(s_vv);
    q.lb[3] = w.mean_w<W>(s_ut);
    q.lb[4] = w.mean_w<W>(s_vt);
    DynStrain s;
    s.xx = sh[DT::kOffS + 0 * DT::kT + idx];
    s.xy = sh[DT::kOffS + 1 * DT::kT + idx];
    s.yy = sh[DT::kOffS + 2 * DT::kT + idx];
    s.mag = sh[DT::kOffS + 3 * DT::kT + idx];
    keep[j] = dyn_terms_node(s, q, a);
  }
  own();
  __syncthreads();
#pragma unroll
  for (int j = 0; j < DT::kTPer; ++j) {
    const int idx = t + j * DT::kNT;
    if (!have[j]) {
      // +0 contractions outside the lattice, for the unclamped box sums
      if (idx < DT::kT) {
#pragma unroll
        for (int m = 0; m < 4; ++m) sh[m * DT::kT + idx] = 0.0;
      }
      continue;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) sh[m * DT::kT + idx] = keep[j].t[m];
    sh[4 * DT::kT + idx] = keep[j].nu_mix_g;
    sh[5 * DT::kT + idx] = keep[j].kap_mix_g;
  }
  __syncthreads();
}

// S: the storage of the populations (the pull, collision and stores run in
// its arithmetic type R); the model is FP64 in every storage, on the pulled
// values widened to FP64 exactly as MODE_MATERIALIZE_F stores them.
template <int S, int GEQ, int W, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_dyn_step(
    Geom geo, Geom tiles, StepCtx ctx, Walls walls,
    const typename StorageT<S>::T* __restrict__ src, typename StorageT<S>::T* __restrict__ dst,
    const typename StorageT<S>::R* __restrict__ fy_in,
    typename StorageT<S>::R* __restrict__ fy_out, typename StorageT<S>::R* __restrict__ nut,
    DynSgsArgs a, DynCoef dc, int write_aux) {
  using ST = StorageT<S>;
  using R = typename ST::R;
  using DT = DynStepTile<W, TX, TY>;
  __shared__ double sh[DT::kLds];
  int bx, by;
  tile_of_block(tiles, bx, by);
  const int nx = geo.nx, ny = geo.ny, pitch = geo.pitch;
  const unsigned P32 = (unsigned)geo.plane;
  const int ox = bx * TX, oy = by * TY;  // first output node of the tile
  const int t = (int)threadIdx.x;
  // every pulled node at least one node inside the walls: no boundary handling
  const bool interior = ox - DT::kRlo >= 1 && ox + TX - 1 + DT::kRhi <= nx - 2 &&
                        oy - DT::kRlo >= 1 && oy + TY - 1 + DT::kRhi <= ny - 2;

  const int rx0 = ox - DT::kRlo, ry0 = oy - DT::kRlo;
  const int tx0 = ox - DT::kTlo, ty0 = oy - DT::kTlo;

  // Pulled node (x, y): its fields on R and, inside T, its grid-level strain.
  auto put_node = [&](int x, int y, const double (&fp)[9], double r, double u, double v,
                      double tt) {
    const int ir = (y - ry0) * DT::kRW + (x - rx0);
    sh[0 * DT::kR + ir] = u;
    sh[1 * DT::kR + ir] = v;
    sh[2 * DT::kR + ir] = tt;
    const int lx = x - tx0, ly = y - ty0;
    if (lx >= 0 && lx < DT::kTW && ly >= 0 && ly < DT::kTH) {
      const DynStrain s = dyn_grid_strain(fp, r, u, v, a.tau);
      const int it = ly * DT::kTW + lx;
      sh[DT::kOffS + 0 * DT::kT + it] = s.xx;
      sh[DT::kOffS + 1 * DT::kT + it] = s.xy;
      sh[DT::kOffS + 2 * DT::kT + it] = s.yy;
      sh[DT::kOffS + 3 * DT::kT + it] = s.mag;
    }
  };

  // +0 fields at a position of R outside the lattice
  auto put_zero = [&](int x, int y) {
    const int ir = (y - ry0) * DT::kRW + (x - rx0);
    sh[0 * DT::kR + ir] = 0.0;
    sh[1 * DT::kR + ir] = 0.0;
    sh[2 * DT::kR + ir] = 0.0;
  };

  // 1. streaming of every node of R that lies in the lattice
  auto pull = [&](int x, int y, R (&pf)[9], R (&pg)[9], R& r, R& u, R& v, R& tt) {
    const R fyp = fy_in[y * pitch + x];
    if (interior)
      pull_node<S, GEQ, true>(geo, walls, src, fyp, x, y, pf, pg, r, u, v, tt);
    else
      pull_node<S, GEQ, false>(geo, walls, src, fyp, x, y, pf, pg, r, u, v, tt);
  };
  auto pull_put = [&](int x, int y, R (&pf)[9], R (&pg)[9], R& r, R& u, R& v, R& tt) {
    pull(x, y, pf, pg, r, u, v, tt);
    double pd[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) pd[i] = (double)pf[i];
    put_node(x, y, pd, (double)r, (double)u, (double)v, (double)tt);
  };
  // FP32 and FP16-offset storage pull the thread's own output node first and
  // keep its post-stream state in registers through steps 2 and 3 (22
  // VGPRs), then the halo nodes (19.6 -> 18.8 us per FP32 step at 256x1024).
  // FP64 (44 VGPRs) does not fit beside the model at 1024 threads: it pulls
  // R in row order (pulling the halo first and the own nodes last, so that
  // more of the second pull's lines stay in L2, measured 26.2 against 25.2
  // us) and pulls its node again (mostly from L2) as soon as the thread's
  // contractions are done, so that waves without a second contraction node
  // stream while waves 0-1 compute theirs (25.9 against 27.2 us).
  constexpr bool kRepull = S == ST_FP64;
  const int ix = ox + t % TX, iy = oy + t / TX;
  const bool mine = ix < nx && iy < ny;
  R f[9], g[9], rho, ux, uy, temp;
  if constexpr (kRepull) {
    for (int idx = t; idx < DT::kR; idx += DT::kNT) {
      const int x = rx0 + idx % DT::kRW, y = ry0 + idx / DT::kRW;
      if (x < 0 || x >= nx || y < 0 || y >= ny) {
        put_zero(x, y);
        continue;
      }
      R pf[9], pg[9], r, u, v, tt;
      pull_put(x, y, pf, pg, r, u, v, tt);
    }
  } else {
    if (mine)
      pull_put(ix, iy, f, g, rho, ux, uy, temp);
    else
      put_zero(ix, iy);
    for (int j = t; j < DT::kHalo; j += DT::kNT) {
      int x, y;
      dyn_halo_node<W, TX, TY>(j, ox, oy, x, y);
      if (x < 0 || x >= nx || y < 0 || y >= ny) {
        put_zero(x, y);
        continue;
      }
      R pf[9], pg[9], r, u, v, tt;
      pull_put(x, y, pf, pg, r, u, v, tt);
    }
  }
  __syncthreads();

  // 2-4. filter, similarity, contractions; FP64 pulls its node again
  dyn_tile_terms<W, TX, TY>(geo, a, sh, ox, oy, [&]() {
    if (kRepull && mine) pull(ix, iy, f, g, rho, ux, uy, temp);
  });

  // 5. coefficient stage (box of the contractions) and the collision with
  //    this node's nu_t / alpha_t (collide_node, solver.rs:1054-1185)
  if (!mine) return;
  const int k = iy * pitch + ix;
  const DynWin w = dyn_window_w<W>(geo, DT::kHL, ix, iy);
  double num[4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
    num[m] = w.mean_w<W>(box_sum_w<W>(LdsPlane{sh + m * DT::kT, tx0, ty0, DT::kTW},
                                      ix - DT::kHL, iy - DT::kHL));
  const int it = (iy - ty0) * DT::kTW + (ix - tx0);
  double cnu, ck, nt, at;
  dyn_coef_vals(num[0], num[1], num[2], num[3], sh[4 * DT::kT + it], sh[5 * DT::kT + it],
                dc.nu_t_min, dc.alpha_t_min, cnu, ck, nt, at);
  R fy_new, nut_new;
  collide_node<R, GEQ, true>(f, g, rho, ux, uy, temp, ctx, R(ctx.cs), fy_new, nut_new, R(nt),
                             R(at));
  fy_out[k] = fy_new;
  if (write_aux) {
    nut[k] = nut_new;
    dc.cnu[k] = cnu;
    dc.ck[k] = ck;
    dc.nut_pre[k] = nt;
    dc.at_pre[k] = at;
  }
  const SrKey rk = ST::key((unsigned)k, f, g);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    ST::store_f(dst, (unsigned)i * P32 + (unsigned)k, i, f[i], rk);
    ST::store_g(dst, (unsigned)(9 + i) * P32 + (unsigned)k, i, g[i], temp, rk);
  }
  if (ST::kTPlane) reinterpret_cast<float*>(dst + (unsigned)18 * P32)[k] = (float)temp;
}

constexpr int kDynTX = 16, kDynTY = 16;

// The model alone on one or two lattices (the octree's hot and cold blocks)
// from their M-form fields: steps 1-3 of the one-launch step with the fields
// read in place of the streaming, then the coefficient stage, which stores
// C_nu, C_kappa and the floored nu_t / alpha_t, and (part) the workgroup's
// sums of nu_t and alpha_t.  Tiles of item 1 follow those of item 0; the
// per-node arithmetic is the three passes' (k_dyn_filter, k_dyn_terms,
// k_dyn_coeffs), bit for bit.
struct DynModelPair {
  Geom geo[2];
  DynSgsArgs a[2];
  int nt0;  // tiles of item 0
};

template <int W, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_dyn_model(DynModelPair p) {
  using DT = DynStepTile<W, TX, TY>;
  __shared__ double sh[DT::kLds];
  const int b = (int)blockIdx.x;
  const int item = b < p.nt0 ? 0 : 1;
  const int tile = item ? b - p.nt0 : b;
  const Geom& geo = p.geo[item];
  const DynSgsArgs& a = p.a[item];
  const int nx = geo.nx, ny = geo.ny, pitch = geo.pitch;
  const long long P = geo.plane;
  const int tbx = (nx + TX - 1) / TX;
  const int ox = (tile % tbx) * TX, oy = (tile / tbx) * TY;
  const int t = (int)threadIdx.x;
  const int rx0 = ox - DT::kRlo, ry0 = oy - DT::kRlo;
  const int tx0 = ox - DT::kTlo, ty0 = oy - DT::kTlo;

  // 1. u, v, T of every lattice node of R; the grid-level strain on T
  for (int idx = t; idx < DT::kR; idx += DT::kNT) {
    const int x = rx0 + idx % DT::kRW, y = ry0 + idx / DT::kRW;
    if (x < 0 || x >= nx || y < 0 || y >= ny) {
      // +0 outside the lattice, for the unclamped box sums
      sh[0 * DT::kR + idx] = 0.0;
      sh[1 * DT::kR + idx] = 0.0;
      sh[2 * DT::kR + idx] = 0.0;
      continue;
    }
    const int k = y * pitch + x;
    const double u = a.ux[k], v = a.uy[k];
    sh[0 * DT::kR + idx] = u;
    sh[1 * DT::kR + idx] = v;
    sh[2 * DT::kR + idx] = a.temp[k];
    const int lx = x - tx0, ly = y - ty0;
    if (lx >= 0 && lx < DT::kTW && ly >= 0 && ly < DT::kTH) {
      double fp[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) fp[i] = a.mpops[i * P + k];
      const DynStrain s = dyn_grid_strain(fp, a.rho[k], u, v, a.tau);
      const int it = ly * DT::kTW + lx;
      sh[DT::kOffS + 0 * DT::kT + it] = s.xx;
      sh[DT::kOffS + 1 * DT::kT + it] = s.xy;
      sh[DT::kOffS + 2 * DT::kT + it] = s.yy;
      sh[DT::kOffS + 3 * DT::kT + it] = s.mag;
    }
  }
  __syncthreads();

  dyn_tile_terms<W, TX, TY>(geo, a, sh, ox, oy, []() {});

  // 4. coefficient stage (box of the contractions) of the output nodes
  const int ix = ox + t % TX, iy = oy + t / TX;
  double nt = 0.0, at = 0.0;
  if (ix < nx && iy < ny) {
    const int k = iy * pitch + ix;
    const DynWin w = dyn_window_w<W>(geo, DT::kHL, ix, iy);
    double num[4];
#pragma unroll
    for (int m = 0; m < 4; ++m)
      num[m] = w.mean_w<W>(box_sum_w<W>(LdsPlane{sh + m * DT::kT, tx0, ty0, DT::kTW},
                                        ix - DT::kHL, iy - DT::kHL));
    const int it = (iy - ty0) * DT::kTW + (ix - tx0);
    double cnu, ck;
    dyn_coef_vals(num[0], num[1], num[2], num[3], sh[4 * DT::kT + it], sh[5 * DT::kT + it],
                  a.nu_t_min, a.alpha_t_min, cnu, ck, nt, at);
    a.cnu[k] = cnu;
    a.ck[k] = ck;
    a.nut_pre[k] = nt;
    a.at_pre[k] = at;
  }
  if (!a.part) return;  // uniform over the workgroup
  constexpr int kWaves = TX * TY / 64;
  double* red = sh;  // the planes are no longer read after the barrier
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    nt += __shfl_xor(nt, off, 64);
    at += __shfl_xor(at, off, 64);
  }
  __syncthreads();
  const int wv = t / 64;
  if (t % 64 == 0) {
    red[wv] = nt;
    red[kWaves + wv] = at;
  }
  __syncthreads();
  if (t == 0) {
    double sn = 0.0, sa = 0.0;
    for (int i = 0; i < kWaves; ++i) {
      sn += red[i];
      sa += red[kWaves + i];
    }
    const int ntiles = tbx * ((ny + TY - 1) / TY);
    a.part[tile] = sn;
    a.part[ntiles + tile] = sa;
  }
}

template <int W>
hipError_t go_dyn_model(const Geom* geo, const DynSgsArgs* a, int count, hipStream_t st) {
  DynModelPair p{};
  int nt = 0;
  for (int i = 0; i < count; ++i) {
    p.geo[i] = geo[i];
    p.a[i] = a[i];
    nt += dyn_model_blocks(geo[i], W);
  }
  p.nt0 = dyn_model_blocks(geo[0], W);
  hipLaunchKernelGGL((k_dyn_model<W, kDynTX, kDynTY>), dim3(nt), dim3(kDynTX * kDynTY), 0, st,
                     p);
  return hipGetLastError();
}

template <int S, int GEQ, int W>
hipError_t go_dyn_step(const Geom& geo, const StepCtx& ctx, const Walls& walls, const void* src,
                       void* dst, const void* fy_in, void* fy_out, void* nut,
                       const DynSgsArgs& a, int write_aux, hipStream_t st) {
  using ST = StorageT<S>;
  using T = typename ST::T;
  using R = typename ST::R;
  // 32x32 tiles: one workgroup of 1024 threads per CU (114 KB of LDS at
  // W = 2, 123 KB at W = 3).  The grown regions the tile pulls and models
  // are 1.27x its nodes at W = 2 against 1.56x for 16x16, so the step moves
  // and recomputes less.  Measured at 256x1024, Ra=1e10, W = 2 (same box,
  // one process per variant, three each): FP64 27.3 us against 28.3 (16x16),
  // 29.3 (32x16) and 29.0 (16x32); FP32 storage 20.3 against 21.3 (32x16).
  constexpr int TX = 32, TY = 32;
  Geom tiles = geo;
  tiles.nbx = (geo.nx + TX - 1) / TX;
  tiles.nby = (geo.ny + TY - 1) / TY;
  hipLaunchKernelGGL((k_dyn_step<S, GEQ, W, TX, TY>), dim3(tiles.nbx * tiles.nby),
                     dim3(TX * TY), 0, st, geo, tiles, ctx, walls,
                     static_cast<const T*>(src), static_cast<T*>(dst),
                     static_cast<const R*>(fy_in), static_cast<R*>(fy_out), static_cast<R*>(nut),
                     a, dyn_coef_of(a), write_aux);
  return hipGetLastError();
}

template <int S, int GEQ>
hipError_t by_width(int width, const Geom& geo, const StepCtx& ctx, const Walls& walls,
                    const void* src, void* dst, const void* fy_in, void* fy_out, void* nut,
                    const DynSgsArgs& a, int write_aux, hipStream_t st) {
  switch (width) {
    case 1: return go_dyn_step<S, GEQ, 1>(geo, ctx, walls, src, dst, fy_in, fy_out, nut, a, write_aux, st);
    case 2: return go_dyn_step<S, GEQ, 2>(geo, ctx, walls, src, dst, fy_in, fy_out, nut, a, write_aux, st);
    case 3: return go_dyn_step<S, GEQ, 3>(geo, ctx, walls, src, dst, fy_in, fy_out, nut, a, write_aux, st);
    default: return hipErrorInvalidValue;
  }
}

template <int S>
hipError_t by_geq(int geq, int width, const Geom& geo, const StepCtx& ctx, const Walls& walls,
                  const void* src, void* dst, const void* fy_in, void* fy_out, void* nut,
                  const DynSgsArgs& a, int write_aux, hipStream_t st) {
  return geq == 1 ? by_width<S, 1>(width, geo, ctx, walls, src, dst, fy_in, fy_out, nut, a,
                                   write_aux, st)
                  : by_width<S, 2>(width, geo, ctx, walls, src, dst, fy_in, fy_out, nut, a,
                                   write_aux, st);
}

}  // namespace

#if CAVLBM_DYN_OBJECT == 64
hipError_t launch_dynamic_sgs(const Geom* geo, const DynSgsArgs* a, int count, bool coeffs,
                              hipStream_t st) {
  if (count < 1 || count > 2) return hipErrorInvalidValue;
  DynSgsPair p{};
  int nb = 0;
  for (int i = 0; i < count; ++i) {
    p.geo[i] = geo[i];
    p.a[i] = a[i];
    nb += nblocks_of(geo[i]);
  }
  p.nb0 = nblocks_of(geo[0]);
  const dim3 grid(nb), block(kBlock);
  hipLaunchKernelGGL(k_dyn_filter, grid, block, 0, st, p);
  hipLaunchKernelGGL(k_dyn_terms, grid, block, 0, st, p);
  if (coeffs) hipLaunchKernelGGL(k_dyn_coeffs, grid, block, 0, st, p);
  return hipGetLastError();
}

int dyn_model_blocks(const Geom& geo, int width) {
  if (!dyn_step_fits(width)) return nblocks_of(geo);
  return ((geo.nx + kDynTX - 1) / kDynTX) * ((geo.ny + kDynTY - 1) / kDynTY);
}

hipError_t launch_dynamic_model(const Geom* geo, const DynSgsArgs* a, int count, hipStream_t st) {
  if (count < 1 || count > 2) return hipErrorInvalidValue;
  const int width = a[0].width;
  if (!dyn_step_fits(width) || (count == 2 && a[1].width != width))
    return launch_dynamic_sgs(geo, a, count, true, st);
  switch (width) {
    case 1: return go_dyn_model<1>(geo, a, count, st);
    case 2: return go_dyn_model<2>(geo, a, count, st);
    default: return go_dyn_model<3>(geo, a, count, st);
  }
}

hipError_t launch_dyn_step(int storage, int geq, const Geom& geo, const StepCtx& ctx,
                           const Walls& walls, const void* src, void* dst, const void* fy_in,
                           void* fy_out, void* nut, const DynSgsArgs& a, int write_aux,
                           hipStream_t st) {
  if (!dyn_step_fits(a.width)) return hipErrorInvalidValue;
  if (storage != ST_FP64)
    return launch_dyn_step_lowp(storage, geq, geo, ctx, walls, src, dst, fy_in, fy_out, nut, a,
                                write_aux, st);
  return by_geq<ST_FP64>(geq, a.width, geo, ctx, walls, src, dst, fy_in, fy_out, nut, a,
                         write_aux, st);
}

#else  // CAVLBM_DYN_OBJECT == 32

hipError_t launch_dyn_step_lowp(int storage, int geq, const Geom& geo, const StepCtx& ctx,
                                const Walls& walls, const void* src, void* dst,
                                const void* fy_in, void* fy_out, void* nut, const DynSgsArgs& a,
                                int write_aux, hipStream_t st) {
  if (!dyn_step_fits(a.width)) return hipErrorInvalidValue;
  switch (storage) {
    case ST_FP32:
      return by_geq<ST_FP32>(geq, a.width, geo, ctx, walls, src, dst, fy_in, fy_out, nut, a,
                             write_aux, st);
    case ST_FP16_OFFSET:
      return by_geq<ST_FP16_OFFSET>(geq, a.width, geo, ctx, walls, src, dst, fy_in, fy_out, nut,
                                    a, write_aux, st);
    default:
      return hipErrorInvalidValue;
  }
}

#endif

}  // namespace cavlbm

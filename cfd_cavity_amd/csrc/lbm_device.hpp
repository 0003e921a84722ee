// lbm_device.hpp — per-node D2Q9 double-distribution physics and the step
// kernel template of the MI355X cavity solver.
//
// Physics follows the reference's uniform solver (bbrangeo/cfd_cavity,
// src/uniform/solver.rs): collide_node :1054-1185, stream :838-913, thermal
// boundary conditions :951-988, update_macroscopic :307-349,
// enforce_wall_velocity :992-1009.  Every floating-point expression keeps the
// reference's operand order; the FP64 instantiation is compiled with
// -ffp-contract=off so that it is bitwise equal to a CPU evaluation in the
// same order.
//
// HBM layout (structure of arrays): a population set holds 18 planes, f_0..f_8
// then g_0..g_8; plane p, node (ix, iy) lives at p*plane + iy*pitch + ix.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cavlbm {

// Direction table: c_i = (0,0) (1,0) (1,1) (0,1) (-1,1) (-1,0) (-1,-1) (0,-1) (1,-1).
constexpr int kCx[9] = {0, 1, 1, 0, -1, -1, -1, 0, 1};
constexpr int kCy[9] = {0, 0, 1, 1, 1, 0, -1, -1, -1};
constexpr int kOpp[9] = {0, 5, 6, 7, 8, 1, 2, 3, 4};
constexpr double kW[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 36.0,
                          1.0 / 9.0,  1.0 / 36.0, 1.0 / 9.0,
                          1.0 / 36.0, 1.0 / 9.0,  1.0 / 36.0};
constexpr double kTHot = 0.5;
constexpr double kTCold = -0.5;
constexpr double kF64Eps = 2.220446049250313e-16;

// Divisors that appear with a constant denominator, evaluated in the same
// order as the reference's expressions, and their correctly rounded
// reciprocals (compile-time IEEE division).
constexpr double kCs2 = 1.0 / 3.0;
constexpr double kDenEu2 = 2.0 * kCs2 * kCs2;  // eu*eu / (2 cs2 cs2)
constexpr double kDenU2 = 2.0 * kCs2;          // u2 / (2 cs2)
constexpr double kDenCs4 = kCs2 * kCs2;        // ei_u*cy / (cs2*cs2)
constexpr double kInvCs2 = 1.0 / kCs2;
constexpr double kInvDenEu2 = 1.0 / kDenEu2;
constexpr double kInvDenU2 = 1.0 / kDenU2;
constexpr double kInvDenCs4 = 1.0 / kDenCs4;

enum Storage { ST_FP64 = 0, ST_FP32 = 1, ST_FP16_OFFSET = 2 };
enum Mode {
  MODE_FUSED = 0,        // C-form -> stream+BC+moments+walls -> collide -> C-form
  MODE_FIRST = 1,        // M-form pops+fields -> collide -> C-form
  MODE_MATERIALIZE = 2,  // C-form -> stream+BC+moments+walls -> M-form
  MODE_COLLIDE_M = 3,    // M-form pops+fields -> collide -> M-form pops (in place)
  MODE_FIRST_PRE = 4,    // MODE_FIRST with per-node nu_t / alpha_t given (dynamic SGS)
  MODE_MATERIALIZE_F = 5,  // MODE_MATERIALIZE writing the f planes and fields only
  MODE_FUSED_DYN = 6,    // MODE_FUSED with nu_t / alpha_t from the model's coefficient stage
  MODE_WINDOW = 7,       // MODE_FUSED reading fy_in, with the Nusselt row term of the pulled state
  MODE_FUSED_PRE = 8     // MODE_FUSED with per-node nu_t / alpha_t given (the dynamic model's
                         // last values, kept after it is switched off)
};

struct Geom {
  int nx, ny, pitch;
  int nbx, nby;       // tile grid: tiles of kTileX x kTileY nodes
  long long plane;    // elements per population plane (18*plane < 2^31)
};

// Per-step scalars computed on the host exactly as collide() does
// (solver.rs:433-454); t_ref/cs/pr_t are constants of the solver and
// inv_pr_t = 1.0 / pr_t (IEEE, host).
struct StepCtx {
  double beta, nu, alpha, tau, t_ref, cs, pr_t, inv_pr_t;
};

struct Walls {
  const double* theta_hot;  // ny values or nullptr (constant kTHot)
};

// ---------------------------------------------------------------------------
// The last stage of the dynamic mixed-scale model as a per-node function: box
// filter of the four Lilly contractions, the regularised coefficients and the
// floored nu_t / alpha_t (solver.rs:787-815, box_filter :1204-1224,
// lilly_coefficient :1251-1261).  Used by the model's coefficient pass and,
// fused, by the collision of a dynamic step.  FP64 in the reference's order.
constexpr double kCDynMax = 0.04;  // clamp of the Lilly coefficients

// Inputs and outputs of the coefficient stage; planes of the solver's
// pitched layout.
struct DynCoef {
  const double* terms;      // 4 planes: L:M, M:M, Lj Mj, Mj Mj
  const double* nu_mix_g;   // grid-level mixed forms
  const double* kap_mix_g;
  double* cnu;              // out: C_nu, C_kappa, nu_t, alpha_t per node
  double* ck;
  double* nut_pre;
  double* at_pre;
  int width, half_left;     // box filter: width nodes, half_left = (width-1)/2
  double nu_t_min, alpha_t_min;
};

// RN(1/b) for 2^-900 <= |b| <= 2^900: the generic FP64 reciprocal sequence
// (v_rcp_f64, two Newton steps, one correction) without its operand scaling,
// which no such b needs (see Arith below).
__device__ __forceinline__ double rcp_f64(double b) {
  double y = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-b, y, 1.0);
  y = __builtin_fma(y, e, y);
  const double r = __builtin_fma(-b, y, 1.0);
  return __builtin_fma(r, y, y);
}

// a / b correctly rounded from y = RN(1/b): q = a*y and two Markstein
// corrections (see Arith below).
__device__ __forceinline__ double quot_f64(double a, double b, double y) {
  double q = a * y;
  double r = __builtin_fma(-b, q, a);
  q = __builtin_fma(r, y, q);
  r = __builtin_fma(-b, q, a);
  return __builtin_fma(r, y, q);
}

// Window [i - half_left, i + width - half_left) clamped to [0, n).  mean(s):
// s / cnt, the correctly rounded quotient from the window's reciprocal (the
// box sums start at +0, so no dividend is -0).
struct DynWin {
  int x0, x1, y0, y1;
  double cnt, rcnt;
  __device__ __forceinline__ double mean(double s) const { return quot_f64(s, cnt, rcnt); }
  // mean(s) of a window of width W: for a power-of-two count W*W, a window
  // that is not clamped divides by multiplying with the exact reciprocal
  // (both round s / (W*W) once), the others as mean().
  template <int W>
  __device__ __forceinline__ double mean_w(double s) const {
    constexpr int kC = W * W;
    if constexpr ((kC & (kC - 1)) == 0) {
      if (cnt == (double)kC) return s * (1.0 / kC);
    }
    return quot_f64(s, cnt, rcnt);
  }
};

__device__ __forceinline__ DynWin dyn_window(const Geom& geo, int width, int half_left, int ix,
                                             int iy) {
  DynWin w;
  w.x0 = ix >= half_left ? ix - half_left : 0;
  w.x1 = min(ix + width - half_left, geo.nx);
  w.y0 = iy >= half_left ? iy - half_left : 0;
  w.y1 = min(iy + width - half_left, geo.ny);
  w.cnt = (double)((w.y1 - w.y0) * (w.x1 - w.x0));
  w.rcnt = rcp_f64(w.cnt);
  return w;
}

// dyn_window for a compile-time width W: the reciprocal of the count is
// formed only for windows that the walls clamp (an unclamped one has the
// count W*W and the constant reciprocal RN(1 / (W*W))).
template <int W>
__device__ __forceinline__ DynWin dyn_window_w(const Geom& geo, int half_left, int ix, int iy) {
  DynWin w;
  w.x0 = ix >= half_left ? ix - half_left : 0;
  w.x1 = min(ix + W - half_left, geo.nx);
  w.y0 = iy >= half_left ? iy - half_left : 0;
  w.y1 = min(iy + W - half_left, geo.ny);
  w.cnt = (double)((w.y1 - w.y0) * (w.x1 - w.x0));
  if (w.cnt == (double)(W * W))
    w.rcnt = 1.0 / (double)(W * W);
  else
    w.rcnt = rcp_f64(w.cnt);
  return w;
}

// Box mean in the reference's (y outer, x inner) order from +0.
__device__ __forceinline__ double dyn_box(const Geom& geo, const DynWin& w, const double* f) {
  double s = 0.0;
  for (int yy = w.y0; yy < w.y1; ++yy)
    for (int xx = w.x0; xx < w.x1; ++xx) s += f[yy * geo.pitch + xx];
  return w.mean(s);
}

// lilly_coefficient (solver.rs:1251-1261).
__device__ __forceinline__ double dyn_lilly(double num, double den) {
  if (!isfinite(den) || !isfinite(num) || den <= 0.0) return 0.0;
  const double c = 0.5 * num / den;
  if (isfinite(c)) return c < 0.0 ? 0.0 : (c > kCDynMax ? kCDynMax : c);
  return 0.0;
}

// The coefficient stage from the boxed contractions of one node: regularised
// Lilly coefficients (solver.rs:787-796), floored turbulent viscosity and
// diffusivity (:810-815).
__device__ __forceinline__ void dyn_coef_vals(double num_v, double den_v, double num_k,
                                              double den_k, double nu_mix_g, double kap_mix_g,
                                              double nu_t_min, double alpha_t_min, double& cnu,
                                              double& ck, double& nut, double& at) {
  cnu = dyn_lilly(num_v, den_v);
  ck = dyn_lilly(num_k, den_k);
  const double nt = cnu * nu_mix_g;
  nut = nt < nu_t_min ? nu_t_min : nt;
  const double a = ck * kap_mix_g;
  at = a < alpha_t_min ? alpha_t_min : a;
}

// nu_t and alpha_t of node (ix, iy) from the contraction planes; with `store`
// the coefficients and the floored values are also written to their planes.
__device__ __forceinline__ void dyn_coef_node(const Geom& geo, const DynCoef& d, int ix, int iy,
                                              bool store, double& nut, double& at) {
  const int k = iy * geo.pitch + ix;
  const long long P = geo.plane;
  const DynWin w = dyn_window(geo, d.width, d.half_left, ix, iy);
  // one box sum per contraction
  const double num_v = dyn_box(geo, w, d.terms + 0 * P);
  const double den_v = dyn_box(geo, w, d.terms + 1 * P);
  const double num_k = dyn_box(geo, w, d.terms + 2 * P);
  const double den_k = dyn_box(geo, w, d.terms + 3 * P);
  double cnu, ck;
  dyn_coef_vals(num_v, den_v, num_k, den_k, d.nu_mix_g[k], d.kap_mix_g[k], d.nu_t_min,
                d.alpha_t_min, cnu, ck, nut, at);
  if (store) {
    d.cnu[k] = cnu;
    d.ck[k] = ck;
    d.nut_pre[k] = nut;
    d.at_pre[k] = at;
  }
}

struct StepBufs {
  const void* src;  // C-form population set (storage type)
  void* dst;        // C-form population set (storage type)
  void* fy;         // fy_node, compute type
  void* nut;        // nu_t, compute type
  double* mpops;    // M-form population set (FP64)
  double* rho;      // M-form fields (FP64)
  double* ux;
  double* uy;
  double* temp;
  const double* nut_pre;  // MODE_FIRST_PRE / MODE_FUSED_PRE: nu_t per node (FP64), else unused
  const double* at_pre;   // MODE_FIRST_PRE / MODE_FUSED_PRE: alpha_t per node (FP64), else unused
  DynCoef dc;             // MODE_FUSED_DYN: the model's coefficient stage
  // MODE_WINDOW (FP64): the previous force plane (may equal fy), and per row
  // iy the accumulated |T(1,iy) - T(2,iy)| and / or its sign, of the state
  // this step pulls (either may be null)
  const void* fy_in;
  double* nu_acc;
  double* nu_sgn;
};

constexpr int kTileX = 64;
constexpr int kTileY = 4;
constexpr int kBlock = kTileX * kTileY;
// The objective window's single steps and its backward and tangent steps run
// one 64-lane wave per workgroup (tiles of kTileX x kRowTileY nodes): their
// lattices are small, so the waves spread over every CU instead of filling a
// fraction of the CUs four waves each.
constexpr int kRowTileY = 1;
constexpr int kRowBlock = kTileX * kRowTileY;

// ---------------------------------------------------------------------------
// Arithmetic policy.
//
// FP64: a / b with a known y = RN(1/b) is evaluated as q0 = a*y followed by
// two Markstein corrections r = fma(-b, q, a), q = fma(r, y, q).  After the
// first correction q is a faithful quotient, and Markstein's theorem (y the
// correctly rounded reciprocal, q within one ulp) makes the second one the
// correctly rounded a/b; the result equals IEEE division for all normal
// operands (a zero dividend may return +0 where division gives -0; every use
// below adds that value to a nonzero term or to +0).  Five FMA-class
// instructions replace the generic division sequence.
// divc: for the constant divisors cs2, 2 cs2^2, 2 cs2 and cs2^2, whose
// rounded reciprocals (3, 4.5, 1.5, 9) satisfy y*b = 1 - 2^-54 exactly, a*y
// is within a relative 2^-54 of a/b and its rounding is already faithful, so
// one correction gives the correctly rounded quotient (three instructions).
//
// div, rcp, sqrtp (device): the generic FP64 division and square root
// sequences without their operand scaling.  The division is v_rcp_f64, two
// Newton steps, the quotient and one correction: the generic sequence with
// the v_div_scale operands left unscaled, v_div_fmas as a plain fma and no
// v_div_fixup; when the numerator is +0 or of magnitude in [2^-900, 2^900]
// and the denominator and the quotient are in that range, no scaling
// condition holds and no intermediate leaves the normal range, so it returns
// what the generic sequence returns.  The square root is v_rsq_f64 and the
// generic Goldschmidt / Newton refinement without the 2^256 scaling of inputs
// below 2^-767, with the +0 / infinity select kept.  Every operand they
// receive lies in these ranges: reciprocals of relaxation times (> 0.5) and of
// densities above f64::EPSILON, 3/16 over tau+ - 0.5 > 0, and |S|'s square
// root of a sum of squares (+0 or above 1e-40) over a positive denominator.
// Refined from any first approximation within 2^-20 these equal IEEE
// division, reciprocal and sqrt on random operands and on quotients next to a
// rounding midpoint; two cases depend on the hardware's first approximation
// instead: divisors with an all-ones significand, 2^k (2 - 2^-52), need it to
// be RN(1/b), and square roots within about 2^-100 relative of a midpoint
// need it within about 2^-24.  v_rcp_f64 returns RN(1/b) for those divisors
// and both approximations are within 2^-24.2, and the helpers themselves
// equal IEEE results on such operands on the GPU.  The host keeps IEEE
// division (relax_factors is evaluated on both sides).
//
// FP32 (precision forms): reciprocal multiplication, division through the
// hardware reciprocal and the hardware square root (about 1 ulp each); every
// FP32-arithmetic kernel shares these, so blocked, single-step and batched
// steps still round alike.
template <class R> inline constexpr bool kLowp = sizeof(R) == 4;  // FP32 arithmetic

template <class R> struct Arith;
template <> struct Arith<double> {
  __device__ static __forceinline__ double divr(double a, double b, double y) {
    return quot_f64(a, b, y);
  }
  __device__ static __forceinline__ double divc(double a, double b, double y) {
    const double q = a * y;
    const double r = __builtin_fma(-b, q, a);
    return __builtin_fma(r, y, q);
  }
  __host__ __device__ static __forceinline__ double div(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    const double q = a * y;
    const double r = __builtin_fma(-b, q, a);
    return __builtin_fma(r, y, q);
#else
    return a / b;
#endif
  }
  __host__ __device__ static __forceinline__ double rcp(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return rcp_f64(b);
#else
    return 1.0 / b;
#endif
  }
  __device__ static __forceinline__ double sqrtp(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return (x == 0.0 || __builtin_isinf(x)) ? x : g;
  }
};
template <> struct Arith<float> {
  __device__ static __forceinline__ float divr(float a, float, float y) { return a * y; }
  __device__ static __forceinline__ float divc(float a, float, float y) { return a * y; }
  __device__ static __forceinline__ float div(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
  }
  __device__ static __forceinline__ float rcp(float b) { return __builtin_amdgcn_rcpf(b); }
  __device__ static __forceinline__ float sqrtp(float x) { return __builtin_amdgcn_sqrtf(x); }
};

// ---------------------------------------------------------------------------
// storage traits; offsets are element offsets within one plane.
template <int S> struct StorageT;
// The random words of a node's stochastically rounded stores (FP16-offset
// storage; unused by the other storages).
struct SrKey {
  uint32_t f, g;
};

template <> struct StorageT<ST_FP64> {
  using T = double;
  using R = double;
  static constexpr bool kTPlane = false;
  __device__ static __forceinline__ SrKey key(unsigned, const R (&)[9], const R (&)[9]) {
    return SrKey{0u, 0u};
  }
  __device__ static __forceinline__ R load_f(const T* p, unsigned o, int) { return p[o]; }
  __device__ static __forceinline__ R load_g(const T* p, unsigned o, int, R) { return p[o]; }
  __device__ static __forceinline__ void store_f(T* p, unsigned o, int, R v, const SrKey&) {
    p[o] = v;
  }
  __device__ static __forceinline__ void store_g(T* p, unsigned o, int, R v, R, const SrKey&) {
    p[o] = v;
  }
};
template <> struct StorageT<ST_FP32> {
  using T = float;
  using R = float;
  static constexpr bool kTPlane = false;
  __device__ static __forceinline__ SrKey key(unsigned, const R (&)[9], const R (&)[9]) {
    return SrKey{0u, 0u};
  }
  __device__ static __forceinline__ R load_f(const T* p, unsigned o, int) { return p[o]; }
  __device__ static __forceinline__ R load_g(const T* p, unsigned o, int, R) { return p[o]; }
  __device__ static __forceinline__ void store_f(T* p, unsigned o, int, R v, const SrKey&) {
    p[o] = v;
  }
  __device__ static __forceinline__ void store_g(T* p, unsigned o, int, R v, R, const SrKey&) {
    p[o] = v;
  }
};
// The FP32 value x as an opaque register value: keeps the compiler from
// fusing the FP32 operation that produced x with a following FP32 -> FP16
// conversion (v_fma_mixlo_f16 rounds the exact result to FP16 once, the
// separate conversion rounds the FP32 result again), so every kernel rounds
// a stored deviation the same way: to FP32, then to FP16.
__device__ __forceinline__ float opaque_f32(float x) {
  asm volatile("" : "+v"(x));
  return x;
}

// A 32-bit mixing function (xor-shift-multiply rounds): every output bit
// depends on every input bit.
__device__ __forceinline__ uint32_t sr_mix(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return h;
}

// FP32 -> FP16 with stochastic rounding (v_cvt_sr_f16_f32): rounds away from
// zero when the top 13 bits of r plus the 13 dropped significand bits carry,
// toward zero otherwise, so the result is one of the two neighbours of x and
// its mean over uniform r is x (normal and subnormal FP16 results).
__device__ __forceinline__ __half sr_half(float x, uint32_t r) {
  typedef short short2v __attribute__((ext_vector_type(2)));
  const short2v z = {0, 0};
  const short2v y = __builtin_amdgcn_cvt_sr_f16_f32(z, x, r, false);
  return __builtin_bit_cast(__half, (unsigned short)y[0]);
}

// FP16 storage of the deviations from each population's local rest state:
// f_i - o_i, with o_i within 2.2e-7 of the weight w_i, and g_i - w_i*T_j with
// T_j the temperature of the node j that wrote the value (kept in one FP32
// plane after the 18 FP16 planes).  The stored deviations are O(1e-2) of the
// populations, so the ~11 significant bits of FP16 act on them rather than on
// the populations themselves.  The density is 1 + sum of the f deviations
// (storage_moments): summing the populations in FP32 would add the offsets'
// constant low bits to every node's density and round it with the same bias
// at every node and step.
//
// The deviations are rounded stochastically.  Rounding to nearest gives a
// slowly varying deviation the same error step after step, so the errors of
// mass, heat and momentum add up over a run; stochastic rounding makes each
// rounding error zero on average.  The random words are a function of the
// node index and of the FP32 bits of the 18 values the node stores (key()),
// and each population takes a rotation of its node's word: identical states
// round identically, so the single-step, multi-step, batched and dynamic
// kernels still give the same bits, and no step counter is needed.
template <> struct StorageT<ST_FP16_OFFSET> {
  using T = __half;
  using R = float;
  static constexpr bool kTPlane = true;
  // Rest-state offsets of f: the weights rounded to multiples of 2^-22
  // (within 2.2e-7 of 4/9, 1/9, 1/36), chosen so that they sum to exactly 1.
  // A stored deviation h is a multiple of 2^-24, so a loaded population
  // h + o is exact in FP32 and f - o gives h back exactly.
  __device__ static __forceinline__ float off(int i) {
    return i == 0 ? 233017.0f / 524288.0f
                  : ((i & 1) ? 233017.0f / 2097152.0f : 29127.0f / 1048576.0f);
  }
  __device__ static __forceinline__ SrKey key(unsigned k, const R (&f)[9], const R (&g)[9]) {
    uint32_t hf = 0u, hg = 0u;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      hf = (hf ^ __float_as_uint(f[i])) * 0x01000193u;  // FNV-style fold
      hg = (hg ^ __float_as_uint(g[i])) * 0x01000193u;
    }
    const uint32_t a = sr_mix(hf ^ sr_mix(k ^ 0x9E3779B9u));
    return SrKey{a, sr_mix(hg ^ a)};
  }
  __device__ static __forceinline__ uint32_t word(uint32_t key, int i) {
    return __builtin_amdgcn_alignbit(key, key, (unsigned)(7 * i) & 31u);
  }
  __device__ static __forceinline__ R load_f(const T* p, unsigned o, int i) {
    return __half2float(p[o]) + off(i);
  }
  __device__ static __forceinline__ R load_g(const T* p, unsigned o, int i, R t_src) {
    return __half2float(p[o]) + (float)kW[i] * t_src;
  }
  __device__ static __forceinline__ void store_f(T* p, unsigned o, int i, R v, const SrKey& rk) {
    p[o] = sr_half(opaque_f32(v - off(i)), word(rk.f, i));
  }
  __device__ static __forceinline__ void store_g(T* p, unsigned o, int i, R v, R t_node,
                                                 const SrKey& rk) {
    p[o] = sr_half(opaque_f32(v - (float)kW[i] * t_node), word(rk.g, i));
  }
};

// ---------------------------------------------------------------------------
// Equilibria (the crate's feq/geq; the feq form is the one written out in
// src/uniform/tests.rs:314-316):
//   w * s * (1 + eu/cs2 + eu*eu/(2 cs2 cs2) - u2/(2 cs2)),  eu = cx*ux + cy*uy.
// EqTerms holds the velocity-only parts shared by f, g and all directions.
template <class R>
struct EqTerms {
  R q1[9];  // eu/cs2
  R q2[9];  // eu*eu/(2 cs2 cs2)
  R t;      // u2/(2 cs2)
};

// The direction products c_i.u are formed exactly as the reference's
// cx*ux + cy*uy up to the sign of zero: multiplying by +-1 is exact, and a
// zero product only changes the sign of an exact zero, which every use below
// absorbs (1 + q1, q2 = eu*eu, or an added nonzero term).  Opposite
// directions give exactly negated dividends, and RN is sign-symmetric, so
// their quotients are shared: q1[opp] = -q1[i], q2[opp] = q2[i].
template <class R>
__device__ __forceinline__ void eq_terms(R ux, R uy, EqTerms<R>& e) {
  using A = Arith<R>;
  const R u2 = ux * ux + uy * uy;
  e.t = A::divc(u2, R(kDenU2), R(kInvDenU2));
  const R eu[5] = {R(0.0), ux, ux + uy, uy, -ux + uy};
  e.q1[0] = R(0.0);
  e.q2[0] = R(0.0);
#pragma unroll
  for (int i = 1; i < 5; ++i) {
    const R q1 = A::divc(eu[i], R(kCs2), R(kInvCs2));
    const R q2 = A::divc(eu[i] * eu[i], R(kDenEu2), R(kInvDenEu2));
    e.q1[i] = q1;
    e.q1[i + 4] = -q1;
    e.q2[i] = q2;
    e.q2[i + 4] = q2;
  }
}

template <class R>
__device__ __forceinline__ R eq_quad(int i, R s, const EqTerms<R>& e) {
  return R(kW[i]) * s * (R(1.0) + e.q1[i] + e.q2[i] - e.t);
}

template <class R, int GEQ>
__device__ __forceinline__ R eq_g(int i, R t, const EqTerms<R>& e) {
  if (GEQ == 1) return R(kW[i]) * t * (R(1.0) + e.q1[i]);
  return eq_quad<R>(i, t, e);
}

// Stand-alone forms used by the initialisation and wall kernels.
template <class R>
__device__ __forceinline__ R feq(int i, R rho, R ux, R uy) {
  EqTerms<R> e;
  eq_terms<R>(ux, uy, e);
  return eq_quad<R>(i, rho, e);
}

template <class R, int GEQ>
__device__ __forceinline__ R geq(int i, R t, R ux, R uy) {
  EqTerms<R> e;
  eq_terms<R>(ux, uy, e);
  return eq_g<R, GEQ>(i, t, e);
}

// ---------------------------------------------------------------------------
// The collision in two parts.  collide_neq: the equilibria (the crate's
// feq/geq) and the non-equilibrium parts f - feq, g - geq.  collide_relax:
// the TRT update of f with the Guo odd/even force split and the BGK update of
// g, given the relaxation factors
//   a_f = (1/tau+ + 1/tau-)/2, b_f = (1/tau+ - 1/tau-)/2,
//   a_g = (1/tau_t + 1/tau_t)/2, b_g = (1/tau_t - 1/tau_t)/2,
//   k_minus = 1 - 1/(2 tau-), k_plus = 1 - 1/(2 tau+)
// (solver.rs:1140-1185).
//
// Terms that are always a zero are left out: b_g = +0 and the odd force
// part s_minus = w_i (0/cs2) fy of the directions with cy = 0, and sums start
// at their first term instead of +0.  A zero term (of either sign) leaves a
// rounded sum or difference with the value it has without that term, and
// changes at most the sign of a result that is zero; adding such a value to
// a nonzero population gives the same population.  So every collision below
// returns the reference's bits for every state whose populations are
// nonzero (a population is w_i rho or w_i T plus O(u) terms), and equal
// values in any case.  Sums whose terms are all zeros, the only ones whose
// sign could change, occur only there.
template <class R, int GEQ>
__device__ __forceinline__ void collide_neq(const R (&f)[9], const R (&g)[9], R rho, R ux, R uy,
                                            R temp, R (&d_f)[9], R (&d_g)[9], EqTerms<R>& e) {
  eq_terms<R>(ux, uy, e);
  R feq_[9], geq_[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    feq_[i] = eq_quad<R>(i, rho, e);
    geq_[i] = eq_g<R, GEQ>(i, temp, e);
  }
  if (kLowp<R>) {
    // FP32 arithmetic only: close the zeroth moments exactly (sum feq = rho,
    // sum geq = T) through the rest direction, so rounding of the weights and
    // of the quadratic terms cannot leak mass or heat at every collision.
    feq_[0] = rho - (feq_[1] + feq_[2] + feq_[3] + feq_[4] + feq_[5] + feq_[6] + feq_[7] + feq_[8]);
    geq_[0] = temp - (geq_[1] + geq_[2] + geq_[3] + geq_[4] + geq_[5] + geq_[6] + geq_[7] + geq_[8]);
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    d_f[i] = f[i] - feq_[i];
    d_g[i] = g[i] - geq_[i];
  }
}

// Relaxation factors of one collision (the per-node part of collide_node, or
// constants of a level with fixed relaxation times).
// b_g = (1/tau_t - 1/tau_t)/2 is +0 for every finite tau_t and is not kept.
template <class R>
struct RelaxF {
  R a_f, b_f, a_g, k_minus, k_plus;
};

// From tau+ and tau_t+ (tau- from the magic parameter Lambda = 3/16), in the
// reference's order: 1/tau (IEEE); 0.5/tau = 0.5*(1/tau) exactly
// (power-of-two scaling).  Host and device evaluate it alike.
template <class R>
__host__ __device__ __forceinline__ RelaxF<R> relax_factors(R tau_plus, R tau_t_plus) {
  using A = Arith<R>;
  const R tau_minus = R(0.5) + A::div(R(3.0 / 16.0), tau_plus - R(0.5));
  const R itp = A::rcp(tau_plus);
  const R itm = A::rcp(tau_minus);
  const R itt = A::rcp(tau_t_plus);  // tau_t_minus = tau_t_plus (BGK for g)
  RelaxF<R> r;
  r.a_f = R(0.5) * (itp + itm);
  r.b_f = R(0.5) * (itp - itm);
  r.a_g = itt;  // (itt + itt)/2: doubling and halving are exact
  r.k_minus = R(1.0) - R(0.5) * itm;
  r.k_plus = R(1.0) - R(0.5) * itp;
  return r;
}

template <class R>
__device__ __forceinline__ void collide_relax(R (&f)[9], R (&g)[9], R ux, R uy, R fy,
                                              const R (&d_f)[9], const R (&d_g)[9],
                                              const RelaxF<R>& r, const EqTerms<R>& e) {
  using A = Arith<R>;
  // Guo source S_i = w_i [(cy - uy)/cs2 + (c.u) cy / cs2^2] fy; its odd part
  // s_minus = w_i (cy/cs2) fy carries the momentum (solver.rs:1168-1179).
  // (cy - uy)/cs2 takes three values; (c.u) cy is +-eu of the directions
  // with cy = 1 (exact negation for cy = -1), and a signed zero for cy = 0,
  // which leaves the bracket's value unchanged.  (0 - uy)/cs2 is the negated
  // quotient uy/cs2 of the equilibria (RN is sign-symmetric; for uy = 0 the
  // two are zeros of possibly different sign, and the term S_i is a zero).
  const R a_p = A::divc(R(1.0) - uy, R(kCs2), R(kInvCs2));
  const R a_0 = -e.q1[3];
  const R a_m = A::divc(R(-1.0) - uy, R(kCs2), R(kInvCs2));
  const R b2 = A::divc(ux + uy, R(kDenCs4), R(kInvDenCs4));
  const R b3 = A::divc(uy, R(kDenCs4), R(kInvDenCs4));
  const R b4 = A::divc(-ux + uy, R(kDenCs4), R(kInvDenCs4));
  const R bracket[9] = {a_0, a_0, a_p + b2, a_p + b3, a_p + b4, a_0, a_m + b2, a_m + b3, a_m + b4};
  R s_i[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) s_i[i] = R(kW[i]) * bracket[i] * fy;
  if (kLowp<R>)  // FP32 arithmetic only: sum_i S_i = 0 exactly
    s_i[0] = -(s_i[1] + s_i[2] + s_i[3] + s_i[4] + s_i[5] + s_i[6] + s_i[7] + s_i[8]);
  // odd force parts of the directions (0, 1) and (1, 1) and their k_minus terms
  const R sm_a = R(kW[3]) * (R(1.0) / R(kCs2)) * fy;
  const R sm_d = R(kW[2]) * (R(1.0) / R(kCs2)) * fy;
  const R km_a = r.k_minus * sm_a;
  const R km_d = r.k_minus * sm_d;
  R fo[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    if (kCy[i] == 0) {  // s_minus is a zero, s_plus = s_i
      fo[i] = f[i] + (-r.a_f * d_f[i] - r.b_f * d_f[kOpp[i]] + r.k_plus * s_i[i]);
    } else if (!kLowp<R>) {
      // w_i (cy/cs2) fy and k_minus times it for cy = -1 are the exact
      // negations of those for cy = +1 (RN is sign-symmetric)
      const R sm = kCx[i] == 0 ? sm_a : sm_d;
      const R km = kCx[i] == 0 ? km_a : km_d;
      const R s_minus = kCy[i] > 0 ? sm : -sm;
      const R s_plus = s_i[i] - s_minus;
      fo[i] = f[i] + (-r.a_f * d_f[i] - r.b_f * d_f[kOpp[i]] + (kCy[i] > 0 ? km : -km) +
                      r.k_plus * s_plus);
    } else {
      const R s_minus = R(kW[i]) * (R(kCy[i]) / R(kCs2)) * fy;
      const R s_plus = s_i[i] - s_minus;
      fo[i] = f[i] + (-r.a_f * d_f[i] - r.b_f * d_f[kOpp[i]] + r.k_minus * s_minus +
                      r.k_plus * s_plus);
    }
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    f[i] = fo[i];
    g[i] = g[i] + -r.a_g * d_g[i];  // b_g d_g[opp] is a zero
  }
}

// collide_node (solver.rs:1054-1185): TRT Lambda = 3/16 with Smagorinsky nu_t
// from pi_neq and the Guo odd/even force split for f; BGK for g.  With PRE the
// turbulent viscosity and diffusivity are the given per-node values of the
// dynamic model (solver.rs:1118-1139) instead of (Cs dx)^2 |S| and nu_t/Pr_t.
template <class R, int GEQ, bool PRE = false>
__device__ __forceinline__ void collide_node(R (&f)[9], R (&g)[9], R rho, R ux, R uy,
                                             R temp, const StepCtx& c, R cs, R& fy_out,
                                             R& nut_out, R nut_pre = R(0.0),
                                             R at_pre = R(0.0)) {
  using A = Arith<R>;
  const R fy = rho * R(c.beta) * (temp - R(c.t_ref));
  fy_out = fy;
  R d_f[9], d_g[9];
  EqTerms<R> e;
  collide_neq<R, GEQ>(f, g, rho, ux, uy, temp, d_f, d_g, e);
  // pi_neq = sum c c (f - feq) (solver.rs:1096-1106).  With c in {-1,0,1} the
  // products are exact and the zero ones are dropped; the sums start at their
  // first term rather than +0, which changes at most the sign of a zero
  // component, and the components are only squared.
  const R pi_xx = d_f[1] + d_f[2] + d_f[4] + d_f[5] + d_f[6] + d_f[8];
  const R pi_xy = d_f[2] - d_f[4] + d_f[6] - d_f[8];
  const R pi_yy = d_f[2] + d_f[3] + d_f[4] + d_f[6] + d_f[7] + d_f[8];
  const R pi2 = pi_xx * pi_xx + R(2.0) * pi_xy * pi_xy + pi_yy * pi_yy;
  const R s_mag = A::div(A::sqrtp(pi2), R(2.0) * R(kCs2) * R(c.tau) * fmax(rho, R(1e-12)));
  const R nu_t = PRE ? nut_pre : cs * cs * s_mag;
  nut_out = nu_t;

  const R tau_plus = R(0.5) + R(3.0) * (R(c.nu) + nu_t);
  const R alpha_t = PRE ? at_pre : A::divr(nu_t, R(c.pr_t), R(c.inv_pr_t));
  const R tau_t_plus = R(0.5) + R(3.0) * (R(c.alpha) + alpha_t);
  collide_relax<R>(f, g, ux, uy, fy, d_f, d_g, relax_factors<R>(tau_plus, tau_t_plus), e);
}

// A collision with relaxation factors fixed in advance (the octree levels'
// RelaxTimes, solver.rs:578-596 of the octree; collide_node_cavity,
// cavity.rs:167-243): the force, the non-equilibrium parts and the update.
template <int GEQ>
__device__ __forceinline__ void collide_node_fixed(double (&f)[9], double (&g)[9], double rho,
                                                   double ux, double uy, double temp,
                                                   double beta, double t_ref,
                                                   const RelaxF<double>& r, double& fy_out) {
  const double fy = rho * beta * (temp - t_ref);
  fy_out = fy;
  double d_f[9], d_g[9];
  EqTerms<double> e;
  collide_neq<double, GEQ>(f, g, rho, ux, uy, temp, d_f, d_g, e);
  collide_relax<double>(f, g, ux, uy, fy, d_f, d_g, r, e);
}

// ---------------------------------------------------------------------------
// Tile index of this workgroup, remapped so that each XCD (workgroups are
// dispatched round-robin over the 8 XCDs) owns a contiguous band of rows and
// the pull stencil's halo rows stay in that XCD's L2.
__device__ __forceinline__ void tile_of_block(const Geom& geo, int& bx, int& by) {
  const int total = geo.nbx * geo.nby;
  const int lin = blockIdx.x;
  const int per = total / 8, rem = total % 8;
  const int xcd = lin % 8, idx = lin / 8;
  const int t = xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
  bx = t % geo.nbx;
  by = t / geo.nbx;
}

// ---------------------------------------------------------------------------
// g of the isothermal columns (solver.rs:951-970): the equilibrium at rest at
// the wall temperature; the hot column takes the profile value of its row,
// the corner rows that of rows 1 / ny-2.
template <class R, int GEQ>
__device__ __forceinline__ void isothermal_g(const Walls& walls, int ix, int iy, int ny,
                                             R (&g)[9]) {
  R th;
  if (ix == 0) {
    const int iyc = iy == 0 ? 1 : (iy == ny - 1 ? ny - 2 : iy);
    th = walls.theta_hot ? R(walls.theta_hot[iyc]) : R(kTHot);
  } else {
    th = R(kTCold);
  }
  EqTerms<R> e0;
  eq_terms<R>(R(0.0), R(0.0), e0);
#pragma unroll
  for (int i = 0; i < 9; ++i) g[i] = eq_g<R, GEQ>(i, th, e0);
}

// update_macroscopic with the Guo correction (solver.rs:319-345) and
// enforce_wall_velocity (solver.rs:992-1009) on the pulled populations.
// Sums in the reference's order j = 0..8; zero products dropped and products
// by -1 turned into subtractions (exact).  The reference's accumulators start
// at +0; these start at the first term, which is the same unless every term
// is a zero (then at most the sign of the zero sum differs).
template <class R>
__device__ __forceinline__ void node_moments(const R (&f)[9], const R (&g)[9], R fyp,
                                             bool wall, R& rho, R& ux, R& uy, R& temp) {
  using A = Arith<R>;
  const R r = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
  R un = f[1] + f[2] - f[4] - f[5] - f[6] + f[8];
  R vn = f[2] + f[3] + f[4] - f[6] - f[7] - f[8];
  const R tn = g[0] + g[1] + g[2] + g[3] + g[4] + g[5] + g[6] + g[7] + g[8];
  if (r > R(kF64Eps)) {
    const R ir = A::rcp(r);
    un = A::divr(un, r, ir);
    vn = A::divr(vn, r, ir);
    vn += A::divr(R(0.5) * fyp, r, ir);
  }
  if (wall) {
    un = R(0.0);
    vn = R(0.0);
  }
  rho = r;
  ux = un;
  uy = vn;
  temp = tn;
}

// The moments of a pulled node for storage S.  FP16-offset storage forms the
// density and the momentum from the deviations h_i = f_i - o_i (exact) and
// rho = 1 + sum h (the offsets sum to 1, and sum c_i o_i = 0): summing the
// populations themselves in FP32 rounds the same way at every node and
// step, because their low bits are the offsets' constant tails, and the
// collision turns that rounding of rho into a mass change of one sign
// (a relative drift of 9e-4 over 20 000 steps at 64x256).
template <int S>
__device__ __forceinline__ void storage_moments(const typename StorageT<S>::R (&f)[9],
                                                const typename StorageT<S>::R (&g)[9],
                                                typename StorageT<S>::R fyp, bool wall,
                                                typename StorageT<S>::R& rho,
                                                typename StorageT<S>::R& ux,
                                                typename StorageT<S>::R& uy,
                                                typename StorageT<S>::R& temp) {
  using ST = StorageT<S>;
  using R = typename ST::R;
  if constexpr (S == ST_FP16_OFFSET) {
    R h[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) h[i] = f[i] - ST::off(i);
    const R r = R(1.0) + (h[0] + h[1] + h[2] + h[3] + h[4] + h[5] + h[6] + h[7] + h[8]);
    R un = h[1] + h[2] - h[4] - h[5] - h[6] + h[8];
    R vn = h[2] + h[3] + h[4] - h[6] - h[7] - h[8];
    const R tn = g[0] + g[1] + g[2] + g[3] + g[4] + g[5] + g[6] + g[7] + g[8];
    if (r > R(kF64Eps)) {
      const R ir = Arith<R>::rcp(r);
      un = Arith<R>::divr(un, r, ir);
      vn = Arith<R>::divr(vn, r, ir);
      vn += Arith<R>::divr(R(0.5) * fyp, r, ir);
    }
    if (wall) {
      un = R(0.0);
      vn = R(0.0);
    }
    rho = r;
    ux = un;
    uy = vn;
    temp = tn;
  } else {
    node_moments<R>(f, g, fyp, wall, rho, ux, uy, temp);
  }
}

// Streaming of one node from a C-form population set in HBM: pull with
// full-way bounce-back for f (solver.rs:890-912), pull for g followed by the
// thermal boundary conditions (solver.rs:951-988: isothermal columns, then
// adiabatic rows copying rows 1 / ny-2), then the node's moments.
// INTERIOR: the caller guarantees 1 <= ix <= nx-2 and 1 <= iy <= ny-2, where
// every pull is inside the lattice and no boundary condition applies; the
// same arithmetic without the checks.
template <int S, int GEQ, bool INTERIOR = false>
__device__ __forceinline__ void pull_node(const Geom& geo, const Walls& walls,
                                          const typename StorageT<S>::T* src,
                                          typename StorageT<S>::R fyp, int ix, int iy,
                                          typename StorageT<S>::R (&f)[9],
                                          typename StorageT<S>::R (&g)[9],
                                          typename StorageT<S>::R& rho,
                                          typename StorageT<S>::R& ux,
                                          typename StorageT<S>::R& uy,
                                          typename StorageT<S>::R& temp) {
  using ST = StorageT<S>;
  using R = typename ST::R;
  const int nx = geo.nx, ny = geo.ny;
  const unsigned P32 = (unsigned)geo.plane;  // 18*plane < 2^31 (checked on the host)
  const int pitch = geo.pitch;
  const int k = iy * pitch + ix;
  // element offsets from the start of the population set fit in 32 bits
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int sx = ix - kCx[i], sy = iy - kCy[i];
    const bool in = INTERIOR || (sx >= 0 && sx < nx && sy >= 0 && sy < ny);
    const unsigned o = in ? (unsigned)i * P32 + (unsigned)(sy * pitch + sx)
                          : (unsigned)kOpp[i] * P32 + (unsigned)k;
    f[i] = ST::load_f(src, o, in ? i : kOpp[i]);
  }
  if (!INTERIOR && (ix == 0 || ix == nx - 1)) {
    isothermal_g<R, GEQ>(walls, ix, iy, ny, g);
  } else {
    const int gy = INTERIOR ? iy : (iy == 0 ? 1 : (iy == ny - 1 ? ny - 2 : iy));
    const float* tplane =
        reinterpret_cast<const float*>(src + (unsigned)18 * P32);  // used by FP16 only
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const unsigned nj = (unsigned)((gy - kCy[i]) * pitch + (ix - kCx[i]));
      const R tj = ST::kTPlane ? R(tplane[nj]) : R(0.0);
      g[i] = ST::load_g(src, (unsigned)(9 + i) * P32 + nj, i, tj);
    }
  }
  const bool wall = !INTERIOR && (ix == 0 || ix == nx - 1 || iy == 0 || iy == ny - 1);
  storage_moments<S>(f, g, fyp, wall, rho, ux, uy, temp);
}

// ---------------------------------------------------------------------------
// The step kernel: one thread per lattice node.
// The work of one lattice node (ix, iy) inside [0, nx) x [0, ny) for one
// step kernel mode.
template <int S, int GEQ, int MODE>
__device__ __forceinline__ void step_node(const Geom& geo, const StepCtx& ctx,
                                          const Walls& walls, const StepBufs& b,
                                          int write_aux, int ix, int iy) {
  using ST = StorageT<S>;
  using T = typename ST::T;
  using R = typename ST::R;
  const long long P = geo.plane;
  const unsigned P32 = (unsigned)geo.plane;
  const int k = iy * geo.pitch + ix;

  R f[9], g[9];
  R rho, ux, uy, temp;

  if (MODE == MODE_FUSED || MODE == MODE_MATERIALIZE || MODE == MODE_MATERIALIZE_F ||
      MODE == MODE_FUSED_DYN || MODE == MODE_WINDOW || MODE == MODE_FUSED_PRE) {
    const R fyp = static_cast<const R*>(MODE == MODE_WINDOW ? b.fy_in : b.fy)[k];
    pull_node<S, GEQ>(geo, walls, static_cast<const T*>(b.src), fyp, ix, iy, f, g, rho, ux, uy,
                      temp);
    if (MODE == MODE_WINDOW && ix == 1) {
      // the hot-wall Nusselt term of row iy (post.rs:68-78) for the pulled
      // state: T(1, iy) is this node's, T(2, iy) the same pull and sum at
      // column 2 (the isothermal value when column 2 is the cold wall)
      const R t1 = temp;
      R t2 = R(0.0);
      if (geo.nx == 3) {
#pragma unroll
        for (int i = 0; i < 9; ++i) t2 += R(kW[i]) * R(kTCold);
      } else {
        const int gy = iy == 0 ? 1 : (iy == geo.ny - 1 ? geo.ny - 2 : iy);
        const T* src = static_cast<const T*>(b.src);
#pragma unroll
        for (int i = 0; i < 9; ++i)
          t2 += R(src[(unsigned)(9 + i) * P32 + (unsigned)((gy - kCy[i]) * geo.pitch + (2 - kCx[i]))]);
      }
      const R d = t1 - t2;
      if (b.nu_acc) b.nu_acc[iy] += fabs((double)d);
      if (b.nu_sgn) b.nu_sgn[iy] = d > R(0.0) ? 1.0 : (d < R(0.0) ? -1.0 : 0.0);
    }
    if (MODE == MODE_MATERIALIZE || MODE == MODE_MATERIALIZE_F) {
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        b.mpops[i * P + k] = (double)f[i];
        if (MODE == MODE_MATERIALIZE) b.mpops[(9 + i) * P + k] = (double)g[i];
      }
      b.rho[k] = (double)rho;
      b.ux[k] = (double)ux;
      b.uy[k] = (double)uy;
      b.temp[k] = (double)temp;
      return;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      f[i] = R(b.mpops[i * P + k]);
      g[i] = R(b.mpops[(9 + i) * P + k]);
    }
    rho = R(b.rho[k]);
    ux = R(b.ux[k]);
    uy = R(b.uy[k]);
    temp = R(b.temp[k]);
  }

  R fy_new, nut_new;
  if (MODE == MODE_FIRST_PRE || MODE == MODE_FUSED_PRE) {
    collide_node<R, GEQ, true>(f, g, rho, ux, uy, temp, ctx, R(ctx.cs), fy_new, nut_new,
                               R(b.nut_pre[k]), R(b.at_pre[k]));
  } else if (MODE == MODE_FUSED_DYN) {
    // the coefficient stage for this node; its planes are stored with the
    // auxiliary outputs (the values the solver keeps for its getters and for
    // later steps without the model)
    double nt, at;
    dyn_coef_node(geo, b.dc, ix, iy, write_aux != 0, nt, at);
    collide_node<R, GEQ, true>(f, g, rho, ux, uy, temp, ctx, R(ctx.cs), fy_new, nut_new, R(nt),
                               R(at));
  } else
    collide_node<R, GEQ>(f, g, rho, ux, uy, temp, ctx, R(ctx.cs), fy_new, nut_new);
  static_cast<R*>(b.fy)[k] = fy_new;
  if (write_aux) static_cast<R*>(b.nut)[k] = nut_new;

  if (MODE == MODE_COLLIDE_M) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      b.mpops[i * P + k] = (double)f[i];
      b.mpops[(9 + i) * P + k] = (double)g[i];
    }
  } else {
    T* dst = static_cast<T*>(b.dst);
    const SrKey rk = ST::key((unsigned)k, f, g);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      ST::store_f(dst, (unsigned)i * P32 + (unsigned)k, i, f[i], rk);
      ST::store_g(dst, (unsigned)(9 + i) * P32 + (unsigned)k, i, g[i], temp, rk);
    }
    if (ST::kTPlane) reinterpret_cast<float*>(dst + (unsigned)18 * P32)[k] = (float)temp;
  }
}

template <int S, int GEQ, int MODE>
__global__ __launch_bounds__(kBlock) void k_step(Geom geo, StepCtx ctx, Walls walls,
                                                  StepBufs b, int write_aux) {
  constexpr int kTY = MODE == MODE_WINDOW ? kRowTileY : kTileY;  // geo: its tile grid
  int bx, by;
  tile_of_block(geo, bx, by);
  const int ix = bx * kTileX + (int)(threadIdx.x % kTileX);
  const int iy = by * kTY + (int)(threadIdx.x / kTileX);
  if (ix >= geo.nx || iy >= geo.ny) return;
  step_node<S, GEQ, MODE>(geo, ctx, walls, b, write_aux, ix, iy);
}

}  // namespace cavlbm

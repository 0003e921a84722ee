// arith_check.hip — test library: runs the FP64 division, reciprocal and
// square-root helpers of the step kernels (cfd_cavity_amd/csrc/lbm_device.hpp)
// on given operands, so that a GPU test can compare them with IEEE results
// computed on the host, and the FP16-offset storage's stochastic rounding
// (sr_half) on given values and random words.  Built with the step kernels' FP64 flags
// (-ffp-contract=off).  Not part of the product library.
#include "../cfd_cavity_amd/csrc/lbm_device.hpp"

namespace {

constexpr int kPlanes = 8;

// Output planes (n values each):
//   0 the hardware's first approximation of 1/b   (v_rcp_f64)
//   1 the hardware's first approximation of 1/sqrt(x) (v_rsq_f64)
//   2 Arith<double>::rcp(b)
//   3 Arith<double>::div(a, b)
//   4 Arith<double>::sqrtp(x)
//   5 Arith<double>::divr(a, b, y)   (y = RN(1/b) given by the caller)
//   6 Arith<double>::divc(a, b, y)   (one correction; valid for the constant divisors)
//   7 DynWin::mean: a / b from the device reciprocal of b (the box means)
__global__ void k_arith(const double* a, const double* b, const double* x, const double* y,
                        double* out, int n) {
  using A = cavlbm::Arith<double>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ai = a[i], bi = b[i], xi = x[i], yi = y[i];
  const size_t P = (size_t)n;
  out[0 * P + i] = __builtin_amdgcn_rcp(bi);
  out[1 * P + i] = __builtin_amdgcn_rsq(xi);
  out[2 * P + i] = A::rcp(bi);
  out[3 * P + i] = A::div(ai, bi);
  out[4 * P + i] = A::sqrtp(xi);
  out[5 * P + i] = A::divr(ai, bi, yi);
  out[6 * P + i] = A::divc(ai, bi, yi);
  cavlbm::DynWin w;
  w.cnt = bi;
  w.rcnt = cavlbm::rcp_f64(bi);
  out[7 * P + i] = w.mean(ai);
}

}  // namespace

// a, b, x, y: n host doubles each; out: kPlanes * n host doubles.  Copies
// the operands to device 0, runs the helpers and copies the results back.
// Returns 0, or 1 for a bad size, or the first failing HIP call's error code.
extern "C" int arith_check(const double* a, const double* b, const double* x, const double* y,
                           double* out, int n) {
  if (n <= 0 || n > (1 << 24)) return 1;
  const size_t bytes = (size_t)n * sizeof(double);
  double* d[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  hipError_t e = hipSetDevice(0);
  for (int k = 0; k < 5 && e == hipSuccess; ++k)
    e = hipMalloc((void**)&d[k], k == 4 ? kPlanes * bytes : bytes);
  const double* src[4] = {a, b, x, y};
  for (int k = 0; k < 4 && e == hipSuccess; ++k)
    e = hipMemcpy(d[k], src[k], bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_arith, dim3((n + 255) / 256), dim3(256), 0, 0, d[0], d[1], d[2], d[3],
                       d[4], n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, d[4], kPlanes * bytes, hipMemcpyDeviceToHost);
  for (int k = 0; k < 5; ++k)
    if (d[k] != nullptr) (void)hipFree(d[k]);
  return (int)e;
}

extern "C" int arith_check_planes() { return kPlanes; }

namespace {

// out[j * nx + i] = the FP16 bits of sr_half(x[i], r[j]).
__global__ void k_sr(const float* x, const unsigned* r, unsigned short* out, int nx, int nr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx || j >= nr) return;
  const __half h = cavlbm::sr_half(x[i], r[j]);
  out[(size_t)j * nx + i] = __builtin_bit_cast(unsigned short, h);
}

}  // namespace

// The FP16-offset storage's stochastic rounding: x: nx host floats, r: nr
// host random words, out: nx * nr host FP16 bit patterns (row j for r[j]).
// Returns 0, or 1 for a bad size, or the first failing HIP call's error code.
extern "C" int sr_check(const float* x, const unsigned* r, unsigned short* out, int nx, int nr) {
  if (nx <= 0 || nr <= 0 || nx > (1 << 20) || nr > 65535 || (size_t)nx * nr > ((size_t)1 << 28))
    return 1;
  const size_t xb = (size_t)nx * sizeof(float), rb = (size_t)nr * sizeof(unsigned);
  const size_t ob = (size_t)nx * nr * sizeof(unsigned short);
  void* d[3] = {nullptr, nullptr, nullptr};
  hipError_t e = hipSetDevice(0);
  if (e == hipSuccess) e = hipMalloc(&d[0], xb);
  if (e == hipSuccess) e = hipMalloc(&d[1], rb);
  if (e == hipSuccess) e = hipMalloc(&d[2], ob);
  if (e == hipSuccess) e = hipMemcpy(d[0], x, xb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d[1], r, rb, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_sr, dim3((nx + 255) / 256, nr), dim3(256), 0, 0, (const float*)d[0],
                       (const unsigned*)d[1], (unsigned short*)d[2], nx, nr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, d[2], ob, hipMemcpyDeviceToHost);
  for (int k = 0; k < 3; ++k)
    if (d[k] != nullptr) (void)hipFree(d[k]);
  return (int)e;
}

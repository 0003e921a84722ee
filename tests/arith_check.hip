// arith_check.hip — test library: runs the FP64 division, reciprocal and
// square-root helpers of the step kernels (cfd_cavity_amd/csrc/lbm_device.hpp)
// on given operands, so that a GPU test can compare them with IEEE results
// computed on the host.  Built with the step kernels' FP64 flags
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

// Small utility kernels: buffer fills (used instead of memset so that a captured stage chain holds
// only kernel nodes) and the device entry of the deterministic-math self test.
#include <hip/hip_runtime.h>

#include "ce_common.h"
#include "ce_eig_grp.h"

namespace ce {

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* p, unsigned long long v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t* p, int32_t v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_detmath(int fn, const double* x, const double* y, double* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r = 0.0;
  switch (fn) {
    case 0: r = cedm::exp(x[i]); break;
    case 1: r = cedm::log(x[i]); break;
    case 2: r = cedm::sin(x[i]); break;
    case 3: r = cedm::cos(x[i]); break;
    case 4: r = cedm::atan2(x[i], y[i]); break;
    case 5: r = cedm::hypot(x[i], y[i]); break;
    case 6: r = cedm::log10(x[i]); break;
    case 7: r = grp_sqrt(x[i]); break;  // the eigensolver's square root
    default: r = cedm::qnan(); break;
  }
  out[i] = r;
}

void launch_fill_u64(unsigned long long* p, unsigned long long v, size_t n, hipStream_t st) {
  if (n) hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
}
void launch_fill_i32(int32_t* p, int32_t v, size_t n, hipStream_t st) {
  if (n) hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
}
void launch_detmath(int fn, const double* x, const double* y, double* out, int n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_detmath, dim3((n + 255) / 256), dim3(256), 0, st, fn, x, y, out, n);
}

}  // namespace ce

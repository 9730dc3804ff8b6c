// ww3-gpu_amd/csrc/devlibm_eval.hip
// ww3gpu_devlibm_eval: evaluates the device math functions of devlibm.hpp over an
// array, so a caller can compare them with the host libm on the same inputs.
#include "devlibm.hpp"
#include "runtime.hpp"

namespace ww3gpu {
namespace {

__global__ void eval_kernel(int fn, const float* __restrict__ x, const float* __restrict__ x2,
                            float* __restrict__ y, long n) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  switch (fn) {
    case 0:
      y[i] = libm::expf(x[i]);
      break;
    default:
      y[i] = 0.0f;
      break;
  }
  (void)x2;
}

}  // namespace
}  // namespace ww3gpu

extern "C" int ww3gpu_devlibm_eval(int fn, const float* x, const float* x2, float* y, long n,
                                   void* stream) {
  if (!ww3gpu::runtime().ready) return WW_KOKKOS_ERR_NOT_INITIALISED;
  if (fn < 0 || fn > 0 || n < 0 || x == nullptr || y == nullptr) return WW_KOKKOS_ERR_BAD_SHAPE;
  if (n == 0) return WW_KOKKOS_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);  // NULL is the default stream
  const long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(ww3gpu::eval_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st,
                     fn, x, x2, y, n);
  return hipGetLastError() == hipSuccess ? WW_KOKKOS_OK : WW_KOKKOS_ERR_KERNEL;
}

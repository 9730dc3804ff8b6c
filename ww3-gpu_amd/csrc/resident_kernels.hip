// ww3-gpu_amd/csrc/resident_kernels.hip
// Column moves on the resident VA(NSPEC, 0:NSEALM): gather listed columns into a packed
// block, scatter a packed block back, zero listed columns.  One workgroup per listed column,
// its NSPEC floats contiguous in both layouts.
#include <hip/hip_runtime.h>

#include "resident.hpp"

namespace ww3gpu {
namespace resident {
namespace {

__global__ __launch_bounds__(256) void gather_kernel(const float* va, int nspec, const int* cols, float* out) {
  const size_t src = static_cast<size_t>(cols[blockIdx.x]) * nspec;
  const size_t dst = static_cast<size_t>(blockIdx.x) * nspec;
  for (int i = threadIdx.x; i < nspec; i += blockDim.x) out[dst + i] = va[src + i];
}

__global__ __launch_bounds__(256) void scatter_kernel(float* va, int nspec, const int* cols, const float* in) {
  const size_t dst = static_cast<size_t>(cols[blockIdx.x]) * nspec;
  const size_t src = static_cast<size_t>(blockIdx.x) * nspec;
  for (int i = threadIdx.x; i < nspec; i += blockDim.x) va[dst + i] = in[src + i];
}

__global__ __launch_bounds__(256) void zero_kernel(float* va, int nspec, const int* cols) {
  const size_t dst = static_cast<size_t>(cols[blockIdx.x]) * nspec;
  for (int i = threadIdx.x; i < nspec; i += blockDim.x) va[dst + i] = 0.f;
}

}  // namespace

hipError_t gather_cols(const float* va, int nspec, const int* cols, int ncol, float* out, hipStream_t st) {
  if (ncol <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_kernel, dim3(ncol), dim3(256), 0, st, va, nspec, cols, out);
  return hipGetLastError();
}

hipError_t scatter_cols(float* va, int nspec, const int* cols, int ncol, const float* in, hipStream_t st) {
  if (ncol <= 0) return hipSuccess;
  hipLaunchKernelGGL(scatter_kernel, dim3(ncol), dim3(256), 0, st, va, nspec, cols, in);
  return hipGetLastError();
}

hipError_t zero_cols(float* va, int nspec, const int* cols, int ncol, hipStream_t st) {
  if (ncol <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_kernel, dim3(ncol), dim3(256), 0, st, va, nspec, cols);
  return hipGetLastError();
}

}  // namespace resident
}  // namespace ww3gpu

// ww3-gpu_amd/csrc/resident.hpp
// The resident spectrum VA(NSPEC, 0:NSEALM) of a grid: the host array it is bound to, its
// device copy and which copy holds the newer bits (HOST, DEVICE, BOTH).
//
// Reference: kokkos/src/fortran_iface/resident_c.hpp:38-103 (the ww_va_* calls, the state
// rules, ww_srce_batch_va).
#pragma once

#include <hip/hip_runtime.h>

namespace ww3gpu {
namespace resident {

/// The device copy of grid igrid's VA, synced for a kernel: binds the grid on its first call
/// (va, nspec, nsealm) or checks them against the binding (BAD_SHAPE), copies the host VA in
/// when the host holds the newer bits, and marks the state DEVICE when `write`.  Returns the
/// status; d_va receives the device pointer.
int device_va(int igrid, float* va, int nspec, int nsealm, bool write, float** d_va, const char* who);

/// Device kernels on VA columns (resident_kernels.hip).  Columns are 0-based JSEA.
hipError_t gather_cols(const float* va, int nspec, const int* cols, int ncol, float* out, hipStream_t st);
hipError_t scatter_cols(float* va, int nspec, const int* cols, int ncol, const float* in, hipStream_t st);
hipError_t zero_cols(float* va, int nspec, const int* cols, int ncol, hipStream_t st);

}  // namespace resident
}  // namespace ww3gpu

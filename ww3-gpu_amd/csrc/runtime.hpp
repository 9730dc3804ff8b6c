// ww3-gpu_amd/csrc/runtime.hpp
// Library-wide state behind the C ABI: the selected device, the library's stream,
// status reporting, run-time flags and a small device-buffer helper.
//
// Reference behaviour kept: status codes are recorded in relaxed atomics read by each
// module's *_last_error() and every failure prints one stderr line
// (kokkos/src/fortran_iface/snl1_shim.cpp:92-110); flags are read once per process
// (kokkos/src/fortran_iface/part_flags.cpp:1-60).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstddef>
#include <functional>

#include "../../include/ww3gpu_c.h"

namespace ww3gpu {

struct Runtime {
  bool ready = false;
  int device = 0;
  hipStream_t stream = nullptr;
};

Runtime& runtime();

/// Record `code` in `slot` and, when it is not OK, print one line to stderr.
void report(std::atomic<int>& slot, int code, const char* what);

/// Register a function that frees a module's grid contexts; ww_kokkos_finalize runs them.
void on_finalize(std::function<void()> fn);

/// Run-time flag resolution (ww_kokkos_flag and its helpers): see runtime.cpp.
int resolve_flag(const char* name);
int preset_status();
int fallback_slot(const char* part);
void fallback_add(int slot);

/// A device allocation that only grows.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n);
  void release();
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

}  // namespace ww3gpu

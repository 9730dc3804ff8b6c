// ww3-gpu_amd/csrc/resident_capi.cpp
// The resident spectrum: the ww_va_* calls of the C ABI and the device copy the _va kernel
// entries work on.
//
// Reference: kokkos/src/fortran_iface/resident_c.hpp:38-103.  One context per grid, bound on
// its first call to the VA address and shape it passes (a different one later is BAD_SHAPE);
// state HOST at the start and after ww_va_host_write, DEVICE after a kernel wrote the device
// copy, BOTH after a sync.  Host calls copy only when the other side holds the newer bits;
// the column forms move only the listed columns, packed.  The reference's diagnostic modes
// (WW_KOKKOS_VA_RESIDENT, _CHECK, _POISON) are not provided; WW_KOKKOS_VA_COUNT=1 prints the
// traffic counters at exit, as there.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ww3gpu_c.h"
#include "resident.hpp"
#include "runtime.hpp"

namespace ww3gpu {
namespace resident {
namespace {

enum State { HOST = 0, DEVICE = 1, BOTH = 2 };

struct Ctx {
  float* va = nullptr;
  int nspec = 0, nsealm = 0;
  DevBuf d, cols, pack;
  State state = HOST;
};

struct Counters {
  long long h2d_calls = 0, d2h_calls = 0, h2d_bytes = 0, d2h_bytes = 0, col_calls = 0;
};

std::map<int, Ctx>& ctxs() {
  static auto* m = new std::map<int, Ctx>();
  return *m;
}
Counters& counters() {
  static Counters c;
  return c;
}
std::atomic<int> g_last{WW_KOKKOS_OK};
std::recursive_mutex g_mutex;

int fail(int code, const std::string& what) {
  report(g_last, code, what.c_str());
  return code;
}

void report_counts() {
  const Counters& c = counters();
  std::fprintf(stderr,
               "ww3gpu: resident VA: %lld host-to-device copies (%lld B), %lld device-to-host copies (%lld B), "
               "%lld column moves, %zu grids\n",
               c.h2d_calls, c.h2d_bytes, c.d2h_calls, c.d2h_bytes, c.col_calls, ctxs().size());
}

[[maybe_unused]] const bool g_count = [] {
  const char* v = std::getenv("WW_KOKKOS_VA_COUNT");
  if (v != nullptr && std::strcmp(v, "1") == 0) std::atexit(report_counts);
  return true;
}();

bool hooked = false;

/// The context of igrid bound to (va, nspec, nsealm), created on the first call.
int bind(int igrid, float* va, int nspec, int nsealm, Ctx** out, const char* who) {
  if (!runtime().ready) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, std::string(who) + " before ww_kokkos_init");
  if (va == nullptr || nspec < 1 || nsealm < 0)
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": null VA or bad shape");
  if (!hooked) {
    hooked = true;
    on_finalize([] {
      for (auto& kv : ctxs()) {
        kv.second.d.release();
        kv.second.cols.release();
        kv.second.pack.release();
      }
      ctxs().clear();
    });
  }
  auto it = ctxs().find(igrid);
  if (it == ctxs().end()) {
    Ctx c;
    c.va = va;
    c.nspec = nspec;
    c.nsealm = nsealm;
    if (c.d.ensure(static_cast<size_t>(nspec) * (nsealm + 1) * sizeof(float)) != hipSuccess)
      return fail(WW_KOKKOS_ERR_RUNTIME, std::string(who) + ": device allocation");
    it = ctxs().emplace(igrid, c).first;
  } else if (it->second.va != va || it->second.nspec != nspec || it->second.nsealm != nsealm) {
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": VA differs from the grid's first call");
  }
  *out = &it->second;
  return WW_KOKKOS_OK;
}

size_t bytes_of(const Ctx& c) { return static_cast<size_t>(c.nspec) * (c.nsealm + 1) * sizeof(float); }

int host_read(Ctx& c, const char* who) {
  if (c.state != DEVICE) return WW_KOKKOS_OK;
  hipStream_t st = runtime().stream;
  if (hipMemcpyAsync(c.va, c.d.p, bytes_of(c), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(WW_KOKKOS_ERR_RUNTIME, std::string(who) + ": device-to-host copy");
  counters().d2h_calls++;
  counters().d2h_bytes += static_cast<long long>(bytes_of(c));
  c.state = BOTH;
  return WW_KOKKOS_OK;
}

int check_cols(const Ctx& c, int ncol, const int* cols, const char* who) {
  if (ncol < 0 || (ncol > 0 && cols == nullptr)) return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": bad column list");
  for (int i = 0; i < ncol; ++i)
    if (cols[i] < 0 || cols[i] > c.nsealm) return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": a column outside VA");
  return WW_KOKKOS_OK;
}

/// The column list on the device and a packed buffer of ncol columns.
int stage_cols(Ctx& c, int ncol, const int* cols, const char* who) {
  const size_t pb = static_cast<size_t>(ncol) * c.nspec * sizeof(float);
  if (c.cols.ensure(static_cast<size_t>(ncol) * sizeof(int)) != hipSuccess || c.pack.ensure(pb) != hipSuccess ||
      hipMemcpyAsync(c.cols.p, cols, static_cast<size_t>(ncol) * sizeof(int), hipMemcpyHostToDevice,
                     runtime().stream) != hipSuccess)
    return fail(WW_KOKKOS_ERR_RUNTIME, std::string(who) + ": device staging");
  counters().col_calls++;
  return WW_KOKKOS_OK;
}

}  // namespace

int device_va(int igrid, float* va, int nspec, int nsealm, bool write, float** d_va, const char* who) {
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  Ctx* c = nullptr;
  int rc = bind(igrid, va, nspec, nsealm, &c, who);
  if (rc != WW_KOKKOS_OK) return rc;
  if (c->state == HOST) {
    hipStream_t st = runtime().stream;
    if (hipMemcpyAsync(c->d.p, c->va, bytes_of(*c), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(WW_KOKKOS_ERR_RUNTIME, std::string(who) + ": host-to-device copy");
    counters().h2d_calls++;
    counters().h2d_bytes += static_cast<long long>(bytes_of(*c));
    c->state = BOTH;
  }
  if (write) c->state = DEVICE;
  *d_va = c->d.as<float>();
  return WW_KOKKOS_OK;
}

}  // namespace resident
}  // namespace ww3gpu

using namespace ww3gpu;
using namespace ww3gpu::resident;

extern "C" {

int ww_va_host_read(int igrid, float* va, int nspec, int nsealm) {
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  Ctx* c = nullptr;
  int rc = bind(igrid, va, nspec, nsealm, &c, "ww_va_host_read");
  if (rc == WW_KOKKOS_OK) rc = host_read(*c, "ww_va_host_read");
  return rc;
}

int ww_va_host_write(int igrid, float* va, int nspec, int nsealm) {
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  Ctx* c = nullptr;
  int rc = bind(igrid, va, nspec, nsealm, &c, "ww_va_host_write");
  if (rc == WW_KOKKOS_OK) rc = host_read(*c, "ww_va_host_write");
  if (rc == WW_KOKKOS_OK) c->state = HOST;
  return rc;
}

int ww_va_host_read_cols(int igrid, float* va, int nspec, int nsealm, int ncol, const int* cols) {
  const char* who = "ww_va_host_read_cols";
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  Ctx* c = nullptr;
  int rc = bind(igrid, va, nspec, nsealm, &c, who);
  if (rc == WW_KOKKOS_OK) rc = check_cols(*c, ncol, cols, who);
  if (rc != WW_KOKKOS_OK || c->state != DEVICE || ncol == 0) return rc;
  if ((rc = stage_cols(*c, ncol, cols, who)) != WW_KOKKOS_OK) return rc;
  hipStream_t st = runtime().stream;
  std::vector<float> h(static_cast<size_t>(ncol) * nspec);
  if (gather_cols(c->d.as<float>(), nspec, c->cols.as<int>(), ncol, c->pack.as<float>(), st) != hipSuccess ||
      hipMemcpyAsync(h.data(), c->pack.p, h.size() * sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(WW_KOKKOS_ERR_KERNEL, std::string(who) + ": column copy");
  for (int i = 0; i < ncol; ++i)
    std::memcpy(va + static_cast<size_t>(cols[i]) * nspec, h.data() + static_cast<size_t>(i) * nspec,
                static_cast<size_t>(nspec) * sizeof(float));
  counters().d2h_bytes += static_cast<long long>(h.size() * sizeof(float));
  return WW_KOKKOS_OK;
}

int ww_va_host_wrote_cols(int igrid, float* va, int nspec, int nsealm, int ncol, const int* cols) {
  const char* who = "ww_va_host_wrote_cols";
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  Ctx* c = nullptr;
  int rc = bind(igrid, va, nspec, nsealm, &c, who);
  if (rc == WW_KOKKOS_OK) rc = check_cols(*c, ncol, cols, who);
  if (rc != WW_KOKKOS_OK || c->state == HOST || ncol == 0) return rc;
  if ((rc = stage_cols(*c, ncol, cols, who)) != WW_KOKKOS_OK) return rc;
  std::vector<float> h(static_cast<size_t>(ncol) * nspec);
  for (int i = 0; i < ncol; ++i)
    std::memcpy(h.data() + static_cast<size_t>(i) * nspec, va + static_cast<size_t>(cols[i]) * nspec,
                static_cast<size_t>(nspec) * sizeof(float));
  hipStream_t st = runtime().stream;
  if (hipMemcpyAsync(c->pack.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess ||
      scatter_cols(c->d.as<float>(), nspec, c->cols.as<int>(), ncol, c->pack.as<float>(), st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(WW_KOKKOS_ERR_KERNEL, std::string(who) + ": column copy");
  counters().h2d_bytes += static_cast<long long>(h.size() * sizeof(float));
  return WW_KOKKOS_OK;
}

int ww_va_host_update(int igrid, float* va, int nspec, int nsealm) {
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  Ctx* c = nullptr;
  return bind(igrid, va, nspec, nsealm, &c, "ww_va_host_update");
}

int ww_va_device_zero_cols(int igrid, float* va, int nspec, int nsealm, int ncol, const int* cols) {
  const char* who = "ww_va_device_zero_cols";
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  Ctx* c = nullptr;
  int rc = bind(igrid, va, nspec, nsealm, &c, who);
  if (rc == WW_KOKKOS_OK) rc = check_cols(*c, ncol, cols, who);
  if (rc != WW_KOKKOS_OK || c->state == HOST || ncol == 0) return rc;
  if ((rc = stage_cols(*c, ncol, cols, who)) != WW_KOKKOS_OK) return rc;
  hipStream_t st = runtime().stream;
  if (zero_cols(c->d.as<float>(), nspec, c->cols.as<int>(), ncol, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(WW_KOKKOS_ERR_KERNEL, std::string(who) + ": zero fill");
  return WW_KOKKOS_OK;
}

int ww_va_device(int igrid, int write) {
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  auto it = ctxs().find(igrid);
  if (it == ctxs().end()) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, "ww_va_device: grid has no bound VA");
  float* d = nullptr;
  const Ctx& c = it->second;
  return device_va(igrid, c.va, c.nspec, c.nsealm, write != 0, &d, "ww_va_device");
}

void ww_va_report(void) { report_counts(); }

/// The state of grid igrid's resident VA: 0 HOST, 1 DEVICE, 2 BOTH, -1 unbound.
int ww3gpu_va_state(int igrid) {
  std::lock_guard<std::recursive_mutex> lock(g_mutex);
  auto it = ctxs().find(igrid);
  return it == ctxs().end() ? -1 : static_cast<int>(it->second.state);
}

}  // extern "C"

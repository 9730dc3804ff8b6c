// Compares ww3gpu::libm (ww3-gpu_amd/csrc/devlibm.hpp, built here as host code) with the
// host glibc on every float input.  Usage: devlibm_host_check <fn> [threads]; prints the
// number of differing inputs and the first one.  Test harness, not part of the library.
#include <math.h>

#include <atomic>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "devlibm.hpp"

static inline uint32_t bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

int main(int argc, char** argv) {
  const char* fn = argc > 1 ? argv[1] : "expf";
  const int nthreads = argc > 2 ? std::atoi(argv[2]) : 8;
  std::atomic<uint64_t> bad{0};
  std::atomic<uint32_t> first{0xffffffffu};
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) {
    th.emplace_back([&, t] {
      uint64_t nb = 0;
      uint32_t f0 = 0xffffffffu;
      for (uint64_t u = t; u <= 0xffffffffull; u += nthreads) {
        float x;
        const uint32_t uu = static_cast<uint32_t>(u);
        std::memcpy(&x, &uu, 4);
        float a = 0, b = 0;
        if (std::strcmp(fn, "expf") == 0) {
          a = ::expf(x);
          b = ww3gpu::libm::expf(x);
        }
        if (bits(a) != bits(b) && !(a != a && b != b)) {
          ++nb;
          if (uu < f0) f0 = uu;
        }
      }
      bad += nb;
      uint32_t cur = first.load();
      while (f0 < cur && !first.compare_exchange_weak(cur, f0)) {
      }
    });
  }
  for (auto& x : th) x.join();
  std::printf("%s bad=%" PRIu64 " first=%08x\n", fn, bad.load(), first.load());
  return bad.load() == 0 ? 0 : 1;
}

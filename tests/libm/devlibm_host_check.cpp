// Compares the library's device math (ww3-gpu_amd/csrc/devlibm.hpp, built here as host
// code) with the oracle's restatement (oracle/wwlibm.h), with the correctly rounded value
// (the float rounding of the host's long double function) and with the host glibc.
// Test harness, not part of the library.
//
// Usage: devlibm_host_check <fn> [threads] [samples]
//   one-argument functions (expf logf log10f tanhf sinf cosf, pow025f = powf(x, 0.25), and the
//   short-chain forms logf_short sinf_short cosf_short pow025_short against the oracle's):
//   every float input;
//   two-argument functions (powf atan2f): `samples` random pairs plus structured ones.
// Prints one line: fn, inputs, dev!=oracle, dev!=cr, dev!=glibc, max ulp vs cr, vs glibc.
// Exit status 1 if the device and oracle definitions ever differ (for expf: if the device
// differs from glibc).
#include <math.h>

#include <atomic>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "devlibm.hpp"
#include "wwlibm.h"

namespace {

uint32_t bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
bool same(float a, float b) { return bits(a) == bits(b) || (a != a && b != b); }
// ulp distance on the monotone integer line of floats
int64_t ord(float f) {
  const int32_t i = static_cast<int32_t>(bits(f));
  return i < 0 ? static_cast<int64_t>(INT32_MIN) - i : i;
}
int64_t ulps(float a, float b) {
  if (a != a || b != b) return (a != a && b != b) ? 0 : INT64_MAX / 4;
  const int64_t d = ord(a) - ord(b);
  return d < 0 ? -d : d;
}

struct Eval {
  float dev, ora, cr, gl;
};

Eval eval1_fast(const std::string& fn, float x) {
  namespace L = ww3gpu::libm;
  if (fn == "expf") return {L::expf(x), ::expf(x), 0, 0};
  if (fn == "logf") return {L::logf(x), wl_logf(x), 0, 0};
  if (fn == "log10f") return {L::log10f(x), wl_log10f(x), 0, 0};
  if (fn == "tanhf") return {L::tanhf(x), wl_tanhf(x), 0, 0};
  if (fn == "pow025f") return {L::powf(x, 0.25f), wl_powf(x, 0.25f), 0, 0};
  if (fn == "logf_short") return {L::logf_short(x), wl_logf(x), 0, 0};
  if (fn == "pow025_short") return {L::pow_quarter(x), wl_powf(x, 0.25f), 0, 0};
  if (fn == "sinf_short" || fn == "cosf_short") {
    float s, c, so, co;
    L::sincosf_short(x, &s, &c);
    wl_sincosf(x, &so, &co);
    return fn == "sinf_short" ? Eval{s, so, 0, 0} : Eval{c, co, 0, 0};
  }
  float s, c, so, co;
  L::sincosf(x, &s, &c);
  wl_sincosf(x, &so, &co);
  if (fn == "sinf") return {s, so, 0, 0};
  return {c, co, 0, 0};
}

Eval eval1(const std::string& fn, float x) {
  namespace L = ww3gpu::libm;
  if (fn == "expf") return {L::expf(x), ::expf(x), (float)expl(x), ::expf(x)};
  if (fn == "logf") return {L::logf(x), wl_logf(x), (float)logl(x), ::logf(x)};
  if (fn == "log10f") return {L::log10f(x), wl_log10f(x), (float)log10l(x), ::log10f(x)};
  if (fn == "tanhf") return {L::tanhf(x), wl_tanhf(x), (float)tanhl(x), ::tanhf(x)};
  if (fn == "pow025f") return {L::powf(x, 0.25f), wl_powf(x, 0.25f), (float)powl(x, 0.25L), ::powf(x, 0.25f)};
  if (fn == "logf_short") return {L::logf_short(x), wl_logf(x), (float)logl(x), ::logf(x)};
  if (fn == "pow025_short") return {L::pow_quarter(x), wl_powf(x, 0.25f), (float)powl(x, 0.25L), ::powf(x, 0.25f)};
  if (fn == "sinf_short" || fn == "cosf_short") {
    float s, c, so, co;
    L::sincosf_short(x, &s, &c);
    wl_sincosf(x, &so, &co);
    return fn == "sinf_short" ? Eval{s, so, (float)sinl(x), ::sinf(x)} : Eval{c, co, (float)cosl(x), ::cosf(x)};
  }
  if (fn == "sinf") {
    float s, c, so, co;
    L::sincosf(x, &s, &c);
    wl_sincosf(x, &so, &co);
    return {s, so, (float)sinl(x), ::sinf(x)};
  }
  if (fn == "cosf") {
    float s, c, so, co;
    L::sincosf(x, &s, &c);
    wl_sincosf(x, &so, &co);
    return {c, co, (float)cosl(x), ::cosf(x)};
  }
  std::fprintf(stderr, "unknown function %s\n", fn.c_str());
  std::exit(2);
}

Eval eval2(const std::string& fn, float x, float y) {
  namespace L = ww3gpu::libm;
  if (fn == "powf") return {L::powf(x, y), wl_powf(x, y), (float)powl(x, y), ::powf(x, y)};
  if (fn == "atan2f") return {L::atan2f(x, y), wl_atan2f(x, y), (float)atan2l(x, y), ::atan2f(x, y)};
  std::fprintf(stderr, "unknown function %s\n", fn.c_str());
  std::exit(2);
}

struct Count {
  uint64_t n = 0, d_o = 0, d_cr = 0, d_gl = 0;
  int64_t u_cr = 0, u_gl = 0;
  uint32_t first_do = 0;
  void add(const Eval& e, uint32_t tag) {
    ++n;
    if (!same(e.dev, e.ora)) {
      if (d_o == 0) first_do = tag;
      ++d_o;
    }
    if (!same(e.dev, e.cr)) {
      ++d_cr;
      const int64_t u = ulps(e.dev, e.cr);
      if (u > u_cr) u_cr = u;
    }
    if (!same(e.dev, e.gl)) {
      ++d_gl;
      const int64_t u = ulps(e.dev, e.gl);
      if (u > u_gl) u_gl = u;
    }
  }
  void merge(const Count& o) {
    if (d_o == 0 && o.d_o) first_do = o.first_do;
    n += o.n; d_o += o.d_o; d_cr += o.d_cr; d_gl += o.d_gl;
    if (o.u_cr > u_cr) u_cr = o.u_cr;
    if (o.u_gl > u_gl) u_gl = o.u_gl;
  }
};

}  // namespace

int main(int argc, char** argv) {
  const std::string fn = argc > 1 ? argv[1] : "expf";
  const int nthreads = argc > 2 ? std::atoi(argv[2]) : 8;
  // two-argument functions: number of samples; one-argument functions: the stride of the
  // subset compared with the long double and glibc references (device vs oracle: all inputs)
  const long samples = argc > 3 ? std::atol(argv[3]) : (1L << 26);
  const uint64_t stride = (argc > 3 && fn != "powf" && fn != "atan2f") ? std::atol(argv[3]) : 1;
  // one-argument functions: visit every `every`-th float input (1 = all 2^32)
  const uint64_t every = argc > 4 ? std::atol(argv[4]) : 1;
  const bool binary = (fn == "powf" || fn == "atan2f");
  std::vector<Count> cnt(nthreads);
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) {
    th.emplace_back([&, t] {
      Count& c = cnt[t];
      if (!binary) {
        const uint64_t step = static_cast<uint64_t>(nthreads) * every;
        for (uint64_t u = static_cast<uint64_t>(t) * every; u <= 0xffffffffull; u += step) {
          float x;
          const uint32_t uu = static_cast<uint32_t>(u);
          std::memcpy(&x, &uu, 4);
          if (stride > 1 && (u / nthreads) % stride != 0) {
            // device vs oracle only (cheap); the long double and glibc references on a subset
            Eval e = eval1_fast(fn, x);
            e.cr = e.dev;
            e.gl = e.dev;
            c.add(e, uu);
          } else {
            c.add(eval1(fn, x), uu);
          }
        }
      } else {
        std::mt19937_64 rng(1234 + t);
        std::uniform_real_distribution<double> ex(-30.0, 30.0), un(-1.0, 1.0);
        for (long i = t; i < samples; i += nthreads) {
          float x, y;
          const int kind = static_cast<int>(i % 4);
          if (fn == "powf") {
            // x in (0, 1e6], y over the range the source terms use and beyond
            x = static_cast<float>(std::exp2(ex(rng) * 0.6));
            if (kind == 0) y = static_cast<float>(un(rng) * 4.0);
            else if (kind == 1) y = static_cast<float>(std::round(un(rng) * 8.0) * 0.25);
            else if (kind == 2) y = static_cast<float>(un(rng) * 30.0);
            else { uint32_t r = static_cast<uint32_t>(rng()); std::memcpy(&y, &r, 4); if (y != y) y = 1.5f; }
          } else {
            uint32_t r1 = static_cast<uint32_t>(rng()), r2 = static_cast<uint32_t>(rng());
            if (kind < 2) {
              x = static_cast<float>(un(rng) * std::exp2(ex(rng) * 0.3));
              y = static_cast<float>(un(rng) * std::exp2(ex(rng) * 0.3));
            } else {
              std::memcpy(&x, &r1, 4);
              std::memcpy(&y, &r2, 4);
            }
          }
          c.add(eval2(fn, x, y), static_cast<uint32_t>(i));
        }
      }
    });
  }
  for (auto& x : th) x.join();
  Count tot;
  for (auto& c : cnt) tot.merge(c);
  std::printf("%s n=%" PRIu64 " dev_ne_oracle=%" PRIu64 " dev_ne_cr=%" PRIu64 " dev_ne_glibc=%" PRIu64
              " maxulp_cr=%" PRId64 " maxulp_glibc=%" PRId64 " first_dev_ne_oracle=%08x\n",
              fn.c_str(), tot.n, tot.d_o, tot.d_cr, tot.d_gl, tot.u_cr, tot.u_gl, tot.first_do);
  if (fn == "expf") return tot.d_gl == 0 ? 0 : 1;
  return tot.d_o == 0 ? 0 : 1;
}

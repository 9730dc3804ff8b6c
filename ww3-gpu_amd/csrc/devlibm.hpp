// ww3-gpu_amd/csrc/devlibm.hpp
// Device math functions that return the host glibc's bits.
//
// The reference holds its float kernels bit-identical to a gfortran/glibc build
// (kokkos/src/ww_kokkos/libm_device.hpp:74-139 names which libm function each kernel
// site calls).  AMD's device libm (OCML) computes expf etc. by other algorithms, so a
// kernel that called it would differ from the host in the last bit on some inputs.
// These functions restate glibc's published algorithms instead, operation for
// operation, with every floating-point operation spelled out (fused multiply-adds are
// explicit fma() calls, everything else is built with -ffp-contract=off).  They are
// __host__ __device__ so the same code can be compared with the host glibc on every
// float input on the CPU, and on the GPU.
//
// expf: glibc 2.27+ sysdeps/ieee754/flt-32/e_expf.c, the ARM optimized-routines design
// (Szabolcs Nagy, 2017): x*N/ln2 = k + r with N = 32, 2^(k/N) from a 32-entry table of
// 2^(i/32) whose exponent bits are pre-shifted, 2^(r/N) by a cubic in double precision,
// one rounding to float at the end.  The table entries below are 2^(i/32) rounded to
// the nearest double, minus i<<47, computed from the mathematics.  glibc's x86-64 build
// runs its FMA variant on CPUs with FMA: the rounding of z = x*N/ln2 is fused into the
// shift-add and into r.  Checked equal to the host glibc 2.35 on all 2^32 inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ww3gpu {
namespace libm {

__host__ __device__ inline uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
__host__ __device__ inline uint64_t d2u(double d) { return __builtin_bit_cast(uint64_t, d); }
__host__ __device__ inline double u2d(uint64_t u) { return __builtin_bit_cast(double, u); }

struct ExpfData {
  uint64_t tab[32];
};

// 2^(i/32) rounded to double, minus (i << 47): adding (k << 47) for k = i mod 32 then
// gives the double 2^(k/32) directly.
__host__ __device__ inline uint64_t expf_tab(int i) {
  constexpr ExpfData d = {{
      0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
      0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
      0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
      0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
      0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
      0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
      0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
      0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull,
  }};
  return d.tab[i];
}

__host__ __device__ inline float expf(float x) {
  constexpr double kN = 32.0;
  constexpr double kInvLn2N = 0x1.71547652b82fep+0 * kN;
  constexpr double kShift = 0x1.8p+52;
  // 2^(r/N) ~ C0 r^3 + C1 r^2 + C2 r + 1, the published minimax coefficients scaled by N.
  constexpr double kC0 = 0x1.c6af84b912394p-5 / kN / kN / kN;
  constexpr double kC1 = 0x1.ebfce50fac4f3p-3 / kN / kN;
  constexpr double kC2 = 0x1.62e42ff0c52d6p-1 / kN;

  const uint32_t ux = f2u(x);
  const uint32_t abstop = (ux >> 20) & 0x7ffu;
  if (abstop >= (f2u(88.0f) >> 20)) {
    if (ux == f2u(-__builtin_inff())) return 0.0f;
    if (abstop >= (f2u(__builtin_inff()) >> 20)) return x + x;  // NaN or +inf
    if (x > 0x1.62e42ep6f) return __builtin_inff();             // overflow
    if (x < -0x1.9fe368p6f) return 0.0f;                        // underflow
  }
  const double xd = static_cast<double>(x);
  double kd = __builtin_fma(kInvLn2N, xd, kShift);
  const uint64_t ki = d2u(kd);
  kd -= kShift;
  const double r = __builtin_fma(kInvLn2N, xd, -kd);
  uint64_t t = expf_tab(static_cast<int>(ki % 32u));
  t += ki << (52 - 5);
  const double s = u2d(t);
  const double z = __builtin_fma(kC0, r, kC1);
  const double r2 = r * r;
  double y = __builtin_fma(kC2, r, 1.0);
  y = __builtin_fma(z, r2, y);
  y = y * s;
  return static_cast<float>(y);
}

}  // namespace libm
}  // namespace ww3gpu

// ww3-gpu_amd/csrc/devlibm.hpp
// The library's math functions, with bits defined independently of the device libm.
//
// The reference holds its float kernels bit-identical to a gfortran/glibc build
// (kokkos/src/ww_kokkos/libm_device.hpp:74-139 names which libm function each kernel
// site calls).  AMD's device libm (OCML) computes expf etc. by other algorithms, so a
// kernel that called it would differ from the host in the last bit on some inputs.
// expf restates glibc's published algorithm operation for operation and returns the
// host glibc's bits; the others (below) are defined here as near-correctly-rounded
// double-precision evaluations.  Every floating-point operation is spelled out (fused
// multiply-adds are explicit fma() calls, everything else is built with
// -ffp-contract=off).  The functions are __host__ __device__ so the same code can be
// compared on the CPU with the host glibc and the correctly rounded value on every float
// input, and on the GPU.
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

// ---------------------------------------------------------------------------------------
// The other float functions of the source-term step (logf, log10f, powf, atan2f, tanhf,
// sincosf).  glibc computes them from internal tables that only its binary holds, so they
// are not restated; instead this library defines them as the float rounding of a double
// precision evaluation accurate to a few double ulps, so the result is the correctly
// rounded float except when the exact value lies within ~2^-50 relative of a rounding
// boundary.  Every operation is a plain IEEE double operation (no fused multiply-add, no
// library call), so a host build of this code computes the same bits as the device.
// Algorithms (textbook, named per function): log by the atanh series of s = (m-1)/(m+1)
// after reduction to m in [sqrt(1/2), sqrt(2)); exp by reduction to r = x - k ln2/32 and a
// degree-7 Taylor polynomial times 2^(k/32) from the table above; expm1 by its Taylor
// series near zero; atan by reduction to |u| <= tan(pi/8) and its Maclaurin series;
// sin/cos by Cody-Waite reduction with a three-part pi/2 and their Taylor series.
// ---------------------------------------------------------------------------------------

namespace dp {

constexpr double kLn2Hi = 0x1.62e42fefa3800p-1;   // ln 2 truncated to 42 bits
constexpr double kLn2Lo = 0x1.ef35793c76730p-45;  // ln 2 - kLn2Hi
constexpr double kSqrt2 = 0x1.6a09e667f3bcdp+0;
constexpr double kInvLn10 = 0x1.bcb7b1526e50ep-2;
constexpr double kPi = 0x1.921fb54442d18p+1;
constexpr double kPio2 = 0x1.921fb54442d18p+0;
constexpr double kPio4 = 0x1.921fb54442d18p-1;
constexpr double kTanPi8 = 0x1.a827999fcef32p-2;
constexpr double kInvPio2 = 0x1.45f306dc9c883p-1;
constexpr double kPio2_1 = 0x1.921fb54400000p+0;   // pi/2, first 33 bits
constexpr double kPio2_2 = 0x1.0b4611a600000p-34;  // next 33 bits
constexpr double kPio2_3 = 0x1.3198a2e037073p-69;  // the rest
constexpr double kShift = 0x1.8p+52;

// 1/(2n+1), n = 1..22, and 1/n!, n = 2..16, each the double nearest the exact value.
struct Coef {
  double inv_odd[22];
  double inv_fact[15];
};
__host__ __device__ inline const Coef& coef() {
  static constexpr Coef c = {
      {0x1.5555555555555p-2, 0x1.999999999999ap-3, 0x1.2492492492492p-3, 0x1.c71c71c71c71cp-4,
       0x1.745d1745d1746p-4, 0x1.3b13b13b13b14p-4, 0x1.1111111111111p-4, 0x1.e1e1e1e1e1e1ep-5,
       0x1.af286bca1af28p-5, 0x1.8618618618618p-5, 0x1.642c8590b2164p-5, 0x1.47ae147ae147bp-5,
       0x1.2f684bda12f68p-5, 0x1.1a7b9611a7b96p-5, 0x1.0842108421084p-5, 0x1.f07c1f07c1f08p-6,
       0x1.d41d41d41d41dp-6, 0x1.bacf914c1bad0p-6, 0x1.a41a41a41a41ap-6, 0x1.8f9c18f9c18fap-6,
       0x1.7d05f417d05f4p-6, 0x1.6c16c16c16c17p-6},
      {0x1.0000000000000p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5, 0x1.1111111111111p-7,
       0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19,
       0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33,
       0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41, 0x1.ae7f3e733b81fp-45}};
  return c;
}
// 1/n! for n >= 2
__host__ __device__ inline double inv_fact(int n) { return coef().inv_fact[n - 2]; }

/// ln(x) for a positive normal double.
__host__ __device__ inline double log(double x) {
  const uint64_t u = d2u(x);
  int k = static_cast<int>((u >> 52) & 0x7ff) - 1023;
  double m = u2d((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
  if (m > kSqrt2) {
    m = m * 0.5;
    k = k + 1;
  }
  const double f = m - 1.0;  // exact
  const double s = f / (2.0 + f);
  const double s2 = s * s;
  double p = coef().inv_odd[21];
  for (int n = 20; n >= 0; --n) p = coef().inv_odd[n] + s2 * p;
  const double lm = 2.0 * s + 2.0 * s * (s2 * p);
  const double kd = static_cast<double>(k);
  return kd * kLn2Hi + (kd * kLn2Lo + lm);
}

/// e^x for -708 < x < 709.
__host__ __device__ inline double exp(double x) {
  constexpr double kInvLn2N = 0x1.71547652b82fep+5;   // 32/ln2
  constexpr double kLn2NHi = 0x1.62e42fefa3800p-6;    // ln2/32, 42 bits
  constexpr double kLn2NLo = 0x1.ef35793c76730p-50;
  double kd = x * kInvLn2N + kShift;
  const uint64_t ki = d2u(kd);
  kd = kd - kShift;
  const double r = (x - kd * kLn2NHi) - kd * kLn2NLo;
  double p = inv_fact(7);
  for (int n = 6; n >= 2; --n) p = inv_fact(n) + r * p;
  const double er = 1.0 + (r + (r * r) * p);
  const double s = u2d(expf_tab(static_cast<int>(ki % 32u)) + (ki << 47));
  return s * er;
}

/// e^x - 1 for |x| < 0.35 by its Taylor series (degree 16), else exp(x) - 1.
__host__ __device__ inline double expm1(double x) {
  if (x > 0.35 || x < -0.35) return dp::exp(x) - 1.0;
  double p = inv_fact(16);
  for (int n = 15; n >= 2; --n) p = inv_fact(n) + x * p;
  return x + (x * x) * p;
}

/// atan(u) for |u| <= tan(pi/8), Maclaurin series to u^45.
__host__ __device__ inline double atan_small(double u) {
  const double u2 = u * u;
  double p = coef().inv_odd[21];
  for (int n = 20; n >= 0; --n) p = (n % 2 == 0 ? -1.0 : 1.0) * coef().inv_odd[n] + u2 * p;
  // p = -1/3 + u2/5 - u2^2/7 + ... ; atan(u) = u + u*u2*p
  return u + u * (u2 * p);
}

/// atan(t) for t >= 0.
__host__ __device__ inline double atan_pos(double t) {
  bool inv = false;
  if (t > 1.0) {
    t = 1.0 / t;
    inv = true;
  }
  double a;
  if (t > kTanPi8) {
    a = kPio4 + atan_small((t - 1.0) / (t + 1.0));
  } else {
    a = atan_small(t);
  }
  return inv ? kPio2 - a : a;
}

/// Reduction of a float |x| >= 2^20 by pi/2 with the bits of 2/pi (Payne-Hanek): returns
/// r in [-pi/4, pi/4] and the quadrant q with |x| = (4n + q) pi/2 + r.
__host__ __device__ inline uint32_t two_over_pi_bits(int pos) {  // bits pos..pos+31, 1-based
  constexpr uint32_t w[12] = {0xa2f9836eu, 0x4e441529u, 0xfc2757d1u, 0xf534ddc0u,
                              0xdb629599u, 0x3c439041u, 0xfe5163abu, 0xdebbc561u,
                              0xb7246e3au, 0x424dd2e0u, 0x06492eeau, 0x09d1921cu};
  const int k = (pos - 1) / 32, o = (pos - 1) % 32;
  return o == 0 ? w[k] : ((w[k] << o) | (w[k + 1] >> (32 - o)));
}
__host__ __device__ inline double reduce_big(float ax, int* q) {
  const uint32_t ix = f2u(ax);
  const int ex = static_cast<int>(ix >> 23) - 150;  // ax = mi * 2^ex
  const uint64_t mi = (ix & 0x7fffffu) | 0x800000u;
  const int j0 = ex - 1 > 1 ? ex - 1 : 1;
  const unsigned __int128 b = (static_cast<unsigned __int128>(two_over_pi_bits(j0)) << 64) |
                              (static_cast<unsigned __int128>(two_over_pi_bits(j0 + 32)) << 32) |
                              static_cast<unsigned __int128>(two_over_pi_bits(j0 + 64));
  const unsigned __int128 p = b * mi;
  const int sh = j0 + 95 - ex;
  int qq = static_cast<int>(static_cast<uint64_t>(p >> sh) & 3u);
  const unsigned __int128 frac = p & ((static_cast<unsigned __int128>(1) << sh) - 1u);
  double f = static_cast<double>(static_cast<uint64_t>(frac >> (sh - 64))) * 0x1p-64;
  if (f >= 0.5) {
    f = f - 1.0;
    qq = qq + 1;
  }
  *q = qq & 3;
  return f * kPio2;
}

/// sin and cos of r in [-pi/4, pi/4] rotated by quadrant q.
__host__ __device__ inline void sincos_q(double r, int q, double* sn, double* cs);

/// sin and cos of x, Cody-Waite reduction by pi/2 for |x| < 2^20, Payne-Hanek above.
__host__ __device__ inline void sincos(double x, double* sn, double* cs) {
  const double ax = x < 0.0 ? -x : x;
  if (ax >= 0x1p20) {
    int q;
    const double r = reduce_big(static_cast<float>(ax), &q);
    double s, c;
    sincos_q(r, q, &s, &c);
    *sn = x < 0.0 ? -s : s;
    *cs = c;
    return;
  }
  double kd = x * kInvPio2 + kShift;
  const uint64_t ki = d2u(kd);
  kd = kd - kShift;
  const double r = ((x - kd * kPio2_1) - kd * kPio2_2) - kd * kPio2_3;
  sincos_q(r, static_cast<int>(ki & 3u), sn, cs);
}

__host__ __device__ inline void sincos_q(double r, int q, double* sn, double* cs) {
  const double r2 = r * r;
  // sin r = r - r^3/3! + ... - r^19/19!;  cos r = 1 - r^2/2! + ... + r^20/20!
  double ps = 0.0, pc = 0.0;
  {
    // odd/even factorials from 1/n! (n up to 16) extended: 1/17! .. 1/20! by division.
    const double f17 = inv_fact(16) / 17.0, f18 = f17 / 18.0, f19 = f18 / 19.0, f20 = f19 / 20.0;
    ps = -f19;
    ps = f17 + r2 * ps;
    for (int n = 15; n >= 3; n -= 2) ps = ((n % 4 == 3) ? -inv_fact(n) : inv_fact(n)) + r2 * ps;
    pc = f20;
    pc = -f18 + r2 * pc;
    for (int n = 16; n >= 2; n -= 2) pc = ((n % 4 == 2) ? -inv_fact(n) : inv_fact(n)) + r2 * pc;
  }
  const double s = r + r * (r2 * ps);
  const double c = 1.0 + r2 * pc;
  switch (q) {
    case 0: *sn = s; *cs = c; break;
    case 1: *sn = c; *cs = -s; break;
    case 2: *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s; break;
  }
}

}  // namespace dp

__host__ __device__ inline bool isnan_f(float x) { return (f2u(x) & 0x7fffffffu) > 0x7f800000u; }
__host__ __device__ inline bool isinf_f(float x) { return (f2u(x) & 0x7fffffffu) == 0x7f800000u; }

__host__ __device__ inline float logf(float x) {
  if (isnan_f(x)) return x + x;
  if (x < 0.0f) return __builtin_nanf("");
  if (x == 0.0f) return -__builtin_inff();
  if (isinf_f(x)) return x;
  return static_cast<float>(dp::log(static_cast<double>(x)));
}

__host__ __device__ inline float log10f(float x) {
  if (isnan_f(x)) return x + x;
  if (x < 0.0f) return __builtin_nanf("");
  if (x == 0.0f) return -__builtin_inff();
  if (isinf_f(x)) return x;
  return static_cast<float>(dp::log(static_cast<double>(x)) * dp::kInvLn10);
}

/// x^y with C99/IEEE special cases.  y == 2 is x*x and y == 0.5 is sqrt(x) for x >= +0,
/// x finite (both correctly rounded); otherwise exp(y ln|x|) in double.
__host__ __device__ inline float powf(float x, float y) {
  if (y == 0.0f || x == 1.0f) return 1.0f;
  if (isnan_f(x) || isnan_f(y)) return x + y;
  const float ay = y < 0.0f ? -y : y;
  // is y an integer, and odd?
  bool yint = false, yodd = false;
  if (ay >= 0x1p24f) {
    yint = true;
  } else if (ay >= 1.0f) {
    const float t = static_cast<float>(static_cast<int>(ay));
    yint = (t == ay);
    yodd = yint && (static_cast<int>(ay) & 1);
  }
  if (isinf_f(y)) {
    const float ax = x < 0.0f ? -x : x;
    if (ax == 1.0f) return 1.0f;
    const bool big = ax > 1.0f;
    return (big == (y > 0.0f)) ? __builtin_inff() : 0.0f;
  }
  if (x == 0.0f) {
    if (y > 0.0f) return yodd ? x : 0.0f;
    return yodd ? (f2u(x) >> 31 ? -__builtin_inff() : __builtin_inff()) : __builtin_inff();
  }
  if (isinf_f(x)) {
    const bool neg = x < 0.0f;
    if (y > 0.0f) return (neg && yodd) ? -__builtin_inff() : __builtin_inff();
    return (neg && yodd) ? -0.0f : 0.0f;
  }
  if (x < 0.0f && !yint) return __builtin_nanf("");
  if (x > 0.0f && y == 2.0f) return x * x;
  if (x > 0.0f && y == 0.5f) return __builtin_sqrtf(x);
  const double lx = dp::log(static_cast<double>(x < 0.0f ? -x : x));
  const double t = static_cast<double>(y) * lx;
  double r;
  if (t > 89.0) {
    r = __builtin_inf();
  } else if (t < -104.0) {
    r = 0.0;
  } else {
    r = dp::exp(t);
  }
  const float fr = static_cast<float>(r);
  return (x < 0.0f && yodd) ? -fr : fr;
}

__host__ __device__ inline float atan2f(float y, float x) {
  if (isnan_f(x) || isnan_f(y)) return x + y;
  const bool yneg = (f2u(y) >> 31) != 0;
  const bool xneg = (f2u(x) >> 31) != 0;
  double a;
  if (y == 0.0f) {
    a = xneg ? dp::kPi : 0.0;
  } else if (x == 0.0f) {
    a = dp::kPio2;
  } else if (isinf_f(x) && isinf_f(y)) {
    a = xneg ? 3.0 * dp::kPio4 : dp::kPio4;
  } else if (isinf_f(x)) {
    a = xneg ? dp::kPi : 0.0;
  } else if (isinf_f(y)) {
    a = dp::kPio2;
  } else {
    const double yd = static_cast<double>(y < 0.0f ? -y : y);
    const double xd = static_cast<double>(x < 0.0f ? -x : x);
    a = dp::atan_pos(yd / xd);
    if (xneg) a = dp::kPi - a;
  }
  const float fa = static_cast<float>(a);
  return yneg ? -fa : fa;
}

__host__ __device__ inline float tanhf(float x) {
  if (isnan_f(x) || x == 0.0f) return x + x;
  const double xd = static_cast<double>(x);
  const double a = xd < 0.0 ? -xd : xd;
  double r;
  if (a > 20.0) {
    r = 1.0;
  } else if (a < 0x1p-28) {
    r = a;
  } else {
    const double t = dp::expm1(2.0 * a);
    r = t / (t + 2.0);
  }
  const float fr = static_cast<float>(r);
  return x < 0.0f ? -fr : fr;
}

__host__ __device__ inline void sincosf(float x, float* s, float* c) {
  if (isnan_f(x) || isinf_f(x)) {
    *s = __builtin_nanf("");
    *c = *s;
    return;
  }
  double sd, cd;
  dp::sincos(static_cast<double>(x), &sd, &cd);
  *s = static_cast<float>(sd);
  *c = static_cast<float>(cd);
}

__host__ __device__ inline float cosf(float x) {
  float s, c;
  libm::sincosf(x, &s, &c);
  return c;
}

// ---------------------------------------------------------------------------------------
// Short-chain forms for the serial chains of the source terms (W3SIN4's stress scan and tail
// integral, CALC_USTAR), where each call's latency adds to a point's sub-step.  Each returns
// the same float as the function it stands for on every input: both are correctly rounded on
// all 2^32 float inputs (tests/test_libm_cpu.py compares them exhaustively with
// WW3GPU_LIBM_EXHAUSTIVE=1 and on every 16th input by default), so both equal the correctly
// rounded value.  The evaluations: ln by the same atanh series cut to s^18/19 (truncation
// below 2^-55) as an fma Horner chain; sin/cos by the same reduction and the Taylor series to
// r^15 / r^16 (below 2^-54) in fma chains; x^(1/4) for x > 0 as sqrt(sqrt(x)) in double.
// ---------------------------------------------------------------------------------------
namespace shortc {

__host__ __device__ inline double log(double x) {
  const uint64_t u = d2u(x);
  int k = static_cast<int>((u >> 52) & 0x7ff) - 1023;
  double m = u2d((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
  if (m > dp::kSqrt2) {
    m = m * 0.5;
    k = k + 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double s2 = s * s;
  double p = dp::coef().inv_odd[8];  // 1/19
#pragma unroll
  for (int n = 7; n >= 0; --n) p = __builtin_fma(p, s2, dp::coef().inv_odd[n]);
  const double s2s = 2.0 * s;
  const double lm = __builtin_fma(s2s, s2 * p, s2s);
  const double kd = static_cast<double>(k);
  return __builtin_fma(kd, dp::kLn2Hi, __builtin_fma(kd, dp::kLn2Lo, lm));
}

}  // namespace shortc

/// logf(x), the same value on every input.
__host__ __device__ inline float logf_short(float x) {
  if (isnan_f(x)) return x + x;
  if (x < 0.0f) return __builtin_nanf("");
  if (x == 0.0f) return -__builtin_inff();
  if (isinf_f(x)) return x;
  return static_cast<float>(shortc::log(static_cast<double>(x)));
}

/// powf(x, 0.25f), the same value on every input.
__host__ __device__ inline float pow_quarter(float x) {
  if (x > 0.0f && !isinf_f(x)) return static_cast<float>(__builtin_sqrt(__builtin_sqrt(static_cast<double>(x))));
  return libm::powf(x, 0.25f);
}

/// sincosf(x), the same values on every input.
__host__ __device__ inline void sincosf_short(float x, float* s, float* c) {
  if (isnan_f(x) || isinf_f(x) || (x < 0.0f ? -x : x) >= 0x1p20f) {
    libm::sincosf(x, s, c);
    return;
  }
  const double xd = static_cast<double>(x);
  double kd = xd * dp::kInvPio2 + dp::kShift;
  const uint64_t ki = d2u(kd);
  kd = kd - dp::kShift;
  const double r = ((xd - kd * dp::kPio2_1) - kd * dp::kPio2_2) - kd * dp::kPio2_3;
  const double r2 = r * r;
  double ps = -dp::inv_fact(15), pc = dp::inv_fact(16);
#pragma unroll
  for (int n = 13; n >= 3; n -= 2) ps = __builtin_fma(ps, r2, (n % 4 == 3) ? -dp::inv_fact(n) : dp::inv_fact(n));
#pragma unroll
  for (int n = 14; n >= 2; n -= 2) pc = __builtin_fma(pc, r2, (n % 4 == 2) ? -dp::inv_fact(n) : dp::inv_fact(n));
  const double sr = __builtin_fma(r, r2 * ps, r);
  const double cr = __builtin_fma(r2, pc, 1.0);
  double sn, cs;
  switch (static_cast<int>(ki & 3u)) {
    case 0: sn = sr; cs = cr; break;
    case 1: sn = cr; cs = -sr; break;
    case 2: sn = -sr; cs = -cr; break;
    default: sn = -cr; cs = sr; break;
  }
  *s = static_cast<float>(sn);
  *c = static_cast<float>(cs);
}

}  // namespace libm
}  // namespace ww3gpu

/*
 * oracle/wwlibm.h -- the oracle's math functions.  TEST INFRASTRUCTURE ONLY.
 *
 * The reference's kernels call the host glibc (kokkos/src/ww_kokkos/libm_device.hpp:74-139
 * lists the function at each site).  The oracle does the same for expf, whose glibc
 * algorithm the GPU library reproduces bit for bit, and for the two double functions of
 * W3SDB1 (exp, tanh).  For logf, log10f, powf, atan2f, tanhf and sincosf the GPU library
 * defines its own functions: the float rounding of a double-precision evaluation (log by
 * the atanh series, exp by 2^(k/32) times a Taylor polynomial, expm1 by Taylor, atan by
 * reduction to |u| <= tan(pi/8) and Maclaurin, sin/cos by Cody-Waite reduction and
 * Taylor).  This file restates those definitions in plain C so the oracle computes the
 * same bits; how far each one is from the correctly rounded value and from host glibc is
 * measured by tests/test_libm_cpu.py (all 2^32 inputs for the one-argument functions).
 *
 * Built with -ffp-contract=off: every operation below is one IEEE double operation.
 */
#ifndef WW3GPU_ORACLE_WWLIBM_H
#define WW3GPU_ORACLE_WWLIBM_H

#include <math.h>
#include <stdint.h>
#include <string.h>

static inline uint64_t wl_d2u(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }
static inline double wl_u2d(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
static inline uint32_t wl_f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* 2^(i/32) rounded to double, less i<<47 (the table of glibc's expf). */
static const uint64_t wl_exp2_32[32] = {
    0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
    0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
    0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
    0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
    0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
    0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
    0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
    0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull};

/* 1/(2n+3), n = 0..21 */
static const double wl_odd[22] = {
    0x1.5555555555555p-2, 0x1.999999999999ap-3, 0x1.2492492492492p-3, 0x1.c71c71c71c71cp-4,
    0x1.745d1745d1746p-4, 0x1.3b13b13b13b14p-4, 0x1.1111111111111p-4, 0x1.e1e1e1e1e1e1ep-5,
    0x1.af286bca1af28p-5, 0x1.8618618618618p-5, 0x1.642c8590b2164p-5, 0x1.47ae147ae147bp-5,
    0x1.2f684bda12f68p-5, 0x1.1a7b9611a7b96p-5, 0x1.0842108421084p-5, 0x1.f07c1f07c1f08p-6,
    0x1.d41d41d41d41dp-6, 0x1.bacf914c1bad0p-6, 0x1.a41a41a41a41ap-6, 0x1.8f9c18f9c18fap-6,
    0x1.7d05f417d05f4p-6, 0x1.6c16c16c16c17p-6};
/* 1/n!, index n = 0..16 (entries 0 and 1 unused) */
static const double wl_fact[17] = {
    1.0, 1.0, 0x1.0000000000000p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5,
    0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16,
    0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29,
    0x1.6124613a86d09p-33, 0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41, 0x1.ae7f3e733b81fp-45};

static inline double wl_log(double x) {
  uint64_t u = wl_d2u(x);
  int k = (int)((u >> 52) & 0x7ff) - 1023;
  double m = wl_u2d((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
  if (m > 0x1.6a09e667f3bcdp+0) { m = m * 0.5; k = k + 1; }
  double f = m - 1.0, s = f / (2.0 + f), s2 = s * s, p = wl_odd[21];
  for (int n = 20; n >= 0; --n) p = wl_odd[n] + s2 * p;
  double lm = 2.0 * s + 2.0 * s * (s2 * p);
  double kd = (double)k;
  return kd * 0x1.62e42fefa3800p-1 + (kd * 0x1.ef35793c76730p-45 + lm);
}

static inline double wl_exp(double x) {
  double kd = x * 0x1.71547652b82fep+5 + 0x1.8p+52;
  uint64_t ki = wl_d2u(kd);
  kd = kd - 0x1.8p+52;
  double r = (x - kd * 0x1.62e42fefa3800p-6) - kd * 0x1.ef35793c76730p-50;
  double p = wl_fact[7];
  for (int n = 6; n >= 2; --n) p = wl_fact[n] + r * p;
  double er = 1.0 + (r + (r * r) * p);
  double s = wl_u2d(wl_exp2_32[ki % 32u] + (ki << 47));
  return s * er;
}

static inline double wl_expm1(double x) {
  if (x > 0.35 || x < -0.35) return wl_exp(x) - 1.0;
  double p = wl_fact[16];
  for (int n = 15; n >= 2; --n) p = wl_fact[n] + x * p;
  return x + (x * x) * p;
}

static inline double wl_atan_small(double u) {
  double u2 = u * u, p = wl_odd[21];
  for (int n = 20; n >= 0; --n) p = ((n & 1) ? wl_odd[n] : -wl_odd[n]) + u2 * p;
  return u + u * (u2 * p);
}

static inline double wl_atan_pos(double t) {
  int inv = 0;
  if (t > 1.0) { t = 1.0 / t; inv = 1; }
  double a = (t > 0x1.a827999fcef32p-2)
                 ? 0x1.921fb54442d18p-1 + wl_atan_small((t - 1.0) / (t + 1.0))
                 : wl_atan_small(t);
  return inv ? 0x1.921fb54442d18p+0 - a : a;
}

static inline uint32_t wl_2opi(int pos) {
  static const uint32_t w[12] = {0xa2f9836eu, 0x4e441529u, 0xfc2757d1u, 0xf534ddc0u,
                                 0xdb629599u, 0x3c439041u, 0xfe5163abu, 0xdebbc561u,
                                 0xb7246e3au, 0x424dd2e0u, 0x06492eeau, 0x09d1921cu};
  int k = (pos - 1) / 32, o = (pos - 1) % 32;
  return o == 0 ? w[k] : ((w[k] << o) | (w[k + 1] >> (32 - o)));
}

/* |x| >= 2^20: x*2/pi mod 4 from the bits of 2/pi (Payne-Hanek). */
static inline double wl_reduce_big(float ax, int* q) {
  uint32_t ix = wl_f2u(ax);
  int ex = (int)(ix >> 23) - 150;
  uint64_t mi = (ix & 0x7fffffu) | 0x800000u;
  int j0 = ex - 1 > 1 ? ex - 1 : 1;
  unsigned __int128 b = ((unsigned __int128)wl_2opi(j0) << 64) | ((unsigned __int128)wl_2opi(j0 + 32) << 32) |
                        (unsigned __int128)wl_2opi(j0 + 64);
  unsigned __int128 p = b * mi;
  int sh = j0 + 95 - ex;
  int qq = (int)((uint64_t)(p >> sh) & 3u);
  unsigned __int128 frac = p & (((unsigned __int128)1 << sh) - 1u);
  double f = (double)(uint64_t)(frac >> (sh - 64)) * 0x1p-64;
  if (f >= 0.5) { f = f - 1.0; qq = qq + 1; }
  *q = qq & 3;
  return f * 0x1.921fb54442d18p+0;
}

static inline void wl_sincos_q(double r, int q, double* sn, double* cs) {
  double r2 = r * r;
  double f17 = wl_fact[16] / 17.0, f18 = f17 / 18.0, f19 = f18 / 19.0, f20 = f19 / 20.0;
  double ps = -f19;
  ps = f17 + r2 * ps;
  for (int n = 15; n >= 3; n -= 2) ps = ((n % 4 == 3) ? -wl_fact[n] : wl_fact[n]) + r2 * ps;
  double pc = f20;
  pc = -f18 + r2 * pc;
  for (int n = 16; n >= 2; n -= 2) pc = ((n % 4 == 2) ? -wl_fact[n] : wl_fact[n]) + r2 * pc;
  double s = r + r * (r2 * ps), c = 1.0 + r2 * pc;
  switch (q) {
    case 0: *sn = s; *cs = c; break;
    case 1: *sn = c; *cs = -s; break;
    case 2: *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s; break;
  }
}

static inline void wl_sincos(double x, double* sn, double* cs) {
  double ax = fabs(x);
  if (ax >= 0x1p20) {
    int q;
    double r = wl_reduce_big((float)ax, &q), s, c;
    wl_sincos_q(r, q, &s, &c);
    *sn = x < 0.0 ? -s : s;
    *cs = c;
    return;
  }
  double kd = x * 0x1.45f306dc9c883p-1 + 0x1.8p+52;
  uint64_t ki = wl_d2u(kd);
  kd = kd - 0x1.8p+52;
  double r = ((x - kd * 0x1.921fb54400000p+0) - kd * 0x1.0b4611a600000p-34) - kd * 0x1.3198a2e037073p-69;
  wl_sincos_q(r, (int)(ki & 3u), sn, cs);
}

static inline int wl_isnan(float x) { return (wl_f2u(x) & 0x7fffffffu) > 0x7f800000u; }
static inline int wl_isinf(float x) { return (wl_f2u(x) & 0x7fffffffu) == 0x7f800000u; }

static inline float wl_logf(float x) {
  if (wl_isnan(x)) return x + x;
  if (x < 0.0f) return NAN;
  if (x == 0.0f) return -INFINITY;
  if (wl_isinf(x)) return x;
  return (float)wl_log((double)x);
}

static inline float wl_log10f(float x) {
  if (wl_isnan(x)) return x + x;
  if (x < 0.0f) return NAN;
  if (x == 0.0f) return -INFINITY;
  if (wl_isinf(x)) return x;
  return (float)(wl_log((double)x) * 0x1.bcb7b1526e50ep-2);
}

static inline float wl_powf(float x, float y) {
  if (y == 0.0f || x == 1.0f) return 1.0f;
  if (wl_isnan(x) || wl_isnan(y)) return x + y;
  float ay = fabsf(y);
  int yint = 0, yodd = 0;
  if (ay >= 0x1p24f) {
    yint = 1;
  } else if (ay >= 1.0f) {
    int iy = (int)ay;
    yint = ((float)iy == ay);
    yodd = yint && (iy & 1);
  }
  if (wl_isinf(y)) {
    float ax = fabsf(x);
    if (ax == 1.0f) return 1.0f;
    return ((ax > 1.0f) == (y > 0.0f)) ? INFINITY : 0.0f;
  }
  if (x == 0.0f) {
    if (y > 0.0f) return yodd ? x : 0.0f;
    return yodd ? ((wl_f2u(x) >> 31) ? -INFINITY : INFINITY) : INFINITY;
  }
  if (wl_isinf(x)) {
    int neg = x < 0.0f;
    if (y > 0.0f) return (neg && yodd) ? -INFINITY : INFINITY;
    return (neg && yodd) ? -0.0f : 0.0f;
  }
  if (x < 0.0f && !yint) return NAN;
  if (x > 0.0f && y == 2.0f) return x * x;
  if (x > 0.0f && y == 0.5f) return sqrtf(x);
  double t = (double)y * wl_log((double)fabsf(x));
  double r = t > 89.0 ? INFINITY : (t < -104.0 ? 0.0 : wl_exp(t));
  float fr = (float)r;
  return (x < 0.0f && yodd) ? -fr : fr;
}

static inline float wl_atan2f(float y, float x) {
  if (wl_isnan(x) || wl_isnan(y)) return x + y;
  int yneg = (wl_f2u(y) >> 31) != 0, xneg = (wl_f2u(x) >> 31) != 0;
  double a;
  if (y == 0.0f) a = xneg ? 0x1.921fb54442d18p+1 : 0.0;
  else if (x == 0.0f) a = 0x1.921fb54442d18p+0;
  else if (wl_isinf(x) && wl_isinf(y)) a = xneg ? 3.0 * 0x1.921fb54442d18p-1 : 0x1.921fb54442d18p-1;
  else if (wl_isinf(x)) a = xneg ? 0x1.921fb54442d18p+1 : 0.0;
  else if (wl_isinf(y)) a = 0x1.921fb54442d18p+0;
  else {
    a = wl_atan_pos((double)fabsf(y) / (double)fabsf(x));
    if (xneg) a = 0x1.921fb54442d18p+1 - a;
  }
  float fa = (float)a;
  return yneg ? -fa : fa;
}

static inline float wl_tanhf(float x) {
  if (wl_isnan(x) || x == 0.0f) return x + x;
  double a = fabs((double)x), r;
  if (a > 20.0) r = 1.0;
  else if (a < 0x1p-28) r = a;
  else { double t = wl_expm1(2.0 * a); r = t / (t + 2.0); }
  float fr = (float)r;
  return x < 0.0f ? -fr : fr;
}

static inline void wl_sincosf(float x, float* s, float* c) {
  if (wl_isnan(x) || wl_isinf(x)) { *s = NAN; *c = NAN; return; }
  double sd, cd;
  wl_sincos((double)x, &sd, &cd);
  *s = (float)sd;
  *c = (float)cd;
}

static inline float wl_cosf(float x) { float s, c; wl_sincosf(x, &s, &c); return c; }

#endif

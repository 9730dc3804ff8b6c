// ww3-gpu_amd/csrc/srce_group.hip
// W3SRCE with the ST4 source terms on gfx950: one 16-lane group (a DPP row) per point.
//
// Every per-point statement of the reference keeps its float32 operation order
// (-ffp-contract=off), its sums in IK/ITH order, and its MIN/MAX operand order:
//   W3SPR4 + CALC_USTAR  kokkos/src/ww_kokkos/st4_spr4.hpp:40-162, st4_calc_ustar.hpp:36-122
//   W3SIN4               kokkos/src/ww_kokkos/st4_sin4.hpp:69-277, 472-627
//   W3SDS4               kokkos/src/ww_kokkos/st4_sds4.hpp:122-331, 574-711
//   W3SLN1, W3SBT1, W3SDB1  srce_ln1.hpp:41-99, srce_bt1.hpp:26-58, srce_db1.hpp:44-149
//   W3SRCE               kokkos/src/ww_kokkos/srce_kernel.hpp:212-583
//
// Lanes and order.  The 16 lanes of a point share its per-bin work.  A sum the reference
// forms in bin order is formed the same way here: each lane computes the terms of its bins,
// then every lane adds the row's terms in bin order, lane k's term arriving through a
// row_newbcast:k DPP operand, so all 16 lanes hold the same, reference-ordered result.
// Per-bin work uses the flat mapping (bin b on lane b % 16) unless a band's directions must
// be visited in order inside one sub-step of a scan (W3SIN4's stress scan, the tail copy),
// which use the band mapping (direction ith on lane ith % 16).
//
// Shared data.  A point's spectrum lives in LDS (X) while its directional gathers run; the
// same buffer then holds BRLAMBDA for the cumulative term, then the updated spectrum.  Lanes
// exchange data through X, the point's small LDS block (W) and the point's global arrays;
// all exchanges stay inside one wavefront, whose LDS and vector-memory operations complete in
// program order, so a compiler barrier (wave_sync) is the only ordering needed.
//
// Sub-step structure (each kernel launch is one of these):
//   k_init  W3SRCE section 1, then the first sub-step's input terms (W3SIN4) and KDMEAN.
//   [DIA]   W3SNL1 of the points still integrating (snl1_*.hip).
//   k_step  W3SLN1, W3SDS4, W3SDB1, W3SBT1, the time-step limit, the update, the fluxes,
//           W3SPR4, FHIGH, the limiter of the mean energy, the tail, W3SIN4.  That last
//           W3SIN4 is the next sub-step's first: nothing it reads changes in between under
//           SINTAILPAR(4) > 0.5; otherwise W3SPR4 runs after it and the next sub-step's
//           W3SIN4 runs again at the end of the same launch.
//   k_finish W3SRCE section 3.
//
// Fused or reordered, same arithmetic: W3SLN1's S and W3SBT1's S and D are recomputed where
// read; the isotropic sums of W3SDS4's cumulative term (SSDSC(3) > 0) are recomputed per
// direction; sincosf(USDIR) of W3SDS4's turbulence loop is computed once; W3SDS4's
// mean-square-slope block is skipped (nothing reads MSSSUM under SDSBCHOICE = 1); the
// whitecap sums run after the cumulative term, which does not touch their inputs.  The
// time-step limit's MIN over bins is formed in lane order: its operands are never NaN or
// -0 (each is AFAC / MAX(., 1E-10) >= +0), so the order cannot change the result.
#include <algorithm>
#include <atomic>

#include "devlibm.hpp"
#include "srce.hpp"

namespace ww3gpu {
namespace srce {
namespace {


constexpr int kTailJ = 50;       // W3SIN4's tail integration steps
constexpr int kTailStride = kTailJ + 2;  // row length of its per-step arrays in LDS

constexpr float kGrav = 9.806f;
constexpr float kDwat = 1000.f;
constexpr float kNuAir = 1.4E-5f;
constexpr float kKappa = 0.40f;
constexpr float kPi = static_cast<float>(3.141592653589793);
constexpr float kTpi = 2.0f * kPi;
constexpr float kTpiInv = 1.0f / kTpi;

__device__ inline float mx(float a, float b) { return a > b ? a : b; }
__device__ inline float mn(float a, float b) { return a < b ? a : b; }
__device__ inline int mxi(int a, int b) { return a > b ? a : b; }
__device__ inline int mni(int a, int b) { return a < b ? a : b; }
__device__ inline float fmax2(float a, float b) { return (b > a) ? b : a; }
__device__ inline float fmin2(float a, float b) { return (b < a) ? b : a; }
__device__ inline int fint(float x) {
  return (x > -2147483904.f && x < 2147483648.f) ? static_cast<int>(x)
                                                 : static_cast<int>(0x80000000u);
}
__device__ inline int nint(float x) {
  const float r = __builtin_roundf(x);
  return (r > -9.2233715e18f && r < 9.2233715e18f) ? static_cast<int>(static_cast<long long>(r)) : 0;
}
__device__ inline int ffloor(float x) {
  const int i = fint(x);
  return static_cast<float>(i) <= x ? i : static_cast<int>(static_cast<unsigned>(i) - 1u);
}

/// Ordered accumulation across a group, one term per lane K in K order.  bc gives every lane
/// of a 16-lane row the value of the row's lane K % 16 (DPP row_newbcast).  A group of 16 is
/// one row, so each step adds the group's lane K.  A group of 32 or 64 lanes accumulates row
/// by row: while row R's terms are added only row R's lanes hold the true running value (the
/// other rows add their own lanes' terms, which is discarded), and row_take then gives every
/// lane of the group row R's running value before row R + 1 (WW_ROWS).
template <int G, int K>
__device__ inline float bc(float v) {
  static_assert(G == 16 || G == 32 || G == 64, "groups are one, two or four DPP rows");
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x150 + (K % 16), 0xF, 0xF, false));
}

/// Row R's value of v in every lane of the group (a 32- or 64-lane group); a 16-lane group's
/// own value.
template <int G>
__device__ inline float row_take(float v, int r) {
  if constexpr (G == 64) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 16 * r));
  } else if constexpr (G == 32) {
    return __shfl(v, 16 * r, 32);
  } else {
    return v;
  }
}
template <int G>
__device__ inline bool row_take(bool v, int r) {
  if constexpr (G == 64) {
    return __builtin_amdgcn_readlane(v ? 1 : 0, 16 * r) != 0;
  } else if constexpr (G == 32) {
    return __shfl(v ? 1 : 0, 16 * r, 32) != 0;
  } else {
    return v;
  }
}

/// Orders the lanes' LDS and global accesses around it for the compiler (the hardware keeps
/// a wavefront's memory operations in order).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#define WW_ROW16(F) F(0) F(1) F(2) F(3) F(4) F(5) F(6) F(7) F(8) F(9) F(10) F(11) F(12) F(13) F(14) F(15)
#define WW_ROW16B(F, o) F(o + 0) F(o + 1) F(o + 2) F(o + 3) F(o + 4) F(o + 5) F(o + 6) F(o + 7) \
  F(o + 8) F(o + 9) F(o + 10) F(o + 11) F(o + 12) F(o + 13) F(o + 14) F(o + 15)
/// F(k) for every lane k of a group (the entries k >= G are discarded at compile time), and
/// S(r) after row r: the row_take of the running values in a group wider than a row.
#define WW_ROWS(F, S) WW_ROW16B(F, 0) S(0) WW_ROW16B(F, 16) S(1) WW_ROW16B(F, 32) S(2) WW_ROW16B(F, 48) S(3)

/// Per-point LDS layout, in floats: X (NSPEC), then the arrays that live through a sub-step
/// (DIRF, WNF, CBETA), then one region that W3SDS4 (BTH0, the saturated-band list, the
/// isotropic band sums and RENEW), W3SPR4 (SF) and W3SIN4 (four factors per band during the
/// stress scan, then the tail integral's per-step terms: 3 arrays, 7 with SINTAILPAR(3) > 0)
/// use in turn.  5.5 KB per point at NK=32, NTH=36: seven 4-point workgroups per CU.
struct Lds {
  int dirf, wnf, cbeta, bth0, sf, pk, sat, vb, rniso, small, stride;
  __host__ __device__ Lds(int nk, int nth, bool tail7) {
    dirf = 0;
    wnf = dirf + nth;
    cbeta = wnf + nk;
    const int un = cbeta + nk;
    bth0 = un;
    sat = bth0 + nk;  // ints: count, then the saturated bands
    vb = sat + nk + 1;
    rniso = vb + nk;
    sf = un;
    pk = un;
    const int tail = (tail7 ? 7 : 3) * kTailStride;
    small = un + (tail > 4 * nk + 1 ? tail : 4 * nk + 1);
    // X and W of the four points start 16 banks apart: the 16 consecutive floats that each
    // 16-lane row reads or writes in one access never share a bank with another row's
    stride = ((nk * nth + small + 31) / 32) * 32 + 16;
  }
};

/// One point's view.
struct Grp {
  int p, l;
  float* X;
  float* W;
  float* spec;
  const float* wn1;
  const float* cg1;
  const float* pin;
  float* specinit;
  float* vsin;
  float* vdin;
  float* vsnl;
  float* vdnl;
  float* vsds;
  float* vdds;
  float* vsdb;
  float* rnw;
  unsigned char* llws;
  Lds o;
  const float* mft;
};

/// The point's scalars, in registers while a kernel runs.
struct Sc {
  float v[N_SCAL];
  int iv[N_ISCAL];
};

template <int G>
__device__ inline Grp make_grp(const State& s, const St4Dev& q, int p, float* lds, int g) {
  const int nk = q.nk, nth = q.nth;
  Grp t{p, static_cast<int>(threadIdx.x) % G, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
        nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
        Lds(nk, nth, q.sintailpar[2] > 0.f), nullptr};
  const size_t nspec = static_cast<size_t>(nk) * nth;
  t.X = lds + static_cast<size_t>(g) * t.o.stride;
  t.W = t.X + nspec;
  const size_t ps = static_cast<size_t>(p) * nspec;
  t.spec = s.spec + ps;
  t.wn1 = s.wn1 + static_cast<size_t>(p) * nk;
  t.cg1 = s.cg1 + static_cast<size_t>(p) * nk;
  t.pin = s.pin + static_cast<size_t>(p) * N_PIN;
  t.specinit = s.specinit + ps;
  t.vsin = s.vsin + ps;
  t.vdin = s.vdin + ps;
  t.vsnl = s.vsnl + ps;
  t.vdnl = s.vdnl + ps;
  t.vsds = s.vsds + ps;
  t.vdds = s.vdds + ps;
  t.vsdb = s.vsdb + ps;
  t.rnw = s.rnw + ps;
  t.llws = s.llws + ps;
  t.mft = s.mft + static_cast<size_t>(p) * nk;
  return t;
}

__device__ inline void load_sc(const State& s, int p, Sc& c) {
  const float* src = s.sc + static_cast<size_t>(p) * N_SCAL;
  for (int i = 0; i < N_SCAL; ++i) c.v[i] = src[i];
  const int* isrc = s.isc + static_cast<size_t>(p) * N_ISCAL;
  for (int i = 0; i < N_ISCAL; ++i) c.iv[i] = isrc[i];
}

__device__ inline void store_sc(const State& s, const Grp& t, const Sc& c) {
  if (t.l != 0) return;
  float* dst = s.sc + static_cast<size_t>(t.p) * N_SCAL;
  for (int i = 0; i < N_SCAL; ++i) dst[i] = c.v[i];
  int* idst = s.isc + static_cast<size_t>(t.p) * N_ISCAL;
  for (int i = 0; i < N_ISCAL; ++i) idst[i] = c.iv[i];
}

// ------------------------------------------------------------------------------------
// CALC_USTAR (iterative branch, SINTAILPAR(1) < 0.5); uniform in the group
// ------------------------------------------------------------------------------------
__device__ __forceinline__ void calc_ustar(const St4Dev& q, float windspeed, float tauw, float& ustar, float& z0,
                           float& charn) {
  const float eps1 = 0.00001f;
  float chath = q.aalpha, ust = 0.f, xmin;
  if (q.capchnk[0] == 1.f) {
    chath = q.capchnk[1] + 0.5f * (q.capchnk[2] - q.capchnk[1]) *
                               (1.f - libm::tanhf((windspeed - q.capchnk[3]) / q.capchnk[4]));
    xmin = 0.15f * (q.capchnk[2] - chath);
  } else {
    xmin = 0.f;
  }
  const float tl = tauw;
  const float wcd = __builtin_sqrtf(0.0012875f);
  const float ustold = windspeed * wcd;
  float tauold = fmax2(ustold * ustold, tl + eps1);
  constexpr float znunum = 0.11f * kNuAir;
  constexpr float xm1 = 0.50f + 1.f;
  for (int it = 1; it <= 10; ++it) {
    const float x = fmax2(tl / tauold, xmin);
    ust = __builtin_sqrtf(tauold);
    float zz00 = chath * tauold / kGrav;
    if (q.zz0max != 0.f) zz00 = fmin2(zz00, q.zz0max);
    const float omx = 1.f - x;
    float zz0 = zz00 / libm::powf(omx, 0.50f);
    const float znu = znunum / fmax2(ust, 1E-6f);
    zz0 = q.sintailpar[4] * znu + zz0;
    const float lg = libm::logf_short(q.zzwnd / zz0);
    const float fval = ust - kKappa * windspeed / lg;
    const float delf = 1.f - kKappa * windspeed / (lg * lg) * 2.f / ust * (1.f - xm1 * x) / omx;
    ust = ust - fval / delf;
    tauold = fmax2(ust * ust, tl + eps1);
  }
  ustar = ust;
  if (ustar > 0.001f) {
    const float s = fmin2(windspeed / ustar, 100.0f);
    z0 = q.zzwnd * libm::expf(-kKappa * s);
    charn = kGrav * z0 / (ustar * ustar);
  } else {
    if (ustar > 0.f) {
      const float s = fmin2(windspeed / ustar, 100.0f);
      z0 = q.zzwnd * libm::expf(-kKappa * s);
    } else {
      z0 = chath * 0.001f * 0.001f / kGrav;
    }
    charn = chath;
  }
  if (q.capchnk[0] == 1.f) {
    charn = fmin2(0.09f, charn);
    if (charn < chath) charn = chath;
  }
}

// ------------------------------------------------------------------------------------
// W3SPR4 (reads the spectrum from X, LLWS from the point's array)
// ------------------------------------------------------------------------------------
template <int G>
__device__ __forceinline__ void spr4(const Grp& t, const St4Dev& q, Sc& c) {
  const int nk = q.nk, nth = q.nth, nspec = q.nspec, l = t.l;
  const float eps = 1.E-7f;
  const float u = t.pin[P_U10ABS], udir = t.pin[P_U10DIR];
  float ustar = mx(0.0001f, c.v[S_USTAR]);
  float emean = 0.f, emeanws = 0.f, fmeanws = 0.f, fmean = 0.f, fmean1 = 0.f, wnm = 0.f;
  float amax = 0.f, elcs = 0.f, elsn = 0.f;
  float* sf = t.W + t.o.sf;
  for (int ik = l; ik < nk; ik += G) sf[ik] = libm::powf(q.sig[ik + 1], q.ssdsc[11]) * q.dden[ik] / t.cg1[ik];
  wave_sync();
  // Section 1: ELCS, ELSN, AMAX over every bin in order.
  for (int b0 = 0; b0 < nspec; b0 += G) {
    const int b = b0 + l;
    float pc = 0.f, ps = 0.f, aa = 0.f;
    if (b < nspec) {
      aa = t.X[b];
      const float f = sf[b / nth];
      pc = aa * q.ecos[b] * f;
      ps = aa * q.esin[b] * f;
    }
    const int cnt = nspec - b0;
#define ST(k) if constexpr ((k) < G)                                \
  if (k < cnt) {                             \
    elcs = elcs + bc<G, k>(pc);                 \
    elsn = elsn + bc<G, k>(ps);                 \
    amax = mx(amax, bc<G, k>(aa));              \
  }
#define SY(r) if constexpr (G > 16 && 16 * (r) < G) if (16 * (r) < cnt) { elcs = row_take<G>(elcs, r); elsn = row_take<G>(elsn, r); amax = row_take<G>(amax, r); }
    WW_ROWS(ST, SY)
#undef SY
#undef ST
  }
  const float dlw = libm::atan2f(elsn, elcs);
  const float tw = 2.f * q.wwnmeanptail;
  // Section 2: band sums per lane, then the band terms in IK order.
  float elast = 0.f, e2last = 0.f;
  for (int base = 0; base < nk; base += G) {
    const int ik = base + l;
    float e = 0.f, e2 = 0.f, tf = 0.f, tf1 = 0.f, twn = 0.f, tfw = 0.f;
    if (ik < nk) {
      float eb = 0.f, eb2 = 0.f;
      const int is0 = ik * nth;
      for (int ith = 0; ith < nth; ++ith) {
        const float aa = t.X[is0 + ith];
        eb = eb + aa;
        if (t.llws[is0 + ith] != 0) eb2 = eb2 + aa;
      }
      const float pws = libm::powf(q.sig[ik + 1], tw);
      const float pwk = libm::powf(t.wn1[ik], q.wwnmeanp);
      const float cgk = t.cg1[ik];
      e = eb * q.dden[ik] / cgk;
      e2 = eb2 * q.dden[ik] / cgk;
      tf = e / q.sig[ik + 1];
      tf1 = e * pws;
      twn = e * pwk;
      tfw = e2 * pws;
    }
    const int cnt = nk - base;
#define ST(k) if constexpr ((k) < G)                                \
  if (k < cnt) {                             \
    emean = emean + bc<G, k>(e);                \
    fmean = fmean + bc<G, k>(tf);               \
    fmean1 = fmean1 + bc<G, k>(tf1);            \
    wnm = wnm + bc<G, k>(twn);                  \
    emeanws = emeanws + bc<G, k>(e2);           \
    fmeanws = fmeanws + bc<G, k>(tfw);          \
  }
#define SY(r) if constexpr (G > 16 && 16 * (r) < G) if (16 * (r) < cnt) { emean = row_take<G>(emean, r); fmean = row_take<G>(fmean, r); fmean1 = row_take<G>(fmean1, r); wnm = row_take<G>(wnm, r); emeanws = row_take<G>(emeanws, r); fmeanws = row_take<G>(fmeanws, r); }
    WW_ROWS(ST, SY)
#undef SY
#undef ST
    if (base + G >= nk) {
      elast = __shfl(e, nk - 1 - base, G);
      e2last = __shfl(e2, nk - 1 - base, G);
    }
  }
  float eband = elast / q.dden[nk - 1];
  emean = emean + eband * q.fte;
  fmean = fmean + eband * q.ftf;
  fmean1 = fmean1 + eband * q.sstxftftail;
  wnm = wnm + eband * q.sstxftwn;
  eband = e2last / q.dden[nk - 1];
  emeanws = emeanws + eband * q.fte;
  fmeanws = fmeanws + eband * q.sstxftftail;
  fmean = kTpiInv * emean / mx(eps, fmean);
  if (fmean1 < eps)
    fmean1 = kTpiInv * q.sig[nk];
  else
    fmean1 = kTpiInv * libm::powf(mx(eps, fmean1) / mx(eps, emean), 1.f / (2.f * q.wwnmeanptail));
  wnm = libm::powf(mx(eps, wnm) / mx(eps, emean), 1.f / q.wwnmeanp);
  if (fmeanws < eps || emeanws < eps)
    fmeanws = kTpiInv * q.sig[nk];
  else
    fmeanws = kTpiInv * libm::powf(mx(eps, fmeanws) / mx(eps, emeanws), 1.f / (2.f * q.wwnmeanptail));
  const float tx = c.v[S_TAUWX], ty = c.v[S_TAUWY];
  const float tauw = __builtin_sqrtf(tx * tx + ty * ty);
  float z0 = 0.f, charn = c.v[S_CHARN];
  calc_ustar(q, u, tauw, ustar, z0, charn);
  const float r = ustar / mx(0.01f, u);
  c.v[S_USTAR] = ustar;
  c.v[S_USTDIR] = udir;
  c.v[S_EMEAN] = emean;
  c.v[S_FMEAN] = fmean;
  c.v[S_FMEAN1] = fmean1;
  c.v[S_WNMEAN] = wnm;
  c.v[S_AMAX] = amax;
  c.v[S_CD] = r * r;
  c.v[S_Z0] = z0;
  c.v[S_CHARN] = charn;
  c.v[S_FMEANWS] = fmeanws;
  c.v[S_DLWMEAN] = dlw;
}

// ------------------------------------------------------------------------------------
// W3SIN4 (W3SRCE's call: A = SPEC, K = WN2, U = U10ABS, USDIR = U10DIR); S, D into VSIN,
// VDIN; TAUWNX/Y into TAUWAX/Y.  Band mapping for the per-bin arrays.
// ------------------------------------------------------------------------------------
template <int G, int NTH = 0>
__device__ __forceinline__ void sin4(const Grp& t, const St4Dev& q, Sc& c) {
  const int nk = q.nk, l = t.l;
  const int nth = NTH > 0 ? NTH : q.nth;  // a compile-time loop bound for the compiled NTH
  const float u = t.pin[P_U10ABS], usdir = t.pin[P_U10DIR], as = t.pin[P_AS];
  const float ustar = c.v[S_USTAR], drat = c.v[S_DRAT], z0 = c.v[S_Z0];
  constexpr float kappa2 = kKappa * kKappa;
  constexpr float z0viscnum = 0.1f * kNuAir;
  constexpr float twonu = 2.f * kNuAir;
  constexpr float grav2tpi = (kGrav * kGrav) * kTpi;
  constexpr float logtpi005 = 7.1361942291259765625f;  // LOG(TPI/0.005) as folded
  constexpr float eps1 = 0.00001f, eps2 = 0.000001f, eps2sq = eps2 * eps2;
  const float const1 = q.bbeta / kappa2;
  float const0 = const1 * drat;
  const float z0visc = z0viscnum / mx(ustar, 0.0001f);
  const float z0noz = mx(z0visc, q.zz0rat * z0);
  float uorb = 0.f, aorb = 0.f;
  for (int base = 0; base < nk; base += G) {
    const int ik = base + l;
    float tu = 0.f, ta = 0.f;
    if (ik < nk) {
      float eb = 0.f;
      for (int ith = 0; ith < nth; ++ith) eb = eb + t.X[ik * nth + ith];
      const float sk = q.sig[ik + 1];
      const float cgk = t.cg1[ik];
      tu = eb * (sk * sk) * q.dden[ik] / cgk;
      ta = eb * q.dden[ik] / cgk;
    }
    const int cnt = nk - base;
#define ST(k) if constexpr ((k) < G)                      \
  if (k < cnt) {                   \
    uorb = uorb + bc<G, k>(tu);       \
    aorb = aorb + bc<G, k>(ta);       \
  }
#define SY(r) if constexpr (G > 16 && 16 * (r) < G) if (16 * (r) < cnt) { uorb = row_take<G>(uorb, r); aorb = row_take<G>(aorb, r); }
    WW_ROWS(ST, SY)
#undef SY
#undef ST
  }
  uorb = 2.f * __builtin_sqrtf(uorb);
  const float aorb1 = 2.f * libm::powf(aorb, 1.f - 0.5f * q.sswellf[5]);
  const float re = 4.f * uorb * aorb1 / kNuAir;
  float pturb, pvisc;
  if (q.sswellf[3] > 0.f) {
    if (q.sswellf[6] > 0.f) {
      const float smooth = 0.5f * libm::tanhf((re - q.sswellf[3]) / q.sswellf[6]);
      pturb = 0.5f + smooth;
      pvisc = 0.5f - smooth;
    } else if (re <= q.sswellf[3]) {
      pturb = 0.f;
      pvisc = 1.f;
    } else {
      pturb = 1.f;
      pvisc = 0.f;
    }
  } else {
    pturb = 1.f;
    pvisc = 1.f;
  }
  float fw, fu, fud;
  if (q.sswellf[1] == 0.f) {
    fw = __builtin_fabsf(q.sswellf[2]);
    fu = 0.f;
    fud = 0.f;
  } else {
    fu = __builtin_fabsf(q.sswellf[2]);
    fud = q.sswellf[1];
    aorb = 2.f * __builtin_sqrtf(aorb);
    const float xi = (libm::log10f(mx(aorb / z0noz, 3.f)) - (-1.f)) / q.delab;
    const int ind = mni(300 - 1, fint(xi));
    const float deli1 = mn(1.f, xi - static_cast<float>(ind));
    const float deli2 = 1.f - deli1;
    fw = q.fwtable[ind] * deli2 + q.fwtable[ind + 1] * deli1;
  }
  float ustarsigma = 0.f;
  if (q.stab3) {
    const float usigma = (as != as) ? 0.f : mx(0.f, -0.025f * as);
    ustarsigma = (1.0f + u / (10.f + u)) * usigma;
  }

  // The scan's per-band factors that do not depend on the running stress and cost a
  // division, a square root or a logarithm, one band per lane: CM, CONST2, ZCN and the
  // viscous swell factor (st4_sin4.hpp:211-219).
  float* pk = t.W + t.o.pk;
  wave_sync();
  for (int ik0 = l; ik0 < nk; ik0 += G) {
    const float kis = t.wn1[ik0];
    const float s2 = q.sig2[ik0 * nth];
    pk[ik0] = kis / s2;
    pk[nk + ik0] = q.dden2[ik0 * nth] / t.cg1[ik0] * kGrav / (q.sig[ik0 + 1] / kis * drat);
    pk[2 * nk + ik0] = libm::logf(kis * z0);
    pk[3 * nk + ik0] = -(q.sswellf[4] * drat * 2.f * kis * __builtin_sqrtf(twonu * s2));
  }
  wave_sync();
  // The stress scan.  DSTAB(1, IS) (STAB3) is kept in VSIN until the average is formed.
  float ust = ustar, taux = 0.f, tauy = 0.f, cosu = 0.f, sinu = 0.f, ustp = 0.f, usdirp = 0.f;
  float s1x = 0.f, s1y = 0.f, n1x = 0.f, n1y = 0.f;  // istab 1
  float s2x = 0.f, s2y = 0.f, n2x = 0.f, n2y = 0.f;  // istab 2
  float s3x = 0.f, s3y = 0.f, n3x = 0.f, n3y = 0.f;  // istab 3
  const int i0 = q.stab3 ? 1 : 3, i1 = q.stab3 ? 2 : 3;
  for (int istab = i0; istab <= i1; ++istab) {
    float* const dst = (istab == 1) ? t.vsin : t.vdin;
    if (istab == 1) ust = ustar * (1.f - ustarsigma);
    if (istab == 2) ust = ustar * (1.f + ustarsigma);
    float cu, su;
    libm::sincosf(usdir, &su, &cu);
    taux = (ust * ust) * cu;
    tauy = (ust * ust) * su;
    float ssx = 0.f, ssy = 0.f, snx = 0.f, sny = 0.f;
    for (int ik = 1; ik <= nk; ++ik) {
      const float taupx = taux - __builtin_fabsf(q.ttauwshelter) * ssx;
      const float taupy = tauy - __builtin_fabsf(q.ttauwshelter) * ssy;
      ustp = mn(libm::pow_quarter(taupx * taupx + taupy * taupy), mx(ust, 0.3f));
      usdirp = libm::atan2f(taupy, taupx);
      libm::sincosf_short(usdirp, &sinu, &cosu);
      const int is1 = 1 + (ik - 1) * nth;
      const float s2 = q.sig2[is1 - 1];
      const float cm = pk[ik - 1];
      const float ucn = ustp * cm + q.zzalp;
      const float const2 = pk[nk + ik - 1];
      const float cnst = s2 * const0;
      const float zcn = pk[2 * nk + ik - 1];
      const float swv = pk[3 * nk + ik - 1];
      const float swt = -(drat * q.sswellf[0] * 16.f * (s2 * s2) / kGrav);
#pragma unroll 1
      for (int base = 0; base < nth; base += G) {
        const int ith = base + l;
        float dv = 0.f, temp2 = 0.f;
        float pc = 0.f, ps = 0.f;
        if (ith < nth) {
          const int is = ith + (ik - 1) * nth;  // 0-based
          const float ec = q.ecos[is], es = q.esin[is];
          const float coswind = (ec * cosu + es * sinu);
          int lw;
          if (coswind > 0.01f) {
            const float x = coswind * ucn;
            const float zarg = kKappa / x;
            const float zlog = zcn + zarg;
            if (zlog < 0.f) {
              const float zl2 = zlog * zlog;
              // COSWIND > 0.01 here, where POWF(COSWIND, 2) is COSWIND*COSWIND
              const float pw = q.ssinthp == 2.f ? coswind * coswind : libm::powf(coswind, q.ssinthp);
              dv = cnst * libm::expf(zlog) * (zl2 * zl2) * ucn * ucn * pw;
              lw = 1;
            } else {
              dv = 0.f;
              lw = 0;
            }
            if (28.f * cm * ustar * coswind >= 1.f) lw = 1;
          } else {
            dv = 0.f;
            lw = 0;
          }
          if ((q.sswellf[0] != 0.f && dv < 1E-7f * s2) || q.sswellf[2] > 0.f) {
            const float dvisc = swv;
            const float dturb = swt * (fw * uorb + (fu + fud * coswind) * ustp);
            dv = dv + pturb * dturb + pvisc * dvisc;
          }
          t.llws[is] = static_cast<unsigned char>(lw);
          dst[is] = dv;
          temp2 = const2 * dv * t.X[is];
          pc = temp2 * ec;
          ps = temp2 * es;
        }
        const int cnt = nth - base;
        // SN (DSTAB < 0) or SS gets TEMP2*ECOS, TEMP2*ESIN of each direction, in ITH order.
        // The sum that a direction does not go to adds -0, which leaves every value unchanged
        // (x + -0 == x for all x, zeros and NaN included), so each step is one add per sum.
        const bool neg = dv < 0.f;
        const float pcn = neg ? pc : -0.f, psn = neg ? ps : -0.f;
        const float pcs = neg ? -0.f : pc, pss = neg ? -0.f : ps;
#define ST(k) if constexpr ((k) < G)                                       \
  if (k < cnt) {                                    \
    snx = snx + bc<G, k>(pcn);                      \
    sny = sny + bc<G, k>(psn);                      \
    ssx = ssx + bc<G, k>(pcs);                      \
    ssy = ssy + bc<G, k>(pss);                      \
  }
#define SY(r) if constexpr (G > 16 && 16 * (r) < G) if (16 * (r) < cnt) { snx = row_take<G>(snx, r); sny = row_take<G>(sny, r); ssx = row_take<G>(ssx, r); ssy = row_take<G>(ssy, r); }
        WW_ROWS(ST, SY)
#undef SY
#undef ST
      }
    }
    if (istab == 1) { s1x = ssx; s1y = ssy; n1x = snx; n1y = sny; }
    if (istab == 2) { s2x = ssx; s2y = ssy; n2x = snx; n2y = sny; }
    if (istab == 3) { s3x = ssx; s3y = ssy; n3x = snx; n3y = sny; }
  }
  float xstress, ystress, tauwnx, tauwny;
  if (!q.stab3) {
    xstress = s3x;
    ystress = s3y;
    tauwnx = n3x;
    tauwny = n3y;
  } else {
    xstress = 0.5f * (s1x + s2x);
    ystress = 0.5f * (s1y + s2y);
    tauwnx = 0.5f * (n1x + n2x);
    tauwny = 0.5f * (n1y + n2y);
  }
  // D (the STAB3 average) and S = D*A, on the lanes that wrote them.
  for (int ik = 0; ik < nk; ++ik)
    for (int ith = l; ith < nth; ith += G) {
      const int is = ith + ik * nth;
      if (q.stab3) t.vdin[is] = 0.5f * (t.vsin[is] + t.vdin[is]);
      t.vsin[is] = t.vdin[is] * t.X[is];
    }
  {
    const float taupx = taux - __builtin_fabsf(q.ttauwshelter) * xstress;
    const float taupy = tauy - __builtin_fabsf(q.ttauwshelter) * ystress;
    ustp = libm::pow_quarter(taupx * taupx + taupy * taupy);
    usdirp = libm::atan2f(taupy, taupx);
  }
  ust = ustp;
  {
    const float sg = q.sig[nk], s2 = sg * sg, s5 = s2 * (s2 * sg);
    const0 = q.dth * s5 / grav2tpi * kTpi * sg / t.cg1[nk - 1];
  }
  float temp = 0.f;
  for (int base = 0; base < nth; base += G) {
    const int ith = base + l;
    float term = 0.f;
    if (ith < nth) {
      const int is = ith + (nk - 1) * nth;
      const float cw = (q.ecos[is] * cosu + q.esin[is] * sinu);
      const float cp = mx(cw, 0.f);
      term = t.X[is] * (cp * (cp * cp));
    }
    const int cnt = nth - base;
#define ST(k) if constexpr ((k) < G) \
  if (k < cnt) temp = temp + bc<G, k>(term);
#define SY(r) if constexpr (G > 16 && 16 * (r) < G) if (16 * (r) < cnt) { temp = row_take<G>(temp, r); }
    WW_ROWS(ST, SY)
#undef SY
#undef ST
  }
  const float levtail0 = const0 * temp;
  // The tail integral: the terms independent of the running u* first, one step per lane.
  const float sg = q.sig[nk];
  const float zinf = libm::logf(sg * sg / kGrav);
  const float dely = (logtpi005 - zinf) / 50.f;
  const float sp2 = q.sintailpar[2];
  constexpr int TS = kTailStride;
  float* tl = t.W + t.o.pk;  // cm, cgtail, g0, c1, fm, a0, a1 (over the scan's factors)
  wave_sync();
  for (int j0 = l; j0 < kTailJ; j0 += G) {
    const int j = j0 + 1;
    const float y = libm::expf(zinf + static_cast<float>(j - 1) * dely);
    const float ykm = (y / 363.f);
    const float ykm2 = ykm * ykm;
    const float tensk = 1.f + ykm2;
    const float cm = __builtin_sqrtf(kGrav * tensk / y);
    tl[0 * TS + j0] = cm;
    tl[1 * TS + j0] = 0.5f * (3.f * ykm2 + 1.f) * __builtin_sqrtf(kGrav / (y * tensk));
    tl[2 * TS + j0] = kGrav * z0 / (cm * cm);
    if (sp2 > 0.f) {
      const float ykm1 = (ykm - 1.f);
      tl[3 * TS + j0] = 0.2325f / cm;
      tl[4 * TS + j0] = libm::expf(-0.25f * (ykm1 * ykm1));
      const float th = libm::tanhf((y - 20.f) / 5.f);
      tl[5 * TS + j0] = levtail0 * 0.5f * (1.f - th);
      tl[6 * TS + j0] = sp2 * 0.5f * (1.f + th);
    }
  }
  wave_sync();
  float tauhf;
  {
    float ustr = ust;
    float taur = ust * ust, tau1 = 0.f;
    for (int j = 1; j <= kTailJ; ++j) {
      const float wj = (j == 1 || j == kTailJ) ? 0.5f : 1.f;
      const float cm = tl[0 * TS + j - 1];
      const float cgtail = tl[1 * TS + j - 1];
      const float zx = ustr / cm + q.zzalp;
      const float zarg = mn(kKappa / zx, 20.f);
      const float zmu = mn(tl[2 * TS + j - 1] * libm::expf(zarg), 1.f);
      const float zlog = mn(libm::logf_short(zmu), 0.f);
      const float zl2 = zlog * zlog;
      const float zbeta = const1 * zmu * (zl2 * zl2);
      float levtail;
      if (sp2 > 0.f) {
        float alpham;
        if (ustr < cm)
          alpham = mx(0.f, 0.01f * (1.f + libm::logf_short(ustr / cm)));
        else
          alpham = 0.01f * (1.f + 3.f * libm::logf_short(ustr / cm));
        const float alphat = alpham * tl[3 * TS + j - 1] * tl[4 * TS + j - 1];
        levtail = tl[5 * TS + j - 1] + tl[6 * TS + j - 1] * alphat;
      } else {
        levtail = levtail0;
      }
      tau1 = tau1 + wj * zbeta * (ustr * ustr) * levtail * dely * cgtail / cm;
      taur = taur - wj * q.sintailpar[1] * (ustr * ustr) * zbeta * levtail * dely * cgtail / cm;
      ustr = __builtin_sqrtf(mx(taur, 0.f));
    }
    tauhf = tau1;
  }
  float cp, sp;
  libm::sincosf(usdirp, &sp, &cp);
  float tauwx = xstress + tauhf * cp;
  float tauwy = ystress + tauhf * sp;
  const float tauw = __builtin_sqrtf(tauwx * tauwx + tauwy * tauwy);
  const float u2b = mx(ustar, eps2);
  const float u2 = u2b * u2b;
  const float tauwb = mn(tauw, mx(u2 - eps1, eps2sq));
  if (tauwb < tauw) {
    tauwx = tauwx * tauwb / tauw;
    tauwy = tauwy * tauwb / tauw;
  }
  c.v[S_TAUWX] = tauwx;
  c.v[S_TAUWY] = tauwy;
  c.v[S_TAUWAX] = tauwnx;
  c.v[S_TAUWAY] = tauwny;
  wave_sync();
}

// ------------------------------------------------------------------------------------
// W3SDS4's anisotropic cumulative term (SSDSC(3) < 0) for a compiled NTH.
//
// RENEW(IS) = sum over saturated IK2 <= IK - DIKCUMUL, in IK2 order, of
// V = sum over ITH2 in order of CUMULW(IS2, IS) * BRLAMBDA(IS2).  For one (IK, IK2), the NTH
// values of V are a Toeplitz product: CUMULW depends on |ITH2 - ITH| only.  Sixteen lanes of
// a group take band pairs IK = base + p and base + 31 - p, p = 0..15 (the two together hold
// about the same number of (IK, IK2) pairs for every lane); a group of G lanes splits each
// pair's NTH output directions into G/16 slices of NI = 16*NTH/G.  A lane keeps the window
// of the mirrored CUMULW row that its slice needs in registers and runs the NTH x NI
// products with every register index fixed at compile time: per pair, NTH + NI - 1 row
// loads, NTH loads of BRLAMBDA and NTH*NI multiply-add pairs on NI independent accumulators.
// RENEW goes to the point's RNW array.  X holds BRLAMBDA; W's sat list holds the saturated
// bands in order.
// ------------------------------------------------------------------------------------
template <int G, int NTH>
__device__ __forceinline__ void cumul_aniso(const Grp& t, const St4Dev& q, int dikcumul) {
  constexpr int MS = kMirStride(NTH);
  constexpr int DS = G / 16;     // lanes per band pair
  constexpr int NI = NTH / DS;   // output directions per lane
  constexpr int NW = NTH + NI - 1;  // the CUMULW window of a slice
  static_assert(NTH % DS == 0, "the direction slices must be equal");
  const int nk = q.nk;
  const int pi = t.l % 16, i0 = (t.l / 16) * NI;
  const int* sat = reinterpret_cast<const int*>(t.W + t.o.sat);
  const int nsat = sat[0];
  for (int base = 0; base < nk; base += 32) {
    const int ika = base + pi, ikb = base + 31 - pi;  // 0-based output bands
    const int lima = ika + 1 - dikcumul, limb = ikb + 1 - dikcumul;
    int ca = 0, cb = 0;
    if (ika < nk && lima > 0)
      while (ca < nsat && sat[1 + ca] <= lima) ++ca;
    if (ikb < nk && limb > 0)
      while (cb < nsat && sat[1 + cb] <= limb) ++cb;
    float rn[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) rn[i] = 0.f;
    const int total = ca + cb;
    for (int m = 0; m < total; ++m) {
      if (m == ca) {  // band a is complete
        if (ika < nk) {
#pragma unroll
          for (int i = 0; i < NI; ++i) t.rnw[ika * NTH + i0 + i] = rn[i];
        }
#pragma unroll
        for (int i = 0; i < NI; ++i) rn[i] = 0.f;
      }
      const bool ina = m < ca;
      const int ik0 = ina ? ika : ikb;
      const int ik2 = sat[1 + (ina ? m : m - ca)];
      const float* row = q.cumulm + (static_cast<size_t>(ik0) * nk + (ik2 - 1)) * MS;
      // kw[w] = CUMULW row entry NTH - NI - i0 + w: output i0 + i takes input j from
      // entry j - (i0 + i) + NTH - 1 = kw[j - i + NI - 1]
      float kw[NW];
      if constexpr (DS == 1) {
        const float4* kr4 = reinterpret_cast<const float4*>(row);
#pragma unroll
        for (int d = 0; d < MS / 4; ++d) {
          const float4 v = kr4[d];
          if (4 * d < NW) kw[4 * d] = v.x;
          if (4 * d + 1 < NW) kw[4 * d + 1] = v.y;
          if (4 * d + 2 < NW) kw[4 * d + 2] = v.z;
          if (4 * d + 3 < NW) kw[4 * d + 3] = v.w;
        }
      } else {
        const float* kr = row + (NTH - NI - i0);
#pragma unroll
        for (int w = 0; w < NW; ++w) kw[w] = kr[w];
      }
      const float* bl = t.X + (ik2 - 1) * NTH;
      float acc[NI];
#pragma unroll
      for (int i = 0; i < NI; ++i) acc[i] = 0.f;
#pragma unroll
      for (int j = 0; j < NTH; ++j) {
        const float b = bl[j];
#pragma unroll
        for (int i = 0; i < NI; ++i) acc[i] = acc[i] + kw[j - i + NI - 1] * b;
      }
#pragma unroll
      for (int i = 0; i < NI; ++i) rn[i] = rn[i] + acc[i];
    }
    if (ca == total) {  // no switch: rn is band a's, band b has no pair
      if (ika < nk) {
#pragma unroll
        for (int i = 0; i < NI; ++i) t.rnw[ika * NTH + i0 + i] = rn[i];
      }
      if (ikb < nk) {
#pragma unroll
        for (int i = 0; i < NI; ++i) t.rnw[ikb * NTH + i0 + i] = 0.f;
      }
    } else if (ikb < nk) {
#pragma unroll
      for (int i = 0; i < NI; ++i) t.rnw[ikb * NTH + i0 + i] = rn[i];
    }
  }
}

// The whitecap's MFT for one band: the 50-step integral of the foam decay over the
// band's period (st4_sds4.hpp:690-707); it depends on K and SIG only, so a call computes it
// once per point instead of every sub-step.
__device__ __forceinline__ float mft_integral(float kk, float sigk) {
  const float tstr = ((0.8f * 2.f) * kPi) / sigk;
  const float tmax = ((5.f * 2.f) * kPi) / sigk;
  const float dt = tmax / 50.f;
  float mft = 0.f;
  for (int it = 1; it <= 50; ++it) {
    const float tt = static_cast<float>(it) * dt;
    if (tt < tstr)
      mft = mft + ((0.4f / (kk * tstr)) * tt) * dt;
    else
      mft = mft + ((0.4f / kk) * libm::expf(-((tt - tstr) / 3.8f))) * dt;
  }
  return mft / tmax;
}

// ------------------------------------------------------------------------------------
// W3SDS4 (W3SRCE's call: A = SPEC, K = WN1, USDIR = USTDIR), SDSBCHOICE = 1.  On entry X
// holds the spectrum; on return X holds BRLAMBDA, VSDS/VDDS hold SRHS/DDIAG, W's bth0
// the saturation maxima.
// ------------------------------------------------------------------------------------
template <int G, int NTH>
__device__ __forceinline__ void sds4(const Grp& t, const St4Dev& q, const StepFlags& f, Sc& c, float* whitecap) {
  const int nk = q.nk, nth = q.nth, nspec = q.nspec, l = t.l;
  const float* cs = q.ssdsc;
  const float ustar = c.v[S_USTAR], usdir = c.v[S_USTDIR], depth = c.v[S_DEPTH];
  const float dair = t.pin[P_DAIR];
  const float facturb = (((cs[4] * (ustar * ustar)) / kGrav) * dair) / kDwat;
  const int dikcumul = nint(q.ssdsbrf1 / (q.xfr - 1.f));
  const int ik1 = q.ig1 ? nint(q.igpars5) + 1 : 1;
  constexpr float twopisq = 2.f * (kPi * kPi);
  const float epsr = __builtin_sqrtf(q.ssdsbr);
  const bool iso = q.ssdsdth >= 180.f;
  float* bth0w = t.W + t.o.bth0;

  if constexpr (NTH > 0) {
    // Saturation, MAXVAL (or the isotropic sum), DDIAG, PB -> BRLAMBDA (into X) and SRHS in
    // three passes with no reduction across lanes inside a band loop.
    //  1. B for every bin of the bands from IK1 up: one direction per lane and pass, its
    //     weights and indices in registers for all bands; B goes to the point's SRHS array
    //     (the lane's own bins), then into X once every band has been read.
    //  2. One band per lane: MAXVAL of B in ITH order (a NaN is skipped, the first non-NaN
    //     value starts the scan, a later value replaces it only when greater: NaN when the
    //     band holds nothing else), which is what the in-order scan of the reference returns;
    //     or the isotropic sum of A.  The band's COEF1, SDIAGISO and COEF2 go to LDS.
    //  3. Every bin: DDIAG, PB, BRLAMBDA, SRHS.
    constexpr int NSAT = 2 * (NTH / 2 - 1) + 1;
    constexpr int NP = (NTH + G - 1) / G;
    const bool fixed = q.nsat == NSAT;
    const int kb = ik1 - 1 < nk ? ik1 - 1 : nk;  // first band of the saturation
    if (!iso) {
#pragma unroll 1
      for (int pp = 0; pp < NP; ++pp) {
        const int ith = pp * G + l;
        if (ith >= NTH) break;
        if (fixed) {
          float wr[NSAT];
          int xr[NSAT];
          const float* w = q.satw + NSAT * ith;
          const int* x = q.sati + NSAT * ith;
#pragma unroll
          for (int j = 0; j < NSAT; ++j) {
            wr[j] = w[j];
            xr[j] = x[j];
          }
          for (int ik0 = kb; ik0 < nk; ++ik0) {
            const float kk = t.wn1[ik0];
            const float facsat = (q.sig[ik0 + 1] * ((kk * kk) * kk)) * q.dth;
            const float* xb = t.X + ik0 * NTH - 1;
            float acc = 0.f;
#pragma unroll
            for (int j = 0; j < NSAT; ++j) acc = acc + wr[j] * xb[xr[j]];
            t.vsds[ik0 * NTH + ith] = acc * facsat;
          }
        } else {
          const float* w = q.satw + q.nsat * ith;
          const int* x = q.sati + q.nsat * ith;
          for (int ik0 = kb; ik0 < nk; ++ik0) {
            const float kk = t.wn1[ik0];
            const float facsat = (q.sig[ik0 + 1] * ((kk * kk) * kk)) * q.dth;
            const float* xb = t.X + ik0 * NTH - 1;
            float acc = 0.f;
            for (int j = 0; j < q.nsat; ++j) acc = acc + w[j] * xb[x[j]];
            t.vsds[ik0 * NTH + ith] = acc * facsat;
          }
        }
      }
      wave_sync();  // every gather of A is done: X takes B
#pragma unroll 1
      for (int pp = 0; pp < NP; ++pp) {
        const int ith = pp * G + l;
        if (ith >= NTH) break;
        for (int ik0 = kb; ik0 < nk; ++ik0) t.X[ik0 * NTH + ith] = t.vsds[ik0 * NTH + ith];
      }
      wave_sync();
    }
    float* c1w = t.W + t.o.vb;     // COEF1 per band (the isotropic band sums come later)
    float* sdw = t.W + t.o.rniso;  // SDIAGISO per band
    float* c2w = t.W + t.o.sat;    // COEF2 per band (the saturated-band list comes later)
    for (int ik0 = l; ik0 < nk; ik0 += G) {
      if (ik0 < kb) {
        bth0w[ik0] = 0.f;
        continue;
      }
      const float* xb = t.X + ik0 * NTH;
      float bth0;
      if (iso) {
        const float kk = t.wn1[ik0];
        const float facsat = (q.sig[ik0 + 1] * ((kk * kk) * kk)) * q.dth;
        float asum = 0.f;
        for (int ith = 0; ith < NTH; ++ith) asum = asum + xb[ith];
        bth0 = asum * facsat;
      } else {
        float lim = -__builtin_inff();
        bool found = false;
        for (int ith = 0; ith < NTH; ++ith) {
          const float bk = xb[ith];
          if (!found) {
            if (bk >= lim) {
              lim = bk;
              found = true;
            }
          } else if (bk > lim) {
            lim = bk;
          }
        }
        bth0 = found ? lim : __builtin_nanf("");
      }
      bth0w[ik0] = bth0;
      float miche;
      if (q.ssdsbm[0] == 1.f) {
        miche = 1.f;
      } else {
        const float x = libm::tanhf(mn(t.wn1[ik0] * depth, 10.f));
        const float tt = x * (q.ssdsbm[1] + x * (q.ssdsbm[2] + x * (q.ssdsbm[3] + x * q.ssdsbm[4])));
        miche = tt * tt;
      }
      const float coef1 = q.ssdsbr * miche;
      const float m = mx(bth0 / coef1 - 1.f, 0.f);
      const float sigk = q.sig[ik0 + 1];
      c1w[ik0] = coef1;
      sdw[ik0] = ((cs[1] * sigk) * cs[5]) * (m * m);
      c2w[ik0] = ((cs[1] * sigk) * (1.f - cs[5])) / (coef1 * coef1);
    }
    wave_sync();
    for (int b = l; b < nspec; b += G) {
      const int ik0 = b / NTH;
      const float a = t.spec[b];  // X held A = SPEC until pass 1
      if (ik0 < kb) {
        t.vdds[b] = 0.f;
        t.vsds[b] = 0.f * a;  // SRHS = DDIAG * A with DDIAG = 0
        t.X[b] = 0.f;         // BTH = 0 there, so PB = MAX(-EPSR, 0)**2 * 28.16 = 0
        continue;
      }
      const float coef1 = c1w[ik0];
      const float bb = iso ? bth0w[ik0] : t.X[b];
      const float xp = mx(bb - coef1, 0.f);  // >= 0 or +inf, where POWF(XP, 2) is XP*XP
      const float dd = sdw[ik0] + c2w[ik0] * (q.ssdsp == 2.f ? xp * xp : libm::powf(xp, q.ssdsp));
      const float mm = mx(__builtin_sqrtf(bb) - epsr, 0.f);
      const float pb = (mm * mm) * 28.16f;
      t.vdds[b] = dd;
      t.X[b] = pb / twopisq;
      t.vsds[b] = dd * a;
    }
    wave_sync();
  } else {
  // Pass 1: B per bin (kept in VSDS), and per band the sum (isotropic case) or MAXVAL.
    float lim = -__builtin_inff(), asum = 0.f;
    bool found = false;
    for (int b0 = 0; b0 < nspec; b0 += G) {
      const int b = b0 + l;
      float av = 0.f, bv = 0.f;
      if (b < nspec) {
        const int ik0 = b / nth;
        if (ik0 + 1 >= ik1) {
          av = t.X[b];
          if (!iso) {
            const int ith = b - ik0 * nth;
            const float kk = t.wn1[ik0];
            const float facsat = (q.sig[ik0 + 1] * ((kk * kk) * kk)) * q.dth;
            const float* w = q.satw + q.nsat * ith;
            const int* x = q.sati + q.nsat * ith;
            const float* xb = t.X + ik0 * nth - 1;
            float acc = 0.f;
            constexpr int NSAT = NTH > 0 ? 2 * (NTH / 2 - 1) + 1 : 0;
            if (NSAT > 0 && q.nsat == NSAT) {
              // all weights and indices first, then the gathers, then the sum in J order
              float wr[NSAT > 0 ? NSAT : 1], ar[NSAT > 0 ? NSAT : 1];
  #pragma unroll
              for (int j = 0; j < NSAT; ++j) wr[j] = w[j];
  #pragma unroll
              for (int j = 0; j < NSAT; ++j) ar[j] = xb[x[j]];
  #pragma unroll
              for (int j = 0; j < NSAT; ++j) acc = acc + wr[j] * ar[j];
            } else {
              for (int j = 0; j < q.nsat; ++j) acc = acc + w[j] * xb[x[j]];
            }
            bv = acc * facsat;
            t.vsds[b] = bv;
          }
        }
      }
      const int cnt = nspec - b0;
  #define ST(k) if constexpr ((k) < G)                                                                   \
    if (k < cnt) {                                                                \
      const int bin = b0 + k;                                                     \
      const int ik0 = bin / nth;                                                  \
      const int ith = bin - ik0 * nth;                                            \
      if (ik0 + 1 >= ik1) {                                                       \
        if (ith == 0) {                                                           \
          lim = -__builtin_inff();                                                \
          found = false;                                                          \
          asum = 0.f;                                                             \
        }                                                                         \
        if (iso) {                                                                \
          asum = asum + bc<G, k>(av);                                                \
        } else {                                                                  \
          const float bk = bc<G, k>(bv);                                             \
          if (!found) {                                                           \
            if (bk >= lim) {                                                      \
              lim = bk;                                                           \
              found = true;                                                       \
            }                                                                     \
          } else if (bk > lim) {                                                  \
            lim = bk;                                                             \
          }                                                                       \
        }                                                                         \
        if (ith == nth - 1 && l == ((k) & 48)) {                                  \
          const float kk = t.wn1[ik0];                                            \
          const float facsat = (q.sig[ik0 + 1] * ((kk * kk) * kk)) * q.dth;       \
          bth0w[ik0] = iso ? asum * facsat : (found ? lim : __builtin_nanf(""));  \
        }                                                                         \
      }                                                                           \
    }
#define SY(r) if constexpr (G > 16 && 16 * (r) < G) if (16 * (r) < cnt) { lim = row_take<G>(lim, r); found = row_take<G>(found, r); asum = row_take<G>(asum, r); }
      WW_ROWS(ST, SY)
  #undef SY
#undef ST
    }
    for (int ik0 = l; ik0 < ik1 - 1 && ik0 < nk; ik0 += G) bth0w[ik0] = 0.f;
    wave_sync();
  
    // Pass 2: DDIAG, PB -> BRLAMBDA (into X), SRHS = DDIAG*A.
    for (int b = l; b < nspec; b += G) {
      const int ik0 = b / nth;
      const float a = t.X[b];
      if (ik0 + 1 < ik1) {
        t.vdds[b] = 0.f;
        t.vsds[b] = 0.f * a;  // SRHS = DDIAG * A with DDIAG = 0
        t.X[b] = 0.f;         // BTH = 0 there, so PB = MAX(-EPSR, 0)**2 * 28.16 = 0
        continue;
      }
      const float kk = t.wn1[ik0];
      const float bth0 = bth0w[ik0];
      float miche;
      if (q.ssdsbm[0] == 1.f) {
        miche = 1.f;
      } else {
        const float x = libm::tanhf(mn(kk * depth, 10.f));
        const float tt = x * (q.ssdsbm[1] + x * (q.ssdsbm[2] + x * (q.ssdsbm[3] + x * q.ssdsbm[4])));
        miche = tt * tt;
      }
      const float coef1 = q.ssdsbr * miche;
      const float m = mx(bth0 / coef1 - 1.f, 0.f);
      const float sigk = q.sig[ik0 + 1];
      const float sdiagiso = ((cs[1] * sigk) * cs[5]) * (m * m);
      const float coef2 = ((cs[1] * sigk) * (1.f - cs[5])) / (coef1 * coef1);
      const float bb = iso ? bth0 : t.vsds[b];
      const float xp = mx(bb - coef1, 0.f);  // >= 0 or +inf, where POWF(XP, 2) is XP*XP
      const float dd = sdiagiso + coef2 * (q.ssdsp == 2.f ? xp * xp : libm::powf(xp, q.ssdsp));
      const float mm = mx(__builtin_sqrtf(bb) - epsr, 0.f);
      const float pb = (mm * mm) * 28.16f;
      t.vdds[b] = dd;
      t.X[b] = pb / twopisq;
      t.vsds[b] = dd * a;
    }
    wave_sync();
  
  }

  // Whitecap parameters (the reference computes them after the cumulative term, which does
  // not touch their inputs).  COEF4/COEF5 are zero above IKC; BRM12 is never set by this
  // routine and stays 0.  Without FLOGRD(5,7|8) WHITECAP keeps its values.
  if (f.flogrd57 || f.flogrd58) {
    float wc0 = 0.f, wc1 = 0.f, wc3 = 0.f;
    const int ikc = ffloor(q.aaircmin) < nk ? ffloor(q.aaircmin) : nk;
    for (int base = 0; base < nk; base += G) {
      const int ik = base + l + 1;
      float coef4 = 0.f, coef5 = 0.f, mft = 0.f;
      if (ik <= nk) {
        const float kk = t.wn1[ik - 1];
        if (ik <= ikc) {
          const float cc = q.sig[ik] / kk;
          const int is0 = (ik - 1) * nth;
          float v = 0.f;
          for (int i = is0; i < is0 + nth; ++i) v = v + t.X[i];
          const float cgk = t.cg1[ik - 1];
          coef4 = ((((((cc * cc) * v) * 2.f) * kPi) / kGrav) * cs[6]) * q.dden[ik - 1] / (q.sig[ik] * cgk);
          const float pw = libm::powf(0.f / cs[1], 3.f / 5.f);
          float ww = 0.f;
          for (int i = is0; i < is0 + nth; ++i) ww = ww + (t.X[i] * (-cs[1])) * pw;
          coef5 = ((((cc * cc) * cc) * ww) * q.aairgb / kGrav) * q.dden[ik - 1] / (q.sig[ik] * cgk);
        }
        if (f.flogrd58 && ik >= ik1) mft = t.mft[ik - 1];
      }
      const int cnt = nk - base;
#define ST(k) if constexpr ((k) < G)                                                \
  if (k < cnt) {                                             \
    const int ikk = base + k + 1;                            \
    const float c4 = bc<G, k>(coef4);                           \
    const float c5 = bc<G, k>(coef5);                           \
    const float mf = bc<G, k>(mft);                             \
    if (f.flogrd57 && ikk >= ik1 && ikk <= ikc) {            \
      wc0 = wc0 + c4 * (1.f - wc0);                          \
      wc3 = wc3 + c5;                                        \
    }                                                        \
    if (f.flogrd58 && ikk >= ik1) wc1 = wc1 + c4 * mf;       \
  }
#define SY(r) if constexpr (G > 16 && 16 * (r) < G) if (16 * (r) < cnt) { wc0 = row_take<G>(wc0, r); wc1 = row_take<G>(wc1, r); wc3 = row_take<G>(wc3, r); }
      WW_ROWS(ST, SY)
#undef SY
#undef ST
    }
    if (l == 0) {
      whitecap[0] = wc0;
      whitecap[1] = wc1;
      whitecap[2] = 0.f;
      whitecap[3] = wc3;
    }
  }

  if (!(cs[2] != 0.f || cs[4] != 0.f || cs[20] != 0.f)) return;

  // The cumulative effect.  Isotropic (SSDSC(3) > 0): one sum of BRLAMBDA per band, then
  // RENEW per output band.  Otherwise RENEW per bin, by the Toeplitz kernel when NTH is one
  // of the compiled ones.
  float* rniso = t.W + t.o.rniso;
  if (cs[2] > 0.f) {
    float* vb = t.W + t.o.vb;
    for (int ik0 = l; ik0 < nk; ik0 += G) {
      float v = 0.f;
      for (int i = ik0 * nth; i < ik0 * nth + nth; ++i) v = v + t.X[i];
      vb[ik0] = v;
    }
    wave_sync();
    for (int ik0 = l; ik0 < nk; ik0 += G) {
      const int ik = ik0 + 1;
      float renew = 0.f;
      if (ik >= ik1 && ik > dikcumul) {
        const float cc = q.sig[ik] / t.wn1[ik0];
        for (int ik2 = ik1; ik2 <= ik - dikcumul; ++ik2) {
          const float c2 = q.sig[ik2] / t.wn1[ik2 - 1];
          const float cumulwiso = ((__builtin_fabsf(c2 - cc) * q.dsip[ik2]) / (c2 * 0.5f)) * q.dth;
          renew = renew - cumulwiso * vb[ik2 - 1];
        }
      }
      rniso[ik0] = renew;
    }
  } else if (cs[2] < 0.f) {
    int* sat = reinterpret_cast<int*>(t.W + t.o.sat);
    if (l == 0) {
      int n = 0;
      for (int ik2 = ik1; ik2 <= nk; ++ik2)
        if (bth0w[ik2 - 1] > q.ssdsbr) sat[1 + n++] = ik2;
      sat[0] = n;
    }
    wave_sync();
    if constexpr (NTH > 0) cumul_aniso<G, NTH>(t, q, dikcumul);
  }
  wave_sync();

  // Wave-turbulence interaction and the cumulative term into DDIAG and SRHS.
  float su, cu;
  libm::sincosf(usdir, &su, &cu);
  for (int b = l; b < nspec; b += G) {
    const int ik0 = b / nth;
    const int ik = ik0 + 1;
    if (ik < ik1) continue;
    const int ith = b - ik0 * nth;
    const float kk = t.wn1[ik0];
    const float facturb2 = -(((q.sig[ik] * 2.f) * kk) * facturb);
    const float dvisc = -(((cs[20] * 4.f) * kk) * kk);
    float renew = 0.f;
    if (cs[2] > 0.f && ik > dikcumul) renew = rniso[ik0];
    if (cs[2] < 0.f && ik > dikcumul) {
      if constexpr (NTH > 0) {
        renew = t.rnw[b];
      } else {
        const int* sat = reinterpret_cast<const int*>(t.W + t.o.sat);
        const int ms = kMirStride(nth);
        for (int m = 0; m < sat[0]; ++m) {
          const int ik2 = sat[1 + m];
          if (ik2 > ik - dikcumul) break;
          const float* bl = t.X + (ik2 - 1) * nth;
          const float* kr = q.cumulm + (static_cast<size_t>(ik0) * nk + (ik2 - 1)) * ms + (nth - 1 - ith);
          float v = 0.f;
          for (int j = 0; j < nth; ++j) v = v + kr[j] * bl[j];
          renew = renew + v;
        }
      }
    }
    const float coswind = q.ecos[ith] * cu + q.esin[ith] * su;
    const float dturb = facturb2 * mx(coswind, 0.f);
    const float diag2 = (cs[2] * renew + dturb) + dvisc;
    t.vdds[b] = t.vdds[b] + diag2;
    t.vsds[b] = t.vsds[b] + t.spec[b] * diag2;
  }
}

// ------------------------------------------------------------------------------------
// W3SLN1 -> DIRF, WNF in W (VSLN(IS) = WNF(IK) * DIRF(ITH))
// ------------------------------------------------------------------------------------
template <int G>
__device__ __forceinline__ void ln1(const Grp& t, const SrceDev& p, const Sc& c) {
  const int nk = p.nk, nth = p.nth, l = t.l;
  const float ustar = c.v[S_USTAR], usdir = t.pin[P_U10DIR], fhigh = c.v[S_FHIGH];
  float cosu = 0.f, sinu = 0.f;
  libm::sincosf(usdir, &sinu, &cosu);
  float* dirf = t.W + t.o.dirf;
  float* wnf = t.W + t.o.wnf;
  for (int ith = l; ith < nth; ith += G) {
    const float m = fmax2(0.f, p.ecos[ith] * cosu + p.esin[ith] * sinu);
    const float m2 = m * m;
    dirf[ith] = m2 * m2;
  }
  const float u2 = ustar * ustar;
  const float fac = p.slnc1 * (u2 * u2);
  const float ff1 = p.fspm * kGrav / (28.f * ustar);
  const float ff2 = p.fshf * fmin2(p.sig[nk - 1], fhigh);
  const float ffilt = fmin2(2.f * p.sig[nk - 1], fmax2(ff2, ff1));
  for (int ik = l; ik < nk; ik += G) {
    const float rfr = p.sig[ik] / ffilt;
    float w;
    if (rfr < 0.5f) {
      w = 0.f;
    } else {
      const float r2 = rfr * rfr;
      const float rm4 = 1.f / (r2 * r2);
      w = fac / t.wn1[ik] * libm::expf(-rm4);
    }
    wnf[ik] = w;
  }
}

// W3SBT1 -> CBETA in W and the cut-off NSCUT: D(IS) = CBETA(MAPWN(IS)) for IS <= NSCUT,
// else 0; S = D*A, except after the early return (NSCUT = -1), where S = +0 on every bin.
__device__ __forceinline__ void sbt1(const Grp& t, const SrceDev& p, Sc& c) {
  const int nk = p.nk, nth = p.nth;
  const float depth = c.v[S_DEPTH];
  if (depth * t.wn1[0] > 6.f) {
    c.iv[I_NSCUT] = -1;
    return;
  }
  const float factor = p.sbtc1 / depth;
  float* cbeta = t.W + t.o.cbeta;
  int ik = 1;
  for (; ik <= nk; ++ik) {
    const float w = t.wn1[ik - 1];
    if (w * depth > 6.f) break;
    const float x = t.cg1[ik - 1] * w / p.sig[ik - 1] - 0.5f;
    if (t.l == 0) cbeta[ik - 1] = factor * (x > 0.f ? x : 0.f);
  }
  c.iv[I_NSCUT] = (ik - 1) * nth;
}

// W3SDB1 -> S_DTDB and VSDB (the early return keeps both).
template <int G>
__device__ __forceinline__ void sdb1(const Grp& t, const SrceDev& p, Sc& c) {
  const int nspec = p.nspec;
  const float emean = c.v[S_EMEAN], fmean = c.v[S_FMEAN], wnmean = c.v[S_WNMEAN];
  const float depth = c.v[S_DEPTH];
  const double thr = static_cast<double>(0x1p-126f);
  if (static_cast<double>(emean) < 0x1p-1022) return;
  double hm;
  if (p.fdonly != 0) {
    hm = static_cast<double>(p.sdbc2) * static_cast<double>(depth);
  } else {
    const float dpos = depth > 0.f ? depth : 0.f;
    const double wnm = static_cast<double>(wnmean);
    hm = static_cast<double>(p.sdbc2) / wnm * ::tanh(wnm * static_cast<double>(dpos));
  }
  const double hrms = __builtin_sqrt(8. * static_cast<double>(emean));
  double bb, b;
  if (hm > thr) {
    bb = hrms * hrms / (hm * hm);
    b = __builtin_sqrt(bb);
  } else {
    bb = 0.;
    b = 0.;
  }
  double q0 = 0.;
  if (b <= 0.5) {
    q0 = 0.;
  } else if (b <= 1.) {
    const double tt = 2. * b - 1.;
    q0 = tt * tt;
  }
  double qb;
  if (b <= 0.2) {
    qb = 0.;
  } else if (b < 1.) {
    const double arg = ::exp((q0 - 1.) / bb);
    qb = q0 - bb * (q0 - arg) / (bb - arg);
    for (int i = 1; i <= 3; ++i) qb = ::exp((qb - 1.) / bb);
  } else {
    qb = static_cast<double>(1.0f - 0x1p-126f);
  }
  double cbj;
  if (b != 1. && bb > thr && __builtin_fabs(bb - qb) > thr)
    cbj = (bb < 1.0) ? 2. * static_cast<double>(p.sdbc1) * qb * static_cast<double>(fmean) / bb
                     : 2. * static_cast<double>(p.sdbc1) * static_cast<double>(fmean) * bb;
  else
    cbj = 0.;
  const float dv = static_cast<float>(-cbj);
  c.v[S_DTDB] = dv;
  for (int is = t.l; is < nspec; is += G) t.vsdb[is] = dv * t.spec[is];
}

// ------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------
template <int G>
__global__ __launch_bounds__(64) void init_kernel(State s, St4Dev q, SrceDev pp, StepFlags f) {
  extern __shared__ float lds[];
  for (int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); i < kMaxIter;
       i += static_cast<int>(gridDim.x * blockDim.x))
    s.count[i] = 0;
  const int g = static_cast<int>(threadIdx.x) / G;
  const int p = static_cast<int>(blockIdx.x) * (64 / G) + g;
  if (p >= s.npts) return;
  const int nk = pp.nk, nspec = pp.nspec;
  const Grp t = make_grp<G>(s, q, p, lds, g);
  const int l = t.l;
  Sc c;
  const float* psc = s.psc + static_cast<size_t>(p) * N_PSC;
  for (int i = 0; i < N_PSC; ++i) c.v[i] = psc[i];
  for (int i = N_PSC; i < N_SCAL; ++i) c.v[i] = 0.f;
  for (int i = 0; i < N_ISCAL; ++i) c.iv[i] = 0;
  float* whitecap = s.whitecap + static_cast<size_t>(p) * 4;
  float* taubbl = s.taubbl + static_cast<size_t>(p) * 2;
  float* tauice = s.tauice + static_cast<size_t>(p) * 2;
  for (int b = l; b < nspec; b += G) {
    const float a = t.spec[b];
    t.X[b] = a;
    t.vsdb[b] = 0.f;
    t.llws[b] = f.it == 0 ? 1 : 0;
  }
  if (l == 0) {
    for (int j = 0; j < 4; ++j) whitecap[j] = 0.f;
    taubbl[0] = taubbl[1] = 0.f;
    tauice[0] = tauice[1] = 0.f;
  }
  if (f.flogrd58) {
    float* mft = s.mft + static_cast<size_t>(p) * nk;
    for (int ik0 = l; ik0 < nk; ik0 += G) mft[ik0] = mft_integral(t.wn1[ik0], q.sig[ik0 + 1]);
  }
  c.v[S_DTDB] = 0.f;
  const float d_inp = t.pin[P_D_INP], dair = t.pin[P_DAIR];
  c.v[S_DEPTH] = mx(d_inp, pp.dmin);
  c.v[S_DRAT] = dair / kDwat;
  c.v[S_DTDYN] = 0.f;
  c.v[S_DTTOT] = 0.f;
  c.v[S_PHIAW] = 0.f;
  c.v[S_CHARN] = 0.f;
  c.v[S_TWS] = 0.f;
  c.v[S_PHIBBL] = 0.f;
  c.v[S_TAUWIX] = 0.f;
  c.v[S_TAUWIY] = 0.f;
  c.v[S_TAUWNX] = 0.f;
  c.v[S_TAUWNY] = 0.f;
  c.v[S_TAUWAX] = 0.f;
  c.v[S_TAUWAY] = 0.f;
  c.v[S_PHICE] = 0.f;
  c.v[S_TAUOCX] = 0.f;
  c.v[S_TAUOCY] = 0.f;
  c.v[S_WNMEAN] = 0.f;
  c.v[S_DLWMEAN] = 0.f;
  wave_sync();
  const float sintail4 = q.sintailpar[3];
  if (sintail4 > 0.5f) {
    c.v[S_TAUWX] = 0.f;
    c.v[S_TAUWY] = 0.f;
  }
  if (f.it == 0) {
    c.v[S_TAUWX] = 0.f;
    c.v[S_TAUWY] = 0.f;
    c.v[S_USTAR] = 0.f;
    c.v[S_USTDIR] = 0.f;
  } else {
    // LLWS is read here before anything sets it; only FMEANWS depends on it and the W3SPR4
    // below overwrites it (TAUWBUG, the only configuration accepted with IT > 0).
    spr4<G>(t, q, c);
    if (sintail4 > 0.5f) sin4<G>(t, q, c);
  }
  spr4<G>(t, q, c);
  c.v[S_TWS] = 1.f / c.v[S_FMEANWS];
  for (int b = l; b < nspec; b += G) t.specinit[b] = t.X[b];
  c.v[S_FHIGH] = mx((pp.ffxfm + 0.f) * mx(c.v[S_FMEANWS], c.v[S_FMEAN1]), pp.ffxpm / c.v[S_USTAR]);
  // the first sub-step's input terms and KDMEAN
  sin4<G>(t, q, c);
  if (l == 0) {
    s.kdmean[p] = c.v[S_WNMEAN] * c.v[S_DEPTH];
    s.list[0][p] = p;
  }
  store_sc(s, t, c);
}

template <int G, int NTH>
__global__ __launch_bounds__(64) void step_kernel(State s, St4Dev q, SrceDev pp, StepFlags f, int iter,
                                                  int first, int n) {
  extern __shared__ float lds[];
  const int g = static_cast<int>(threadIdx.x) / G;
  const int slot = first + static_cast<int>(blockIdx.x) * (64 / G) + g;
  if (slot >= n) return;
  const int p = s.list[iter & 1][slot];
  const int nk = pp.nk, nth = pp.nth, nspec = pp.nspec;
  const Grp t = make_grp<G>(s, q, p, lds, g);
  const int l = t.l;
  Sc c;
  load_sc(s, p, c);
  float* whitecap = s.whitecap + static_cast<size_t>(p) * 4;
  for (int b = l; b < nspec; b += G) t.X[b] = t.spec[b];
  c.iv[I_NSTEPS] = c.iv[I_NSTEPS] + 1;
  if (pp.ln1) ln1<G>(t, pp, c);
  wave_sync();
  sds4<G, NTH>(t, q, f, c, whitecap);
  sdb1<G>(t, pp, c);
  sbt1(t, pp, c);
  wave_sync();

  const float ice = t.pin[P_ICE];
  const float isln = mx(mn(1.f - ice * pp.icescales[0], 1.f), 0.f);
  const float isin = mx(mn(1.f - ice * pp.icescales[1], 1.f), 0.f);
  const float isnl = mx(mn(1.f - ice * pp.icescales[2], 1.f), 0.f);
  const float isds = mx(0.f, mn(1.f - ice * pp.icescales[3], 1.f));
  const bool scale = f.inflags24 != 0;
  const float dtg = f.dtg;
  float fhigh = c.v[S_FHIGH];
  int nkh = mni(nk, fint(pp.facti1 * libm::logf(mx(1.E-7f, fhigh)) + pp.facti2));
  const int nkh1 = mni(nk, nkh + 1);
  // NSPECH = NKH1*NTH with the wrap of the compiled reference (NKH = INT_MIN when FHIGH is
  // +Inf, i.e. USTAR = 0 at U10 = 0, gives NSPECH = NTH for even NTH); never above NSPEC.
  const int nspech_w = static_cast<int>(static_cast<unsigned>(nkh1) * static_cast<unsigned>(nth));
  const int nspech = nspech_w < nspec ? nspech_w : nspec;
  float dttot = c.v[S_DTTOT];
  float dt = mn(pp.dtmax, dtg - dttot);
  const float wl = t.wn1[nk - 1];
  const float damlast = pp.facp / (pp.sig[nk - 1] * (wl * (wl * wl)));
  const float afilt = mx(damlast, pp.xflt * c.v[S_AMAX]);
  const int nscut = c.iv[I_NSCUT];
  const float* dirf = t.W + t.o.dirf;
  const float* wnf = t.W + t.o.wnf;
  const float* cbeta = t.W + t.o.cbeta;

  // The time-step limit over the bins below NKH1: each lane's MIN, then the row's.  The
  // loads of a lane's next bin are issued before the current one is computed.
  {
    struct Lim {
      float vsnl, vdnl, vsin, vdin, vsds, vdds, a, ai;
      int mw;
    };
    const auto ld = [&](int is) {
      Lim v{};
      if (is < nspech) {
        v.vsnl = t.vsnl[is];
        v.vdnl = t.vdnl[is];
        v.vsin = t.vsin[is];
        v.vdin = t.vdin[is];
        v.vsds = t.vsds[is];
        v.vdds = t.vdds[is];
        v.a = t.spec[is];
        v.ai = t.specinit[is];
        v.mw = is < nscut ? pp.mapwn[is] : 1;
      }
      return v;
    };
    float dtl = dt;
    Lim nx = ld(l);
    for (int is = l; is < nspech; is += G) {
      const Lim cu = nx;
      nx = ld(is + G);
      const int ik0 = is / nth;
      const int ith = is - ik0 * nth;
      const float w = t.wn1[ik0];
      const float dam = pp.facp / (pp.sig[ik0] * (w * (w * w)));
      float vsln = pp.ln1 ? wnf[ik0] * dirf[ith] : 0.f;
      float vsnl = cu.vsnl, vdnl = cu.vdnl, vsin = cu.vsin, vdin = cu.vdin;
      float vsds = cu.vsds, vdds = cu.vdds;
      if (scale) {
        vsnl = isnl * vsnl;
        vdnl = isnl * vdnl;
        vsln = isln * vsln;
        vsin = isin * vsin;
        vdin = isin * vdin;
        vsds = isds * vsds;
        vdds = isds * vdds;
      }
      const float a = cu.a;
      const float vdbt = is < nscut ? cbeta[cu.mw - 1] : 0.f;
      const float vsbt = nscut < 0 ? 0.f : vdbt * a;
      const float vs = (((vsln + vsin) + vsnl) + vsds) + vsbt;
      const float vd = ((vdin + vdnl) + vdds) + vdbt;
      const float damax = mn(mx(pp.xrel * cu.ai, afilt), dam);
      const float afac = 1.f / mx(__builtin_fabsf(vs / damax), 1.E-10f);
      dtl = mn(dtl, afac / mx(mn(vd, 0.f) * afac + 1.f, 1.E-10f));
    }
#define ST(k) if constexpr ((k) < G) dt = mn(dt, bc<G, k>(dtl));
#define SY(r) if constexpr (G > 16 && 16 * (r) < G) { dt = row_take<G>(dt, r); }
    WW_ROWS(ST, SY)
#undef SY
#undef ST
  }
  dt = mx(dt, 0.5f);
  c.v[S_DTDYN] = c.v[S_DTDYN] + dt;
  const float rest = dtg - dttot;
  const int idt = 1 + fint((rest * 0.99f) / dt);
  dt = rest / static_cast<float>(idt);
  const bool shave = dt < pp.dtmin && dt < rest;
  c.iv[I_SHAVE] = shave ? 1 : 0;
  dt = mx(mn(pp.dtmin, rest), dt);
  const float hdt = dt;
  dttot = dttot + dt;
  c.v[S_DTTOT] = dttot;

  // Update (bins below NKH1), W3SDB1 update, flux sums (all bins), in bin order.
  const float vddb = c.v[S_DTDB];
  float phiaw = c.v[S_PHIAW], phibbl = c.v[S_PHIBBL], wc2 = 0.f;
  struct Upd {
    float a0, vsin, vdin, vsnl, vdnl, vsds, vdds, vsdb;
    int mw;
  };
  // a lane's next bin is loaded before the current one is computed (the stores below touch
  // the current bin only)
  const auto ld = [&](int is) {
    Upd v{};
    if (is < nspec) {
      v.a0 = t.spec[is];
      v.vsin = t.vsin[is];
      v.vdin = t.vdin[is];
      v.mw = is < nscut ? pp.mapwn[is] : 1;
      if (is < nspech) {
        v.vsnl = t.vsnl[is];
        v.vdnl = t.vdnl[is];
        v.vsds = t.vsds[is];
        v.vdds = t.vdds[is];
        v.vsdb = t.vsdb[is];
      }
    }
    return v;
  };
  Upd nx = ld(l);
  for (int b0 = 0; b0 < nspec; b0 += G) {
    const int is = b0 + l;
    const Upd cu = nx;
    nx = ld(is + G);
    float tpa = 0.f, tpb = 0.f, twc = 0.f, vsin = 0.f;
    if (is < nspec) {
      const int ik0 = is / nth;
      const int ith = is - ik0 * nth;
      const float a0 = cu.a0;
      const float vdbt = is < nscut ? cbeta[cu.mw - 1] : 0.f;
      const float vsbt = nscut < 0 ? 0.f : vdbt * a0;
      vsin = cu.vsin;
      float vdin = cu.vdin;
      float a = a0;
      if (is < nspech) {
        const float w = t.wn1[ik0];
        const float dam = pp.facp / (pp.sig[ik0] * (w * (w * w)));
        float vsln = pp.ln1 ? wnf[ik0] * dirf[ith] : 0.f;
        float vsnl = cu.vsnl, vdnl = cu.vdnl, vsds = cu.vsds, vdds = cu.vdds;
        if (scale) {
          vsnl = isnl * vsnl;
          vdnl = isnl * vdnl;
          vsln = isln * vsln;
          vsin = isin * vsin;
          vdin = isin * vdin;
          vsds = isds * vsds;
          vdds = isds * vdds;
        }
        const float vs = (((vsln + vsin) + vsnl) + vsds) + vsbt;
        const float vd = ((vdin + vdnl) + vdds) + vdbt;
        const float e1 = (dt * vs) / mx(1.f - hdt * vd, 1.f);
        if (shave) {
          const float e2 = __builtin_copysignf(mn(__builtin_fabsf(e1), dam), e1);
          a = mx(a + e2, 0.f);
        } else {
          a = mx(a + e1, 0.f);
        }
        const float e3 = (dt * cu.vsdb) / mx(1.f - hdt * vddb, 1.f);
        a = mx(a + e3, 0.f);
        t.spec[is] = a;
      }
      t.X[is] = a;
      const float factor = pp.dden[ik0] / t.cg1[ik0];
      tpa = ((vsin * dt) * factor) / mx(1.f - vdin * dt, 1.f);
      tpb = ((vsbt * dt) * factor) / mx(1.f - vdbt * dt, 1.f);
      // only growing bins add to the whitecap sum; the others add -0 (no change)
      twc = vsin > 0.f ? a * factor : -0.f;
    }
    const int cnt = nspec - b0;
#define ST(k) if constexpr ((k) < G)                                   \
  if (k < cnt) {                                \
    phiaw = bc<G, k>(tpa) + phiaw;                 \
    phibbl = phibbl - bc<G, k>(tpb);               \
    wc2 = wc2 + bc<G, k>(twc);                     \
  }
#define SY(r) if constexpr (G > 16 && 16 * (r) < G) if (16 * (r) < cnt) { phiaw = row_take<G>(phiaw, r); phibbl = row_take<G>(phibbl, r); wc2 = row_take<G>(wc2, r); }
    WW_ROWS(ST, SY)
#undef SY
#undef ST
  }
  wave_sync();
  c.v[S_PHIAW] = phiaw;
  c.v[S_PHIBBL] = phibbl;
  if (l == 0) whitecap[2] = __builtin_sqrtf(wc2) * 4.f;
  const float drat = c.v[S_DRAT];
  c.v[S_TAUWIX] = c.v[S_TAUWIX] + (c.v[S_TAUWX] * drat) * dt;
  c.v[S_TAUWIY] = c.v[S_TAUWIY] + (c.v[S_TAUWY] * drat) * dt;
  c.v[S_TAUWNX] = c.v[S_TAUWNX] + (drat * c.v[S_TAUWAX]) * dt;
  c.v[S_TAUWNY] = c.v[S_TAUWNY] + (drat * c.v[S_TAUWAY]) * dt;

  spr4<G>(t, q, c);
  const float u10abs = t.pin[P_U10ABS];
  const float fage = pp.ffxfa * libm::tanhf((((u10abs * 0.3f) * c.v[S_FMEANWS]) * kTpi) / kGrav);
  const float fh1 = (pp.ffxfm + fage) * c.v[S_FMEAN1];
  const float fh2 = pp.ffxpm / c.v[S_USTAR];
  fhigh = mn(mx(fh1, fh2), pp.sig[nk - 1]);
  c.v[S_FHIGH] = fhigh;
  nkh = mxi(mni(nkh1, fint(pp.facti1 * libm::logf(mx(1.E-7f, fhigh)) + pp.facti2)), 2);
  if (dttot >= dtg * 0.9999f) {
    const float wnmean = c.v[S_WNMEAN];
    const float hm = (pp.fhmax * libm::tanhf(mx(0.f, t.pin[P_D_INP]) * wnmean)) /
                     mx(wnmean, 9.99999974737875163555145263671875e-5f);
    const float em = (hm * hm) * 6.25e-2f;
    const float emean = c.v[S_EMEAN];
    if (emean > em && emean > 1.E-30f) {
      for (int is = l; is < nspec; is += G) {
        const float a = em * (t.X[is] / emean);
        t.X[is] = a;
        t.spec[is] = a;
      }
      c.v[S_EMEAN] = em;
    }
  }
  wave_sync();
  for (int ik = nkh; ik < nk; ++ik)
    for (int ith = l; ith < nth; ith += G) {
      const float a = t.X[ith + (ik - 1) * nth] * pp.fachfa + 0.f;
      t.X[ith + ik * nth] = a;
      t.spec[ith + ik * nth] = a;
    }
  wave_sync();
  sin4<G, NTH>(t, q, c);
  const bool done = dttot >= 0.9999f * dtg;
  if (q.sintailpar[3] < 0.5f) {
    spr4<G>(t, q, c);
    // the next sub-step's input terms see the u* and z0 of this W3SPR4
    if (!done) sin4<G, NTH>(t, q, c);
  }
  c.iv[I_DONE] = done ? 1 : 0;
  if (l == 0) {
    s.kdmean[p] = c.v[S_WNMEAN] * c.v[S_DEPTH];
    if (!done) {
      const int at = atomicAdd(&s.count[iter], 1);
      s.list[(iter + 1) & 1][at] = p;
    }
  }
  store_sc(s, t, c);
}

// ------------------------------------------------------------------------------------
// The per-routine entries (ww_spr4, ww_sin4, ww_sds4): one routine per point, on the same
// device functions W3SRCE uses, the point's arrays taken from the call.
// ------------------------------------------------------------------------------------
template <int NTH>
__global__ __launch_bounds__(64) void routine_kernel(St4Call io, St4Dev q) {
  constexpr int G = 16;
  extern __shared__ float lds[];
  const int g = static_cast<int>(threadIdx.x) / G, l = static_cast<int>(threadIdx.x) % G;
  const int p = static_cast<int>(blockIdx.x) * (64 / G) + g;
  if (p >= io.npts) return;
  const int nk = q.nk, nth = q.nth, nspec = q.nspec;
  const size_t ps = static_cast<size_t>(p) * nspec, pk = static_cast<size_t>(p) * nk;
  const Lds o(nk, nth, q.sintailpar[2] > 0.f);
  float* X = lds + static_cast<size_t>(g) * o.stride;
  const bool sin = io.which == 1, sds = io.which == 2;
  Grp t{p, l, X, X + nspec, const_cast<float*>(io.a + ps),
        sin ? io.kband + pk : io.k + pk, io.cg + pk, io.pin + static_cast<size_t>(p) * N_PIN, nullptr,
        sin ? io.s + ps : nullptr, sin ? io.d + ps : nullptr, nullptr, nullptr,
        sds ? io.s + ps : nullptr, sds ? io.d + ps : nullptr, nullptr, io.rnw + ps, io.llws8 + ps, o,
        io.mft + pk};
  for (int b = l; b < nspec; b += G) X[b] = io.a[ps + b];
  if (io.which == 0)
    for (int b = l; b < nspec; b += G) t.llws[b] = io.llws[ps + b] != 0 ? 1 : 0;
  if (sin)  // W3SIN4 reads K at the first bin of each band only (st4_sin4.hpp:212)
    for (int ik = l; ik < nk; ik += G) io.kband[pk + ik] = io.k[ps + static_cast<size_t>(ik) * nth];
  if (sds && io.flogrd58)
    for (int ik = l; ik < nk; ik += G) io.mft[pk + ik] = mft_integral(io.k[pk + ik], q.sig[ik + 1]);
  Sc c;
  for (int i = 0; i < N_SCAL; ++i) c.v[i] = 0.f;
  for (int i = 0; i < N_ISCAL; ++i) c.iv[i] = 0;
  float* pin = io.pin + static_cast<size_t>(p) * N_PIN;
  if (io.which == 0) {
    if (l == 0) {
      pin[P_U10ABS] = io.in[0][p];
      pin[P_U10DIR] = io.in[1][p];
    }
    c.v[S_TAUWX] = io.in[2][p];
    c.v[S_TAUWY] = io.in[3][p];
    c.v[S_USTAR] = io.out[0][p];
    c.v[S_CHARN] = io.out[9][p];
  } else if (sin) {
    if (l == 0) {
      pin[P_U10ABS] = io.in[0][p];
      pin[P_AS] = io.in[3][p];
      pin[P_U10DIR] = io.in[4][p];
    }
    c.v[S_USTAR] = io.in[1][p];
    c.v[S_DRAT] = io.in[2][p];
    c.v[S_Z0] = io.in[5][p];
  } else {
    if (l == 0) pin[P_DAIR] = io.in[3][p];
    c.v[S_USTAR] = io.in[0][p];
    c.v[S_USTDIR] = io.in[1][p];
    c.v[S_DEPTH] = io.in[2][p];
  }
  StepFlags f{1, 0, io.flogrd57, io.flogrd58, 0.f};
  wave_sync();
  if (io.which == 0) {
    spr4<G>(t, q, c);
    if (l == 0) {
      const int ids[12] = {S_USTAR, S_USTDIR, S_EMEAN, S_FMEAN, S_FMEAN1, S_WNMEAN, S_AMAX, S_CD,
                           S_Z0, S_CHARN, S_FMEANWS, S_DLWMEAN};
      for (int i = 0; i < 12; ++i) io.out[i][p] = c.v[ids[i]];
    }
  } else if (sin) {
    sin4<G, NTH>(t, q, c);
    for (int b = l; b < nspec; b += G) io.llws[ps + b] = t.llws[b];
    if (l == 0) {
      io.out[0][p] = c.v[S_TAUWX];
      io.out[1][p] = c.v[S_TAUWY];
      io.out[2][p] = c.v[S_TAUWAX];
      io.out[3][p] = c.v[S_TAUWAY];
    }
  } else {
    sds4<G, NTH>(t, q, f, c, io.whitecap + static_cast<size_t>(p) * 4);
    wave_sync();
    for (int b = l; b < nspec; b += G) io.brlambda[ps + b] = X[b];
  }
}

template <int G>
__global__ __launch_bounds__(64) void finish_kernel(State s, St4Dev q, SrceDev pp, StepFlags f) {
  const int g = static_cast<int>(threadIdx.x) / G;
  const int p = static_cast<int>(blockIdx.x) * (64 / G) + g;
  if (p >= s.npts) return;
  const int nk = pp.nk, nth = pp.nth, nspec = pp.nspec;
  const int l = static_cast<int>(threadIdx.x) % G;
  Sc c;
  load_sc(s, p, c);
  const size_t ps = static_cast<size_t>(p) * nspec;
  float* spec = s.spec + ps;
  const float* specinit = s.specinit + ps;
  const float* wn1 = s.wn1 + static_cast<size_t>(p) * nk;
  const float* cg1 = s.cg1 + static_cast<size_t>(p) * nk;
  const float* pin = s.pin + static_cast<size_t>(p) * N_PIN;
  float* taubbl = s.taubbl + static_cast<size_t>(p) * 2;
  float* tauice = s.tauice + static_cast<size_t>(p) * 2;
  const float dtg = f.dtg;
  c.v[S_DTDYN] = c.v[S_DTDYN] / static_cast<float>(mxi(c.iv[I_NSTEPS], 1));
  c.v[S_FCUT] = c.v[S_FHIGH] * kTpiInv;
  float efinish = 0.f, mwx = 0.f, mwy = 0.f;
  for (int base = 0; base < nk; base += G) {
    const int ik = base + l;
    float te = 0.f, tx = 0.f, ty = 0.f;
    if (ik < nk) {
      float eband = 0.f, a1 = 0.f, b1 = 0.f;
      for (int ith = 0; ith < nth; ++ith) {
        const float diff = specinit[ith + ik * nth] - spec[ith + ik * nth];
        eband = diff + eband;
        a1 = pp.ecos[ith] * diff + a1;
        b1 = pp.esin[ith] * diff + b1;
      }
      const float cgk = cg1[ik], w = wn1[ik];
      te = (pp.dden[ik] * eband) / cgk;
      tx = (((pp.dden[ik] * a1) / cgk) * w) / pp.sig[ik];
      ty = (w * ((pp.dden[ik] * b1) / cgk)) / pp.sig[ik];
    }
    const int cnt = nk - base;
#define ST(k) if constexpr ((k) < G)                          \
  if (k < cnt) {                       \
    efinish = bc<G, k>(te) + efinish;     \
    mwx = bc<G, k>(tx) + mwx;             \
    mwy = bc<G, k>(ty) + mwy;             \
  }
#define SY(r) if constexpr (G > 16 && 16 * (r) < G) if (16 * (r) < cnt) { efinish = row_take<G>(efinish, r); mwx = row_take<G>(mwx, r); mwy = row_take<G>(mwy, r); }
    WW_ROWS(ST, SY)
#undef SY
#undef ST
  }
  wave_sync();
  const float tauwix = c.v[S_TAUWIX], tauwiy = c.v[S_TAUWIY];
  const float tbx = taubbl[0], tby = taubbl[1];
  const float tauox = ((tauwix + mwx * kGrav) - tbx) / dtg;
  const float tauoy = ((tauwiy + mwy * kGrav) - tby) / dtg;
  c.v[S_TAUOX] = tauox;
  c.v[S_TAUOY] = tauoy;
  const float twix = tauwix / dtg, twiy = tauwiy / dtg;
  c.v[S_TAUWIX] = twix;
  c.v[S_TAUWIY] = twiy;
  c.v[S_TAUWNX] = c.v[S_TAUWNX] / dtg;
  c.v[S_TAUWNY] = c.v[S_TAUWNY] / dtg;
  {
    float sn, cs;
    libm::sincosf(c.v[S_USTDIR], &sn, &cs);
    const float ustar = c.v[S_USTAR], coef = pin[P_COEF], dair = pin[P_DAIR];
    const float kk = ustar * (ustar * (coef * (dair * coef)));
    c.v[S_TAUOCX] = kk * cs + (tauox - twix) * kDwat;
    c.v[S_TAUOCY] = kk * sn + (tauoy - twiy) * kDwat;
  }
  constexpr float dg = kDwat * kGrav;
  const float phiaw = c.v[S_PHIAW], phibbl = c.v[S_PHIBBL];
  c.v[S_PHIOC] = (((phiaw + efinish) - phibbl) * dg) / dtg;
  c.v[S_PHIAW] = (phiaw * dg) / dtg;
  c.v[S_PHIBBL] = (phibbl * dg) / dtg;
  const float ice = pin[P_ICE];
  float ti0 = tauice[0], ti1 = tauice[1];
  if (f.inflags24 && ice > 0.f && !pp.ic_numerics) {
    // Without ice dispersion ATT = 1: the spectrum is unchanged and the sums are of zeros.
    float phice = 0.f;
    ti0 = 0.f;
    ti1 = 0.f;
    for (int b0 = 0; b0 < nspec; b0 += G) {
      const int is = b0 + l;
      float d = 0.f, d0 = 0.f, d1 = 0.f;
      if (is < nspec) {
        const int ik0 = is / nth;
        const float factor = pp.dden[ik0] / cg1[ik0];
        const float factor2 = ((factor * kGrav) * wn1[ik0]) / pp.sig[ik0];
        const float a2 = spec[is];
        const float a = static_cast<float>(1.0 * static_cast<double>(a2));
        spec[is] = a;
        d = (a - a2) * factor;
        d0 = ((a - a2) * factor2) * pp.ecos[is];
        d1 = ((a - a2) * factor2) * pp.esin[is];
      }
      const int cnt = nspec - b0;
#define ST(k) if constexpr ((k) < G)                      \
  if (k < cnt) {                   \
    phice = phice + bc<G, k>(d);      \
    ti0 = ti0 - bc<G, k>(d0);         \
    ti1 = ti1 - bc<G, k>(d1);         \
  }
#define SY(r) if constexpr (G > 16 && 16 * (r) < G) if (16 * (r) < cnt) { phice = row_take<G>(phice, r); ti0 = row_take<G>(ti0, r); ti1 = row_take<G>(ti1, r); }
      WW_ROWS(ST, SY)
#undef SY
#undef ST
    }
    c.v[S_PHICE] = (((-1.f * kDwat) * kGrav) * phice) / dtg;
    ti0 = ti0 / dtg;
    ti1 = ti1 / dtg;
  }
  wave_sync();
  const int nlast = (nspec % 4 == 3) ? nspec - 1 : nspec;
  for (int is = l; is < nspec; is += G) {
    float a = f.it == 0 ? specinit[is] : spec[is];
    a = is < nlast ? mx(a, 0.f) : mx(0.f, a);
    spec[is] = a;
  }
  if (l == 0) {
    taubbl[0] = tbx / dtg;
    taubbl[1] = tby / dtg;
    tauice[0] = ti0;
    tauice[1] = ti1;
    float* psc = s.psc + static_cast<size_t>(p) * N_PSC;
    for (int i = 0; i < N_PSC; ++i) psc[i] = c.v[i];
    s.shave[p] = c.iv[I_SHAVE];
  }
}

template <int G>
inline dim3 groups(int n) { return dim3(static_cast<unsigned>((n + 64 / G - 1) / (64 / G))); }

template <int G>
inline unsigned lds_of(const St4Dev& q) {
  const Lds o(q.nk, q.nth, q.sintailpar[2] > 0.f);
  return static_cast<unsigned>((64 / G) * static_cast<size_t>(o.stride) * sizeof(float));
}

}  // namespace

size_t group_lds_bytes(int nk, int nth) {
  return 4 * static_cast<size_t>(Lds(nk, nth, true).stride) * sizeof(float);
}

size_t state_bytes(int nk, int nth, int npts) {
  const size_t n = static_cast<size_t>(npts);
  const size_t nspec = static_cast<size_t>(nk) * nth;
  const auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  size_t b = 0;
  b += 9 * al(nspec * n * 4);                       // specinit .. vsdb, rnw
  b += al(nspec * n);                               // llws
  b += al(static_cast<size_t>(nk) * n * 4);         // mft
  b += al(static_cast<size_t>(N_SCAL) * n * 4);
  b += al(static_cast<size_t>(N_ISCAL) * n * 4);
  b += al(n * 4);                                   // kdmean
  b += 2 * al(n * 4);                               // lists
  b += al(static_cast<size_t>(kMaxIter) * 4);       // counts
  return b;
}

State carve_state(void* base, int nk, int nth, int npts) {
  State s{};
  s.npts = npts;
  const size_t n = static_cast<size_t>(npts);
  const size_t nspec = static_cast<size_t>(nk) * nth;
  const auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  char* cur = static_cast<char*>(base);
  const auto take = [&](size_t bytes) {
    char* r = cur;
    cur += al(bytes);
    return r;
  };
  s.specinit = reinterpret_cast<float*>(take(nspec * n * 4));
  s.vsin = reinterpret_cast<float*>(take(nspec * n * 4));
  s.vdin = reinterpret_cast<float*>(take(nspec * n * 4));
  s.vsnl = reinterpret_cast<float*>(take(nspec * n * 4));
  s.vdnl = reinterpret_cast<float*>(take(nspec * n * 4));
  s.vsds = reinterpret_cast<float*>(take(nspec * n * 4));
  s.vdds = reinterpret_cast<float*>(take(nspec * n * 4));
  s.vsdb = reinterpret_cast<float*>(take(nspec * n * 4));
  s.rnw = reinterpret_cast<float*>(take(nspec * n * 4));
  s.llws = reinterpret_cast<unsigned char*>(take(nspec * n));
  s.mft = reinterpret_cast<float*>(take(static_cast<size_t>(nk) * n * 4));
  s.sc = reinterpret_cast<float*>(take(static_cast<size_t>(N_SCAL) * n * 4));
  s.isc = reinterpret_cast<int*>(take(static_cast<size_t>(N_ISCAL) * n * 4));
  s.kdmean = reinterpret_cast<float*>(take(n * 4));
  s.list[0] = reinterpret_cast<int*>(take(n * 4));
  s.list[1] = reinterpret_cast<int*>(take(n * 4));
  s.count = reinterpret_cast<int*>(take(static_cast<size_t>(kMaxIter) * 4));
  return s;
}

hipError_t k_init(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f, hipStream_t st) {
  hipLaunchKernelGGL(init_kernel<16>, groups<16>(s.npts), dim3(64), lds_of<16>(q), st, s, q, p, f);
  return hipGetLastError();
}

/// Step-kernel launch for list slots [first, n).
template <int G>
hipError_t launch_step(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f, int iter, int first,
                       int n, hipStream_t st) {
  const unsigned lds = lds_of<G>(q);
  const dim3 grid = groups<G>(n - first);
  if (p.nth == 36)
    hipLaunchKernelGGL((step_kernel<G, 36>), grid, dim3(64), lds, st, s, q, p, f, iter, first, n);
  else if (p.nth == 24)
    hipLaunchKernelGGL((step_kernel<G, 24>), grid, dim3(64), lds, st, s, q, p, f, iter, first, n);
  else
    hipLaunchKernelGGL((step_kernel<G, 0>), grid, dim3(64), lds, st, s, q, p, f, iter, first, n);
  return hipGetLastError();
}

namespace {
std::atomic<int> g_group_lanes{0};  // set_group_lanes
std::atomic<int> g_half_below{kHalfBelow}, g_wide_below{kWideBelow};  // set_width_thresholds
}  // namespace

/// Points that one round of 16-lane-row workgroups holds on the current device: the CUs times
/// the workgroups one CU takes (by LDS, at most eight: the step kernel's two wavefronts per
/// SIMD) times four; 0 if the device cannot be queried.
int rows_round(const St4Dev& q) {
  int dev = 0, cus = 0, lds_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
      cus <= 0 || lds_cu <= 0)
    return 0;
  const int per_cu = std::min(8, lds_cu / static_cast<int>(lds_of<16>(q)));
  return cus * per_cu * 4;
}

hipError_t k_step(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f, int iter, int n,
                  hipStream_t st) {
  // Many points: a 16-lane row per point (four per wavefront) for throughput.  Fewer: two
  // rows per point, then (the late sub-steps) a wavefront per point, which shorten each
  // point's sub-step while every wavefront still fits the GPU in one round.  Past one round
  // of rows, the points left over (at most what one round of two-row groups holds) go to a
  // second launch of two-row groups, which finishes sooner than a second round of rows.
  const int lanes = g_group_lanes.load(std::memory_order_relaxed);
  if (lanes == 16) return launch_step<16>(s, q, p, f, iter, 0, n, st);
  if (lanes == 32) return launch_step<32>(s, q, p, f, iter, 0, n, st);
  if (lanes == 64) return launch_step<64>(s, q, p, f, iter, 0, n, st);
  const int half_below = g_half_below.load(std::memory_order_relaxed);
  const int wide_below = g_wide_below.load(std::memory_order_relaxed);
  if (n > half_below) {
    const int round16 = rows_round(q);
    if (round16 > 0 && n > round16 && n - round16 <= half_below) {
      const hipError_t e = launch_step<16>(s, q, p, f, iter, 0, round16, st);
      if (e != hipSuccess) return e;
      return launch_step<32>(s, q, p, f, iter, round16, n, st);
    }
    return launch_step<16>(s, q, p, f, iter, 0, n, st);
  }
  if (n > wide_below) return launch_step<32>(s, q, p, f, iter, 0, n, st);
  return launch_step<64>(s, q, p, f, iter, 0, n, st);
}

void set_group_lanes(int lanes) { g_group_lanes.store(lanes, std::memory_order_relaxed); }

void set_width_thresholds(int half_below, int wide_below) {
  g_half_below.store(half_below, std::memory_order_relaxed);
  g_wide_below.store(wide_below, std::memory_order_relaxed);
}

hipError_t k_routine(const St4Call& io, const St4Dev& q, hipStream_t st) {
  if (io.npts <= 0) return hipSuccess;
  const dim3 grid((io.npts + 3) / 4);
  const unsigned lds = lds_of<16>(q);
  if (q.nth == 36)
    hipLaunchKernelGGL((routine_kernel<36>), grid, dim3(64), lds, st, io, q);
  else if (q.nth == 24)
    hipLaunchKernelGGL((routine_kernel<24>), grid, dim3(64), lds, st, io, q);
  else
    hipLaunchKernelGGL((routine_kernel<0>), grid, dim3(64), lds, st, io, q);
  return hipGetLastError();
}

hipError_t k_finish(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f, hipStream_t st) {
  hipLaunchKernelGGL(finish_kernel<16>, groups<16>(s.npts), dim3(64), 0, st, s, q, p, f);
  return hipGetLastError();
}

}  // namespace srce
}  // namespace ww3gpu

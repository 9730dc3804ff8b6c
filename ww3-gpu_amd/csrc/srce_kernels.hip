// ww3-gpu_amd/csrc/srce_kernels.hip
// W3SRCE with the ST4 source terms on gfx950: one sea point per lane.
//
// Every per-point statement of the reference keeps its float32 operation order
// (-ffp-contract=off), its sequential sums in IK/ITH order, and its MIN/MAX operand order:
//   W3SPR4 + CALC_USTAR  kokkos/src/ww_kokkos/st4_spr4.hpp:40-162, st4_calc_ustar.hpp:36-122
//   W3SIN4               kokkos/src/ww_kokkos/st4_sin4.hpp:69-277, 472-627
//   W3SDS4               kokkos/src/ww_kokkos/st4_sds4.hpp:122-331, 574-711
//   W3SLN1, W3SBT1, W3SDB1  srce_ln1.hpp:41-99, srce_bt1.hpp:26-58, srce_db1.hpp:44-149
//   W3SRCE               kokkos/src/ww_kokkos/srce_kernel.hpp:212-583
// The DIA of each sub-step runs between k_step_a and k_step_c as the workgroup-per-point
// kernel of snl1_*.hip on the same resident arrays.
//
// State is bin-major (srce.hpp): array X of point p, bin i is X[i*P + p], so the lanes of
// a wavefront (consecutive points) load one contiguous line per bin.
//
// Fused, same arithmetic: W3SRCE's three passes after the time-step limit (spectrum update,
// W3SDB1 update, flux sums) run as one pass over the bins, each bin's statements in the
// reference order (bins do not interact); W3SLN1's S(IK,ITH) = WNF(IK)*DIRF(ITH) and
// W3SBT1's D are recomputed where read instead of being stored; sincosf(USDIR) of W3SDS4's
// turbulence loop is computed once per point (the reference calls it per direction with the
// same argument).  W3SDS4's mean-square-slope block (SSDSC(8), (11), (18) > 0) is skipped:
// under SDSBCHOICE = 1 nothing reads MSSSUM.
#include "devlibm.hpp"
#include "srce.hpp"

namespace ww3gpu {
namespace srce {
namespace {

constexpr float kGrav = 9.806f;
constexpr float kDwat = 1000.f;
constexpr float kNuAir = 1.4E-5f;
constexpr float kKappa = 0.40f;
constexpr float kPi = static_cast<float>(3.141592653589793);
constexpr float kTpi = 2.0f * kPi;
constexpr float kTpiInv = 1.0f / kTpi;
constexpr int kBlock = 64;

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

/// One point's view of the bin-major state.
struct Pt {
  int p, P;
  __device__ float& operator()(float* x, int i) const { return x[static_cast<size_t>(i) * P + p]; }
  __device__ int& operator()(int* x, int i) const { return x[static_cast<size_t>(i) * P + p]; }
};

// ------------------------------------------------------------------------------------
// CALC_USTAR (iterative branch, SINTAILPAR(1) < 0.5)
// ------------------------------------------------------------------------------------
__device__ void calc_ustar(const St4Dev& q, float windspeed, float tauw, float& ustar, float& z0,
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
    const float lg = libm::logf(q.zzwnd / zz0);
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
// W3SPR4
// ------------------------------------------------------------------------------------
__device__ void spr4(const State& s, const St4Dev& q, const Pt& t) {
  const int nk = q.nk, nth = q.nth;
  const float eps = 1.E-7f;
  float* sc = s.sc;
  const float u = t(s.pin, P_U10ABS), udir = t(s.pin, P_U10DIR);
  float ustar = mx(0.0001f, t(sc, S_USTAR));
  float emean = 0.f, emeanws = 0.f, fmeanws = 0.f, fmean = 0.f, fmean1 = 0.f, wnm = 0.f;
  float amax = 0.f, elcs = 0.f, elsn = 0.f;
  // Section 1 (ELCS, ELSN, AMAX over every bin) must end before DLWMEAN; the band sums
  // EB(IK), EB2(IK) it also forms feed only section 2, so section 2 forms them again, in
  // the same order, instead of keeping NK values per lane.
  for (int ik = 1; ik <= nk; ++ik) {
    const float sigfac = libm::powf(q.sig[ik], q.ssdsc[11]) * q.dden[ik - 1] / t(s.cg1, ik - 1);
    for (int ith = 1; ith <= nth; ++ith) {
      const int is = ith + (ik - 1) * nth;
      const float aa = t(s.spec, is - 1);
      elcs = elcs + aa * q.ecos[is - 1] * sigfac;
      elsn = elsn + aa * q.esin[is - 1] * sigfac;
      amax = mx(amax, aa);
    }
  }
  const float dlw = libm::atan2f(elsn, elcs);
  const float tw = 2.f * q.wwnmeanptail;
  float elast = 0.f, e2last = 0.f;
  for (int ik = 1; ik <= nk; ++ik) {
    float eb = 0.f, eb2 = 0.f;
    for (int ith = 1; ith <= nth; ++ith) {
      const int is = ith + (ik - 1) * nth;
      const float aa = t(s.spec, is - 1);
      eb = eb + aa;
      if (t(s.llws, is - 1) != 0) eb2 = eb2 + aa;
    }
    const float pws = libm::powf(q.sig[ik], tw);
    const float pwk = libm::powf(t(s.wn1, ik - 1), q.wwnmeanp);
    const float cgk = t(s.cg1, ik - 1);
    const float e = eb * q.dden[ik - 1] / cgk;
    const float e2 = eb2 * q.dden[ik - 1] / cgk;
    elast = e;
    e2last = e2;
    emean = emean + e;
    fmean = fmean + e / q.sig[ik];
    fmean1 = fmean1 + e * pws;
    wnm = wnm + e * pwk;
    emeanws = emeanws + e2;
    fmeanws = fmeanws + e2 * pws;
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
  const float tx = t(sc, S_TAUWX), ty = t(sc, S_TAUWY);
  const float tauw = __builtin_sqrtf(tx * tx + ty * ty);
  float z0 = 0.f, charn = t(sc, S_CHARN);
  calc_ustar(q, u, tauw, ustar, z0, charn);
  const float r = ustar / mx(0.01f, u);
  t(sc, S_USTAR) = ustar;
  t(sc, S_USTDIR) = udir;
  t(sc, S_EMEAN) = emean;
  t(sc, S_FMEAN) = fmean;
  t(sc, S_FMEAN1) = fmean1;
  t(sc, S_WNMEAN) = wnm;
  t(sc, S_AMAX) = amax;
  t(sc, S_CD) = r * r;
  t(sc, S_Z0) = z0;
  t(sc, S_CHARN) = charn;
  t(sc, S_FMEANWS) = fmeanws;
  t(sc, S_DLWMEAN) = dlw;
}

// ------------------------------------------------------------------------------------
// W3SIN4 (W3SRCE's call: A = SPEC, K = WN2 (WN1 per frequency), U = U10ABS,
// USDIR = U10DIR); S, D into VSIN, VDIN; TAUWNX/Y into TAUWAX/Y.
// ------------------------------------------------------------------------------------
__device__ void sin4(const State& s, const St4Dev& q, const Pt& t) {
  const int nk = q.nk, nth = q.nth, nspec = q.nspec;
  float* sc = s.sc;
  const float u = t(s.pin, P_U10ABS), usdir = t(s.pin, P_U10DIR), as = t(s.pin, P_AS);
  const float ustar = t(sc, S_USTAR), drat = t(sc, S_DRAT), z0 = t(sc, S_Z0);
  constexpr float kappa2 = kKappa * kKappa;
  constexpr float z0viscnum = 0.1f * kNuAir;
  constexpr float twonu = 2.f * kNuAir;
  constexpr float grav2tpi = (kGrav * kGrav) * kTpi;
  constexpr float logtpi005 = 7.1361942291259765625f;  // LOG(TPI/0.005) as folded
  constexpr float eps1 = 0.00001f, eps2 = 0.000001f, eps2sq = eps2 * eps2;
  float ss[3][2] = {{0, 0}, {0, 0}, {0, 0}}, sn[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  const float const1 = q.bbeta / kappa2;
  float const0 = const1 * drat;
  const float z0visc = z0viscnum / mx(ustar, 0.0001f);
  const float z0noz = mx(z0visc, q.zz0rat * z0);
  float uorb = 0.f, aorb = 0.f;
  for (int ik = 1; ik <= nk; ++ik) {
    float eb = 0.f;
    for (int ith = 1; ith <= nth; ++ith) eb = eb + t(s.spec, ith + (ik - 1) * nth - 1);
    const float sk = q.sig[ik];
    const float cgk = t(s.cg1, ik - 1);
    uorb = uorb + eb * (sk * sk) * q.dden[ik - 1] / cgk;
    aorb = aorb + eb * q.dden[ik - 1] / cgk;
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

  // The stress scan.  DSTAB(ISTAB, IS): the last pass (ISTAB = 3, or 2 under STAB3) goes to
  // VDIN; under STAB3 the first pass is kept in VSIN until the average is formed.
  float ust = ustar, taux = 0.f, tauy = 0.f, cosu = 0.f, sinu = 0.f, ustp = 0.f, usdirp = 0.f;
  const int i0 = q.stab3 ? 1 : 3, i1 = q.stab3 ? 2 : 3;
  for (int istab = i0; istab <= i1; ++istab) {
    float* const dst = (istab == 1) ? s.vsin : s.vdin;
    if (istab == 1) ust = ustar * (1.f - ustarsigma);
    if (istab == 2) ust = ustar * (1.f + ustarsigma);
    float cu, su;
    libm::sincosf(usdir, &su, &cu);
    taux = (ust * ust) * cu;
    tauy = (ust * ust) * su;
    ss[istab - 1][0] = ss[istab - 1][1] = 0.f;
    sn[istab - 1][0] = sn[istab - 1][1] = 0.f;
    for (int ik = 1; ik <= nk; ++ik) {
      const float taupx = taux - __builtin_fabsf(q.ttauwshelter) * ss[istab - 1][0];
      const float taupy = tauy - __builtin_fabsf(q.ttauwshelter) * ss[istab - 1][1];
      ustp = mn(libm::powf((taupx * taupx + taupy * taupy), 0.25f), mx(ust, 0.3f));
      usdirp = libm::atan2f(taupy, taupx);
      libm::sincosf(usdirp, &sinu, &cosu);
      const int is1 = 1 + (ik - 1) * nth;
      const float kis = t(s.wn1, ik - 1);
      const float s2 = q.sig2[is1 - 1];
      const float cm = kis / s2;
      const float ucn = ustp * cm + q.zzalp;
      const float const2 = q.dden2[is1 - 1] / t(s.cg1, ik - 1) * kGrav / (q.sig[ik] / kis * drat);
      const float cnst = s2 * const0;
      const float zcn = libm::logf(kis * z0);
      const float swv = -(q.sswellf[4] * drat * 2.f * kis * __builtin_sqrtf(twonu * s2));
      const float swt = -(drat * q.sswellf[0] * 16.f * (s2 * s2) / kGrav);
      for (int ith = 1; ith <= nth; ++ith) {
        const int is = ith + (ik - 1) * nth;
        const float ec = q.ecos[is - 1], es = q.esin[is - 1];
        const float coswind = (ec * cosu + es * sinu);
        float dv;
        int lw;
        if (coswind > 0.01f) {
          const float x = coswind * ucn;
          const float zarg = kKappa / x;
          const float zlog = zcn + zarg;
          if (zlog < 0.f) {
            const float zl2 = zlog * zlog;
            dv = cnst * libm::expf(zlog) * (zl2 * zl2) * ucn * ucn * libm::powf(coswind, q.ssinthp);
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
        t(s.llws, is - 1) = lw;
        t(dst, is - 1) = dv;
        const float temp2 = const2 * dv * t(s.spec, is - 1);
        if (dv < 0.f) {
          sn[istab - 1][0] = sn[istab - 1][0] + temp2 * ec;
          sn[istab - 1][1] = sn[istab - 1][1] + temp2 * es;
        } else {
          ss[istab - 1][0] = ss[istab - 1][0] + temp2 * ec;
          ss[istab - 1][1] = ss[istab - 1][1] + temp2 * es;
        }
      }
    }
  }
  float xstress, ystress, tauwnx, tauwny;
  if (!q.stab3) {
    xstress = ss[2][0];
    ystress = ss[2][1];
    tauwnx = sn[2][0];
    tauwny = sn[2][1];
  } else {
    for (int is = 0; is < nspec; ++is) t(s.vdin, is) = 0.5f * (t(s.vsin, is) + t(s.vdin, is));
    xstress = 0.5f * (ss[0][0] + ss[1][0]);
    ystress = 0.5f * (ss[0][1] + ss[1][1]);
    tauwnx = 0.5f * (sn[0][0] + sn[1][0]);
    tauwny = 0.5f * (sn[0][1] + sn[1][1]);
  }
  for (int is = 0; is < nspec; ++is) t(s.vsin, is) = t(s.vdin, is) * t(s.spec, is);
  {
    const float taupx = taux - __builtin_fabsf(q.ttauwshelter) * xstress;
    const float taupy = tauy - __builtin_fabsf(q.ttauwshelter) * ystress;
    ustp = libm::powf((taupx * taupx + taupy * taupy), 0.25f);
    usdirp = libm::atan2f(taupy, taupx);
  }
  ust = ustp;
  {
    const float sg = q.sig[nk], s2 = sg * sg, s5 = s2 * (s2 * sg);
    const0 = q.dth * s5 / grav2tpi * kTpi * sg / t(s.cg1, nk - 1);
  }
  float temp = 0.f;
  for (int ith = 1; ith <= nth; ++ith) {
    const int is = ith + (nk - 1) * nth;
    const float cw = (q.ecos[is - 1] * cosu + q.esin[is - 1] * sinu);
    const float cp = mx(cw, 0.f);
    temp = temp + t(s.spec, is - 1) * (cp * (cp * cp));
  }
  const float levtail0 = const0 * temp;
  float tauhf;
  {
    float ustr = ust;
    const float zz0 = z0, sg = q.sig[nk];
    const float zinf = libm::logf(sg * sg / kGrav);
    const float dely = (logtpi005 - zinf) / 50.f;
    float taur = ust * ust, tau1 = 0.f;
    for (int j = 1; j <= 50; ++j) {
      const float wj = (j == 1 || j == 50) ? 0.5f : 1.f;
      const float y = libm::expf(zinf + static_cast<float>(j - 1) * dely);
      const float ykm = (y / 363.f);
      const float ykm2 = ykm * ykm;
      const float tensk = 1.f + ykm2;
      const float cm = __builtin_sqrtf(kGrav * tensk / y);
      const float cgtail = 0.5f * (3.f * ykm2 + 1.f) * __builtin_sqrtf(kGrav / (y * tensk));
      const float zx = ustr / cm + q.zzalp;
      const float zarg = mn(kKappa / zx, 20.f);
      const float zmu = mn(kGrav * zz0 / (cm * cm) * libm::expf(zarg), 1.f);
      const float zlog = mn(libm::logf(zmu), 0.f);
      const float zl2 = zlog * zlog;
      const float zbeta = const1 * zmu * (zl2 * zl2);
      float levtail;
      if (q.sintailpar[2] > 0.f) {
        float alpham;
        if (ustr < cm)
          alpham = mx(0.f, 0.01f * (1.f + libm::logf(ustr / cm)));
        else
          alpham = 0.01f * (1.f + 3.f * libm::logf(ustr / cm));
        const float ykm1 = (ykm - 1.f);
        const float fm = libm::expf(-0.25f * (ykm1 * ykm1));
        const float alphat = alpham * (0.2325f / cm) * fm;
        const float th = libm::tanhf((y - 20.f) / 5.f);
        levtail = levtail0 * 0.5f * (1.f - th) + q.sintailpar[2] * 0.5f * (1.f + th) * alphat;
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
  t(sc, S_TAUWX) = tauwx;
  t(sc, S_TAUWY) = tauwy;
  t(sc, S_TAUWAX) = tauwnx;
  t(sc, S_TAUWAY) = tauwny;
}

// ------------------------------------------------------------------------------------
// W3SDS4 (W3SRCE's call: A = SPEC, K = WN1, USDIR = USTDIR), SDSBCHOICE = 1.
// ------------------------------------------------------------------------------------
__device__ void sds4(const State& s, const St4Dev& q, const StepFlags& f, const Pt& t) {
  const int nk = q.nk, nth = q.nth, nspec = q.nspec;
  const float* c = q.ssdsc;
  float* sc = s.sc;
  const float ustar = t(sc, S_USTAR), usdir = t(sc, S_USTDIR), depth = t(sc, S_DEPTH);
  const float dair = t(s.pin, P_DAIR);
  const float facturb = (((c[4] * (ustar * ustar)) / kGrav) * dair) / kDwat;
  const int dikcumul = nint(q.ssdsbrf1 / (q.xfr - 1.f));
  const int ik1 = q.ig1 ? nint(q.igpars5) + 1 : 1;
  constexpr float twopisq = 2.f * (kPi * kPi);
  const float epsr = __builtin_sqrtf(q.ssdsbr);
  // Saturation, spontaneous breaking, PB, BRLAMBDA, SRHS = DDIAG*A; bins below IK1 are 0.
  for (int is = 0; is < ik1 - 1 && is < nk; ++is) t(s.bth0, is) = 0.f;
  for (int i = 0; i < (ik1 - 1) * nth && i < nspec; ++i) {
    t(s.vdds, i) = 0.f;
    t(s.vsds, i) = 0.f * t(s.spec, i);  // SRHS = DDIAG * A with DDIAG = 0
    t(s.brlambda, i) = 0.f;  // BTH = 0 there, so PB = MAX(-EPSR, 0)**2 * 28.16 = 0
  }
  for (int ik = ik1; ik <= nk; ++ik) {
    const float kk = t(s.wn1, ik - 1);
    const float facsat = (q.sig[ik] * ((kk * kk) * kk)) * q.dth;
    const int is0 = (ik - 1) * nth;
    float asum = 0.f;
    for (int ith = 0; ith < nth; ++ith) asum = asum + t(s.spec, is0 + ith);
    float bth0 = asum * facsat;
    float miche;
    if (q.ssdsbm[0] == 1.f) {
      miche = 1.f;
    } else {
      const float x = libm::tanhf(mn(kk * depth, 10.f));
      const float tt = x * (q.ssdsbm[1] + x * (q.ssdsbm[2] + x * (q.ssdsbm[3] + x * q.ssdsbm[4])));
      miche = tt * tt;
    }
    const float coef1 = q.ssdsbr * miche;
    if (q.ssdsdth >= 180.f) {
      // BTH = BTH0 on every direction
      const float m = mx(bth0 / coef1 - 1.f, 0.f);
      const float sdiagiso = ((c[1] * q.sig[ik]) * c[5]) * (m * m);
      const float coef2 = ((c[1] * q.sig[ik]) * (1.f - c[5])) / (coef1 * coef1);
      for (int ith = 0; ith < nth; ++ith) {
        const int is = is0 + ith;
        const float dd = sdiagiso + coef2 * libm::powf(mx(bth0 - coef1, 0.f), q.ssdsp);
        const float mm = mx(__builtin_sqrtf(bth0) - epsr, 0.f);
        const float pb = (mm * mm) * 28.16f;
        t(s.vdds, is) = dd;
        t(s.brlambda, is) = pb / twopisq;
        t(s.vsds, is) = dd * t(s.spec, is);
      }
    } else {
      // BTH(IS) = FACSAT * SUM_J SATWEIGHTS(J,ITH)*A(SATINDICES(J,ITH)), then MAXVAL; BTH
      // is kept in BRLAMBDA until the MAXVAL is known.
      float lim = -__builtin_inff();
      bool found = false;
      for (int ith = 0; ith < nth; ++ith) {
        float acc = 0.f;
        const float* w = q.satw + q.nsat * ith;
        const int* x = q.sati + q.nsat * ith;
        for (int j = 0; j < q.nsat; ++j) acc = acc + w[j] * t(s.spec, is0 + x[j] - 1);
        const float b = acc * facsat;
        t(s.brlambda, is0 + ith) = b;
        if (!found) {
          if (b >= lim) {
            lim = b;
            found = true;
          }
        } else if (b > lim) {
          lim = b;
        }
      }
      bth0 = found ? lim : __builtin_nanf("");
      const float m = mx(bth0 / coef1 - 1.f, 0.f);
      const float sdiagiso = ((c[1] * q.sig[ik]) * c[5]) * (m * m);
      const float coef2 = ((c[1] * q.sig[ik]) * (1.f - c[5])) / (coef1 * coef1);
      for (int ith = 0; ith < nth; ++ith) {
        const int is = is0 + ith;
        const float b = t(s.brlambda, is);
        const float dd = sdiagiso + coef2 * libm::powf(mx(b - coef1, 0.f), q.ssdsp);
        const float mm = mx(__builtin_sqrtf(b) - epsr, 0.f);
        const float pb = (mm * mm) * 28.16f;
        t(s.vdds, is) = dd;
        t(s.brlambda, is) = pb / twopisq;
        t(s.vsds, is) = dd * t(s.spec, is);
      }
    }
    t(s.bth0, ik - 1) = bth0;
  }

  // Wave-turbulence interaction and the cumulative effect.
  if (c[2] != 0.f || c[4] != 0.f || c[20] != 0.f) {
    float su, cu;
    libm::sincosf(usdir, &su, &cu);
    for (int ik = ik1; ik <= nk; ++ik) {
      const float kk = t(s.wn1, ik - 1);
      const float facturb2 = -(((q.sig[ik] * 2.f) * kk) * facturb);
      const float dvisc = -(((c[20] * 4.f) * kk) * kk);
      const float cc = q.sig[ik] / kk;
      float renew = 0.f;
      if (c[2] > 0.f && ik > dikcumul) {
        for (int ik2 = ik1; ik2 <= ik - dikcumul; ++ik2) {
          const float c2 = q.sig[ik2] / t(s.wn1, ik2 - 1);
          const int is2 = (ik2 - 1) * nth;
          const float cumulwiso = ((__builtin_fabsf(c2 - cc) * q.dsip[ik2]) / (c2 * 0.5f)) * q.dth;
          float v = 0.f;
          for (int i = is2; i < is2 + nth; ++i) v = v + t(s.brlambda, i);
          renew = renew - cumulwiso * v;
        }
      }
      for (int ith = 1; ith <= nth; ++ith) {
        const int is = ith + (ik - 1) * nth;
        if (c[2] < 0.f && ik > dikcumul) {
          renew = 0.f;
          for (int ik2 = ik1; ik2 <= ik - dikcumul; ++ik2) {
            if (t(s.bth0, ik2 - 1) > q.ssdsbr) {
              const int is2 = (ik2 - 1) * nth;
              const float* cw = q.cumulw + static_cast<size_t>(nspec) * (is - 1) + is2;
              float v = 0.f;
              for (int j = 0; j < nth; ++j) v = v + cw[j] * t(s.brlambda, is2 + j);
              renew = renew + v;
            }
          }
        }
        const float coswind = q.ecos[ith - 1] * cu + q.esin[ith - 1] * su;
        const float dturb = facturb2 * mx(coswind, 0.f);
        const float diag2 = (c[2] * renew + dturb) + dvisc;
        t(s.vdds, is - 1) = t(s.vdds, is - 1) + diag2;
        t(s.vsds, is - 1) = t(s.vsds, is - 1) + t(s.spec, is - 1) * diag2;
      }
    }
  }

  // Whitecap parameters (early return keeps WHITECAP).  COEF4/COEF5 are zero above IKC;
  // BRM12 is never set by this routine and stays 0.
  if (!(f.flogrd57 || f.flogrd58)) return;
  float wc0 = 0.f, wc1 = 0.f, wc3 = 0.f;
  t(s.whitecap, 2) = 0.f;
  const int ikc = ffloor(q.aaircmin) < nk ? ffloor(q.aaircmin) : nk;
  const float tstrfac = (0.8f * 2.f) * kPi, tmaxfac = (5.f * 2.f) * kPi;
  for (int ik = 1; ik <= nk; ++ik) {
    float coef4 = 0.f, coef5 = 0.f;
    const float kk = t(s.wn1, ik - 1);
    if (ik <= ikc) {
      const float cc = q.sig[ik] / kk;
      const int is0 = (ik - 1) * nth;
      float v = 0.f;
      for (int i = is0; i < is0 + nth; ++i) v = v + t(s.brlambda, i);
      const float cgk = t(s.cg1, ik - 1);
      coef4 = ((((((cc * cc) * v) * 2.f) * kPi) / kGrav) * c[6]) * q.dden[ik - 1] / (q.sig[ik] * cgk);
      const float pw = libm::powf(0.f / c[1], 3.f / 5.f);
      float ww = 0.f;
      for (int i = is0; i < is0 + nth; ++i) ww = ww + (t(s.brlambda, i) * (-c[1])) * pw;
      coef5 = ((((cc * cc) * cc) * ww) * q.aairgb / kGrav) * q.dden[ik - 1] / (q.sig[ik] * cgk);
    }
    if (f.flogrd57 && ik >= ik1 && ik <= ikc) {
      wc0 = wc0 + coef4 * (1.f - wc0);
      wc3 = wc3 + coef5;
    }
    if (f.flogrd58 && ik >= ik1) {
      const float tstr = tstrfac / q.sig[ik];
      const float tmax = tmaxfac / q.sig[ik];
      const float dt = tmax / 50.f;
      float mft = 0.f;
      for (int it = 1; it <= 50; ++it) {
        const float tt = static_cast<float>(it) * dt;
        if (tt < tstr)
          mft = mft + ((0.4f / (kk * tstr)) * tt) * dt;
        else
          mft = mft + ((0.4f / kk) * libm::expf(-((tt - tstr) / 3.8f))) * dt;
      }
      mft = mft / tmax;
      wc1 = wc1 + coef4 * mft;
    }
  }
  t(s.whitecap, 0) = wc0;
  t(s.whitecap, 1) = wc1;
  t(s.whitecap, 3) = wc3;
}

// ------------------------------------------------------------------------------------
// W3SLN1 -> DIRF, WNF (VSLN(IS) = WNF(IK) * DIRF(ITH))
// ------------------------------------------------------------------------------------
__device__ void sln1(const State& s, const SrceDev& p, const Pt& t) {
  const int nk = p.nk, nth = p.nth;
  const float ustar = t(s.sc, S_USTAR), usdir = t(s.pin, P_U10DIR), fhigh = t(s.sc, S_FHIGH);
  float cosu = 0.f, sinu = 0.f;
  libm::sincosf(usdir, &sinu, &cosu);
  for (int ith = 1; ith <= nth; ++ith) {
    const float m = fmax2(0.f, p.ecos[ith - 1] * cosu + p.esin[ith - 1] * sinu);
    const float m2 = m * m;
    t(s.dirf, ith - 1) = m2 * m2;
  }
  const float u2 = ustar * ustar;
  const float fac = p.slnc1 * (u2 * u2);
  const float ff1 = p.fspm * kGrav / (28.f * ustar);
  const float ff2 = p.fshf * fmin2(p.sig[nk - 1], fhigh);
  const float ffilt = fmin2(2.f * p.sig[nk - 1], fmax2(ff2, ff1));
  for (int ik = 1; ik <= nk; ++ik) {
    const float rfr = p.sig[ik - 1] / ffilt;
    float w;
    if (rfr < 0.5f) {
      w = 0.f;
    } else {
      const float r2 = rfr * rfr;
      const float rm4 = 1.f / (r2 * r2);
      w = fac / t(s.wn1, ik - 1) * libm::expf(-rm4);
    }
    t(s.wnf, ik - 1) = w;
  }
}

// W3SBT1 -> CBETA and the cut-off NSCUT: D(IS) = CBETA(MAPWN(IS)) for IS <= NSCUT, else 0;
// S = D*A, except after the early return (NSCUT = -1), where S = +0 on every bin.
__device__ void sbt1(const State& s, const SrceDev& p, const Pt& t) {
  const int nk = p.nk, nth = p.nth;
  const float depth = t(s.sc, S_DEPTH);
  if (depth * t(s.wn1, 0) > 6.f) {
    t(s.isc, I_NSCUT) = -1;  // D = 0 and S = +0 on every bin
    return;
  }
  const float factor = p.sbtc1 / depth;
  int ik = 1;
  for (; ik <= nk; ++ik) {
    const float w = t(s.wn1, ik - 1);
    if (w * depth > 6.f) break;
    const float x = t(s.cg1, ik - 1) * w / p.sig[ik - 1] - 0.5f;
    t(s.cbeta, ik - 1) = factor * (x > 0.f ? x : 0.f);
  }
  t(s.isc, I_NSCUT) = (ik - 1) * nth;
}

// W3SDB1 -> VSDB (and its constant D in S_DTDB); the early return keeps both.
__device__ void sdb1(const State& s, const SrceDev& p, const Pt& t) {
  const int nk = p.nk, nth = p.nth, nspec = p.nspec;
  const float emean = t(s.sc, S_EMEAN), fmean = t(s.sc, S_FMEAN), wnmean = t(s.sc, S_WNMEAN);
  const float depth = t(s.sc, S_DEPTH);
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
  (void)nk;
  (void)nth;
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
  t(s.sc, S_DTDB) = dv;
  for (int is = 0; is < nspec; ++is) t(s.vsdb, is) = dv * t(s.spec, is);
}

// ------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------
__device__ inline bool lane_point(const State& s, Pt& t) {
  const int p = static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x);
  if (p >= s.npts) return false;
  t.p = p;
  t.P = s.P;
  return true;
}

__global__ __launch_bounds__(kBlock) void init_kernel(State s, St4Dev q, SrceDev p, StepFlags f) {
  Pt t;
  if (!lane_point(s, t)) return;
  const int nk = p.nk, nth = p.nth, nspec = p.nspec;
  float* sc = s.sc;
  for (int is = 0; is < nspec; ++is) {
    t(s.vsin, is) = 0.f;
    t(s.vdin, is) = 0.f;
    t(s.vsnl, is) = 0.f;
    t(s.vdnl, is) = 0.f;
    t(s.vsds, is) = 0.f;
    t(s.vdds, is) = 0.f;
    t(s.vsdb, is) = 0.f;
    t(s.brlambda, is) = 0.f;
  }
  t(sc, S_DTDB) = 0.f;
  for (int ik = 0; ik < nk; ++ik) {
    t(s.wnf, ik) = 0.f;
    t(s.cbeta, ik) = 0.f;
  }
  for (int ith = 0; ith < nth; ++ith) t(s.dirf, ith) = 0.f;
  t(s.isc, I_NSCUT) = 0;
  const float d_inp = t(s.pin, P_D_INP), ice = t(s.pin, P_ICE), dair = t(s.pin, P_DAIR);
  (void)ice;
  t(sc, S_DEPTH) = mx(d_inp, p.dmin);
  t(sc, S_DRAT) = dair / kDwat;
  t(sc, S_DTDYN) = 0.f;
  t(sc, S_DTTOT) = 0.f;
  t(s.isc, I_NSTEPS) = 0;
  t(s.isc, I_DONE) = 0;
  t(sc, S_PHIAW) = 0.f;
  t(sc, S_CHARN) = 0.f;
  t(sc, S_TWS) = 0.f;
  t(sc, S_PHIBBL) = 0.f;
  t(sc, S_TAUWIX) = 0.f;
  t(sc, S_TAUWIY) = 0.f;
  t(sc, S_TAUWNX) = 0.f;
  t(sc, S_TAUWNY) = 0.f;
  t(sc, S_TAUWAX) = 0.f;
  t(sc, S_TAUWAY) = 0.f;
  t(s.taubbl, 0) = 0.f;
  t(s.taubbl, 1) = 0.f;
  t(s.tauice, 0) = 0.f;
  t(s.tauice, 1) = 0.f;
  t(sc, S_PHICE) = 0.f;
  t(sc, S_TAUOCX) = 0.f;
  t(sc, S_TAUOCY) = 0.f;
  t(sc, S_WNMEAN) = 0.f;
  t(sc, S_DLWMEAN) = 0.f;
  for (int j = 0; j < 4; ++j) t(s.whitecap, j) = 0.f;
  const float sintail4 = q.sintailpar[3];
  if (sintail4 > 0.5f) {
    t(sc, S_TAUWX) = 0.f;
    t(sc, S_TAUWY) = 0.f;
  }
  if (f.it == 0) {
    for (int is = 0; is < nspec; ++is) t(s.llws, is) = 1;
    t(sc, S_TAUWX) = 0.f;
    t(sc, S_TAUWY) = 0.f;
    t(sc, S_USTAR) = 0.f;
    t(sc, S_USTDIR) = 0.f;
  } else {
    // LLWS is read here before anything sets it; only FMEANWS depends on it and the
    // W3SPR4 below overwrites it (TAUWBUG, the only configuration accepted with IT > 0).
    for (int is = 0; is < nspec; ++is) t(s.llws, is) = 0;
    spr4(s, q, t);
    if (sintail4 > 0.5f) sin4(s, q, t);
  }
  spr4(s, q, t);
  t(sc, S_TWS) = 1.f / t(sc, S_FMEANWS);
  for (int is = 0; is < nspec; ++is) t(s.specinit, is) = t(s.spec, is);
  t(sc, S_FHIGH) = mx((p.ffxfm + 0.f) * mx(t(sc, S_FMEANWS), t(sc, S_FMEAN1)),
                      p.ffxpm / t(sc, S_USTAR));
}

__global__ __launch_bounds__(kBlock) void step_a_kernel(State s, St4Dev q, SrceDev p, StepFlags f) {
  Pt t;
  if (!lane_point(s, t)) return;
  if (t(s.isc, I_DONE)) return;
  t(s.isc, I_NSTEPS) = t(s.isc, I_NSTEPS) + 1;
  if (p.ln1) sln1(s, p, t);
  sin4(s, q, t);
  s.kdmean[t.p] = t(s.sc, S_WNMEAN) * t(s.sc, S_DEPTH);
  (void)f;
}

__global__ __launch_bounds__(kBlock) void step_c_kernel(State s, St4Dev q, SrceDev p, StepFlags f,
                                                        int iter) {
  Pt t;
  if (!lane_point(s, t)) return;
  if (t(s.isc, I_DONE)) return;
  const int nk = p.nk, nth = p.nth, nspec = p.nspec;
  float* sc = s.sc;
  sds4(s, q, f, t);
  sdb1(s, p, t);
  sbt1(s, p, t);
  const float ice = t(s.pin, P_ICE);
  const float isln = mx(mn(1.f - ice * p.icescales[0], 1.f), 0.f);
  const float isin = mx(mn(1.f - ice * p.icescales[1], 1.f), 0.f);
  const float isnl = mx(mn(1.f - ice * p.icescales[2], 1.f), 0.f);
  const float isds = mx(0.f, mn(1.f - ice * p.icescales[3], 1.f));
  const float dtg = f.dtg;
  float fhigh = t(sc, S_FHIGH);
  int nkh = mni(nk, fint(p.facti1 * libm::logf(mx(1.E-7f, fhigh)) + p.facti2));
  const int nkh1 = mni(nk, nkh + 1);
  // NSPECH = NKH1*NTH with the wrap of the compiled reference (NKH = INT_MIN when FHIGH is
  // +Inf, i.e. USTAR = 0 at U10 = 0, gives NSPECH = NTH for even NTH); never above NSPEC.
  const int nspech_w = static_cast<int>(static_cast<unsigned>(nkh1) * static_cast<unsigned>(nth));
  const int nspech = nspech_w < nspec ? nspech_w : nspec;
  float dttot = t(sc, S_DTTOT);
  float dt = mn(p.dtmax, dtg - dttot);
  // DAM(IS) = FACP / (SIG(IK) * WN1(IK)**3) for the bin's frequency.
  const float wl = t(s.wn1, nk - 1);
  const float damlast = p.facp / (p.sig[nk - 1] * (wl * (wl * wl)));
  const float afilt = mx(damlast, p.xflt * t(sc, S_AMAX));
  if (f.inflags24) {
    for (int is = 0; is < nspech; ++is) {
      t(s.vsnl, is) = isnl * t(s.vsnl, is);
      t(s.vdnl, is) = isnl * t(s.vdnl, is);
      t(s.vsin, is) = isin * t(s.vsin, is);
      t(s.vdin, is) = isin * t(s.vdin, is);
      t(s.vsds, is) = isds * t(s.vsds, is);
      t(s.vdds, is) = isds * t(s.vdds, is);
    }
  }
  const int nscut = t(s.isc, I_NSCUT);
  const float xrel = p.xrel;
  // The time-step limit over the bins below NKH1 (VS, VD as W3SRCE sums them).
  for (int ik = 0; ik < nk && ik * nth < nspech; ++ik) {
    const float w = t(s.wn1, ik);
    const float dam = p.facp / (p.sig[ik] * (w * (w * w)));
    const float wnf = p.ln1 ? t(s.wnf, ik) : 0.f;
    for (int ith = 0; ith < nth && ik * nth + ith < nspech; ++ith) {
      const int is = ik * nth + ith;
      const float vsln = (p.ln1 ? wnf * t(s.dirf, ith) : 0.f) * (f.inflags24 ? isln : 1.f);
      const float a = t(s.spec, is);
      const float vdbt = is < nscut ? t(s.cbeta, p.mapwn[is] - 1) : 0.f;
      const float vsbt = nscut < 0 ? 0.f : vdbt * a;
      const float vs = (((vsln + t(s.vsin, is)) + t(s.vsnl, is)) + t(s.vsds, is)) + vsbt;
      const float vd = ((t(s.vdin, is) + t(s.vdnl, is)) + t(s.vdds, is)) + vdbt;
      const float damax = mn(mx(xrel * t(s.specinit, is), afilt), dam);
      const float afac = 1.f / mx(__builtin_fabsf(vs / damax), 1.E-10f);
      dt = mn(dt, afac / mx(mn(vd, 0.f) * afac + 1.f, 1.E-10f));
    }
  }
  dt = mx(dt, 0.5f);
  t(sc, S_DTDYN) = t(sc, S_DTDYN) + dt;
  const float rest = dtg - dttot;
  const int idt = 1 + fint((rest * 0.99f) / dt);
  dt = rest / static_cast<float>(idt);
  const bool shave = dt < p.dtmin && dt < rest;
  t(s.isc, I_SHAVE) = shave ? 1 : 0;
  dt = mx(mn(p.dtmin, rest), dt);
  const float hdt = dt;
  dttot = dttot + dt;
  t(sc, S_DTTOT) = dttot;

  // Update (bins below NKH1), W3SDB1 update, flux sums (all bins), in one pass.
  const float vddb = t(sc, S_DTDB);
  float phiaw = t(sc, S_PHIAW), phibbl = t(sc, S_PHIBBL), wc2 = 0.f;
  for (int ik = 0; ik < nk; ++ik) {
    const float w = t(s.wn1, ik);
    const float dam = p.facp / (p.sig[ik] * (w * (w * w)));
    const float wnf = p.ln1 ? t(s.wnf, ik) : 0.f;
    const float factor = p.dden[ik] / t(s.cg1, ik);
    for (int ith = 0; ith < nth; ++ith) {
      const int is = ik * nth + ith;
      const float a0 = t(s.spec, is);
      const float vdbt = is < nscut ? t(s.cbeta, p.mapwn[is] - 1) : 0.f;
      const float vsbt = nscut < 0 ? 0.f : vdbt * a0;
      const float vsin = t(s.vsin, is), vdin = t(s.vdin, is);
      float a = a0;
      if (is < nspech) {
        const float vsln = (p.ln1 ? wnf * t(s.dirf, ith) : 0.f) * (f.inflags24 ? isln : 1.f);
        const float vs = (((vsln + vsin) + t(s.vsnl, is)) + t(s.vsds, is)) + vsbt;
        const float vd = ((vdin + t(s.vdnl, is)) + t(s.vdds, is)) + vdbt;
        const float e1 = (dt * vs) / mx(1.f - hdt * vd, 1.f);
        if (shave) {
          const float e2 = __builtin_copysignf(mn(__builtin_fabsf(e1), dam), e1);
          a = mx(a + e2, 0.f);
        } else {
          a = mx(a + e1, 0.f);
        }
        const float e3 = (dt * t(s.vsdb, is)) / mx(1.f - hdt * vddb, 1.f);
        a = mx(a + e3, 0.f);
        t(s.spec, is) = a;
      }
      phiaw = ((vsin * dt) * factor) / mx(1.f - vdin * dt, 1.f) + phiaw;
      phibbl = phibbl - ((vsbt * dt) * factor) / mx(1.f - vdbt * dt, 1.f);
      if (vsin > 0.f) wc2 = wc2 + a * factor;
    }
  }
  t(sc, S_PHIAW) = phiaw;
  t(sc, S_PHIBBL) = phibbl;
  t(s.whitecap, 2) = __builtin_sqrtf(wc2) * 4.f;
  const float drat = t(sc, S_DRAT);
  t(sc, S_TAUWIX) = t(sc, S_TAUWIX) + (t(sc, S_TAUWX) * drat) * dt;
  t(sc, S_TAUWIY) = t(sc, S_TAUWIY) + (t(sc, S_TAUWY) * drat) * dt;
  t(sc, S_TAUWNX) = t(sc, S_TAUWNX) + (drat * t(sc, S_TAUWAX)) * dt;
  t(sc, S_TAUWNY) = t(sc, S_TAUWNY) + (drat * t(sc, S_TAUWAY)) * dt;

  spr4(s, q, t);
  const float u10abs = t(s.pin, P_U10ABS);
  const float fage = p.ffxfa * libm::tanhf((((u10abs * 0.3f) * t(sc, S_FMEANWS)) * kTpi) / kGrav);
  const float fh1 = (p.ffxfm + fage) * t(sc, S_FMEAN1);
  const float fh2 = p.ffxpm / t(sc, S_USTAR);
  fhigh = mn(mx(fh1, fh2), p.sig[nk - 1]);
  t(sc, S_FHIGH) = fhigh;
  nkh = mxi(mni(nkh1, fint(p.facti1 * libm::logf(mx(1.E-7f, fhigh)) + p.facti2)), 2);
  if (dttot >= dtg * 0.9999f) {
    const float wnmean = t(sc, S_WNMEAN);
    const float hm = (p.fhmax * libm::tanhf(mx(0.f, t(s.pin, P_D_INP)) * wnmean)) /
                     mx(wnmean, 9.99999974737875163555145263671875e-5f);
    const float em = (hm * hm) * 6.25e-2f;
    const float emean = t(sc, S_EMEAN);
    if (emean > em && emean > 1.E-30f) {
      for (int is = 0; is < nspec; ++is) t(s.spec, is) = em * (t(s.spec, is) / emean);
      t(sc, S_EMEAN) = em;
    }
  }
  for (int ik = nkh; ik < nk; ++ik)
    for (int ith = 0; ith < nth; ++ith)
      t(s.spec, ith + ik * nth) = t(s.spec, ith + (ik - 1) * nth) * p.fachfa + 0.f;
  sin4(s, q, t);
  if (q.sintailpar[3] < 0.5f) spr4(s, q, t);
  const int done = dttot >= 0.9999f * dtg ? 1 : 0;
  t(s.isc, I_DONE) = done;
  if (!done) atomicAdd(&s.active[iter], 1);
}

__global__ __launch_bounds__(kBlock) void finish_kernel(State s, St4Dev q, SrceDev p, StepFlags f) {
  Pt t;
  if (!lane_point(s, t)) return;
  const int nk = p.nk, nth = p.nth, nspec = p.nspec;
  float* sc = s.sc;
  const float dtg = f.dtg;
  t(sc, S_DTDYN) = t(sc, S_DTDYN) / static_cast<float>(mxi(t(s.isc, I_NSTEPS), 1));
  t(sc, S_FCUT) = t(sc, S_FHIGH) * kTpiInv;
  float efinish = 0.f, mwx = 0.f, mwy = 0.f;
  for (int ik = 0; ik < nk; ++ik) {
    float eband = 0.f, a1 = 0.f, b1 = 0.f;
    for (int ith = 0; ith < nth; ++ith) {
      const float diff = t(s.specinit, ith + ik * nth) - t(s.spec, ith + ik * nth);
      eband = diff + eband;
      a1 = p.ecos[ith] * diff + a1;
      b1 = p.esin[ith] * diff + b1;
    }
    const float cgk = t(s.cg1, ik), w = t(s.wn1, ik);
    efinish = (p.dden[ik] * eband) / cgk + efinish;
    mwx = (((p.dden[ik] * a1) / cgk) * w) / p.sig[ik] + mwx;
    mwy = (w * ((p.dden[ik] * b1) / cgk)) / p.sig[ik] + mwy;
  }
  const float tauwix = t(sc, S_TAUWIX), tauwiy = t(sc, S_TAUWIY);
  const float tauox = ((tauwix + mwx * kGrav) - t(s.taubbl, 0)) / dtg;
  const float tauoy = ((tauwiy + mwy * kGrav) - t(s.taubbl, 1)) / dtg;
  t(sc, S_TAUOX) = tauox;
  t(sc, S_TAUOY) = tauoy;
  const float twix = tauwix / dtg, twiy = tauwiy / dtg;
  t(sc, S_TAUWIX) = twix;
  t(sc, S_TAUWIY) = twiy;
  t(sc, S_TAUWNX) = t(sc, S_TAUWNX) / dtg;
  t(sc, S_TAUWNY) = t(sc, S_TAUWNY) / dtg;
  t(s.taubbl, 0) = t(s.taubbl, 0) / dtg;
  t(s.taubbl, 1) = t(s.taubbl, 1) / dtg;
  {
    float sn, cs;
    libm::sincosf(t(sc, S_USTDIR), &sn, &cs);
    const float ustar = t(sc, S_USTAR), coef = t(s.pin, P_COEF), dair = t(s.pin, P_DAIR);
    const float kk = ustar * (ustar * (coef * (dair * coef)));
    t(sc, S_TAUOCX) = kk * cs + (tauox - twix) * kDwat;
    t(sc, S_TAUOCY) = kk * sn + (tauoy - twiy) * kDwat;
  }
  constexpr float dg = kDwat * kGrav;
  const float phiaw = t(sc, S_PHIAW), phibbl = t(sc, S_PHIBBL);
  t(sc, S_PHIOC) = (((phiaw + efinish) - phibbl) * dg) / dtg;
  t(sc, S_PHIAW) = (phiaw * dg) / dtg;
  t(sc, S_PHIBBL) = (phibbl * dg) / dtg;
  const float ice = t(s.pin, P_ICE);
  if (f.inflags24 && ice > 0.f && !p.ic_numerics) {
    // Without ice dispersion ATT = 1: the spectrum is unchanged and the sums are of zeros.
    float phice = 0.f, ti0 = 0.f, ti1 = 0.f;
    for (int ik = 0; ik < nk; ++ik) {
      const float factor = p.dden[ik] / t(s.cg1, ik);
      const float factor2 = ((factor * kGrav) * t(s.wn1, ik)) / p.sig[ik];
      for (int ith = 0; ith < nth; ++ith) {
        const int is = ith + ik * nth;
        const float a2 = t(s.spec, is);
        const float a = static_cast<float>(1.0 * static_cast<double>(a2));
        t(s.spec, is) = a;
        phice = phice + (a - a2) * factor;
        ti0 = ti0 - ((a - a2) * factor2) * p.ecos[is];
        ti1 = ti1 - ((a - a2) * factor2) * p.esin[is];
      }
    }
    t(sc, S_PHICE) = (((-1.f * kDwat) * kGrav) * phice) / dtg;
    t(s.tauice, 0) = ti0 / dtg;
    t(s.tauice, 1) = ti1 / dtg;
  }
  if (f.it == 0)
    for (int is = 0; is < nspec; ++is) t(s.spec, is) = t(s.specinit, is);
  const int nlast = (nspec % 4 == 3) ? nspec - 1 : nspec;
  for (int is = 0; is < nlast; ++is) t(s.spec, is) = mx(t(s.spec, is), 0.f);
  if (nlast < nspec) t(s.spec, nlast) = mx(0.f, t(s.spec, nlast));
}

// Point-major C ABI arrays <-> bin-major state, one lane per (point, row).
__global__ void pack_kernel(State s, int nk, int nth, const float* spec, const float* wn1,
                            const float* cg1, const float* pin, const float* psc,
                            const float* whitecap, const float* taubbl, const float* tauice) {
  const int p = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (p >= s.npts) return;
  const size_t P = static_cast<size_t>(s.P);
  const int nspec = nk * nth;
  if (r < nk) {
    const float* src = spec + static_cast<size_t>(p) * nspec + static_cast<size_t>(r) * nth;
    for (int ith = 0; ith < nth; ++ith) s.spec[(static_cast<size_t>(r) * nth + ith) * P + p] = src[ith];
    s.wn1[r * P + p] = wn1[static_cast<size_t>(p) * nk + r];
    s.cg1[r * P + p] = cg1[static_cast<size_t>(p) * nk + r];
  } else if (r == nk) {
    for (int j = 0; j < N_PIN; ++j) s.pin[j * P + p] = pin[static_cast<size_t>(p) * N_PIN + j];
    for (int j = 0; j < N_PSC; ++j) s.sc[j * P + p] = psc[static_cast<size_t>(p) * N_PSC + j];
    for (int j = 0; j < 4; ++j) s.whitecap[j * P + p] = whitecap[static_cast<size_t>(p) * 4 + j];
    for (int j = 0; j < 2; ++j) s.taubbl[j * P + p] = taubbl[static_cast<size_t>(p) * 2 + j];
    for (int j = 0; j < 2; ++j) s.tauice[j * P + p] = tauice[static_cast<size_t>(p) * 2 + j];
  }
}

__global__ void unpack_kernel(State s, int nk, int nth, float* spec, float* psc, float* whitecap,
                              float* taubbl, float* tauice, int* shave) {
  const int p = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (p >= s.npts) return;
  const size_t P = static_cast<size_t>(s.P);
  const int nspec = nk * nth;
  if (r < nk) {
    float* dst = spec + static_cast<size_t>(p) * nspec + static_cast<size_t>(r) * nth;
    for (int ith = 0; ith < nth; ++ith) dst[ith] = s.spec[(static_cast<size_t>(r) * nth + ith) * P + p];
  } else if (r == nk) {
    for (int j = 0; j < N_PSC; ++j) psc[static_cast<size_t>(p) * N_PSC + j] = s.sc[j * P + p];
    for (int j = 0; j < 4; ++j) whitecap[static_cast<size_t>(p) * 4 + j] = s.whitecap[j * P + p];
    for (int j = 0; j < 2; ++j) taubbl[static_cast<size_t>(p) * 2 + j] = s.taubbl[j * P + p];
    for (int j = 0; j < 2; ++j) tauice[static_cast<size_t>(p) * 2 + j] = s.tauice[j * P + p];
    shave[p] = s.isc[I_SHAVE * P + p];
  }
}

__global__ void zero_ints(int* x, int n) {
  const int i = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0;
}

inline dim3 lanes(const State& s) { return dim3(static_cast<unsigned>((s.npts + kBlock - 1) / kBlock)); }

}  // namespace

size_t state_bytes(int nk, int nth, int npts) {
  const size_t P = static_cast<size_t>((npts + 63) / 64 * 64);
  const size_t nspec = static_cast<size_t>(nk) * nth;
  const auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  size_t b = 0;
  b += 10 * al(nspec * P * 4);                              // spec .. brlambda
  b += al(nspec * P * 4);                                   // llws
  b += 5 * al(static_cast<size_t>(nk) * P * 4);             // wn1 cg1 bth0 wnf cbeta
  b += al(static_cast<size_t>(nth) * P * 4);                // dirf
  b += al(static_cast<size_t>(N_SCAL) * P * 4);
  b += al(static_cast<size_t>(N_ISCAL) * P * 4);
  b += al(static_cast<size_t>(N_PIN) * P * 4);
  b += al(4 * P * 4) + 2 * al(2 * P * 4) + al(P * 4);      // whitecap taubbl tauice kdmean
  b += al(static_cast<size_t>(kMaxIter) * 4);               // active
  return b;
}

State carve_state(void* base, int nk, int nth, int npts) {
  State s{};
  s.npts = npts;
  s.P = (npts + 63) / 64 * 64;
  const size_t P = static_cast<size_t>(s.P);
  const size_t nspec = static_cast<size_t>(nk) * nth;
  const auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  char* c = static_cast<char*>(base);
  const auto take = [&](size_t bytes) {
    char* r = c;
    c += al(bytes);
    return r;
  };
  s.spec = reinterpret_cast<float*>(take(nspec * P * 4));
  s.specinit = reinterpret_cast<float*>(take(nspec * P * 4));
  s.vsin = reinterpret_cast<float*>(take(nspec * P * 4));
  s.vdin = reinterpret_cast<float*>(take(nspec * P * 4));
  s.vsnl = reinterpret_cast<float*>(take(nspec * P * 4));
  s.vdnl = reinterpret_cast<float*>(take(nspec * P * 4));
  s.vsds = reinterpret_cast<float*>(take(nspec * P * 4));
  s.vdds = reinterpret_cast<float*>(take(nspec * P * 4));
  s.vsdb = reinterpret_cast<float*>(take(nspec * P * 4));
  s.brlambda = reinterpret_cast<float*>(take(nspec * P * 4));
  s.llws = reinterpret_cast<int*>(take(nspec * P * 4));
  s.wn1 = reinterpret_cast<float*>(take(static_cast<size_t>(nk) * P * 4));
  s.cg1 = reinterpret_cast<float*>(take(static_cast<size_t>(nk) * P * 4));
  s.bth0 = reinterpret_cast<float*>(take(static_cast<size_t>(nk) * P * 4));
  s.wnf = reinterpret_cast<float*>(take(static_cast<size_t>(nk) * P * 4));
  s.cbeta = reinterpret_cast<float*>(take(static_cast<size_t>(nk) * P * 4));
  s.dirf = reinterpret_cast<float*>(take(static_cast<size_t>(nth) * P * 4));
  s.sc = reinterpret_cast<float*>(take(static_cast<size_t>(N_SCAL) * P * 4));
  s.isc = reinterpret_cast<int*>(take(static_cast<size_t>(N_ISCAL) * P * 4));
  s.pin = reinterpret_cast<float*>(take(static_cast<size_t>(N_PIN) * P * 4));
  s.whitecap = reinterpret_cast<float*>(take(4 * P * 4));
  s.taubbl = reinterpret_cast<float*>(take(2 * P * 4));
  s.tauice = reinterpret_cast<float*>(take(2 * P * 4));
  s.kdmean = reinterpret_cast<float*>(take(P * 4));
  s.active = reinterpret_cast<int*>(take(static_cast<size_t>(kMaxIter) * 4));
  return s;
}

hipError_t k_pack(const State& s, int nk, int nth, const float* spec, const float* wn1,
                  const float* cg1, const float* pin, const float* psc, const float* whitecap,
                  const float* taubbl, const float* tauice, hipStream_t st) {
  hipLaunchKernelGGL(pack_kernel, dim3(static_cast<unsigned>((s.npts + 255) / 256), nk + 1), dim3(256),
                     0, st, s, nk, nth, spec, wn1, cg1, pin, psc, whitecap, taubbl, tauice);
  hipLaunchKernelGGL(zero_ints, dim3((kMaxIter + 255) / 256), dim3(256), 0, st, s.active, kMaxIter);
  return hipGetLastError();
}

hipError_t k_unpack(const State& s, int nk, int nth, float* spec, float* psc, float* whitecap,
                    float* taubbl, float* tauice, int* shave, hipStream_t st) {
  hipLaunchKernelGGL(unpack_kernel, dim3(static_cast<unsigned>((s.npts + 255) / 256), nk + 1), dim3(256),
                     0, st, s, nk, nth, spec, psc, whitecap, taubbl, tauice, shave);
  return hipGetLastError();
}

hipError_t k_init(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f,
                  hipStream_t st) {
  hipLaunchKernelGGL(init_kernel, lanes(s), dim3(kBlock), 0, st, s, q, p, f);
  return hipGetLastError();
}

hipError_t k_step_a(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f,
                    int iter, hipStream_t st) {
  (void)iter;
  hipLaunchKernelGGL(step_a_kernel, lanes(s), dim3(kBlock), 0, st, s, q, p, f);
  return hipGetLastError();
}

hipError_t k_step_c(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f,
                    int iter, hipStream_t st) {
  hipLaunchKernelGGL(step_c_kernel, lanes(s), dim3(kBlock), 0, st, s, q, p, f, iter);
  return hipGetLastError();
}

hipError_t k_finish(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f,
                    hipStream_t st) {
  hipLaunchKernelGGL(finish_kernel, lanes(s), dim3(kBlock), 0, st, s, q, p, f);
  return hipGetLastError();
}

}  // namespace srce
}  // namespace ww3gpu

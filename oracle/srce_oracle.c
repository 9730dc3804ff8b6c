/*
 * oracle/srce_oracle.c -- CPU oracle for the ST4 source terms and the W3SRCE integration.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the GPU library; never linked into it).
 *
 * A scalar, point-at-a-time restatement of the reference's translation of WW3, in the
 * Fortran's loop order, float32 throughout (float64 in W3SDB1, as the Fortran), built with
 * -ffp-contract=off.  Each routine cites the reference function it restates:
 *   W3SPR4      kokkos/src/ww_kokkos/st4_spr4.hpp:40-162
 *   CALC_USTAR  kokkos/src/ww_kokkos/st4_calc_ustar.hpp:36-122
 *   W3SIN4      kokkos/src/ww_kokkos/st4_sin4.hpp:69-169 (prep), 172-277 (scan),
 *               472-627 (tail), 633-644
 *   W3SDS4      kokkos/src/ww_kokkos/st4_sds4.hpp:122-172, 175-209, 212-276, 279-331,
 *               574-638, 641-711, 719-746
 *   W3SLN1      kokkos/src/ww_kokkos/srce_ln1.hpp:41-99
 *   W3SBT1      kokkos/src/ww_kokkos/srce_bt1.hpp:26-58
 *   W3SDB1      kokkos/src/ww_kokkos/srce_db1.hpp:44-149
 *   W3SRCE      kokkos/src/ww_kokkos/srce_kernel.hpp:212-302 (init), 306-337 (step),
 *               349-488 (after W3SDS4), 491-583 (finish), 588-610 (driver)
 *   INSIN4      kokkos/src/ww_kokkos/st4_insin4.cpp:110-138, 239-262 (saturation and
 *               cumulative tables)
 *
 * Supported configuration (the product refuses the rest): ST4 with SDSBCHOICE = 1, the
 * analytic high-frequency tail and iterative u* (SINTAILPAR(1) < 0.5), FWTABLE friction
 * factor when SWELLF(2) /= 0, with or without W3_STAB3 and W3_IG1; W3SRCE with the lab
 * switch set (NL1 LN1/LN0 BT1 DB1 MLIM), SEED and FLD2 off, TAUWBUG on or IT = 0.
 *
 * Math: expf is glibc's; W3SDB1's exp and tanh are glibc's; logf, log10f, powf, atan2f,
 * tanhf and sincosf are the library's definitions restated in wwlibm.h.  At the five
 * loops gfortran -O3 vectorises with libmvec (reference st4_libmvec.hpp:1-16) the oracle
 * calls the scalar function, as the GPU library does.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "st4_params.h"
#include "wwlibm.h"

#ifdef _OPENMP
#include <omp.h>
#endif

size_t or_snl1_work_floats(int nk, int nth, float xfr, float lam);
void or_w3snl1_one(int nk, int nth, float xfr, float lam, float snlc1, float kdcon, float kdmn,
                   float snls1, float snls2, float snls3, float fachfe, const float* scal,
                   const int* ipim, const int* ic, const float* af11, const float* sig,
                   const float* a, const float* cg, float kdmean, float* s, float* d, float* work);

/* ---------------- constants (CONSTANTS, each operation rounded to float) ---------------- */
#define GRAV 9.806f
#define DWAT 1000.f
#define NU_AIR 1.4E-5f
#define KAPPA 0.40f
static const float PI_F = (float)3.141592653589793;
#define TPI_F (2.0f * PI_F)
#define TPIINV_F (1.0f / TPI_F)

static inline float mx(float a, float b) { return a > b ? a : b; }
static inline float mn(float a, float b) { return a < b ? a : b; }
static inline int mxi(int a, int b) { return a > b ? a : b; }
static inline int mni(int a, int b) { return a < b ? a : b; }
/* the b-first forms of CALC_USTAR and W3SLN1 */
static inline float fmax2(float a, float b) { return (b > a) ? b : a; }
static inline float fmin2(float a, float b) { return (b < a) ? b : a; }

/* INT() as cvttss2si: truncation, INT_MIN for NaN or out of range. */
static inline int fint(float x) {
  return (x > -2147483904.f && x < 2147483648.f) ? (int)x : (int)0x80000000u;
}
/* NINT(): round half away from zero, 0 for NaN or out of range. */
static inline int nint(float x) {
  float r = roundf(x);
  return (r > -9.2233715e18f && r < 9.2233715e18f) ? (int)(long long)r : 0;
}
/* FLOOR() as -O3 computes it: i = INT(x); (float)i <= x ? i : i-1 (wrapping). */
static inline int ffloor(float x) {
  int i = fint(x);
  return (float)i <= x ? i : (int)((unsigned)i - 1u);
}

/* ---------------- configuration ---------------- */
typedef struct {
  int nk, nth, nspec, stab3, ig1;
  float dth, xfr, fte, ftf, aaircmin, aairgb, igpars5, wwnmeanptail, sstxftftail;
  float aalpha, bbeta, zz0max, zz0rat, zzalp, ssinthp, ttauwshelter;
  float sswellf[7], ssdsc[21], ssdsbr, sintailpar[5], ssdsp, wwnmeanp, sstxftwn;
  float ssdsbrf1, ssdsbint, ssdsbck, ssdshck, ssdsabk, ssdspbk, capchnk[10];
  float zzwnd, ssdscos, ssdsdth, ssdsbt, ssdsbm[5], delab;
  /* grid arrays, 0-based pointers at the Fortran lower bound */
  const float* sig;   /* SIG(0:NK+1) */
  const float* sig2;  /* SIG2(1:NSPEC) */
  const float* dsip;  /* DSIP(0:NK+1) */
  const float* dden;  /* DDEN(1:NK) */
  const float* dden2; /* DDEN2(1:NSPEC) */
  const float* esin;  /* ESIN(1:NSPEC+NTH) */
  const float* ecos;  /* ECOS(1:NSPEC+NTH) */
  const float* th;    /* TH(1:NTH) */
  const float* satw;  /* SATWEIGHTS(1:NSAT,1:NTH) */
  const int* sati;    /* SATINDICES(1:NSAT,1:NTH) */
  int nsat;
  const float* cumulw; /* CUMULW(1:NSPEC,1:NSPEC) */
  const float* fwtable; /* FWTABLE(0:300) or NULL */
} St4;

typedef struct {
  int nk, nth, nspec, ic_numerics, fdonly, ln1;
  float dmin, dtmax, dtmin, facti1, facti2, fachfa, facp, xflt, xrel, fhmax, ffxfm, ffxpm, ffxfa;
  float icescales[4], slnc1, fspm, fshf, sbtc1, sdbc1, sdbc2;
  const float *sig, *dden, *ecos, *esin; /* SIG(1:NK), DDEN(1:NK), ECOS/ESIN(1:NSPEC) */
  const int* mapwn;
} Srce;

typedef struct {
  float xfr, lam, snlc1, kdcon, kdmn, snls1, snls2, snls3, fachfe;
  const float *scal, *af11, *sig;
  const int *ipim, *ic;
} Dia;

static void st4_unpack(St4* q, const float* r, const int* i) {
  memset(q, 0, sizeof *q);
  q->nk = i[ORSI_NK]; q->nth = i[ORSI_NTH]; q->nspec = i[ORSI_NSPEC];
  q->stab3 = i[ORSI_STAB3]; q->ig1 = i[ORSI_IG1];
  q->dth = r[ORS_DTH]; q->xfr = r[ORS_XFR]; q->fte = r[ORS_FTE]; q->ftf = r[ORS_FTF];
  q->aaircmin = r[ORS_AAIRCMIN]; q->aairgb = r[ORS_AAIRGB]; q->igpars5 = r[ORS_IGPARS5];
  q->wwnmeanptail = r[ORS_WWNMEANPTAIL]; q->sstxftftail = r[ORS_SSTXFTFTAIL];
  q->aalpha = r[ORS_AALPHA]; q->bbeta = r[ORS_BBETA]; q->zz0max = r[ORS_ZZ0MAX];
  q->zz0rat = r[ORS_ZZ0RAT]; q->zzalp = r[ORS_ZZALP]; q->ssinthp = r[ORS_SSINTHP];
  q->ttauwshelter = r[ORS_TTAUWSHELTER];
  for (int j = 0; j < 7; ++j) q->sswellf[j] = r[ORS_SSWELLF + j];
  for (int j = 0; j < 21; ++j) q->ssdsc[j] = r[ORS_SSDSC + j];
  q->ssdsbr = r[ORS_SSDSBR];
  for (int j = 0; j < 5; ++j) q->sintailpar[j] = r[ORS_SINTAILPAR + j];
  q->ssdsp = r[ORS_SSDSP]; q->wwnmeanp = r[ORS_WWNMEANP]; q->sstxftwn = r[ORS_SSTXFTWN];
  q->ssdsbrf1 = r[ORS_SSDSBRF1]; q->ssdsbint = r[ORS_SSDSBINT]; q->ssdsbck = r[ORS_SSDSBCK];
  q->ssdshck = r[ORS_SSDSHCK]; q->ssdsabk = r[ORS_SSDSABK]; q->ssdspbk = r[ORS_SSDSPBK];
  for (int j = 0; j < 10; ++j) q->capchnk[j] = r[ORS_CAPCHNK + j];
  q->zzwnd = r[ORS_ZZWND]; q->ssdscos = r[ORS_SSDSCOS]; q->ssdsdth = r[ORS_SSDSDTH];
  q->ssdsbt = r[ORS_SSDSBT];
  for (int j = 0; j < 5; ++j) q->ssdsbm[j] = r[ORS_SSDSBM + j];
  q->delab = r[ORS_DELAB];
}

static void srce_unpack(Srce* p, const float* r, const int* i) {
  memset(p, 0, sizeof *p);
  p->nk = i[ORRI_NK]; p->nth = i[ORRI_NTH]; p->nspec = i[ORRI_NSPEC];
  p->ic_numerics = i[ORRI_IC_NUMERICS]; p->fdonly = i[ORRI_FDONLY]; p->ln1 = i[ORRI_LN1];
  p->dmin = r[ORR_DMIN]; p->dtmax = r[ORR_DTMAX]; p->dtmin = r[ORR_DTMIN];
  p->facti1 = r[ORR_FACTI1]; p->facti2 = r[ORR_FACTI2]; p->fachfa = r[ORR_FACHFA];
  p->facp = r[ORR_FACP]; p->xflt = r[ORR_XFLT]; p->xrel = r[ORR_XREL]; p->fhmax = r[ORR_FHMAX];
  p->ffxfm = r[ORR_FFXFM]; p->ffxpm = r[ORR_FFXPM]; p->ffxfa = r[ORR_FFXFA];
  for (int j = 0; j < 4; ++j) p->icescales[j] = r[ORR_ICESCALES + j];
  p->slnc1 = r[ORR_SLNC1]; p->fspm = r[ORR_FSPM]; p->fshf = r[ORR_FSHF]; p->sbtc1 = r[ORR_SBTC1];
  p->sdbc1 = r[ORR_SDBC1]; p->sdbc2 = r[ORR_SDBC2];
}

/* ================================ CALC_USTAR ================================ */
static void calc_ustar(const St4* q, float windspeed, float tauw, float* ustar, float* z0,
                       float* charn) {
  const float eps1 = 0.00001f;
  float chath = q->aalpha, ust = 0.f;
  float xmin;
  if (q->capchnk[0] == 1.f) {
    chath = q->capchnk[1] + 0.5f * (q->capchnk[2] - q->capchnk[1]) *
                                (1.f - wl_tanhf((windspeed - q->capchnk[3]) / q->capchnk[4]));
    xmin = 0.15f * (q->capchnk[2] - chath);
  } else {
    xmin = 0.f;
  }
  const float tl = tauw;
  const float wcd = sqrtf(0.0012875f);
  const float ustold = windspeed * wcd;
  float tauold = fmax2(ustold * ustold, tl + eps1);
  const float znunum = 0.11f * NU_AIR;
  const float xm1 = 0.50f + 1.f;
  for (int it = 1; it <= 10; ++it) {
    const float x = fmax2(tl / tauold, xmin);
    ust = sqrtf(tauold);
    float zz00 = chath * tauold / GRAV;
    if (q->zz0max != 0.f) zz00 = fmin2(zz00, q->zz0max);
    const float omx = 1.f - x;
    float zz0 = zz00 / wl_powf(omx, 0.50f);
    const float znu = znunum / fmax2(ust, 1E-6f);
    zz0 = q->sintailpar[4] * znu + zz0;
    const float lg = wl_logf(q->zzwnd / zz0);
    const float fval = ust - KAPPA * windspeed / lg;
    const float delf = 1.f - KAPPA * windspeed / (lg * lg) * 2.f / ust * (1.f - xm1 * x) / omx;
    ust = ust - fval / delf;
    tauold = fmax2(ust * ust, tl + eps1);
  }
  *ustar = ust;
  if (*ustar > 0.001f) {
    const float s = fmin2(windspeed / *ustar, 100.0f);
    *z0 = q->zzwnd * expf(-KAPPA * s);
    *charn = GRAV * *z0 / (*ustar * *ustar);
  } else {
    if (*ustar > 0.f) {
      const float s = fmin2(windspeed / *ustar, 100.0f);
      *z0 = q->zzwnd * expf(-KAPPA * s);
    } else {
      *z0 = chath * 0.001f * 0.001f / GRAV;
    }
    *charn = chath;
  }
  if (q->capchnk[0] == 1.f) {
    *charn = fmin2(0.09f, *charn);
    if (*charn < chath) *charn = chath;
  }
}

/* ================================ W3SPR4 ================================ */
typedef struct {
  float emean, fmean, fmean1, amax, cd, z0, fmeanws, dlwmean;
} Spr4Out;

static void w3spr4(const St4* q, const float* a, const float* cg, const float* wn, float u,
                   float udir, float tauwx, float tauwy, const int* llws, float* ustar,
                   float* usdir, float* wnmean, float* charn, Spr4Out* o, float* w /*4*NK*/) {
  const int nk = q->nk, nth = q->nth;
  const float eps = 1.E-7f;
  float* eb = w;
  float* eb2 = w + nk;
  float* pws = w + 2 * nk;
  float* pwk = w + 3 * nk;
  *ustar = mx(0.0001f, *ustar);
  float emean = 0.f, emeanws = 0.f, fmeanws = 0.f, fmean = 0.f, fmean1 = 0.f, wnm = 0.f;
  float amax = 0.f, elcs = 0.f, elsn = 0.f;
  for (int ik = 1; ik <= nk; ++ik) {
    float e = 0.f, e2 = 0.f;
    const float sigfac = wl_powf(q->sig[ik], q->ssdsc[11]) * q->dden[ik - 1] / cg[ik - 1];
    for (int ith = 1; ith <= nth; ++ith) {
      const int is = ith + (ik - 1) * nth;
      const float aa = a[is - 1];
      e = e + aa;
      elcs = elcs + aa * q->ecos[is - 1] * sigfac;
      elsn = elsn + aa * q->esin[is - 1] * sigfac;
      if (llws[is - 1] != 0) e2 = e2 + aa;
      amax = mx(amax, aa);
    }
    eb[ik - 1] = e;
    eb2[ik - 1] = e2;
  }
  const float dlw = wl_atan2f(elsn, elcs);
  const float tw = 2.f * q->wwnmeanptail;
  for (int ik = 1; ik <= nk; ++ik) {
    pws[ik - 1] = wl_powf(q->sig[ik], tw);
    pwk[ik - 1] = wl_powf(wn[ik - 1], q->wwnmeanp);
  }
  for (int ik = 1; ik <= nk; ++ik) {
    const float e = eb[ik - 1] * q->dden[ik - 1] / cg[ik - 1];
    const float e2 = eb2[ik - 1] * q->dden[ik - 1] / cg[ik - 1];
    eb[ik - 1] = e;
    eb2[ik - 1] = e2;
    emean = emean + e;
    fmean = fmean + e / q->sig[ik];
    fmean1 = fmean1 + e * pws[ik - 1];
    wnm = wnm + e * pwk[ik - 1];
    emeanws = emeanws + e2;
    fmeanws = fmeanws + e2 * pws[ik - 1];
  }
  float eband = eb[nk - 1] / q->dden[nk - 1];
  emean = emean + eband * q->fte;
  fmean = fmean + eband * q->ftf;
  fmean1 = fmean1 + eband * q->sstxftftail;
  wnm = wnm + eband * q->sstxftwn;
  eband = eb2[nk - 1] / q->dden[nk - 1];
  emeanws = emeanws + eband * q->fte;
  fmeanws = fmeanws + eband * q->sstxftftail;
  fmean = TPIINV_F * emean / mx(eps, fmean);
  if (fmean1 < eps)
    fmean1 = TPIINV_F * q->sig[nk];
  else
    fmean1 = TPIINV_F * wl_powf(mx(eps, fmean1) / mx(eps, emean), 1.f / (2.f * q->wwnmeanptail));
  wnm = wl_powf(mx(eps, wnm) / mx(eps, emean), 1.f / q->wwnmeanp);
  if (fmeanws < eps || emeanws < eps)
    fmeanws = TPIINV_F * q->sig[nk];
  else
    fmeanws = TPIINV_F * wl_powf(mx(eps, fmeanws) / mx(eps, emeanws), 1.f / (2.f * q->wwnmeanptail));
  const float tauw = sqrtf(tauwx * tauwx + tauwy * tauwy);
  float z0 = 0.f;
  calc_ustar(q, u, tauw, ustar, &z0, charn);
  const float unz = mx(0.01f, u);
  const float r = *ustar / unz;
  o->emean = emean;
  o->fmean = fmean;
  o->fmean1 = fmean1;
  *wnmean = wnm;
  o->amax = amax;
  o->cd = r * r;
  o->z0 = z0;
  o->fmeanws = fmeanws;
  o->dlwmean = dlw;
  *usdir = udir;
}

/* ================================ W3SIN4 ================================ */
/* work: DSTAB(3,NSPEC) */
static void w3sin4(const St4* q, const float* a, const float* cg, const float* k, float u,
                   float ustar, float drat, float as, float usdir, float z0, float* tauwx,
                   float* tauwy, float* tauwnx, float* tauwny, float* s, float* d, int* llws,
                   float* dstab) {
  const int nk = q->nk, nth = q->nth, nspec = q->nspec;
  const float kappa2 = KAPPA * KAPPA;
  const float z0viscnum = 0.1f * NU_AIR;
  const float twonu = 2.f * NU_AIR;
  const float grav2tpi = (GRAV * GRAV) * TPI_F;
  const float logtpi005 = 7.1361942291259765625f; /* LOG(TPI/0.005), folded */
  const float eps1 = 0.00001f, eps2 = 0.000001f, eps2sq = eps2 * eps2;
#define DSTAB(i, is) dstab[((i) - 1) * nspec + ((is) - 1)]
  for (int i = 0; i < 3 * nspec; ++i) dstab[i] = 0.f;
  float ss[3][2] = {{0, 0}, {0, 0}, {0, 0}}, sn[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  const float const1 = q->bbeta / kappa2;
  float const0 = const1 * drat;
  const float z0visc = z0viscnum / mx(ustar, 0.0001f);
  const float z0noz = mx(z0visc, q->zz0rat * z0);
  float uorb = 0.f, aorb = 0.f;
  for (int ik = 1; ik <= nk; ++ik) {
    float eb = 0.f;
    for (int ith = 1; ith <= nth; ++ith) eb = eb + a[ith + (ik - 1) * nth - 1];
    const float sk = q->sig[ik];
    uorb = uorb + eb * (sk * sk) * q->dden[ik - 1] / cg[ik - 1];
    aorb = aorb + eb * q->dden[ik - 1] / cg[ik - 1];
  }
  uorb = 2.f * sqrtf(uorb);
  const float aorb1 = 2.f * wl_powf(aorb, 1.f - 0.5f * q->sswellf[5]);
  const float re = 4.f * uorb * aorb1 / NU_AIR;
  float pturb, pvisc;
  if (q->sswellf[3] > 0.f) {
    if (q->sswellf[6] > 0.f) {
      const float smooth = 0.5f * wl_tanhf((re - q->sswellf[3]) / q->sswellf[6]);
      pturb = 0.5f + smooth;
      pvisc = 0.5f - smooth;
    } else if (re <= q->sswellf[3]) {
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
  if (q->sswellf[1] == 0.f) {
    fw = fabsf(q->sswellf[2]);
    fu = 0.f;
    fud = 0.f;
  } else {
    fu = fabsf(q->sswellf[2]);
    fud = q->sswellf[1];
    aorb = 2.f * sqrtf(aorb);
    const float xi = (wl_log10f(mx(aorb / z0noz, 3.f)) - (-1.f)) / q->delab;
    const int ind = mni(300 - 1, fint(xi));
    const float deli1 = mn(1.f, xi - (float)ind);
    const float deli2 = 1.f - deli1;
    fw = q->fwtable[ind] * deli2 + q->fwtable[ind + 1] * deli1;
  }
  float ustarsigma = 0.f;
  if (q->stab3) {
    const float usigma = isnan(as) ? 0.f : mx(0.f, -0.025f * as);
    ustarsigma = (1.0f + u / (10.f + u)) * usigma;
  }

  /* section 2: the stress scan */
  float ust = ustar, taux = 0.f, tauy = 0.f, cosu = 0.f, sinu = 0.f, ustp = 0.f, usdirp = 0.f;
  const int i0 = q->stab3 ? 1 : 3, i1 = q->stab3 ? 2 : 3;
  for (int istab = i0; istab <= i1; ++istab) {
    if (istab == 1) ust = ustar * (1.f - ustarsigma);
    if (istab == 2) ust = ustar * (1.f + ustarsigma);
    float cu, su;
    wl_sincosf(usdir, &su, &cu);
    taux = (ust * ust) * cu;
    tauy = (ust * ust) * su;
    ss[istab - 1][0] = ss[istab - 1][1] = 0.f;
    sn[istab - 1][0] = sn[istab - 1][1] = 0.f;
    for (int ik = 1; ik <= nk; ++ik) {
      const float taupx = taux - fabsf(q->ttauwshelter) * ss[istab - 1][0];
      const float taupy = tauy - fabsf(q->ttauwshelter) * ss[istab - 1][1];
      ustp = mn(wl_powf((taupx * taupx + taupy * taupy), 0.25f), mx(ust, 0.3f));
      usdirp = wl_atan2f(taupy, taupx);
      wl_sincosf(usdirp, &sinu, &cosu);
      const int is1 = 1 + (ik - 1) * nth;
      const float kis = k[is1 - 1];
      const float s2 = q->sig2[is1 - 1];
      const float cm = kis / s2;
      const float ucn = ustp * cm + q->zzalp;
      const float const2 = q->dden2[is1 - 1] / cg[ik - 1] * GRAV / (q->sig[ik] / kis * drat);
      const float cnst = s2 * const0;
      const float zcn = wl_logf(kis * z0);
      const float swv = -(q->sswellf[4] * drat * 2.f * kis * sqrtf(twonu * s2));
      const float swt = -(drat * q->sswellf[0] * 16.f * (s2 * s2) / GRAV);
      for (int ith = 1; ith <= nth; ++ith) {
        const int is = ith + (ik - 1) * nth;
        const float ec = q->ecos[is - 1], es = q->esin[is - 1];
        const float coswind = (ec * cosu + es * sinu);
        if (coswind > 0.01f) {
          const float x = coswind * ucn;
          const float zarg = KAPPA / x;
          const float zlog = zcn + zarg;
          if (zlog < 0.f) {
            const float zl2 = zlog * zlog;
            DSTAB(istab, is) = cnst * expf(zlog) * (zl2 * zl2) * ucn * ucn * wl_powf(coswind, q->ssinthp);
            llws[is - 1] = 1;
          } else {
            DSTAB(istab, is) = 0.f;
            llws[is - 1] = 0;
          }
          if (28.f * cm * ustar * coswind >= 1.f) llws[is - 1] = 1;
        } else {
          DSTAB(istab, is) = 0.f;
          llws[is - 1] = 0;
        }
        if ((q->sswellf[0] != 0.f && DSTAB(istab, is) < 1E-7f * s2) || q->sswellf[2] > 0.f) {
          const float dvisc = swv;
          const float dturb = swt * (fw * uorb + (fu + fud * coswind) * ustp);
          DSTAB(istab, is) = DSTAB(istab, is) + pturb * dturb + pvisc * dvisc;
        }
        const float temp2 = const2 * DSTAB(istab, is) * a[is - 1];
        if (DSTAB(istab, is) < 0.f) {
          sn[istab - 1][0] = sn[istab - 1][0] + temp2 * ec;
          sn[istab - 1][1] = sn[istab - 1][1] + temp2 * es;
        } else {
          ss[istab - 1][0] = ss[istab - 1][0] + temp2 * ec;
          ss[istab - 1][1] = ss[istab - 1][1] + temp2 * es;
        }
      }
    }
  }

  /* the rest: D, S, tail stress, stress limit */
  float xstress, ystress;
  if (!q->stab3) {
    for (int is = 1; is <= nspec; ++is) d[is - 1] = DSTAB(3, is);
    xstress = ss[2][0];
    ystress = ss[2][1];
    *tauwnx = sn[2][0];
    *tauwny = sn[2][1];
  } else {
    for (int is = 1; is <= nspec; ++is) d[is - 1] = 0.5f * (DSTAB(1, is) + DSTAB(2, is));
    xstress = 0.5f * (ss[0][0] + ss[1][0]);
    ystress = 0.5f * (ss[0][1] + ss[1][1]);
    *tauwnx = 0.5f * (sn[0][0] + sn[1][0]);
    *tauwny = 0.5f * (sn[0][1] + sn[1][1]);
  }
  for (int is = 1; is <= nspec; ++is) s[is - 1] = d[is - 1] * a[is - 1];
  {
    const float taupx = taux - fabsf(q->ttauwshelter) * xstress;
    const float taupy = tauy - fabsf(q->ttauwshelter) * ystress;
    ustp = wl_powf((taupx * taupx + taupy * taupy), 0.25f);
    usdirp = wl_atan2f(taupy, taupx);
  }
  ust = ustp;
  {
    const float sg = q->sig[nk], s2 = sg * sg, s5 = s2 * (s2 * sg);
    const0 = q->dth * s5 / grav2tpi * TPI_F * sg / cg[nk - 1];
  }
  float temp = 0.f;
  for (int ith = 1; ith <= nth; ++ith) {
    const int is = ith + (nk - 1) * nth;
    const float cw = (q->ecos[is - 1] * cosu + q->esin[is - 1] * sinu);
    const float cp = mx(cw, 0.f);
    temp = temp + a[is - 1] * (cp * (cp * cp));
  }
  const float levtail0 = const0 * temp;
  float tauhf;
  {
    float w[50];
    for (int j = 2; j <= 49; ++j) w[j - 1] = 1.f;
    w[0] = 0.5f;
    w[49] = 0.5f;
    float ustr = ust;
    const float zz0 = z0, sg = q->sig[nk];
    const float zinf = wl_logf(sg * sg / GRAV);
    const float dely = (logtpi005 - zinf) / 50.f;
    float taur = ust * ust, tau1 = 0.f;
    for (int j = 1; j <= 50; ++j) {
      const float y = expf(zinf + (float)(j - 1) * dely);
      const float ykm = (y / 363.f);
      const float ykm2 = ykm * ykm;
      const float tensk = 1.f + ykm2;
      const float cm = sqrtf(GRAV * tensk / y);
      const float cgtail = 0.5f * (3.f * ykm2 + 1.f) * sqrtf(GRAV / (y * tensk));
      const float zx = ustr / cm + q->zzalp;
      const float zarg = mn(KAPPA / zx, 20.f);
      const float zmu = mn(GRAV * zz0 / (cm * cm) * expf(zarg), 1.f);
      const float zlog = mn(wl_logf(zmu), 0.f);
      const float zl2 = zlog * zlog;
      const float zbeta = const1 * zmu * (zl2 * zl2);
      float levtail;
      if (q->sintailpar[2] > 0.f) {
        float alpham;
        if (ustr < cm)
          alpham = mx(0.f, 0.01f * (1.f + wl_logf(ustr / cm)));
        else
          alpham = 0.01f * (1.f + 3.f * wl_logf(ustr / cm));
        const float ykm1 = (ykm - 1.f);
        const float fm = expf(-0.25f * (ykm1 * ykm1));
        const float alphat = alpham * (0.2325f / cm) * fm;
        const float th = wl_tanhf((y - 20.f) / 5.f);
        levtail = levtail0 * 0.5f * (1.f - th) + q->sintailpar[2] * 0.5f * (1.f + th) * alphat;
      } else {
        levtail = levtail0;
      }
      tau1 = tau1 + w[j - 1] * zbeta * (ustr * ustr) * levtail * dely * cgtail / cm;
      taur = taur - w[j - 1] * q->sintailpar[1] * (ustr * ustr) * zbeta * levtail * dely * cgtail / cm;
      ustr = sqrtf(mx(taur, 0.f));
    }
    tauhf = tau1;
  }
  float cp, sp;
  wl_sincosf(usdirp, &sp, &cp);
  *tauwx = xstress + tauhf * cp;
  *tauwy = ystress + tauhf * sp;
  const float tauw = sqrtf(*tauwx * *tauwx + *tauwy * *tauwy);
  const float u2b = mx(ustar, eps2);
  const float u2 = u2b * u2b;
  const float tauwb = mn(tauw, mx(u2 - eps1, eps2sq));
  if (tauwb < tauw) {
    *tauwx = *tauwx * tauwb / tauw;
    *tauwy = *tauwy * tauwb / tauw;
  }
#undef DSTAB
}

/* ================================ W3SDS4 ================================ */
/* work: 18*NK + 3*NSPEC + NTH floats (COEF4 COEF5 ... MSSSUM(NK,5) BTH PB PB2 ...) */
static void w3sds4(const St4* q, const float* a, const float* k, const float* cg, float ustar,
                   float usdir, float depth, float dair, int flogrd57, int flogrd58, float* srhs,
                   float* ddiag, float* brlambda, float* whitecap, float* w) {
  const int nk = q->nk, nth = q->nth, nspec = q->nspec;
  const float* c = q->ssdsc; /* SSDSC(i) = c[i-1] */
  float* coef4 = w;
  float* coef5 = coef4 + nk;
  float* brm12 = coef5 + nk;
  float* bth0 = brm12 + nk;
  float* mss = bth0 + nk; /* MSSSUM(IK,J) = mss[(J-1)*nk + IK-1] */
  float* bth = mss + 5 * nk;
  float* pb = bth + nspec;
  for (int i = 0; i < nk; ++i) { coef4[i] = 0.f; coef5[i] = 0.f; brm12[i] = 0.f; bth0[i] = 0.f; }
  for (int i = 0; i < nspec; ++i) { bth[i] = 0.f; ddiag[i] = 0.f; srhs[i] = 0.f; pb[i] = 0.f; }
  for (int i = 0; i < 5 * nk; ++i) mss[i] = 0.f;
  const float facturb = (((c[4] * (ustar * ustar)) / GRAV) * dair) / DWAT;
  const int dikcumul = nint(q->ssdsbrf1 / (q->xfr - 1.f));
  const int ik1 = q->ig1 ? nint(q->igpars5) + 1 : 1;

  /* 1.b mean-square slope (only its own MSSSUM is written; nothing later reads it here) */
  if (c[7] > 0.f || c[10] > 0.f || c[17] > 0.f) {
#define MSS(ik, j) mss[((j) - 1) * nk + ((ik) - 1)]
    for (int ik = 1; ik <= nk; ++ik) {
      float mssp = 0.f, mc = 0.f, ms = 0.f;
      const float kp = wl_powf(k[ik - 1], c[19]);
      for (int ith = 1; ith <= nth; ++ith) {
        const int is = ith + (ik - 1) * nth;
        const float ml = ((kp * a[is - 1]) * q->dden[ik - 1]) / cg[ik - 1];
        mc = mc + ml * q->ecos[ith - 1];
        ms = ms + ml * q->esin[ith - 1];
        mssp = mssp + ml;
      }
      for (int i = ik; i <= nk; ++i) {
        MSS(i, 1) = MSS(i, 1) + mssp;
        MSS(i, 3) = MSS(i, 3) + mc;
        MSS(i, 4) = MSS(i, 4) + ms;
      }
      float md = wl_atan2f(MSS(ik, 4), MSS(ik, 3));
      if (md < 0.f) md = md + PI_F;
      MSS(ik, 2) = md;
    }
#undef MSS
  }

  /* case 1: saturation, spontaneous breaking, PB, BRLAMBDA, SRHS */
  for (int ik = ik1; ik <= nk; ++ik) {
    const float kk = k[ik - 1];
    const float facsat = (q->sig[ik] * ((kk * kk) * kk)) * q->dth;
    const int is0 = (ik - 1) * nth;
    bth[is0] = 0.f;
    float asum = 0.f;
    for (int ith = 0; ith < nth; ++ith) asum = asum + a[is0 + ith];
    bth0[ik - 1] = asum * facsat;
    if (q->ssdsdth >= 180.f) {
      for (int i = is0; i < is0 + nth; ++i) bth[i] = bth0[ik - 1];
    } else {
      for (int ith = 0; ith < nth; ++ith) {
        float acc = 0.f;
        for (int j = 0; j < q->nsat; ++j)
          acc = acc + q->satw[j + q->nsat * ith] * a[is0 + q->sati[j + q->nsat * ith] - 1];
        bth[is0 + ith] = acc * facsat;
      }
      /* MAXVAL: the first element >= -Inf, then strictly greater ones */
      float lim = -INFINITY;
      int i = is0, found = 0;
      for (; i < is0 + nth; ++i)
        if (bth[i] >= lim) { lim = bth[i]; ++i; found = 1; break; }
      if (!found) lim = NAN;
      for (; i < is0 + nth; ++i)
        if (bth[i] > lim) lim = bth[i];
      bth0[ik - 1] = lim;
    }
  }
  const float epsr = sqrtf(q->ssdsbr);
  for (int ik = ik1; ik <= nk; ++ik) {
    float miche;
    if (q->ssdsbm[0] == 1.f) {
      miche = 1.f;
    } else {
      const float x = wl_tanhf(mn(k[ik - 1] * depth, 10.f));
      const float t = x * (q->ssdsbm[1] + x * (q->ssdsbm[2] + x * (q->ssdsbm[3] + x * q->ssdsbm[4])));
      miche = t * t;
    }
    const float coef1 = q->ssdsbr * miche;
    const float m = mx(bth0[ik - 1] / coef1 - 1.f, 0.f);
    const float sdiagiso = ((c[1] * q->sig[ik]) * c[5]) * (m * m);
    const float coef2 = ((c[1] * q->sig[ik]) * (1.f - c[5])) / (coef1 * coef1);
    const int is0 = (ik - 1) * nth;
    for (int i = 0; i < nth; ++i)
      ddiag[is0 + i] = sdiagiso + coef2 * wl_powf(mx(bth[is0 + i] - coef1, 0.f), q->ssdsp);
  }
  for (int i = 0; i < nspec; ++i) {
    const float m = mx(sqrtf(bth[i]) - epsr, 0.f);
    pb[i] = m * m;
  }
  for (int i = 0; i < nspec; ++i) pb[i] = pb[i] * 28.16f;
  const float twopisq = 2.f * (PI_F * PI_F);
  for (int i = 0; i < nspec; ++i) brlambda[i] = pb[i] / twopisq;
  for (int i = 0; i < nspec; ++i) srhs[i] = ddiag[i] * a[i];

  /* wave-turbulence interaction and cumulative effect */
  if (c[2] != 0.f || c[4] != 0.f || c[20] != 0.f) {
    float cu, su;
    for (int ik = ik1; ik <= nk; ++ik) {
      float renew = 0.f;
      const float facturb2 = -(((q->sig[ik] * 2.f) * k[ik - 1]) * facturb);
      const float dvisc = -(((c[20] * 4.f) * k[ik - 1]) * k[ik - 1]);
      const float cc = q->sig[ik] / k[ik - 1];
      if (c[2] > 0.f && ik > dikcumul) {
        for (int ik2 = ik1; ik2 <= ik - dikcumul; ++ik2) {
          const float c2 = q->sig[ik2] / k[ik2 - 1];
          const int is2 = (ik2 - 1) * nth;
          const float cumulwiso = ((fabsf(c2 - cc) * q->dsip[ik2]) / (c2 * 0.5f)) * q->dth;
          float v = 0.f;
          for (int i = is2; i < is2 + nth; ++i) v = v + brlambda[i];
          renew = renew - cumulwiso * v;
        }
      }
      for (int ith = 1; ith <= nth; ++ith) {
        const int is = ith + (ik - 1) * nth;
        if (c[2] < 0.f && ik > dikcumul) {
          renew = 0.f;
          for (int ik2 = ik1; ik2 <= ik - dikcumul; ++ik2) {
            if (bth0[ik2 - 1] > q->ssdsbr) {
              const int is2 = (ik2 - 1) * nth;
              float v = 0.f;
              for (int j = 1; j <= nth; ++j)
                v = v + q->cumulw[(size_t)(is2 + j - 1) + (size_t)nspec * (size_t)(is - 1)] *
                            brlambda[is2 + j - 1];
              renew = renew + v;
            }
          }
        }
        wl_sincosf(usdir, &su, &cu);
        const float coswind = q->ecos[ith - 1] * cu + q->esin[ith - 1] * su;
        const float dturb = facturb2 * mx(coswind, 0.f);
        const float diag2 = (c[2] * renew + dturb) + dvisc;
        ddiag[is - 1] = ddiag[is - 1] + diag2;
        srhs[is - 1] = srhs[is - 1] + a[is - 1] * diag2;
      }
    }
  }

  /* whitecap parameters; early return keeps WHITECAP as it was */
  if (!(flogrd57 || flogrd58)) return;
  for (int i = 0; i < 4; ++i) whitecap[i] = 0.f;
  const int ikc = ffloor(q->aaircmin) < nk ? ffloor(q->aaircmin) : nk;
  for (int ik = 1; ik <= ikc; ++ik) {
    const float cc = q->sig[ik] / k[ik - 1];
    const int is0 = (ik - 1) * nth;
    float v = 0.f;
    for (int i = is0; i < is0 + nth; ++i) v = v + brlambda[i];
    coef4[ik - 1] = ((((((cc * cc) * v) * 2.f) * PI_F) / GRAV) * c[6]) * q->dden[ik - 1] /
                    (q->sig[ik] * cg[ik - 1]);
    const float pw = wl_powf(brm12[ik - 1] / c[1], 3.f / 5.f);
    float ww = 0.f;
    for (int i = is0; i < is0 + nth; ++i) ww = ww + (brlambda[i] * (-c[1])) * pw;
    coef5[ik - 1] = ((((cc * cc) * cc) * ww) * q->aairgb / GRAV) * q->dden[ik - 1] /
                    (q->sig[ik] * cg[ik - 1]);
  }
  for (int ik = ikc + 1; ik <= nk; ++ik) {
    coef4[ik - 1] = 0.f;
    coef5[ik - 1] = 0.f;
  }
  if (flogrd57) {
    for (int ik = ik1; ik <= ikc; ++ik) {
      whitecap[0] = whitecap[0] + coef4[ik - 1] * (1.f - whitecap[0]);
      whitecap[3] = whitecap[3] + coef5[ik - 1];
    }
  }
  if (flogrd58) {
    const float tstrfac = (0.8f * 2.f) * PI_F, tmaxfac = (5.f * 2.f) * PI_F;
    for (int ik = ik1; ik <= nk; ++ik) {
      const float tstr = tstrfac / q->sig[ik];
      const float tmax = tmaxfac / q->sig[ik];
      const float dt = tmax / 50.f;
      float mft = 0.f;
      for (int it = 1; it <= 50; ++it) {
        const float t = (float)it * dt;
        if (t < tstr)
          mft = mft + ((0.4f / (k[ik - 1] * tstr)) * t) * dt;
        else
          mft = mft + ((0.4f / k[ik - 1]) * expf(-((t - tstr) / 3.8f))) * dt;
      }
      mft = mft / tmax;
      whitecap[1] = whitecap[1] + coef4[ik - 1] * mft;
    }
  }
}

/* ================================ W3SLN1, W3SBT1, W3SDB1 ================================ */
static void w3sln1(const Srce* p, const float* k, float fhigh, float ustar, float usdir,
                   float* s, float* dirf, float* wnf) {
  const int nk = p->nk, nth = p->nth;
  float cosu = 0.f, sinu = 0.f;
  wl_sincosf(usdir, &sinu, &cosu);
  for (int ith = 1; ith <= nth; ++ith) {
    const float m = fmax2(0.f, p->ecos[ith - 1] * cosu + p->esin[ith - 1] * sinu);
    const float m2 = m * m;
    dirf[ith - 1] = m2 * m2;
  }
  const float u2 = ustar * ustar;
  const float fac = p->slnc1 * (u2 * u2);
  const float ff1 = p->fspm * GRAV / (28.f * ustar);
  const float ff2 = p->fshf * fmin2(p->sig[nk - 1], fhigh);
  const float ffilt = fmin2(2.f * p->sig[nk - 1], fmax2(ff2, ff1));
  for (int ik = 1; ik <= nk; ++ik) {
    const float rfr = p->sig[ik - 1] / ffilt;
    if (rfr < 0.5f) {
      wnf[ik - 1] = 0.f;
    } else {
      const float r2 = rfr * rfr;
      const float rm4 = 1.f / (r2 * r2);
      wnf[ik - 1] = fac / k[ik - 1] * expf(-rm4);
    }
  }
  for (int ik = 1; ik <= nk; ++ik)
    for (int ith = 1; ith <= nth; ++ith) s[(ik - 1) * nth + ith - 1] = wnf[ik - 1] * dirf[ith - 1];
}

static void w3sbt1(const Srce* p, const float* a, const float* cg, const float* wn, float depth,
                   float* s, float* d, float* cbeta) {
  const int nk = p->nk, nth = p->nth, nspec = p->nspec;
  if (depth * wn[0] > 6.f) {
    for (int is = 0; is < nspec; ++is) { d[is] = 0.f; s[is] = 0.f; }
    return;
  }
  const float factor = p->sbtc1 / depth;
  int ik = 1;
  for (; ik <= nk; ++ik) {
    if (wn[ik - 1] * depth > 6.f) break;
    const float x = cg[ik - 1] * wn[ik - 1] / p->sig[ik - 1] - 0.5f;
    cbeta[ik - 1] = factor * (x > 0.f ? x : 0.f);
  }
  const int nscut = (ik - 1) * nth;
  for (int is = 1; is <= nscut; ++is) d[is - 1] = cbeta[p->mapwn[is - 1] - 1];
  for (int is = nscut + 1; is <= nspec; ++is) d[is - 1] = 0.f;
  for (int is = 0; is < nspec; ++is) s[is] = d[is] * a[is];
}

static void w3sdb1(const Srce* p, const float* a, float depth, float emean, float fmean,
                   float wnmean, const float* cg, float* s, float* d, double* eb) {
  const int nk = p->nk, nth = p->nth, nspec = p->nspec;
  const double thr = (double)FLT_MIN;
  if ((double)emean < DBL_MIN) return;
  for (int ik = 1; ik <= nk; ++ik) {
    eb[ik - 1] = 0.;
    for (int ith = 1; ith <= nth; ++ith) eb[ik - 1] = eb[ik - 1] + (double)a[ith + (ik - 1) * nth - 1];
  }
  for (int ik = 1; ik <= nk; ++ik)
    eb[ik - 1] = eb[ik - 1] * (double)p->dden[ik - 1] / (double)cg[ik - 1];
  double hm;
  if (p->fdonly != 0) {
    hm = (double)p->sdbc2 * (double)depth;
  } else {
    const float dpos = depth > 0.f ? depth : 0.f;
    const double wnm = (double)wnmean;
    hm = (double)p->sdbc2 / wnm * tanh(wnm * (double)dpos);
  }
  const double hrms = sqrt(8. * (double)emean);
  double bb, b;
  if (hm > thr) {
    bb = hrms * hrms / (hm * hm);
    b = sqrt(bb);
  } else {
    bb = 0.;
    b = 0.;
  }
  double q0 = 0.;
  if (b <= 0.5) {
    q0 = 0.;
  } else if (b <= 1.) {
    const double t = 2. * b - 1.;
    q0 = t * t;
  }
  double qb;
  if (b <= 0.2) {
    qb = 0.;
  } else if (b < 1.) {
    const double arg = exp((q0 - 1.) / bb);
    qb = q0 - bb * (q0 - arg) / (bb - arg);
    for (int i = 1; i <= 3; ++i) qb = exp((qb - 1.) / bb);
  } else {
    qb = (double)(1.0f - FLT_MIN);
  }
  double cbj;
  if (b != 1. && bb > thr && fabs(bb - qb) > thr)
    cbj = (bb < 1.0) ? 2. * (double)p->sdbc1 * qb * (double)fmean / bb
                     : 2. * (double)p->sdbc1 * (double)fmean * bb;
  else
    cbj = 0.;
  const float dv = (float)(-cbj);
  for (int is = 0; is < nspec; ++is) d[is] = dv;
  for (int is = 0; is < nspec; ++is) s[is] = d[is] * a[is];
}

/* ================================ W3SRCE ================================ */
typedef struct {
  int it, inflags24, flogrd57, flogrd58;
  float dtg;
} StepFlags;

/* per-point work: 18 NSPEC arrays, NTH + 2 NK, ST4 work, DIA work; ints: LLWS; doubles: EB */
static size_t srce_work_floats(const St4* q, size_t dia_floats) {
  const size_t nspec = (size_t)q->nspec, nk = (size_t)q->nk, nth = (size_t)q->nth;
  size_t st4w = 9 * nk + 2 * nspec; /* W3SDS4 (the largest of the three) */
  if (3 * nspec > st4w) st4w = 3 * nspec;
  return 18 * nspec + nth + 2 * nk + st4w + dia_floats;
}

static void srce_point(const Srce* p, const St4* q, const Dia* g, const StepFlags* fl,
                       const float* pin, float* spec, const float* wn1, const float* cg1,
                       float* psc, float* whitecap, float* taubbl, float* tauice, int* shave_out,
                       float* rw, int* llws, double* dw) {
  const int nk = p->nk, nth = p->nth, nspec = p->nspec;
  float* dam = rw;
  float* wn2 = dam + nspec;
  float* vsln = wn2 + nspec;
  float* vsin = vsln + nspec;
  float* vdin = vsin + nspec;
  float* vsnl = vdin + nspec;
  float* vdnl = vsnl + nspec;
  float* vsds = vdnl + nspec;
  float* vdds = vsds + nspec;
  float* vsbt = vdds + nspec;
  float* vdbt = vsbt + nspec;
  float* vs = vdbt + nspec;
  float* vd = vs + nspec;
  float* vsdb = vd + nspec;
  float* vddb = vsdb + nspec;
  float* specinit = vddb + nspec;
  float* spec2 = specinit + nspec;
  float* brlambda = spec2 + nspec;
  float* dirf = brlambda + nspec;
  float* wnf = dirf + nth;
  float* cbeta = wnf + nk;
  float* st4w = cbeta + nk;
  size_t st4n = 9 * (size_t)nk + 2 * (size_t)nspec;
  if (3 * (size_t)nspec > st4n) st4n = 3 * (size_t)nspec;
  float* diaw = st4w + st4n;

  const float d_inp = pin[ORP_D_INP], u10abs = pin[ORP_U10ABS], u10dir = pin[ORP_U10DIR];
  const float as = pin[ORP_AS], ice = pin[ORP_ICE], dair = pin[ORP_DAIR], coef = pin[ORP_COEF];
  const float dtg = fl->dtg;
  float ustar = psc[ORC_USTAR], ustdir = psc[ORC_USTDIR], tauwx = psc[ORC_TAUWX];
  float tauwy = psc[ORC_TAUWY], dtdyn, fcut, tauox, tauoy, tauwix, tauwiy, tauwnx, tauwny;
  float phiaw, charn = psc[ORC_CHARN], tws, phioc, phibbl, phice, tauocx, tauocy, wnmean;
  int shave = *shave_out;
  Spr4Out sp;
  memset(&sp, 0, sizeof sp);
  float tauwax, tauway;
  const float sintail4 = q->sintailpar[3];

  /* 1. initialisation */
  for (int is = 0; is < nspec; ++is) {
    vd[is] = vs[is] = vsbt[is] = vdbt[is] = vsln[is] = vsin[is] = vdin[is] = 0.f;
    vsnl[is] = vdnl[is] = vsds[is] = vdds[is] = vsdb[is] = vddb[is] = 0.f;
  }
  const float depth = mx(d_inp, p->dmin);
  const float drat = dair / DWAT;
  const float isln = mx(mn(1.f - ice * p->icescales[0], 1.f), 0.f);
  const float isin = mx(mn(1.f - ice * p->icescales[1], 1.f), 0.f);
  const float isnl = mx(mn(1.f - ice * p->icescales[2], 1.f), 0.f);
  const float isds = mx(0.f, mn(1.f - ice * p->icescales[3], 1.f));
  for (int ik = 0; ik < nk; ++ik) {
    const float w = wn1[ik];
    dam[ik * nth] = p->facp / (p->sig[ik] * (w * (w * w)));
    wn2[ik * nth] = wn1[ik];
  }
  for (int ik = 0; ik < nk; ++ik)
    for (int ith = 1; ith < nth; ++ith) {
      dam[ith + ik * nth] = dam[ik * nth];
      wn2[ith + ik * nth] = wn2[ik * nth];
    }
  dtdyn = 0.f;
  float dttot = 0.f;
  int nsteps = 0;
  phiaw = 0.f;
  charn = 0.f;
  tws = 0.f;
  phibbl = 0.f;
  tauwix = tauwiy = tauwnx = tauwny = 0.f;
  tauwax = tauway = 0.f;
  taubbl[0] = taubbl[1] = 0.f;
  tauice[0] = tauice[1] = 0.f;
  phice = tauocx = tauocy = 0.f;
  wnmean = 0.f;
  tauox = psc[ORC_TAUOX];
  tauoy = psc[ORC_TAUOY];
  phioc = psc[ORC_PHIOC];
  fcut = psc[ORC_FCUT];
  for (int is = 0; is < nspec; ++is) brlambda[is] = 0.f;
  for (int j = 0; j < 4; ++j) whitecap[j] = 0.f;
  if (sintail4 > 0.5f) { tauwx = 0.f; tauwy = 0.f; }
#define SPR4() w3spr4(q, spec, cg1, wn1, u10abs, u10dir, tauwx, tauwy, llws, &ustar, &ustdir, \
                      &wnmean, &charn, &sp, st4w)
#define SIN4() w3sin4(q, spec, cg1, wn2, u10abs, ustar, drat, as, u10dir, sp.z0, &tauwx, &tauwy, \
                      &tauwax, &tauway, vsin, vdin, llws, st4w)
  if (fl->it == 0) {
    for (int is = 0; is < nspec; ++is) llws[is] = 1;
    tauwx = tauwy = ustar = ustdir = 0.f;
  } else {
    SPR4();
    if (sintail4 > 0.5f) SIN4();
  }
  SPR4();
  tws = 1.f / sp.fmeanws;
  for (int is = 0; is < nspec; ++is) specinit[is] = spec[is];
  float fage = 0.f;
  float fhigh = mx((p->ffxfm + 0.f) * mx(sp.fmeanws, sp.fmean1), p->ffxpm / ustar);
  float hdt, dt;
  int nkh, nkh1;

  /* 2. the integration loop */
  for (;;) {
    nsteps = nsteps + 1;
    if (p->ln1) w3sln1(p, wn1, fhigh, ustar, u10dir, vsln, dirf, wnf);
    SIN4();
    or_w3snl1_one(nk, nth, g->xfr, g->lam, g->snlc1, g->kdcon, g->kdmn, g->snls1, g->snls2,
                  g->snls3, g->fachfe, g->scal, g->ipim, g->ic, g->af11, g->sig, spec, cg1,
                  wnmean * depth, vsnl, vdnl, diaw);
    w3sds4(q, spec, wn1, cg1, ustar, ustdir, depth, dair, fl->flogrd57, fl->flogrd58, vsds, vdds,
           brlambda, whitecap, st4w);
    w3sdb1(p, spec, depth, sp.emean, sp.fmean, wnmean, cg1, vsdb, vddb, dw);
    w3sbt1(p, spec, cg1, wn1, depth, vsbt, vdbt, cbeta);
    nkh = mni(nk, fint(p->facti1 * wl_logf(mx(1.E-7f, fhigh)) + p->facti2));
    nkh1 = mni(nk, nkh + 1);
    /* NKH1*NTH with the wrap of the compiled reference (NKH = INT_MIN when FHIGH = +Inf),
       never above NSPEC */
    const int nspech_w = (int)((unsigned)nkh1 * (unsigned)nth);
    const int nspech = nspech_w < nspec ? nspech_w : nspec;
    dt = mn(p->dtmax, dtg - dttot);
    const float afilt = mx(dam[nspec - 1], p->xflt * sp.amax);
    if (fl->inflags24) {
      for (int is = 0; is < nspech; ++is) vsnl[is] = isnl * vsnl[is];
      for (int is = 0; is < nspech; ++is) vdnl[is] = isnl * vdnl[is];
      for (int is = 0; is < nspech; ++is) vsln[is] = isln * vsln[is];
      for (int is = 0; is < nspech; ++is) vsin[is] = isin * vsin[is];
      for (int is = 0; is < nspech; ++is) vdin[is] = isin * vdin[is];
      for (int is = 0; is < nspech; ++is) vsds[is] = isds * vsds[is];
      for (int is = 0; is < nspech; ++is) vdds[is] = isds * vdds[is];
    }
    for (int is = 0; is < nspech; ++is) {
      vs[is] = (((vsln[is] + vsin[is]) + vsnl[is]) + vsds[is]) + vsbt[is];
      vd[is] = ((vdin[is] + vdnl[is]) + vdds[is]) + vdbt[is];
      const float damax = mn(mx(p->xrel * specinit[is], afilt), dam[is]);
      const float afac = 1.f / mx(fabsf(vs[is] / damax), 1.E-10f);
      dt = mn(dt, afac / mx(mn(vd[is], 0.f) * afac + 1.f, 1.E-10f));
    }
    dt = mx(dt, 0.5f);
    dtdyn = dtdyn + dt;
    const float rest = dtg - dttot;
    const int idt = 1 + fint((rest * 0.99f) / dt);
    dt = rest / (float)idt;
    const int shv = dt < p->dtmin && dt < rest;
    shave = shv;
    dt = mx(mn(p->dtmin, rest), dt);
    hdt = dt;
    dttot = dttot + dt;
    if (shv) {
      for (int is = 0; is < nspech; ++is) {
        const float e1 = (dt * vs[is]) / mx(1.f - hdt * vd[is], 1.f);
        const float e2 = copysignf(mn(fabsf(e1), dam[is]), e1);
        spec[is] = mx(spec[is] + e2, 0.f);
      }
    } else {
      for (int is = 0; is < nspech; ++is) {
        const float e1 = (dt * vs[is]) / mx(1.f - hdt * vd[is], 1.f);
        spec[is] = mx(spec[is] + e1, 0.f);
      }
    }
    for (int is = 0; is < nspech; ++is) {
      const float e1 = (dt * vsdb[is]) / mx(1.f - hdt * vddb[is], 1.f);
      spec[is] = mx(spec[is] + e1, 0.f);
    }
    whitecap[2] = 0.f;
    float hstot = 0.f;
    for (int ik = 0; ik < nk; ++ik) {
      const float factor = p->dden[ik] / cg1[ik];
      for (int ith = 0; ith < nth; ++ith) {
        const int is = ik * nth + ith;
        phiaw = ((vsin[is] * dt) * factor) / mx(1.f - vdin[is] * dt, 1.f) + phiaw;
        phibbl = phibbl - ((vsbt[is] * dt) * factor) / mx(1.f - vdbt[is] * dt, 1.f);
        if (vsin[is] > 0.f) whitecap[2] = whitecap[2] + spec[is] * factor;
        hstot = hstot + spec[is] * factor;
      }
    }
    whitecap[2] = sqrtf(whitecap[2]) * 4.f;
    (void)hstot;
    tauwix = tauwix + (tauwx * drat) * dt;
    tauwiy = tauwiy + (tauwy * drat) * dt;
    tauwnx = tauwnx + (drat * tauwax) * dt;
    tauwny = tauwny + (drat * tauway) * dt;
    SPR4();
    fage = p->ffxfa * wl_tanhf((((u10abs * 0.3f) * sp.fmeanws) * TPI_F) / GRAV);
    const float fh1 = (p->ffxfm + fage) * sp.fmean1;
    const float fh2 = p->ffxpm / ustar;
    fhigh = mn(mx(fh1, fh2), p->sig[nk - 1]);
    nkh = mxi(mni(nkh1, fint(p->facti1 * wl_logf(mx(1.E-7f, fhigh)) + p->facti2)), 2);
    if (dttot >= dtg * 0.9999f) {
      const float hm = (p->fhmax * wl_tanhf(mx(0.f, d_inp) * wnmean)) /
                       mx(wnmean, 9.99999974737875163555145263671875e-5f);
      const float em = (hm * hm) * 6.25e-2f;
      if (sp.emean > em && sp.emean > 1.E-30f) {
        for (int is = 0; is < nspec; ++is) spec[is] = em * (spec[is] / sp.emean);
        sp.emean = em;
      }
    }
    for (int ik = nkh; ik < nk; ++ik)
      for (int ith = 0; ith < nth; ++ith)
        spec[ith + ik * nth] = spec[ith + (ik - 1) * nth] * p->fachfa + 0.f;
    SIN4();
    if (sintail4 < 0.5f) SPR4();
    if (dttot >= 0.9999f * dtg) break;
  }
#undef SPR4
#undef SIN4

  /* 3. after the loop */
  dtdyn = dtdyn / (float)mxi(nsteps, 1);
  fcut = fhigh * TPIINV_F;
  float efinish = 0.f, mwx = 0.f, mwy = 0.f;
  for (int ik = 0; ik < nk; ++ik) {
    float eband = 0.f, a1 = 0.f, b1 = 0.f;
    for (int ith = 0; ith < nth; ++ith) {
      const float diff = specinit[ith + ik * nth] - spec[ith + ik * nth];
      eband = diff + eband;
      a1 = p->ecos[ith] * diff + a1;
      b1 = p->esin[ith] * diff + b1;
    }
    efinish = (p->dden[ik] * eband) / cg1[ik] + efinish;
    mwx = (((p->dden[ik] * a1) / cg1[ik]) * wn1[ik]) / p->sig[ik] + mwx;
    mwy = (wn1[ik] * ((p->dden[ik] * b1) / cg1[ik])) / p->sig[ik] + mwy;
  }
  tauox = ((tauwix + mwx * GRAV) - taubbl[0]) / dtg;
  tauoy = ((tauwiy + mwy * GRAV) - taubbl[1]) / dtg;
  tauwix = tauwix / dtg;
  tauwiy = tauwiy / dtg;
  tauwnx = tauwnx / dtg;
  tauwny = tauwny / dtg;
  taubbl[0] = taubbl[0] / dtg;
  taubbl[1] = taubbl[1] / dtg;
  {
    float s, c;
    wl_sincosf(ustdir, &s, &c);
    const float kk = ustar * (ustar * (coef * (dair * coef)));
    tauocx = kk * c + (tauox - tauwix) * DWAT;
    tauocy = kk * s + (tauoy - tauwiy) * DWAT;
  }
  const float dg = DWAT * GRAV;
  phioc = (((phiaw + efinish) - phibbl) * dg) / dtg;
  phiaw = (phiaw * dg) / dtg;
  phibbl = (phibbl * dg) / dtg;
  if (fl->inflags24 && ice > 0.f && !p->ic_numerics) {
    for (int is = 0; is < nspec; ++is) spec2[is] = spec[is];
    tauice[0] = tauice[1] = 0.f;
    phice = 0.f;
    for (int ik = 0; ik < nk; ++ik) {
      const double att = 1.;
      for (int ith = 0; ith < nth; ++ith)
        spec[ith + ik * nth] = (float)(att * (double)spec2[ith + ik * nth]);
      const float factor = p->dden[ik] / cg1[ik];
      const float factor2 = ((factor * GRAV) * wn1[ik]) / p->sig[ik];
      for (int ith = 0; ith < nth; ++ith) {
        const int is = ith + ik * nth;
        phice = phice + (spec[is] - spec2[is]) * factor;
        tauice[0] = tauice[0] - ((spec[is] - spec2[is]) * factor2) * p->ecos[is];
        tauice[1] = tauice[1] - ((spec[is] - spec2[is]) * factor2) * p->esin[is];
      }
    }
    phice = (((-1.f * DWAT) * GRAV) * phice) / dtg;
    tauice[0] = tauice[0] / dtg;
    tauice[1] = tauice[1] / dtg;
  }
  if (fl->it == 0)
    for (int is = 0; is < nspec; ++is) spec[is] = specinit[is];
  const int nlast = (nspec % 4 == 3) ? nspec - 1 : nspec;
  for (int is = 0; is < nlast; ++is) spec[is] = mx(spec[is], 0.f);
  if (nlast < nspec) spec[nlast] = mx(0.f, spec[nlast]);

  psc[ORC_USTAR] = ustar;
  psc[ORC_USTDIR] = ustdir;
  psc[ORC_TAUWX] = tauwx;
  psc[ORC_TAUWY] = tauwy;
  psc[ORC_DTDYN] = dtdyn;
  psc[ORC_FCUT] = fcut;
  psc[ORC_TAUOX] = tauox;
  psc[ORC_TAUOY] = tauoy;
  psc[ORC_TAUWIX] = tauwix;
  psc[ORC_TAUWIY] = tauwiy;
  psc[ORC_TAUWNX] = tauwnx;
  psc[ORC_TAUWNY] = tauwny;
  psc[ORC_PHIAW] = phiaw;
  psc[ORC_CHARN] = charn;
  psc[ORC_TWS] = tws;
  psc[ORC_PHIOC] = phioc;
  psc[ORC_PHIBBL] = phibbl;
  psc[ORC_PHICE] = phice;
  psc[ORC_TAUOCX] = tauocx;
  psc[ORC_TAUOCY] = tauocy;
  psc[ORC_WNMEAN] = wnmean;
  *shave_out = shave;
}

/*
 * W3SRCE for npts points, the reference's ww_srce_batch argument list
 * (kokkos/src/fortran_iface/ww_kokkos_c.hpp:267-270) plus the tables the GPU library holds
 * from its init calls: ST4 packets and arrays, W3SRCE packets and arrays, and the DIA
 * configuration with its INSNL1 tables.  Returns 0, or -1 on an unsupported configuration.
 */
int or_srce_batch(const float* st4_rpar, const int* st4_ipar, const float* sig0, const float* sig2,
                  const float* dsip0, const float* dden, const float* dden2, const float* esin,
                  const float* ecos, const float* th, const float* satw, const int* sati, int nsat,
                  const float* cumulw, const float* fwtable, const float* srce_rpar,
                  const int* srce_ipar, const int* mapwn, const float* dia_cfg9,
                  const float* dia_scal, const int* dia_ipim, const int* dia_ic,
                  const float* dia_af11, int it, int inflags24, int flogrd57, int flogrd58,
                  float dtg, int npts, float* spec, const float* wn1, const float* cg1,
                  const float* pin, float* psc, float* whitecap, float* taubbl, float* tauice,
                  int* shave, int nthreads) {
  St4 q;
  Srce p;
  Dia g;
  st4_unpack(&q, st4_rpar, st4_ipar);
  srce_unpack(&p, srce_rpar, srce_ipar);
  if (q.sintailpar[0] > 0.5f || nint(q.ssdsc[0]) != 1 || srce_ipar[ORRI_SEED] ||
      srce_ipar[ORRI_FLD2] || srce_ipar[ORRI_LAB_SWITCH] != 1 || (it != 0 && q.sintailpar[3] <= 0.5f) ||
      (q.sswellf[1] != 0.f && fwtable == NULL))
    return -1;
  q.sig = sig0;
  q.sig2 = sig2;
  q.dsip = dsip0;
  q.dden = dden;
  q.dden2 = dden2;
  q.esin = esin;
  q.ecos = ecos;
  q.th = th;
  q.satw = satw;
  q.sati = sati;
  q.nsat = nsat;
  q.cumulw = cumulw;
  q.fwtable = fwtable;
  p.sig = sig0 + 1;
  p.dden = dden;
  p.ecos = ecos;
  p.esin = esin;
  p.mapwn = mapwn;
  g.xfr = dia_cfg9[0]; g.lam = dia_cfg9[1]; g.snlc1 = dia_cfg9[2]; g.kdcon = dia_cfg9[3];
  g.kdmn = dia_cfg9[4]; g.snls1 = dia_cfg9[5]; g.snls2 = dia_cfg9[6]; g.snls3 = dia_cfg9[7];
  g.fachfe = dia_cfg9[8];
  g.scal = dia_scal; g.ipim = dia_ipim; g.ic = dia_ic; g.af11 = dia_af11; g.sig = sig0 + 1;
  const StepFlags fl = {it, inflags24, flogrd57, flogrd58, dtg};
  const int nk = p.nk, nspec = p.nspec;
  const size_t wf = srce_work_floats(&q, or_snl1_work_floats(nk, p.nth, g.xfr, g.lam));
  int err = 0;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(| : err)
#endif
  {
    float* rw = (float*)malloc(sizeof(float) * wf);
    int* iw = (int*)malloc(sizeof(int) * (size_t)nspec);
    double* dw = (double*)malloc(sizeof(double) * (size_t)nk);
    if (!rw || !iw || !dw) {
      err = 1;
    } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
      for (int i = 0; i < npts; ++i) {
        /* LLWS is read (by W3SPR4 at IT > 0) before it is set only under TAUWBUG = 0, which
           is refused above; start it at 0 so no garbage is read. */
        for (int is = 0; is < nspec; ++is) iw[is] = 0;
        srce_point(&p, &q, &g, &fl, pin + (size_t)i * ORP_N, spec + (size_t)i * nspec,
                   wn1 + (size_t)i * nk, cg1 + (size_t)i * nk, psc + (size_t)i * ORC_N,
                   whitecap + (size_t)i * 4, taubbl + (size_t)i * 2, tauice + (size_t)i * 2,
                   shave + i, rw, iw, dw);
      }
    }
    free(rw);
    free(iw);
    free(dw);
  }
  (void)nthreads;
  return err ? -1 : 0;
}

/*
 * INSIN4's saturation and cumulative tables (kokkos/src/ww_kokkos/st4_insin4.cpp:110-138,
 * 239-262) from the registered grid arrays, host libm: satw/sati (nsat x NTH),
 * cumulw (NSPEC x NSPEC).  nsat = 2*(NTH/2-1)+1.
 */
void or_insin4_tables(const float* st4_rpar, const int* st4_ipar, const float* sig0,
                      const float* dsip0, const float* ecos, const float* th, float* satw,
                      int* sati, float* cumulw) {
  St4 q;
  st4_unpack(&q, st4_rpar, st4_ipar);
  const int nk = q.nk, nth = q.nth, nspec = q.nspec;
  const int nsat = 2 * (nth / 2 - 1) + 1;
  const float rade = 180.f / PI_F;
  for (int i = 0; i < nsat * nth; ++i) { sati[i] = 1; satw[i] = 0.f; }
  if (q.ssdsdth < 180.f) {
    int sdsnth = nint(q.ssdsdth / (q.dth * rade));
    if (sdsnth > nth / 2 - 1) sdsnth = nth / 2 - 1;
    for (int ith = 1; ith <= nth; ++ith)
      for (int ii = ith - sdsnth; ii <= ith + sdsnth; ++ii) {
        int jj = ii;
        if (ii < 1) jj = ii + nth;
        if (ii > nth) jj = ii - nth;
        const int row = ii - (ith - sdsnth) + 1;
        sati[(row - 1) + nsat * (ith - 1)] = jj;
        satw[(row - 1) + nsat * (ith - 1)] = powf(cosf(th[ith - 1] - th[jj - 1]), q.ssdscos);
      }
  } else {
    for (int i = 0; i < nsat * nth; ++i) satw[i] = 1.f;
  }
  memset(cumulw, 0, sizeof(float) * (size_t)nspec * (size_t)nspec);
  if (q.ssdsc[2] < 0.f) {
    const int dik = nint(q.ssdsbrf1 / (q.xfr - 1.f));
    for (int ik = 1; ik <= nk; ++ik) {
      const float c = GRAV / sig0[ik];
      for (int ith = 1; ith <= nth; ++ith) {
        const int is = ith + (ik - 1) * nth;
        for (int ik2 = 1; ik2 <= ik - dik; ++ik2) {
          const float c2 = GRAV / sig0[ik2];
          for (int ith2 = 1; ith2 <= nth; ++ith2) {
            const int is2 = ith2 + (ik2 - 1) * nth;
            cumulw[(size_t)(is2 - 1) + (size_t)nspec * (size_t)(is - 1)] =
                sqrtf(c * c + c2 * c2 - 2.f * c * c2 * ecos[abs(ith2 - ith)]) * dsip0[ik2] /
                (0.5f * c2) * q.dth;
          }
        }
      }
    }
  }
}

/*
 * The per-routine entries ww_spr4, ww_sin4, ww_sds4 (kokkos/src/fortran_iface/ww_kokkos_c.hpp:
 * 177-196): each routine on npts points, every per-point scalar an array of npts.  The ST4
 * grid arrays as or_srce_batch takes them.  Returns -1 for a configuration the library refuses.
 */
static int st4_setup(St4* q, const float* st4_rpar, const int* st4_ipar, const float* sig0,
                     const float* sig2, const float* dsip0, const float* dden,
                     const float* dden2, const float* esin, const float* ecos, const float* th,
                     const float* satw, const int* sati, int nsat, const float* cumulw,
                     const float* fwtable) {
  st4_unpack(q, st4_rpar, st4_ipar);
  if (q->sintailpar[0] > 0.5f || nint(q->ssdsc[0]) != 1 || (q->sswellf[1] != 0.f && fwtable == NULL))
    return -1;
  q->sig = sig0;
  q->sig2 = sig2;
  q->dsip = dsip0;
  q->dden = dden;
  q->dden2 = dden2;
  q->esin = esin;
  q->ecos = ecos;
  q->th = th;
  q->satw = satw;
  q->sati = sati;
  q->nsat = nsat;
  q->cumulw = cumulw;
  q->fwtable = fwtable;
  return 0;
}
#define OR_ST4_TABLES                                                                          \
  const float *st4_rpar, const int *st4_ipar, const float *sig0, const float *sig2,            \
      const float *dsip0, const float *dden, const float *dden2, const float *esin,            \
      const float *ecos, const float *th, const float *satw, const int *sati, int nsat,        \
      const float *cumulw, const float *fwtable
#define OR_ST4_SETUP                                                                           \
  St4 q;                                                                                       \
  if (st4_setup(&q, st4_rpar, st4_ipar, sig0, sig2, dsip0, dden, dden2, esin, ecos, th, satw, \
                sati, nsat, cumulw, fwtable) != 0)                                             \
    return -1;                                                                                 \
  const int nk = q.nk, nspec = q.nspec;

int or_spr4_batch(OR_ST4_TABLES, int npts, const float* a, const float* cg, const float* wn,
                  const float* u, const float* udir, const float* tauwx, const float* tauwy,
                  const int* llws, float* ustar, float* usdir, float* emean, float* fmean,
                  float* fmean1, float* wnmean, float* amax, float* cd, float* z0, float* charn,
                  float* fmeanws, float* dlwmean) {
  OR_ST4_SETUP
  float* w = (float*)malloc(sizeof(float) * 4 * (size_t)nk);
  for (int i = 0; i < npts; ++i) {
    Spr4Out o;
    w3spr4(&q, a + (size_t)i * nspec, cg + (size_t)i * nk, wn + (size_t)i * nk, u[i], udir[i],
           tauwx[i], tauwy[i], llws + (size_t)i * nspec, &ustar[i], &usdir[i], &wnmean[i],
           &charn[i], &o, w);
    emean[i] = o.emean;
    fmean[i] = o.fmean;
    fmean1[i] = o.fmean1;
    amax[i] = o.amax;
    cd[i] = o.cd;
    z0[i] = o.z0;
    fmeanws[i] = o.fmeanws;
    dlwmean[i] = o.dlwmean;
  }
  free(w);
  return 0;
}

int or_sin4_batch(OR_ST4_TABLES, int npts, const float* a, const float* cg, const float* k,
                  const float* u, const float* ustar, const float* drat, const float* as,
                  const float* usdir, const float* z0, const float* cd, float* tauwx,
                  float* tauwy, float* tauwnx, float* tauwny, float* s, float* d, int* llws) {
  OR_ST4_SETUP
  (void)cd;  /* an argument of W3SIN4 that it never reads */
  float* w = (float*)malloc(sizeof(float) * 3 * (size_t)nspec);
  for (int i = 0; i < npts; ++i)
    w3sin4(&q, a + (size_t)i * nspec, cg + (size_t)i * nk, k + (size_t)i * nspec, u[i], ustar[i],
           drat[i], as[i], usdir[i], z0[i], &tauwx[i], &tauwy[i], &tauwnx[i], &tauwny[i],
           s + (size_t)i * nspec, d + (size_t)i * nspec, llws + (size_t)i * nspec, w);
  free(w);
  return 0;
}

int or_sds4_batch(OR_ST4_TABLES, int npts, const float* a, const float* k, const float* cg,
                  const float* ustar, const float* usdir, const float* depth, const float* dair,
                  const float* dlwmean, int flogrd57, int flogrd58, float* srhs, float* ddiag,
                  float* brlambda, float* whitecap) {
  OR_ST4_SETUP
  (void)dlwmean;  /* read only under SDSBCHOICE = 2, which is refused */
  float* w = (float*)malloc(sizeof(float) * (9 * (size_t)nk + 2 * (size_t)nspec));
  for (int i = 0; i < npts; ++i)
    w3sds4(&q, a + (size_t)i * nspec, k + (size_t)i * nk, cg + (size_t)i * nk, ustar[i], usdir[i],
           depth[i], dair[i], flogrd57, flogrd58, srhs + (size_t)i * nspec, ddiag + (size_t)i * nspec,
           brlambda + (size_t)i * nspec, whitecap + (size_t)i * 4, w);
  free(w);
  return 0;
}

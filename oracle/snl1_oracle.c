/*
 * oracle/snl1_oracle.c -- CPU oracle for the DIA nonlinear interactions (W3SNL1) and
 * its set-up (INSNL1).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is compared
 * against; it is never linked into the library.  Only tests/, the smoke check and the
 * CPU-baseline leg of bench.py load it.
 *
 * It is a plain, scalar, point-at-a-time restatement of the reference algorithm in the
 * Fortran's own index convention (1-based addresses, the zero row at 1-NTH..0), written
 * to be read next to:
 *   INSNL1 : kokkos/src/ww_kokkos/snl1_tables.cpp:82-329  (WW3 w3snl1md.F90:483-786)
 *   W3SNL1 : kokkos/src/ww_kokkos/snl1_team.hpp:352-497   (WW3 w3snl1md.F90:115-473)
 *   TPIINV : kokkos/src/ww_kokkos/snl1_config.hpp:50-51, real.hpp:24
 * Every expression keeps the reference's float32 operation order; the file is built with
 * -ffp-contract=off so no multiply-add is fused.  Libm calls are the host glibc's.
 *
 * Pinning: the INSNL1 tables and the W3SNL1 outputs are checked bit for bit against the
 * Fortran answers in tests/golden/snl1_nk25_nth24.bin (the reference's fixture,
 * kokkos/tests/fixtures/README.md:40-68).  That fixture was written by a -O0 build whose
 * AF11 uses the x^3*x^8 power chain; -O3 WW3 uses x^5*x^6 (snl1_tables.cpp:43-59), so
 * or_insnl1() takes the chain as an argument.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* PI rounded to float first, then doubled and inverted (snl1_config.hpp:50-51). */
static const float kPi = (float)3.141592653589793;
#define TPI (2.0f * kPi)
#define TPIINV (1.0f / TPI)

/* libgcc __powisf2: Fortran real**integer with a run-time exponent. */
static float powi_f(float x, int m) {
  unsigned n = m < 0 ? (unsigned)(-m) : (unsigned)m;
  float y = (n & 1u) ? x : 1.0f;
  while ((n >>= 1) != 0u) {
    x = x * x;
    if (n & 1u) y = y * x;
  }
  return m < 0 ? 1.0f / y : y;
}

/* x**11 as -O3 gfortran expands it (x5*x6) or as __powisf2 computes it (x3*x8). */
static float pow11_f(float x, int chain) {
  if (chain == 0) {
    float x2 = x * x, x3 = x2 * x, x5 = x2 * x3, x6 = x3 * x3;
    return x5 * x6;
  }
  return powi_f(x, 11);
}

/* Fortran MAX(a,b) for non-NaN operands. */
static float fmax_f(float a, float b) { return a > b ? a : b; }

/* The extents INSNL1 derives: dims[0..4] = NFR, NFRHGH, NFRCHG, NSPECX, NSPECY, and
 * the frequency offsets dims[5..8] = IFRP, IFRP1, IFRM, IFRM1. */
void or_snl1_dims(int nk, int nth, float xfr, float lam, int* dims) {
  const float xfrln = logf(xfr);
  const int ifrp = (int)(logf(1.0f + lam) / xfrln);
  const int ifrm = (int)(logf(1.0f - lam) / xfrln);
  const int ifrp1 = ifrp + 1, ifrm1 = ifrm - 1;
  dims[0] = nk;
  dims[1] = nk + ifrp1 - ifrm1;
  dims[2] = nk - ifrm1;
  dims[3] = dims[2] * nth;
  dims[4] = dims[1] * nth;
  dims[5] = ifrp;
  dims[6] = ifrp1;
  dims[7] = ifrm;
  dims[8] = ifrm1;
}

/*
 * INSNL1.  Outputs in the fixture's order and Fortran convention:
 *   scal[0..2] = DAL1, DAL2, DAL3; scal[3..10] = AWG1..8; scal[11..18] = SWG1..8
 *   ipim[16][NSPECX] = IP11..IP14, IM11..IM14, IP21..IP24, IM21..IM24 (1-based)
 *   ic[16][NSPEC]    = IC11, IC21, .., IC81, IC12, IC22, .., IC82    (1-based)
 *   af11[NSPECX]
 * Returns 0, or -1 on a bad argument.
 */
int or_insnl1(int nk, int nth, float xfr, float dth, float lam, const float* sig, int chain,
              float* scal, int* ipim, int* ic, float* af11) {
  if (nk <= 0 || nth <= 0 || sig == NULL) return -1;
  int dims[9];
  or_snl1_dims(nk, nth, xfr, lam, dims);
  const int nfr = dims[0], nfrchg = dims[2], nspecx = dims[3];
  const int ifrp = dims[5], ifrp1 = dims[6], ifrm = dims[7], ifrm1 = dims[8];
  const int nspec = nk * nth;

  /* 1. quadruplet angles */
  const float lamm2 = (1.0f - lam) * (1.0f - lam);
  const float lamp2 = (1.0f + lam) * (1.0f + lam);
  const float delth3 = acosf((lamm2 * lamm2 + 4.0f - lamp2 * lamp2) / (4.0f * lamm2));
  const float delth4 = asinf(-sinf(delth3) * lamm2 / lamp2);

  /* 2. lambda weights: (1+LAM)**4 as (x*x)*(x*x) */
  const float dal1 = 1.0f / (lamp2 * lamp2);
  const float dal2 = 1.0f / (lamm2 * lamm2);
  const float dal3 = 2.0f * dal1 * dal2;

  /* 3. directional offsets and weights */
  const float cthp = fabsf(delth4 / dth);
  const int ithp = (int)cthp, ithp1 = ithp + 1;
  const float wthp = cthp - (float)ithp, wthp1 = 1.0f - wthp;
  const float cthm = fabsf(delth3 / dth);
  const int ithm = (int)cthm, ithm1 = ithm + 1;
  const float wthm = cthm - (float)ithm, wthm1 = 1.0f - wthm;

  /* 4. frequency weights */
  const float wfrp = (1.0f + lam - powi_f(xfr, ifrp)) / (powi_f(xfr, ifrp1) - powi_f(xfr, ifrp));
  const float wfrp1 = 1.0f - wfrp;
  const float wfrm = (powi_f(xfr, ifrm) - (1.0f - lam)) / (powi_f(xfr, ifrm) - powi_f(xfr, ifrm1));
  const float wfrm1 = 1.0f - wfrm;

  /* 7. addresses: IFn(IFR) and ITn(ITH), 1-based, n = 1..8 */
  int* ifn = (int*)malloc(sizeof(int) * 9 * (size_t)(nfrchg + 1));
  int* itn = (int*)malloc(sizeof(int) * 9 * (size_t)(nth + 1));
  if (!ifn || !itn) { free(ifn); free(itn); return -1; }
#define IFN(n, r) ifn[(n) * (nfrchg + 1) + (r)]
#define ITN(n, t) itn[(n) * (nth + 1) + (t)]
  for (int r = 1; r <= nfrchg; ++r) {
    IFN(1, r) = r + ifrp;
    IFN(2, r) = r + ifrp1;
    IFN(3, r) = r + ifrm > 0 ? r + ifrm : 0;
    IFN(4, r) = r + ifrm1 > 0 ? r + ifrm1 : 0;
    IFN(5, r) = r - ifrp > 0 ? r - ifrp : 0;
    IFN(6, r) = r - ifrp1 > 0 ? r - ifrp1 : 0;
    IFN(7, r) = r - ifrm;
    IFN(8, r) = r - ifrm1;
  }
  for (int t = 1; t <= nth; ++t) {
    ITN(1, t) = t + ithp;
    ITN(2, t) = t + ithp1;
    ITN(3, t) = t + ithm;
    ITN(4, t) = t + ithm1;
    ITN(5, t) = t - ithp;
    ITN(6, t) = t - ithp1;
    ITN(7, t) = t - ithm;
    ITN(8, t) = t - ithm1;
    for (int n = 1; n <= 4; ++n)
      if (ITN(n, t) > nth) ITN(n, t) -= nth;
    for (int n = 5; n <= 8; ++n)
      if (ITN(n, t) < 1) ITN(n, t) += nth;
  }
  /* (it, if) pairs of the 16 NSPECX tables, fixture order */
  static const int kx[16][2] = {{2, 2}, {1, 2}, {2, 1}, {1, 1}, {8, 4}, {7, 4}, {8, 3}, {7, 3},
                                {6, 2}, {5, 2}, {6, 1}, {5, 1}, {4, 4}, {3, 4}, {4, 3}, {3, 3}};
  /* (it, if) pairs of IC11..IC81 then IC12..IC82 */
  static const int kc[16][2] = {{6, 6}, {5, 6}, {6, 5}, {5, 5}, {4, 8}, {3, 8}, {4, 7}, {3, 7},
                                {2, 6}, {1, 6}, {2, 5}, {1, 5}, {8, 8}, {7, 8}, {8, 7}, {7, 7}};
  for (int isp = 1; isp <= nspecx; ++isp) {
    const int r = 1 + (isp - 1) / nth, t = 1 + (isp - 1) % nth;
    for (int j = 0; j < 16; ++j)
      ipim[j * nspecx + (isp - 1)] = ITN(kx[j][0], t) + (IFN(kx[j][1], r) - 1) * nth;
  }
  for (int isp = 1; isp <= nspec; ++isp) {
    const int r = 1 + (isp - 1) / nth, t = 1 + (isp - 1) % nth;
    for (int j = 0; j < 16; ++j)
      ic[j * nspec + (isp - 1)] = ITN(kc[j][0], t) + (IFN(kc[j][1], r) - 1) * nth;
  }
#undef IFN
#undef ITN
  free(ifn);
  free(itn);

  /* 8. AF11 = f**11 per frequency row, the extension continued by FR = FR*XFR */
  for (int r = 1; r <= nfr; ++r) {
    const float v = pow11_f(sig[r - 1] * TPIINV, chain);
    for (int t = 1; t <= nth; ++t) af11[t - 1 + (r - 1) * nth] = v;
  }
  float fr = sig[nfr - 1] * TPIINV;
  for (int r = nfr + 1; r <= nfrchg; ++r) {
    fr = fr * xfr;
    const float v = pow11_f(fr, chain);
    for (int t = 1; t <= nth; ++t) af11[t - 1 + (r - 1) * nth] = v;
  }

  /* 9. interpolation weights */
  float awg[8];
  awg[0] = wthp * wfrp;
  awg[1] = wthp1 * wfrp;
  awg[2] = wthp * wfrp1;
  awg[3] = wthp1 * wfrp1;
  awg[4] = wthm * wfrm;
  awg[5] = wthm1 * wfrm;
  awg[6] = wthm * wfrm1;
  awg[7] = wthm1 * wfrm1;
  scal[0] = dal1;
  scal[1] = dal2;
  scal[2] = dal3;
  for (int j = 0; j < 8; ++j) {
    scal[3 + j] = awg[j];
    scal[11 + j] = awg[j] * awg[j];
  }
  return 0;
}

/*
 * W3SNL1 for one point, on the Fortran-ordered tables of or_insnl1().  The working
 * arrays are indexed like the Fortran's (lower bound 1-NTH): element J lives at
 * w[J - 1 + NTH].  `work` holds (NSPECY+NTH) + 8*(NSPECX+NTH) + NSPEC floats.
 */
static void w3snl1_point(int nk, int nth, const int* dims, float snlc1, float kdcon, float kdmn,
                         float snls1, float snls2, float snls3, float fachfe, const float* scal,
                         const int* ipim, const int* ic, const float* af11, const float* sig,
                         const float* a, const float* cg, float kdmean, float* s, float* d,
                         float* work) {
  const int nfrhgh = dims[1], nspecx = dims[3], nspecy = dims[4];
  const int nspec = nk * nth;
  const float dal1 = scal[0], dal2 = scal[1], dal3 = scal[2];
  const float* awg = scal + 3;
  const float* swg = scal + 11;

  float* ue = work;
  float* sa1 = ue + nspecy + nth;
  float* sa2 = sa1 + nspecx + nth;
  float* da1c = sa2 + nspecx + nth;
  float* da1p = da1c + nspecx + nth;
  float* da1m = da1p + nspecx + nth;
  float* da2c = da1m + nspecx + nth;
  float* da2p = da2c + nspecx + nth;
  float* da2m = da2p + nspecx + nth;
  float* con = da2m + nspecx + nth;
#define F(arr, j) arr[(j) - 1 + nth]

  /* 1. proportionality constant */
  const float x = fmax_f(kdcon * kdmean, kdmn);
  const float x2 = fmax_f(-1.e15f, snls3 * x);
  const float cons = snlc1 * (1.0f + snls1 / x * (1.0f - snls2 * x) * expf(x2));

  /* 2. auxiliary spectrum */
  for (int r = 1; r <= nk; ++r) {
    const float conx = TPIINV / sig[r - 1] * cg[r - 1];
    for (int t = 1; t <= nth; ++t) {
      const int isp = t + (r - 1) * nth;
      F(ue, isp) = a[isp - 1] / conx;
      con[isp - 1] = conx;
    }
  }
  for (int isp = 1 - nth; isp <= 0; ++isp) {
    F(ue, isp) = 0.0f;
    F(sa1, isp) = 0.0f;
    F(sa2, isp) = 0.0f;
    F(da1c, isp) = 0.0f;
    F(da1p, isp) = 0.0f;
    F(da1m, isp) = 0.0f;
    F(da2c, isp) = 0.0f;
    F(da2p, isp) = 0.0f;
    F(da2m, isp) = 0.0f;
  }
  for (int isp = nspec + 1; isp <= nfrhgh * nth; ++isp) F(ue, isp) = F(ue, isp - nth) * fachfe;

  /* 3. interactions over the extended spectrum */
  const int* ip11 = ipim + 0 * nspecx;
  const int* ip12 = ipim + 1 * nspecx;
  const int* ip13 = ipim + 2 * nspecx;
  const int* ip14 = ipim + 3 * nspecx;
  const int* im11 = ipim + 4 * nspecx;
  const int* im12 = ipim + 5 * nspecx;
  const int* im13 = ipim + 6 * nspecx;
  const int* im14 = ipim + 7 * nspecx;
  const int* ip21 = ipim + 8 * nspecx;
  const int* ip22 = ipim + 9 * nspecx;
  const int* ip23 = ipim + 10 * nspecx;
  const int* ip24 = ipim + 11 * nspecx;
  const int* im21 = ipim + 12 * nspecx;
  const int* im22 = ipim + 13 * nspecx;
  const int* im23 = ipim + 14 * nspecx;
  const int* im24 = ipim + 15 * nspecx;
  for (int isp = 1; isp <= nspecx; ++isp) {
    const int i = isp - 1;
    const float e00 = F(ue, isp);
    const float ep1 = awg[0] * F(ue, ip11[i]) + awg[1] * F(ue, ip12[i]) + awg[2] * F(ue, ip13[i]) +
                      awg[3] * F(ue, ip14[i]);
    const float em1 = awg[4] * F(ue, im11[i]) + awg[5] * F(ue, im12[i]) + awg[6] * F(ue, im13[i]) +
                      awg[7] * F(ue, im14[i]);
    const float ep2 = awg[0] * F(ue, ip21[i]) + awg[1] * F(ue, ip22[i]) + awg[2] * F(ue, ip23[i]) +
                      awg[3] * F(ue, ip24[i]);
    const float em2 = awg[4] * F(ue, im21[i]) + awg[5] * F(ue, im22[i]) + awg[6] * F(ue, im23[i]) +
                      awg[7] * F(ue, im24[i]);
    const float factor = cons * af11[i] * e00;
    const float sa1a = e00 * (ep1 * dal1 + em1 * dal2);
    const float sa1b = sa1a - ep1 * em1 * dal3;
    const float sa2a = e00 * (ep2 * dal1 + em2 * dal2);
    const float sa2b = sa2a - ep2 * em2 * dal3;
    F(sa1, isp) = factor * sa1b;
    F(sa2, isp) = factor * sa2b;
    F(da1c, isp) = cons * af11[i] * (sa1a + sa1b);
    F(da1p, isp) = factor * (dal1 * e00 - dal3 * em1);
    F(da1m, isp) = factor * (dal2 * e00 - dal3 * ep1);
    F(da2c, isp) = cons * af11[i] * (sa2a + sa2b);
    F(da2p, isp) = factor * (dal1 * e00 - dal3 * em2);
    F(da2m, isp) = factor * (dal2 * e00 - dal3 * ep2);
  }

  /* 4. source and diagonal terms */
  const int* c1[8];
  const int* c2[8];
  for (int j = 0; j < 8; ++j) {
    c1[j] = ic + j * nspec;
    c2[j] = ic + (8 + j) * nspec;
  }
  for (int isp = 1; isp <= nspec; ++isp) {
    const int i = isp - 1;
    s[i] = con[i] * (-2.0f * (F(sa1, isp) + F(sa2, isp)) +
                     awg[0] * (F(sa1, c1[0][i]) + F(sa2, c2[0][i])) +
                     awg[1] * (F(sa1, c1[1][i]) + F(sa2, c2[1][i])) +
                     awg[2] * (F(sa1, c1[2][i]) + F(sa2, c2[2][i])) +
                     awg[3] * (F(sa1, c1[3][i]) + F(sa2, c2[3][i])) +
                     awg[4] * (F(sa1, c1[4][i]) + F(sa2, c2[4][i])) +
                     awg[5] * (F(sa1, c1[5][i]) + F(sa2, c2[5][i])) +
                     awg[6] * (F(sa1, c1[6][i]) + F(sa2, c2[6][i])) +
                     awg[7] * (F(sa1, c1[7][i]) + F(sa2, c2[7][i])));
    d[i] = -2.0f * (F(da1c, isp) + F(da2c, isp)) +
           swg[0] * (F(da1p, c1[0][i]) + F(da2p, c2[0][i])) +
           swg[1] * (F(da1p, c1[1][i]) + F(da2p, c2[1][i])) +
           swg[2] * (F(da1p, c1[2][i]) + F(da2p, c2[2][i])) +
           swg[3] * (F(da1p, c1[3][i]) + F(da2p, c2[3][i])) +
           swg[4] * (F(da1m, c1[4][i]) + F(da2m, c2[4][i])) +
           swg[5] * (F(da1m, c1[5][i]) + F(da2m, c2[5][i])) +
           swg[6] * (F(da1m, c1[6][i]) + F(da2m, c2[6][i])) +
           swg[7] * (F(da1m, c1[7][i]) + F(da2m, c2[7][i]));
  }
#undef F
}

/* Floats of per-point working storage w3snl1_point() needs. */
static size_t work_floats(int nth, const int* dims) {
  return (size_t)(dims[4] + nth) + 8u * (size_t)(dims[3] + nth) + (size_t)dims[0] * (size_t)nth;
}

size_t or_snl1_work_floats(int nk, int nth, float xfr, float lam) {
  int dims[9];
  or_snl1_dims(nk, nth, xfr, lam, dims);
  return work_floats(nth, dims);
}

/* W3SNL1 for one point (the call W3SRCE makes), on Fortran-ordered tables. */
void or_w3snl1_one(int nk, int nth, float xfr, float lam, float snlc1, float kdcon, float kdmn,
                   float snls1, float snls2, float snls3, float fachfe, const float* scal,
                   const int* ipim, const int* ic, const float* af11, const float* sig,
                   const float* a, const float* cg, float kdmean, float* s, float* d, float* work) {
  int dims[9];
  or_snl1_dims(nk, nth, xfr, lam, dims);
  w3snl1_point(nk, nth, dims, snlc1, kdcon, kdmn, snls1, snls2, snls3, fachfe, scal, ipim, ic,
               af11, sig, a, cg, kdmean, s, d, work);
}

/*
 * W3SNL1 over `npts` points, column-major (NSPEC, npts) a/s/d and (NK, npts) cg, with the
 * tables passed in (the fixture's or or_insnl1()'s).  `nthreads` > 1 splits the points
 * over OpenMP threads; each point is computed by one thread, so the bits do not depend
 * on the thread count.  Returns 0, or -1 on a bad argument / allocation failure.
 */
int or_w3snl1_batch(int nk, int nth, float xfr, float lam, float snlc1, float kdcon, float kdmn,
                    float snls1, float snls2, float snls3, float fachfe, const float* scal,
                    const int* ipim, const int* ic, const float* af11, const float* sig, int npts,
                    const float* a, const float* cg, const float* kdmean, float* s, float* d,
                    int nthreads) {
  if (nk <= 0 || nth <= 0 || npts < 0) return -1;
  int dims[9];
  or_snl1_dims(nk, nth, xfr, lam, dims);
  const int nspec = nk * nth;
  const size_t wf = work_floats(nth, dims);
  int err = 0;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(| : err)
#endif
  {
    float* work = (float*)malloc(sizeof(float) * wf);
    if (work == NULL) {
      err = 1;
    } else {
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
      for (int p = 0; p < npts; ++p)
        w3snl1_point(nk, nth, dims, snlc1, kdcon, kdmn, snls1, snls2, snls3, fachfe, scal, ipim,
                     ic, af11, sig, a + (size_t)p * nspec, cg + (size_t)p * nk, kdmean[p],
                     s + (size_t)p * nspec, d + (size_t)p * nspec, work);
      free(work);
    }
  }
  (void)nthreads;
  return err ? -1 : 0;
}

/* Host glibc expf, exposed so tests can compare the device expf with it. */
void or_expf_many(const float* x, float* y, long n) {
  for (long i = 0; i < n; ++i) y[i] = expf(x[i]);
}

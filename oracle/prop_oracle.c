/*
 * oracle/prop_oracle.c -- CPU oracle for the PR3/UQ propagation: W3XYP3 with W3QCK3 (spatial)
 * and W3KTP3 with W3QCK1/W3QCK2 (spectral), and the batch entries that replace W3WAVE's loops.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the GPU library; never linked into it).
 *
 * A scalar restatement of the reference's translation of WW3, every loop in the Fortran's
 * order on one thread, float32 throughout, built with -ffp-contract=off.  Each routine cites
 * the reference code it restates:
 *   W3QCK3         kokkos/src/ww_kokkos/prop_qck3.hpp:58-214
 *   W3XYP3         kokkos/src/ww_kokkos/prop_xyp3.hpp:156-584 (corners :108-131,
 *                  MINVAL/MAXVAL :64-88)
 *   W3QCK1         kokkos/src/ww_kokkos/prop_qck1.hpp:54-177
 *   W3QCK2         kokkos/src/ww_kokkos/prop_qck2.hpp:60-193
 *   W3KTP3         kokkos/src/ww_kokkos/prop_ktp3.hpp:143-353
 *   ww_xyp3_batch  kokkos/src/fortran_iface/prop3_shim.cpp:470-660 (W3GATH, W3XYP3, W3SCAT
 *                  per component with on /= 0)
 *   ww_ktp3_batch  kokkos/src/fortran_iface/prop3_shim.cpp:673-750 (NTIMES calls per point)
 * The argument lists are those of the C ABI (include/ww3gpu_c.h), so a test passes the same
 * arrays to the oracle and to the library.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define GRAV 9.806f
static const float PI_F = (float)3.141592653589793;

static inline float mx(float a, float b) { return a > b ? a : b; }
static inline float mn(float a, float b) { return a < b ? a : b; }
/* INT() as cvttss2si: truncation, INT_MIN for NaN or out of range. */
static inline int fint(float x) {
  return (x > -2147483904.f && x < 2147483648.f) ? (int)x : (int)0x80000000u;
}
/* NINT(): round half away from zero (lroundf), 0 for NaN or out of range. */
static inline long long nintl(float x) {
  float r = roundf(x);
  return (r > -9.2233715e18f && r < 9.2233715e18f) ? (long long)r : 0;
}
/* SIGN(A, B) on integers: |A| with the sign of B, B = 0 positive. */
static inline int isign(int a, int b) { return ((a ^ b) < 0) ? -a : a; }

/* CONSTANTS as folded to float: 0.575*GRAV, DERA, RADIUS, DERA*RADIUS. */
float or_prop_const(int which) {
  const float tpi = 2.0f * PI_F;
  const float tpiinv = 1.0f / tpi;
  const float dera = PI_F / 180.f;
  const float radius = 4.E7f * tpiinv;
  switch (which) {
    case 0: return 0.575f * GRAV;
    case 1: return dera;
    case 2: return radius;
    default: return dera * radius;
  }
}

/* ------------------------------------------------------------------------------------------ */
/* W3QCK3 (prop_qck3.hpp:58-214).  q: Q(1-MY:MY*(MX+2)) at q[i+MY-1]; cfll: CFLL(1:MY*(MX+1));  */
/* trans: TRANS(MY*MX,-1:1); maps 1-based IXY; fla: FLA(1-MY:MY*MX) work.                       */
/* ------------------------------------------------------------------------------------------ */
static void qck3(int mx_, int my, int nx, int ny, float* cfll, const float* trans, float* q,
                 int close, int inc, const int* mapact, int nact, const int* mapbou, int nb0,
                 int nb1, int nb2, float* fla) {
  const int qoff = my - 1;
  const long mymx = (long)my * mx_;
  const int nfla = my * (mx_ + 1);
  for (int i = 0; i < nfla; ++i) fla[i] = 0.f;
  const int iad00 = -my, iad02 = my, iadn0 = iad00 + my * nx, iadn1 = my * nx,
            iadn2 = iad02 + my * nx;
  if (close) {
    for (int iy = 1; iy <= ny; ++iy) {
      q[iy + iad00 + qoff] = q[iy + iadn0 + qoff];
      q[iy + iadn1 + qoff] = q[iy + qoff];
      q[iy + iadn2 + qoff] = q[iy + iad02 + qoff];
      cfll[iy + iadn1 - 1] = cfll[iy - 1];
    }
  }
  for (int ip = 1; ip <= nb0; ++ip) {
    const int ixy = mapbou[ip - 1];
    const float cfl = (cfll[ixy + inc - 1] + cfll[ixy - 1]) * 0.5f;
    const float s = copysignf(1.1f, cfl);
    const int ixyc = ixy - inc * fint(mn(s, 0.f));
    const float c2 = cfl * cfl;
    const float t1 = (1.f - cfl) * q[ixy + inc + qoff];
    const float t2 = (cfl + 1.f) * q[ixy + qoff];
    const float h1 = (t1 + t2) * 0.5f;
    const float h2 = (1.f - c2) / 6.f;
    const float qixyc = q[ixyc + qoff];
    const float d2 = (q[ixyc - inc + qoff] - qixyc * 2.f) + q[ixyc + inc + qoff];
    const float qb = h1 - h2 * d2;
    const int ixyu = ixyc - inc * fint(s);
    const int ixyd = 2 * ixyc - ixyu;
    const float qixyu = q[ixyu + qoff];
    const float dq = q[ixyd + qoff] - qixyu;
    const float dqnz = copysignf(mx(fabsf(dq), 1.E-15f), dq);
    const float qcn = mn(mx((qixyc - qixyu) / dqnz, -0.1f), 1.1f);
    const float qbn0 = mx((qb - qixyu) / dqnz, qcn);
    const float u = mn(qcn / mx(fabsf(cfl), 1.E-10f), 1.f);
    const float qbn = mn(qbn0, u);
    const float qbr = qixyu + qbn * dq;
    const float cfac = (float)fint(fabsf(qcn - 0.5f) * 2.f);
    const float qb2 = qixyc * cfac + (1.f - cfac) * qbr;
    fla[ixy + qoff] = cfl * qb2;
  }
  for (int ip = nb0 + 1; ip <= nb1; ++ip) {
    const int ixy = mapbou[ip - 1];
    const float cfl = cfll[ixy - 1];
    const int ixyc = ixy - inc * fint(mn(copysignf(1.1f, cfl), 0.f));
    fla[ixy + qoff] = q[ixyc + qoff] * cfl;
  }
  for (int ip = nb1 + 1; ip <= nb2; ++ip) {
    const int ixy = mapbou[ip - 1];
    const float cfl = cfll[ixy + inc - 1];
    const int ixyc = ixy - inc * fint(mn(copysignf(1.1f, cfl), 0.f));
    fla[ixy + qoff] = q[ixyc + qoff] * cfl;
  }
  if (close)
    for (int iy = 1; iy <= ny; ++iy) fla[iy + iad00 + qoff] = fla[iy + iadn0 + qoff];
  for (int ip = 1; ip <= nact; ++ip) {
    const int ixy = mapact[ip - 1];
    const float fm = fla[ixy - inc + qoff];
    const float fp = fla[ixy + qoff];
    const int jn = (fm > 0.f) ? -1 : 0;
    const int jp = (fp < 0.f) ? 1 : 0;
    const float t = q[ixy + qoff] + fm * trans[(long)(jn + 1) * mymx + ixy - 1];
    const float u = t - fp * trans[(long)(jp + 1) * mymx + ixy - 1];
    q[ixy + qoff] = mx(u, 0.f);
  }
}

/* ------------------------------------------------------------------------------------------ */
/* W3XYP3 for component isp (prop_xyp3.hpp:156-584).                                          */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
  int nx, ny, nsea, nk, nth, nspec, iclose, flcx, flcy, flagll, flcur;
  float dtcfl, dth, xfr, wdcg, wdth;
  int nmx0, nmx1, nmx2, nmy0, nmy1, nmy2, nact, ncent;
  const int *mapsf, *mapx2, *mapy2, *mapaxy, *mapcxy, *maptrn, *mapsta, *isbpi;
  const float *clats, *sig, *ecos, *esin, *dpdx, *dpdy, *dqdx, *dqdy, *gsqrt, *atrnx, *atrny;
  const float *cg, *cx, *cy, *bbpi0, *bbpin;
  float dsec_day, dsec_bpi0, dsec_bpin;
  int flbpi, nbi;
} Xyp3;

/* gfortran's in-order MINVAL/MAXVAL (prop_xyp3.hpp:64-88) */
static float minval_io(const float* a, int n) {
  if (n <= 0) return 3.40282347e+38f;
  int i = 0;
  while (i < n && !(a[i] <= INFINITY)) ++i;
  if (i == n) return NAN;
  float lim = a[i];
  for (++i; i < n; ++i) lim = (a[i] < lim) ? a[i] : lim;
  return lim;
}
static float maxval_io(const float* a, int n) {
  if (n <= 0) return -3.40282347e+38f;
  int i = 0;
  while (i < n && !(a[i] >= -INFINITY)) ++i;
  if (i == n) return NAN;
  float lim = a[i];
  for (++i; i < n; ++i) lim = (a[i] > lim) ? a[i] : lim;
  return lim;
}

/* one corner of the averaging stencil (prop_xyp3.hpp:108-131) */
static void corner(float dx, float dy, int ny, int* iixa, int* iixb, float* rda, float* rdb) {
  const int ixc = (dx < 0.f) ? -ny : ny;
  const int iyc = (dy < 0.f) ? -1 : 1;
  *iixa = ixc + iyc;
  if (fabsf(dx) > fabsf(dy)) {
    *iixb = ixc;
    *rda = fabsf(dy / dx);
    *rdb = fabsf(dx);
  } else {
    *iixb = iyc;
    *rda = (fabsf(dy) > 1.E-5f) ? fabsf(dx / dy) : 1.f;
    *rdb = fabsf(dy);
  }
}

#define G2(x, iy, ix) (x)[((ix) - 1) * ny + (iy) - 1]

/* The number of sub-steps W3XYP3 takes for component isp (prop_xyp3.hpp:200-256). */
static int xyp3_ntloc(const Xyp3* P, int isp, float dtg, float* cg0o) {
  const float cg0fac = or_prop_const(0);
  const int ith = 1 + (isp - 1) % P->nth;
  const int ik = 1 + (isp - 1) / P->nth;
  const float cg0 = cg0fac / P->sig[1];
  const float cga = cg0fac / P->sig[ik];
  const float cgx = P->ecos[ith - 1] * cga;
  const float cgy = P->esin[ith - 1] * cga;
  float acgx, acgy, cxc, cyc;
  if (P->flcur) {
    const float cxmin = minval_io(P->cx + 1, P->nsea), cxmax = maxval_io(P->cx + 1, P->nsea);
    const float cymin = minval_io(P->cy + 1, P->nsea), cymax = maxval_io(P->cy + 1, P->nsea);
    acgx = mx(fabsf(cgx + cxmin), fabsf(cgx + cxmax));
    acgy = mx(fabsf(cgy + cymin), fabsf(cgy + cymax));
    cxc = mx(fabsf(cxmin), fabsf(cxmax));
    cyc = mx(fabsf(cymin), fabsf(cymax));
  } else {
    acgx = fabsf(cgx);
    acgy = fabsf(cgy);
    cxc = 0.f;
    cyc = 0.f;
  }
  const float t1 = mx(acgx, acgy);
  const float t2 = mx(cxc, cyc);
  const float t3 = mx(cg0 * 0.001f, t2);
  const float cgn = mx(t1, t3);
  if (cg0o) *cg0o = cg0;
  return 1 + fint(dtg / ((P->dtcfl * cg0) / cgn));
}

/* vq: VQ(1-NY:NY*(NX+2)) at vq[i+NY-1], the gathered field; wr/wi: work. */
static void w3xyp3(const Xyp3* P, int isp, float dtg, float* vq, float* wr, int* wi) {
  const int nx = P->nx, ny = P->ny, nsea = P->nsea;
  const int nxy = nx * ny, nxy1 = (nx + 1) * ny, nq = ny * (nx + 3);
  const int qoff = ny - 1, moff = 2 * ny - 1, cgld = P->nk + 2;
  float* rdi1 = wr;
  float* rdi2 = rdi1 + nsea;
  float* rdi3 = rdi2 + nsea;
  float* rdi4 = rdi3 + nsea;
  float* vlcflx = rdi4 + nsea;
  float* vlcfly = vlcflx + nxy1;
  float* cxtot = vlcfly + nxy1;
  float* cytot = cxtot + nxy1;
  float* aq = cytot + nxy;
  float* fla = aq + nq;
  int* iixy1 = wi;
  int* iixy2 = iixy1 + nsea;
  int* iixy3 = iixy2 + nsea;
  int* iixy4 = iixy3 + nsea;
  int* mapstx = iixy4 + nsea;
  const int* mapsf = P->mapsf;

  const int global = P->iclose != -1;
  const int ith = 1 + (isp - 1) % P->nth;
  const int ik = 1 + (isp - 1) / P->nth;
  const float ecosi = P->ecos[ith - 1], esini = P->esin[ith - 1];
  const int ntloc = xyp3_ntloc(P, isp, dtg, NULL);
  const float rntloc = (float)ntloc;
  const float dtloc = dtg / rntloc;
  float dtrad = dtloc;
  if (P->flagll) dtrad = dtloc / or_prop_const(3);
  const int dttst = fint(P->dsec_day);
  int yfirst = (nintl((float)dttst / dtg) % 2) == 0;

  for (int i = 0; i < nxy1; ++i) vlcflx[i] = vlcfly[i] = cxtot[i] = 0.f;
  for (int i = 0; i < nxy; ++i) cytot[i] = 0.f;
  for (int i = 1; i <= nxy; ++i) mapstx[i + moff] = P->mapsta[i - 1];
  for (int j = 0; j < 2 * ny; ++j) {
    mapstx[(1 - 2 * ny + j) + moff] = global ? P->mapsta[(nx - 2) * ny + j] : 0;
    mapstx[(nxy + 1 + j) + moff] = global ? P->mapsta[j] : 0;
  }
  float fth = (P->wdth * P->dth) * dtloc;
  float fcg = ((P->wdcg * 0.5f) * (P->xfr - (1.f / P->xfr))) * dtloc;
  if (P->flagll) {
    fth = (fth / or_prop_const(2)) / or_prop_const(1);
    fcg = (fcg / or_prop_const(2)) / or_prop_const(1);
  }
  fcg = fcg / rntloc;
  const float fthx = -(esini * fth), fthy = ecosi * fth;
  const float fcgx = ecosi * fcg, fcgy = esini * fcg;

  for (int isea = 1; isea <= nsea; ++isea) {
    const int ix = mapsf[isea - 1], iy = mapsf[nsea + isea - 1];
    const float clat = P->clats[isea];
    const float dpdx = G2(P->dpdx, iy, ix), dpdy = G2(P->dpdy, iy, ix);
    const float dqdx = G2(P->dqdx, iy, ix), dqdy = G2(P->dqdy, iy, ix);
    float tmpx = fthx / clat, tmpy = fthy;
    const float dxcgn = dpdx * tmpx + dpdy * tmpy;
    const float dycgn = dqdx * tmpx + dqdy * tmpy;
    tmpx = fcgx / clat;
    tmpy = fcgy;
    const float dxcgs = dpdx * tmpx + dpdy * tmpy;
    const float dycgs = dqdx * tmpx + dqdy * tmpy;
    corner(dxcgn + dxcgs, dycgn + dycgs, ny, &iixy1[isea - 1], &iixy2[isea - 1], &rdi1[isea - 1],
           &rdi2[isea - 1]);
    corner(dxcgn - dxcgs, dycgn - dycgs, ny, &iixy3[isea - 1], &iixy4[isea - 1], &rdi3[isea - 1],
           &rdi4[isea - 1]);
  }
  for (int isea = 1; isea <= nsea; ++isea) {
    const int ixy = mapsf[2 * nsea + isea - 1];
    const float cgi = P->cg[isea * cgld + ik];
    const float clat = P->clats[isea];
    vq[ixy + qoff] = (vq[ixy + qoff] / cgi) * clat;
    cxtot[ixy - 1] = (ecosi * cgi) / clat;
    cytot[ixy - 1] = esini * cgi;
  }
  if (P->flcur) {
    for (int isea = 1; isea <= nsea; ++isea) {
      const int ixy = mapsf[2 * nsea + isea - 1];
      cxtot[ixy - 1] = cxtot[ixy - 1] + (P->cx[isea] / P->clats[isea]);
      cytot[ixy - 1] = cytot[ixy - 1] + P->cy[isea];
    }
  }
  for (int isea = 1; isea <= nsea; ++isea) {
    const int ix = mapsf[isea - 1], iy = mapsf[nsea + isea - 1];
    const int ixy = mapsf[2 * nsea + isea - 1];
    const float cxt = cxtot[ixy - 1], cyt = cytot[ixy - 1];
    const float cp = cxt * G2(P->dpdx, iy, ix) + cyt * G2(P->dpdy, iy, ix);
    const float cq = cxt * G2(P->dqdx, iy, ix) + cyt * G2(P->dqdy, iy, ix);
    vlcflx[ixy - 1] = dtrad * cp;
    vlcfly[ixy - 1] = dtrad * cq;
  }

  for (int itloc = 1; itloc <= ntloc; ++itloc) {
    for (int i = 0; i < nq; ++i) aq[i] = vq[i];
    for (int i = 0; i < nq; ++i) vq[i] = 0.f;
    /* central points: centre, +1, +2, +3, +4, -1, -2, -3, -4 */
    for (int ip = 1; ip <= P->ncent; ++ip) {
      const int isea = P->mapcxy[ip - 1];
      const int ixy = mapsf[2 * nsea + isea - 1];
      if (P->maptrn[ixy - 1] != 0) {
        vq[ixy + qoff] = aq[ixy + qoff];
        continue;
      }
      const float cgi = P->cg[isea * cgld + ik];
      const float rd1 = rdi1[isea - 1], rd2 = mn(rdi2[isea - 1] * cgi, 1.f);
      const float rd3 = rdi3[isea - 1], rd4 = mn(cgi * rdi4[isea - 1], 1.f);
      const float a = aq[ixy + qoff];
      const int i1 = iixy1[isea - 1], i2 = iixy2[isea - 1], i3 = iixy3[isea - 1],
                i4 = iixy4[isea - 1];
      vq[ixy + qoff] = vq[ixy + qoff] + (((3.f - rd2) - rd4) * a) / 3.f;
      const float c1 = ((rd2 * a) * rd1) / 6.f;
      const float c2 = (((1.f - rd1) * a) * rd2) / 6.f;
      const float c3 = ((rd4 * a) * rd3) / 6.f;
      const float c4 = (((1.f - rd3) * a) * rd4) / 6.f;
      vq[ixy + i1 + qoff] += c1;
      vq[ixy + i2 + qoff] += c2;
      vq[ixy + i3 + qoff] += c3;
      vq[ixy + i4 + qoff] += c4;
      vq[ixy - i1 + qoff] += c1;
      vq[ixy - i2 + qoff] += c2;
      vq[ixy - i3 + qoff] += c3;
      vq[ixy - i4 + qoff] += c4;
    }
    /* near-coast points: centre, +1, -1, +2, -2, +3, -3, +4, -4, redirected to the centre */
    for (int ip = P->ncent + 1; ip <= nsea; ++ip) {
      const int isea = P->mapcxy[ip - 1];
      const int ixy = mapsf[2 * nsea + isea - 1];
      if (P->mapsta[ixy - 1] <= 0) continue;
      if (P->maptrn[ixy - 1] != 0) {
        vq[ixy + qoff] = aq[ixy + qoff];
        continue;
      }
      const float cgi = P->cg[isea * cgld + ik];
      const float rd1 = rdi1[isea - 1], rd3 = rdi3[isea - 1];
      const float rd2 = mn(rdi2[isea - 1] * cgi, 1.f), rd4 = mn(cgi * rdi4[isea - 1], 1.f);
      const float a = aq[ixy + qoff];
      const int i1 = iixy1[isea - 1], i2 = iixy2[isea - 1], i3 = iixy3[isea - 1],
                i4 = iixy4[isea - 1];
      vq[ixy + qoff] = vq[ixy + qoff] + (((3.f - rd2) - rd4) * a) / 3.f;
      const float c1 = ((rd2 * a) * rd1) / 6.f;
      const float c2 = (((1.f - rd1) * a) * rd2) / 6.f;
      const float c3 = ((rd4 * a) * rd3) / 6.f;
      const float c4 = (((1.f - rd3) * a) * rd4) / 6.f;
#define ST(o) mapstx[(o) + moff]
      int ixc = isign(ny, i1), iyc = i1 - ixc, t;
      t = (ST(ixy + i1) >= 1 && !(ST(ixy + ixc) <= 0 && ST(ixy + iyc) <= 0)) ? ixy + i1 : ixy;
      vq[t + qoff] += c1;
      t = (ST(ixy - i1) >= 1 && !(ST(ixy - ixc) <= 0 && ST(ixy - iyc) <= 0)) ? ixy - i1 : ixy;
      vq[t + qoff] += c1;
      t = ST(ixy + i2) >= 1 ? ixy + i2 : ixy;
      vq[t + qoff] += c2;
      t = ST(ixy - i2) >= 1 ? ixy - i2 : ixy;
      vq[t + qoff] += c2;
      ixc = isign(ny, i3);
      iyc = i3 - ixc;
      t = (ST(ixy + i3) >= 1 && !(ST(ixy + ixc) <= 0 && ST(ixy + iyc) <= 0)) ? ixy + i3 : ixy;
      vq[t + qoff] += c3;
      t = (ST(ixy - i3) >= 1 && !(ST(ixy - ixc) <= 0 && ST(ixy - iyc) <= 0)) ? ixy - i3 : ixy;
      vq[t + qoff] += c3;
      t = ST(ixy + i4) >= 1 ? ixy + i4 : ixy;
      vq[t + qoff] += c4;
      t = ST(ixy - i4) >= 1 ? ixy - i4 : ixy;
      vq[t + qoff] += c4;
#undef ST
    }
    for (int ixy = 1; ixy <= nxy; ++ixy)
      if (P->mapsta[ixy - 1] == 2) vq[ixy + qoff] = aq[ixy + qoff];
    if (global) {
      for (int iy = 1; iy <= ny; ++iy) {
        vq[iy + qoff] = vq[iy + qoff] + vq[nxy + iy + qoff];
        vq[(nx - 1) * ny + iy + qoff] = vq[(nx - 1) * ny + iy + qoff] + vq[iy - ny + qoff];
      }
    }
    for (int isea = 1; isea <= nsea; ++isea) {
      const int ix = mapsf[isea - 1], iy = mapsf[nsea + isea - 1];
      const int ixy = mapsf[2 * nsea + isea - 1];
      vq[ixy + qoff] = vq[ixy + qoff] * G2(P->gsqrt, iy, ix);
    }
    for (int pass = 0; pass < 2; ++pass) {
      const int ydir = (pass == 0) == (yfirst != 0);
      if (ydir && P->flcy)
        qck3(nx, ny, nx, ny, vlcfly, P->atrny, vq, 0, 1, P->mapaxy, P->nact, P->mapy2, P->nmy0,
             P->nmy1, P->nmy2, fla);
      if (!ydir && P->flcx)
        qck3(nx, ny, nx, ny, vlcflx, P->atrnx, vq, global, ny, P->mapaxy, P->nact, P->mapx2,
             P->nmx0, P->nmx1, P->nmx2, fla);
    }
    for (int isea = 1; isea <= nsea; ++isea) {
      const int ix = mapsf[isea - 1], iy = mapsf[nsea + isea - 1];
      const int ixy = mapsf[2 * nsea + isea - 1];
      vq[ixy + qoff] = vq[ixy + qoff] / G2(P->gsqrt, iy, ix);
    }
    if (P->flbpi) {
      float rd1 = P->dsec_bpi0 - ((dtg * (float)(ntloc - itloc)) / rntloc);
      float rd2 = P->dsec_bpin;
      if (rd2 > 0.001f) {
        rd2 = mn(mx(rd1 / rd2, 0.f), 1.f);
        rd1 = 1.f - rd2;
      } else {
        rd1 = 0.f;
        rd2 = 1.f;
      }
      for (int ibi = 1; ibi <= P->nbi; ++ibi) {
        const int ib = P->isbpi[ibi - 1];
        const int ixy = mapsf[2 * nsea + ib - 1];
        const float b0 = P->bbpi0[(long)ibi * P->nspec + isp - 1];
        const float bn = P->bbpin[(long)ibi * P->nspec + isp - 1];
        vq[ixy + qoff] = (((b0 * rd1) + (bn * rd2)) / P->cg[ib * cgld + ik]) * P->clats[ib];
      }
    }
    yfirst = !yfirst;
  }
  for (int isea = 1; isea <= nsea; ++isea) {
    const int ixy = mapsf[2 * nsea + isea - 1];
    if (P->mapsta[ixy - 1] > 0) {
      const float x = (P->cg[isea * cgld + ik] / P->clats[isea]) * vq[ixy + qoff];
      vq[ixy + qoff] = mx(x, 0.f);
    }
  }
}
#undef G2

/* ww_xyp3_batch's arguments (include/ww3gpu_c.h); returns 0. Components run on nthreads. */
int or_xyp3_batch(const float* rpar, const int* ipar, const int* imaps, const float* tscal,
                  float dtg, const int* mapsf, const float* clats, const float* sig,
                  const float* ecos, const float* esin, const float* dpdx, const float* dpdy,
                  const float* dqdx, const float* dqdy, const float* gsqrt, const float* atrnx,
                  const float* atrny, const int* mapx2, const int* mapy2, const int* mapaxy,
                  const int* mapcxy, const int* maptrn, const int* mapsta, const float* cg,
                  const float* cx, const float* cy, int flbpi, int nbi, const int* isbpi,
                  const float* bbpi0, const float* bbpin, const int* on, int ncol, float* va,
                  int nthreads) {
  Xyp3 P;
  memset(&P, 0, sizeof P);
  P.dtcfl = rpar[0];
  P.dth = rpar[1];
  P.xfr = rpar[2];
  P.wdcg = rpar[3];
  P.wdth = rpar[4];
  P.nx = ipar[0];
  P.ny = ipar[1];
  P.nsea = ipar[2];
  P.nk = ipar[3];
  P.nth = ipar[4];
  P.nspec = ipar[5];
  P.iclose = ipar[6];
  P.flcx = ipar[7];
  P.flcy = ipar[8];
  P.flagll = ipar[9];
  P.flcur = ipar[10];
  P.nmx0 = imaps[0];
  P.nmx1 = imaps[1];
  P.nmx2 = imaps[2];
  P.nmy0 = imaps[3];
  P.nmy1 = imaps[4];
  P.nmy2 = imaps[5];
  P.nact = imaps[6];
  P.ncent = imaps[7];
  P.mapsf = mapsf;
  P.mapx2 = mapx2;
  P.mapy2 = mapy2;
  P.mapaxy = mapaxy;
  P.mapcxy = mapcxy;
  P.maptrn = maptrn;
  P.mapsta = mapsta;
  P.isbpi = isbpi;
  P.clats = clats;
  P.sig = sig;
  P.ecos = ecos;
  P.esin = esin;
  P.dpdx = dpdx;
  P.dpdy = dpdy;
  P.dqdx = dqdx;
  P.dqdy = dqdy;
  P.gsqrt = gsqrt;
  P.atrnx = atrnx;
  P.atrny = atrny;
  P.cg = cg;
  P.cx = cx;
  P.cy = cy;
  P.bbpi0 = bbpi0;
  P.bbpin = bbpin;
  P.dsec_day = tscal[0];
  P.flbpi = flbpi != 0;
  P.nbi = flbpi ? nbi : 0;
  P.dsec_bpi0 = flbpi ? tscal[1] : 0.f;
  P.dsec_bpin = flbpi ? tscal[2] : 0.f;
  (void)ncol;
  const int nx = P.nx, ny = P.ny, nsea = P.nsea, nspec = P.nspec;
  const long nf = (long)ny * (nx + 3);
  const long nwr = 4L * nsea + 3L * (nx + 1) * ny + (long)nx * ny + (long)ny * (nx + 3) +
                   (long)ny * (nx + 1);
  const long nwi = 4L * nsea + (long)ny * (nx + 4);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    float* field = (float*)malloc(sizeof(float) * (nf + nwr));
    int* wi = (int*)malloc(sizeof(int) * nwi);
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
    for (int isp = 1; isp <= nspec; ++isp) {
      if (on[isp - 1] == 0) continue;
      for (long k = 0; k < nf; ++k) field[k] = 0.f;
      for (int isea = 1; isea <= nsea; ++isea)
        field[mapsf[2 * nsea + isea - 1] + ny - 1] = va[(long)isea * nspec + isp - 1];
      w3xyp3(&P, isp, dtg, field, field + nf, wi);
      for (int isea = 1; isea <= nsea; ++isea) {
        const int ixy = mapsf[2 * nsea + isea - 1];
        if (mapsta[ixy - 1] != 0) va[(long)isea * nspec + isp - 1] = field[ixy + ny - 1];
      }
    }
    free(field);
    free(wi);
  }
  return 0;
}

/* NTLOC of every component (the sub-step counts), for the tests. */
void or_xyp3_ntloc(const float* rpar, const int* ipar, float dtg, const float* sig,
                   const float* ecos, const float* esin, const float* cx, const float* cy,
                   int* out) {
  Xyp3 P;
  memset(&P, 0, sizeof P);
  P.dtcfl = rpar[0];
  P.nsea = ipar[2];
  P.nk = ipar[3];
  P.nth = ipar[4];
  P.nspec = ipar[5];
  P.flcur = ipar[10];
  P.sig = sig;
  P.ecos = ecos;
  P.esin = esin;
  P.cx = cx;
  P.cy = cy;
  for (int isp = 1; isp <= P.nspec; ++isp) out[isp - 1] = xyp3_ntloc(&P, isp, dtg, NULL);
}

/* ------------------------------------------------------------------------------------------ */
/* W3QCK1 (prop_qck1.hpp:54-177) and W3QCK2 (prop_qck2.hpp:60-193).                           */
/* ------------------------------------------------------------------------------------------ */
static void qck1(int mx_, int my, int nx, int ny, float* cfll, float* q, int close, int inc,
                 const int* mapact, int nact, const int* mapbou, int nb0, int nb1, int nb2,
                 float* fla) {
  const int qoff = my - 1;
  const int nfla = my * (mx_ + 1);
  for (int i = 0; i < nfla; ++i) fla[i] = 0.f;
  const int iad00 = -my, iad02 = my, iadn0 = iad00 + my * nx, iadn1 = my * nx,
            iadn2 = iad02 + my * nx;
  if (close) {
    for (int iy = 1; iy <= ny; ++iy) {
      q[iy + iad00 + qoff] = q[iy + iadn0 + qoff];
      q[iy + iadn1 + qoff] = q[iy + qoff];
      q[iy + iadn2 + qoff] = q[iy + iad02 + qoff];
      cfll[iy + iadn1 - 1] = cfll[iy - 1];
    }
  }
  for (int ip = 1; ip <= nb0; ++ip) {
    const int ixy = mapbou[ip - 1];
    const float cfl = (cfll[ixy + inc - 1] + cfll[ixy - 1]) * 0.5f;
    const float s = copysignf(1.1f, cfl);
    const int ixyc = ixy - inc * fint(mn(s, 0.f));
    const float c2 = cfl * cfl;
    const float t2 = (1.f - cfl) * q[ixy + inc + qoff];
    const float t4 = (cfl + 1.f) * q[ixy + qoff];
    const float t6 = (t4 + t2) * 0.5f;
    const float t8 = (1.f - c2) / 6.f;
    const float qixyc = q[ixyc + qoff];
    const float t11 = q[ixyc + inc + qoff] + (q[ixyc - inc + qoff] - qixyc * 2.f);
    const float qb = t6 - t11 * t8;
    const int ixyu = ixyc - inc * fint(s);
    const int ixyd = 2 * ixyc - ixyu;
    const float qixyu = q[ixyu + qoff];
    const float dq = q[ixyd + qoff] - qixyu;
    const float dqnz = copysignf(mx(fabsf(dq), 1.E-15f), dq);
    const float qcn = mn(mx((qixyc - qixyu) / dqnz, -0.1f), 1.1f);
    const float qbn0 = mx((qb - qixyu) / dqnz, qcn);
    const float t13 = mn(qcn / mx(fabsf(cfl), 1.E-10f), 1.f);
    const float qbn = mn(qbn0, t13);
    const float qbr = qbn * dq + qixyu;
    const float cfac = (float)fint(fabsf(qcn - 0.5f) * 2.f);
    const float qb2 = cfac * qixyc + (1.f - cfac) * qbr;
    fla[ixy + qoff] = qb2 * cfl;
  }
  for (int ip = nb0 + 1; ip <= nb1; ++ip) {
    const int ixy = mapbou[ip - 1];
    const float cfl = cfll[ixy - 1];
    const int ixyc = ixy - inc * fint(mn(copysignf(1.1f, cfl), 0.f));
    fla[ixy + qoff] = q[ixyc + qoff] * cfl;
  }
  for (int ip = nb1 + 1; ip <= nb2; ++ip) {
    const int ixy = mapbou[ip - 1];
    const float cfl = cfll[ixy + inc - 1];
    const int ixyc = ixy - inc * fint(mn(copysignf(1.1f, cfl), 0.f));
    fla[ixy + qoff] = q[ixyc + qoff] * cfl;
  }
  if (close)
    for (int iy = 1; iy <= ny; ++iy) fla[iy + iad00 + qoff] = fla[iy + iadn0 + qoff];
  for (int ip = 1; ip <= nact; ++ip) {
    const int ixy = mapact[ip - 1];
    const float u = (q[ixy + qoff] + fla[ixy - inc + qoff]) - fla[ixy + qoff];
    q[ixy + qoff] = mx(u, 0.f);
  }
}

static void qck2(int mx_, int my, int nx, int ny, float* velo, float dt, float* dx1, float* dx2,
                 float* q, int close, int inc, const int* mapact, int nact, const int* mapbou,
                 int nb0, int nb1, int nb2, float* fla) {
  const int qoff = my - 1;
  const int nfla = my * (mx_ + 1);
  for (int i = 0; i < nfla; ++i) fla[i] = 0.f;
  const int iad00 = -my, iad02 = my, iadn0 = iad00 + my * nx, iadn1 = my * nx,
            iadn2 = iad02 + my * nx;
  if (close) {
    for (int iy = 1; iy <= ny; ++iy) {
      q[iy + iad00 + qoff] = q[iy + iadn0 + qoff];
      q[iy + iadn1 + qoff] = q[iy + qoff];
      q[iy + iadn2 + qoff] = q[iy + iad02 + qoff];
      velo[iy + iadn1 - 1] = velo[iy - 1];
      dx1[iy + iadn1 - 1] = dx1[iy - 1];
      dx2[iy + iad00 + qoff] = dx1[iy + iadn0 - 1];
      dx2[iy + iadn1 + qoff] = dx1[iy - 1];
    }
  }
  for (int ip = 1; ip <= nb0; ++ip) {
    const int ixy = mapbou[ip - 1];
    const float vel = (velo[ixy + inc - 1] + velo[ixy - 1]) * 0.5f;
    const float dx2ixy = dx2[ixy + qoff];
    const float cfl = (dt * vel) / dx2ixy;
    const float s = copysignf(1.1f, cfl);
    const int ixyc = ixy - inc * fint(mn(s, 0.f));
    const float d2 = dx2ixy * dx2ixy;
    const float c2 = cfl * cfl;
    const float t2 = q[ixy + inc + qoff] * (1.f - cfl);
    const float t4 = q[ixy + qoff] * (cfl + 1.f);
    const float t6 = (t4 + t2) * 0.5f;
    const float t8 = ((1.f - c2) * (d2 / dx1[ixyc - 1])) / 6.f;
    const float qixyc = q[ixyc + qoff];
    const float u1 = (q[ixyc + inc + qoff] - qixyc) / dx2[ixyc + qoff];
    const float u2 = (qixyc - q[ixyc - inc + qoff]) / dx2[ixyc - inc + qoff];
    const float qb = t6 - (u1 - u2) * t8;
    const int ixyu = ixyc - inc * fint(s);
    const int ixyd = 2 * ixyc - ixyu;
    const float qixyu = q[ixyu + qoff];
    const float dq = q[ixyd + qoff] - qixyu;
    const float dqnz = copysignf(mx(fabsf(dq), 1.E-15f), dq);
    const float qcn = mn(mx((qixyc - qixyu) / dqnz, -0.1f), 1.1f);
    const float qbn0 = mx((qb - qixyu) / dqnz, qcn);
    const float t13 = mn(qcn / mx(fabsf(cfl), 1.E-10f), 1.f);
    const float qbn = mn(qbn0, t13);
    const float qbr = qbn * dq + qixyu;
    const float cfac = (float)fint(fabsf(qcn - 0.5f) * 2.f);
    const float qb2 = cfac * qixyc + (1.f - cfac) * qbr;
    fla[ixy + qoff] = qb2 * vel;
  }
  for (int ip = nb0 + 1; ip <= nb1; ++ip) {
    const int ixy = mapbou[ip - 1];
    const float vel = velo[ixy - 1];
    const int ixyc = ixy - inc * fint(mn(copysignf(1.1f, vel), 0.f));
    fla[ixy + qoff] = q[ixyc + qoff] * vel;
  }
  for (int ip = nb1 + 1; ip <= nb2; ++ip) {
    const int ixy = mapbou[ip - 1];
    const float vel = velo[ixy + inc - 1];
    const int ixyc = ixy - inc * fint(mn(copysignf(1.1f, vel), 0.f));
    fla[ixy + qoff] = q[ixyc + qoff] * vel;
  }
  if (close)
    for (int iy = 1; iy <= ny; ++iy) fla[iy + iad00 + qoff] = fla[iy + iadn0 + qoff];
  for (int ip = 1; ip <= nact; ++ip) {
    const int ixy = mapact[ip - 1];
    const float f = dt / dx1[ixy - 1];
    const float u = q[ixy + qoff] + f * (fla[ixy - inc + qoff] - fla[ixy + qoff]);
    q[ixy + qoff] = mx(u, 0.f);
  }
}

/* ------------------------------------------------------------------------------------------ */
/* W3KTP3 for one point (prop_ktp3.hpp:143-353).                                               */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
  int nk, nth, nspec, nk2, flcth, flck, flcur;
  float dmin, ctmax, fachfa;
  const float *sig, *dsip, *ecos, *esin, *ec2, *esc, *es2;
  const int *mapwn, *mapth2, *mapwn2;
} Ktp3;

static long ktp3_work_floats(int nk, int nth) {
  const long nk2 = nk + 2;
  return (nk + 2) + nk + nk + nth + nk2 * (nth + 3) + nk2 * (nth + 1) + nk2 * (nth + 2) +
         3 * nk2 * (nth + 1);
}

static void w3ktp3(const Ktp3* p, int itstep, float facth, float fack, float cthg0,
                   const float* cg, const float* wn, float dw, float dddx, float dddy, float cx,
                   float cy, float dcxdx, float dcxdy, float dcydx, float dcydy, float* va,
                   float* cflthmax, float* cflkmax, float* work) {
  const int nk = p->nk, nth = p->nth, nspec = p->nspec, nk2 = p->nk2;
  const int vqoff = nk2 - 1;
  float* dsdd = work;
  float* frk = dsdd + (nk + 2);
  float* frg = frk + nk;
  float* fkc = frg + nk;
  float* vq = fkc + nth;
  float* db = vq + nk2 * (nth + 3);
  float* dm = db + nk2 * (nth + 1);
  float* vcflt = dm + nk2 * (nth + 2);
  float* cflk = vcflt + nk2 * (nth + 1);
  float* fla = cflk + nk2 * (nth + 1);
  /* DB(:,NTH+1), DM(:,0), DM(:,NTH+1) are never written by the Fortran and never read */
  for (int i = 0; i < nk2 * (nth + 1); ++i) db[i] = 0.f;
  for (int i = 0; i < nk2 * (nth + 2); ++i) dm[i] = 0.f;

  const float depth = mx(p->dmin, dw);
  for (int i = 0; i < nk2 * (nth + 3); ++i) vq[i] = 0.f;
  if (p->flcth)
    for (int i = 0; i < nk2 * (nth + 1); ++i) vcflt[i] = 0.f;
  if (p->flck)
    for (int i = 0; i < nk2 * (nth + 1); ++i) cflk[i] = 0.f;
  float acc_th = 0.f, acc_k = 0.f;
  for (int ik = 0; ik <= nk + 1; ++ik) {
    if (wn[ik] * depth < 5.f) {
      const float t = wn[ik] * cg[ik] - p->sig[ik] * 0.5f;
      dsdd[ik] = mx(t, 0.f) / depth;
    } else {
      dsdd[ik] = 0.f;
    }
  }
  for (int isp = 1; isp <= nspec; ++isp) vq[p->mapth2[isp - 1] + vqoff] = va[isp - 1];
  if (p->flcth) {
    const float fdg = cthg0 * facth;
    for (int ik = 1; ik <= nk; ++ik) {
      frk[ik - 1] = (facth * dsdd[ik]) / wn[ik];
      frg[ik - 1] = fdg * cg[ik];
    }
    if (p->flcur) {
      const float dcyx = dcydx * facth;
      const float dcxxyy = (dcxdx - dcydy) * facth;
      const float dcxy = dcxdy * facth;
      for (int isp = 1; isp <= nspec; ++isp)
        vcflt[p->mapth2[isp - 1] - 1] =
            ((p->es2[isp - 1] * dcyx) + (p->esc[isp - 1] * dcxxyy)) - (p->ec2[isp - 1] * dcxy);
    } else {
      for (int i = 0; i < nk2 * (nth + 1); ++i) vcflt[i] = 0.f;
    }
    for (int isp = 1; isp <= nspec; ++isp) {
      const int m2 = p->mapth2[isp - 1], mw = p->mapwn[isp - 1];
      const float v = (vcflt[m2 - 1] + frg[mw - 1] * p->ecos[isp - 1]) +
                      frk[mw - 1] * ((p->esin[isp - 1] * dddx) - (p->ecos[isp - 1] * dddy));
      const float av = fabsf(v);
      acc_th = mx(acc_th, av);
      vcflt[m2 - 1] = copysignf(mn(av, p->ctmax), v);
    }
  }
  if (p->flck) {
    const float dcxx = -dcxdx, dcxyyx = -(dcxdy + dcydx), dcyy = -dcydy;
    const float fkd = (cx * dddx) + (cy * dddy);
    for (int ith = 1; ith <= nth; ++ith)
      fkc[ith - 1] = ((p->ec2[ith - 1] * dcxx) + (p->esc[ith - 1] * dcxyyx)) +
                     (p->es2[ith - 1] * dcyy);
    for (int ik = 0; ik <= nk; ++ik) {
      db[ik] = p->dsip[ik] / cg[ik];
      dm[ik + nk2] = wn[ik + 1] - wn[ik];
    }
    db[nk + 1] = p->dsip[nk + 1] / cg[nk + 1];
    dm[nk + 1 + nk2] = 0.f;
    for (int ith = 2; ith <= nth; ++ith)
      for (int ik = 1; ik <= nk + 2; ++ik) {
        db[(ik - 1) + (ith - 1) * nk2] = db[ik - 1];
        dm[(ik - 1) + ith * nk2] = dm[(ik - 1) + nk2];
      }
    for (int ik = 0; ik <= nk + 1; ++ik) {
      const float fkd0 = (fkd / cg[ik]) * dsdd[ik];
      const float velfac = fack / db[ik];
#define VEL(ith) (((wn[ik] * fkc[(ith) - 1]) + fkd0) * velfac)
      for (int ith = 1; ith <= nth; ++ith) {
        const float v = VEL(ith);
        cflk[ik + (ith - 1) * nk2] = copysignf(mn(p->ctmax, fabsf(v)), v) / velfac;
      }
      /* CFLKMAX as the vectorised reduction groups it (prop_ktp3.hpp:277-310) */
      int ith = 1;
      float m0 = acc_k, m1 = acc_k;
      if (nth >= 4) {
        float l0 = acc_k, l1 = acc_k, l2 = acc_k, l3 = acc_k;
        const int n4 = (nth / 4) * 4;
        for (; ith <= n4; ith += 4) {
          l0 = mx(l0, fabsf(VEL(ith)));
          l1 = mx(l1, fabsf(VEL(ith + 1)));
          l2 = mx(l2, fabsf(VEL(ith + 2)));
          l3 = mx(l3, fabsf(VEL(ith + 3)));
        }
        m0 = mx(l2, l0);
        m1 = mx(l3, l1);
        acc_k = mx(m1, m0);
      }
      if (nth - ith + 1 >= 2) {
        m0 = mx(m0, fabsf(VEL(ith)));
        m1 = mx(m1, fabsf(VEL(ith + 1)));
        ith += 2;
        acc_k = mx(m1, m0);
      }
      if (ith <= nth) acc_k = mx(acc_k, fabsf(VEL(ith)));
#undef VEL
    }
  }
  *cflthmax = acc_th;
  *cflkmax = acc_k;
  for (int pass = 0; pass < 2; ++pass) {
    const int kpass = (pass == 0) == (itstep % 2 == 0);
    if (kpass && p->flck) {
      for (int ith = 1; ith <= nth; ++ith)
        vq[nk + 2 + (ith - 1) * nk2 + vqoff] = p->fachfa * vq[nk + 1 + (ith - 1) * nk2 + vqoff];
      qck2(nth, nk2, nth, nk2, cflk, fack, db, dm, vq, 0, 1, p->mapth2, nspec, p->mapwn2,
           nspec - nth, nspec, nspec + nth, fla);
    }
    if (!kpass && p->flcth)
      qck1(nth, nk2, nth, nk2, vcflt, vq, 1, nk2, p->mapth2, nspec, p->mapth2, nspec, nspec,
           nspec, fla);
  }
  for (int isp = 1; isp <= nspec; ++isp) va[isp - 1] = vq[p->mapth2[isp - 1] + vqoff];
}

/* ww_ktp3_batch's arguments (include/ww3gpu_c.h); points run on nthreads. */
int or_ktp3_batch(const float* rpar, const int* ipar, const float* sig, const float* dsip,
                  const float* ecos, const float* esin, const float* ec2, const float* esc,
                  const float* es2, const int* mapwn, const int* mapth2, const int* mapwn2,
                  int itstep, int ntimes, float facth, float fack, int nsea, const float* cg,
                  const float* wn, int npts, const int* jsea, const int* isea,
                  const float* pscal, int ncol, float* va, float* cflthmax, float* cflkmax,
                  int nthreads) {
  Ktp3 p;
  p.dmin = rpar[0];
  p.ctmax = rpar[1];
  p.fachfa = rpar[2];
  p.nk = ipar[0];
  p.nth = ipar[1];
  p.nspec = ipar[2];
  p.nk2 = ipar[3];
  p.flcth = ipar[4];
  p.flck = ipar[5];
  p.flcur = ipar[6];
  p.sig = sig;
  p.dsip = dsip;
  p.ecos = ecos;
  p.esin = esin;
  p.ec2 = ec2;
  p.esc = esc;
  p.es2 = es2;
  p.mapwn = mapwn;
  p.mapth2 = mapth2;
  p.mapwn2 = mapwn2;
  (void)nsea;
  (void)ncol;
  const long nw = ktp3_work_floats(p.nk, p.nth);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    float* w = (float*)malloc(sizeof(float) * nw);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int j = 0; j < npts; ++j) {
      const int js = jsea[j], is = isea[j];
      const float* ps = pscal + 10L * j;
      float* vcol = va + (long)js * p.nspec;
      const float* cgc = cg + (long)is * p.nk2;
      const float* wnc = wn + (long)is * p.nk2;
      float cth = 0.f, ck = 0.f;
      for (int t = 0; t < ntimes; ++t)
        w3ktp3(&p, itstep, facth, fack, ps[0], cgc, wnc, ps[1], ps[2], ps[3], ps[4], ps[5], ps[6],
               ps[7], ps[8], ps[9], vcol, &cth, &ck, w);
      cflthmax[js - 1] = cth;
      cflkmax[js - 1] = ck;
    }
    free(w);
  }
  return 0;
}

/* W3QCK3 and W3QCK1 on one field, for the closed-form tests of the sweeps. */
void or_qck3(int mx_, int my, int nx, int ny, float* cfll, const float* trans, float* q,
             int close, int inc, const int* mapact, int nact, const int* mapbou, int nb0,
             int nb1, int nb2, float* fla) {
  qck3(mx_, my, nx, ny, cfll, trans, q, close, inc, mapact, nact, mapbou, nb0, nb1, nb2, fla);
}
void or_qck1(int mx_, int my, int nx, int ny, float* cfll, float* q, int close, int inc,
             const int* mapact, int nact, const int* mapbou, int nb0, int nb1, int nb2,
             float* fla) {
  qck1(mx_, my, nx, ny, cfll, q, close, inc, mapact, nact, mapbou, nb0, nb1, nb2, fla);
}

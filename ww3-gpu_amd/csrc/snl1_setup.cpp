// ww3-gpu_amd/csrc/snl1_setup.cpp
// INSNL1 on the host: the DIA's quadruplet geometry, weights and AF11, for one grid.
//
// Reference: kokkos/src/ww_kokkos/snl1_tables.cpp:82-329 (WW3 w3snl1md.F90:483-786).
// The float32 expressions and their order are the reference's (they fix the bits of the
// weights and AF11); the output form is this library's own: eight row offsets and
// eight direction shifts (snl1.hpp) instead of 32 gather tables.  The Fortran tables are
// still built, for export and to prove the offset form equal to them.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "snl1.hpp"

namespace ww3gpu {
namespace snl1 {
namespace {

// libgcc's __powisf2, which gfortran calls for REAL**INTEGER with a run-time exponent.
float powi(float x, int m) {
  unsigned n = m < 0 ? static_cast<unsigned>(-m) : static_cast<unsigned>(m);
  float y = (n & 1u) ? x : 1.0f;
  for (n >>= 1; n != 0u; n >>= 1) {
    x = x * x;
    if (n & 1u) y = y * x;
  }
  return m < 0 ? 1.0f / y : y;
}

// x**11 as gfortran -O3 expands a literal power (x^5 * x^6).
float pow11(float x) {
  const float x2 = x * x;
  const float x3 = x2 * x;
  return (x2 * x3) * (x3 * x3);
}

}  // namespace

size_t lds_bytes(const Geom& g) {
  // UE, then SA1, SA2, DA1P, DA1M, DA2P, DA2M, each row of w floats.
  return sizeof(float) * static_cast<size_t>(g.w) *
         (static_cast<size_t>(g.ue_rows) + 6u * static_cast<size_t>(g.sa_rows));
}

bool build_tables(const Config& c, const float* sig, Tables& t) {
  if (c.nk <= 0 || c.nth <= 0 || sig == nullptr) return false;
  const int nk = c.nk, nth = c.nth;

  // Quadruplet angles and lambda weights.
  const float lm = (1.0f - c.lam) * (1.0f - c.lam);
  const float lp = (1.0f + c.lam) * (1.0f + c.lam);
  const float delth3 = std::acos((lm * lm + 4.0f - lp * lp) / (4.0f * lm));
  const float delth4 = std::asin(-std::sin(delth3) * lm / lp);
  const float dal1 = 1.0f / (lp * lp);
  const float dal2 = 1.0f / (lm * lm);
  const float dal3 = 2.0f * dal1 * dal2;

  // Direction offsets: truncation toward zero, as INT().
  const float cthp = std::fabs(delth4 / c.dth);
  const int ithp = static_cast<int>(cthp);
  const float wthp = cthp - static_cast<float>(ithp);
  const float cthm = std::fabs(delth3 / c.dth);
  const int ithm = static_cast<int>(cthm);
  const float wthm = cthm - static_cast<float>(ithm);

  // Frequency offsets.
  const float lnx = std::log(c.xfr);
  const int ifrp = static_cast<int>(std::log(1.0f + c.lam) / lnx);
  const int ifrm = static_cast<int>(std::log(1.0f - c.lam) / lnx);
  const float wfrp = (1.0f + c.lam - powi(c.xfr, ifrp)) / (powi(c.xfr, ifrp + 1) - powi(c.xfr, ifrp));
  const float wfrm = (powi(c.xfr, ifrm) - (1.0f - c.lam)) / (powi(c.xfr, ifrm) - powi(c.xfr, ifrm - 1));

  const int nfrchg = nk - (ifrm - 1);
  const int nfrhgh = nk + (ifrp + 1) - (ifrm - 1);
  if (nfrchg <= 0 || nfrhgh < nfrchg) return false;

  Geom& g = t.g;
  g.nk = nk;
  g.nth = nth;
  g.nspec = nk * nth;
  g.nfrchg = nfrchg;
  g.nfrhgh = nfrhgh;
  g.nspecx = nfrchg * nth;
  const int ifo[8] = {ifrp, ifrp + 1, ifrm, ifrm - 1, -ifrp, -(ifrp + 1), -ifrm, -(ifrm - 1)};
  const int ito[8] = {ithp, ithp + 1, ithm, ithm + 1, -ithp, -(ithp + 1), -ithm, -(ithm + 1)};
  int h = 0, lowest = 1;
  for (int m = 0; m < 8; ++m) {
    g.drow[m] = ifo[m];
    g.sh[m] = ito[m];
    h = std::max(h, std::abs(ito[m]));
    lowest = std::min(lowest, 1 + ifo[m]);
  }
  g.h = h;
  g.w = nth + 2 * h;
  g.zr = 1 - lowest;  // rows lowest..0 are zero rows
  if (g.zr < 1) g.zr = 1;
  g.ue_rows = g.zr + nfrhgh;
  g.sa_rows = g.zr + nfrchg;

  float awg[8];
  awg[0] = wthp * wfrp;
  awg[1] = (1.0f - wthp) * wfrp;
  awg[2] = wthp * (1.0f - wfrp);
  awg[3] = (1.0f - wthp) * (1.0f - wfrp);
  awg[4] = wthm * wfrm;
  awg[5] = (1.0f - wthm) * wfrm;
  awg[6] = wthm * (1.0f - wfrm);
  awg[7] = (1.0f - wthm) * (1.0f - wfrm);
  g.dal1 = dal1;
  g.dal2 = dal2;
  g.dal3 = dal3;
  t.scal[0] = dal1;
  t.scal[1] = dal2;
  t.scal[2] = dal3;
  for (int j = 0; j < 8; ++j) {
    g.awg[j] = awg[j];
    g.swg[j] = awg[j] * awg[j];
    t.scal[3 + j] = g.awg[j];
    t.scal[11 + j] = g.swg[j];
  }
  g.snlc1 = c.snlc1;
  g.kdcon = c.kdcon;
  g.kdmn = c.kdmn;
  g.snls1 = c.snls1;
  g.snls2 = c.snls2;
  g.snls3 = c.snls3;
  g.fachfe = c.fachfe;
  const float pi = static_cast<float>(3.141592653589793);
  g.tpiinv = 1.0f / (2.0f * pi);
  t.ifr_off[0] = ifrp;
  t.ifr_off[1] = ifrp + 1;
  t.ifr_off[2] = ifrm;
  t.ifr_off[3] = ifrm - 1;
  t.ith_off[0] = ithp;
  t.ith_off[1] = ithp + 1;
  t.ith_off[2] = ithm;
  t.ith_off[3] = ithm + 1;

  // AF11 per row; rows beyond NK continue FR = FR*XFR from SIG(NK)/(2 pi).
  t.sig.assign(sig, sig + nk);
  t.af11_row.resize(static_cast<size_t>(nfrchg));
  for (int r = 0; r < nk; ++r) t.af11_row[r] = pow11(sig[r] * g.tpiinv);
  float fr = sig[nk - 1] * g.tpiinv;
  for (int r = nk; r < nfrchg; ++r) {
    fr = fr * c.xfr;
    t.af11_row[r] = pow11(fr);
  }

  // The Fortran tables: IF_m(IFR) clamped at 0, IT_n(ITH) wrapped once.
  const auto ifm = [&](int m, int r) {  // m, r 1-based
    const int v = r + ifo[m - 1];
    return (m >= 3 && m <= 6) ? std::max(0, v) : v;
  };
  const auto itn = [&](int n, int th) {  // n, th 1-based
    int v = th + ito[n - 1];
    if (n <= 4 && v > nth) v -= nth;
    if (n >= 5 && v < 1) v += nth;
    return v;
  };
  static const int kx[16][2] = {{2, 2}, {1, 2}, {2, 1}, {1, 1}, {8, 4}, {7, 4}, {8, 3}, {7, 3},
                                {6, 2}, {5, 2}, {6, 1}, {5, 1}, {4, 4}, {3, 4}, {4, 3}, {3, 3}};
  static const int kc[16][2] = {{6, 6}, {5, 6}, {6, 5}, {5, 5}, {4, 8}, {3, 8}, {4, 7}, {3, 7},
                                {2, 6}, {1, 6}, {2, 5}, {1, 5}, {8, 8}, {7, 8}, {8, 7}, {7, 7}};
  t.ipim.assign(16u * static_cast<size_t>(g.nspecx), 0);
  t.ic.assign(16u * static_cast<size_t>(g.nspec), 0);
  for (int i = 0; i < g.nspecx; ++i)
    for (int j = 0; j < 16; ++j)
      t.ipim[j * g.nspecx + i] = itn(kx[j][0], 1 + i % nth) + (ifm(kx[j][1], 1 + i / nth) - 1) * nth;
  for (int i = 0; i < g.nspec; ++i)
    for (int j = 0; j < 16; ++j)
      t.ic[j * g.nspec + i] = itn(kc[j][0], 1 + i % nth) + (ifm(kc[j][1], 1 + i / nth) - 1) * nth;

  // Prove the offset form: for each entry, the kernel reads row r+drow[m] (a zero row when
  // <= 0, which must lie within the zr zero rows) and column c+sh[n] (within the halo, equal
  // to the wrapped direction), and the Fortran address is that row/direction, or a zero-row
  // address when its row was clamped.
  bool ok = h < nth;
  const auto same = [&](int addr, int n, int m, int i, int max_row) {
    const int r = 1 + i / nth, c0 = i % nth;
    const int row = r + ifo[m - 1];
    const int col = c0 + ito[n - 1];
    if (row > max_row || row < 1 - g.zr + 0 || col < -h || col >= nth + h) return false;
    const int wrapped = ((col % nth) + nth) % nth;
    if (row <= 0) return addr >= 1 - nth && addr <= 0;  // zero row either way
    return addr == wrapped + 1 + (row - 1) * nth;
  };
  for (int i = 0; ok && i < g.nspecx; ++i)
    for (int j = 0; ok && j < 16; ++j) ok = same(t.ipim[j * g.nspecx + i], kx[j][0], kx[j][1], i, nfrhgh);
  for (int i = 0; ok && i < g.nspec; ++i)
    for (int j = 0; ok && j < 16; ++j) ok = same(t.ic[j * g.nspec + i], kc[j][0], kc[j][1], i, nfrchg);
  t.structured = ok;
  for (int j = 0; j < 16; ++j) {
    g.o3[j] = ifo[kx[j][1] - 1] * g.w + ito[kx[j][0] - 1];
    g.o4[j] = ifo[kc[j][1] - 1] * g.w + ito[kc[j][0] - 1];
  }
  return true;
}

}  // namespace snl1
}  // namespace ww3gpu

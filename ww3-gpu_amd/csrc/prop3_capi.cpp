// ww3-gpu_amd/csrc/prop3_capi.cpp
// The C ABI of the PR3/UQ propagation: ww_xyp3_batch (W3GATH + W3XYP3 + W3SCAT for every
// component of a step) and ww_ktp3_batch (NTIMES W3KTP3 calls on every listed sea point).
//
// Reference: kokkos/src/fortran_iface/prop3_c.hpp:17-94 (declarations, packet lengths),
// prop3_shim.cpp:765-833 (the checks of xyp3_entry), :836-882 (ktp3_entry), :470-660 and
// :673-750 (what a call computes).  Status codes and "nothing written on a refused call" as
// there.  The MPI forms are refused (one GPU, no exchange).
//
// Host work per call: the per-component scalars of W3XYP3 (NTLOC, DTRAD, the averaging
// factors; prop_xyp3.hpp:200-275, in the reference's operation order) and, when the maps
// change, the tables the kernels index by cell: the sea point of each cell, each sea point's
// place in MAPCXY, the flux class of each cell in the two sweeps, and for every cell the
// neighbours whose averaging stores reach it, in the order those stores happen.  Read-only
// inputs are copied to the device only when their bytes differ from the last copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ww3gpu_c.h"
#include "prop3.hpp"
#include "resident.hpp"
#include "runtime.hpp"

namespace ww3gpu {
namespace prop3h {
namespace {

using prop::Comp;

std::atomic<int> g_last{WW_KOKKOS_OK};
std::mutex g_mutex;

int fail(int code, const std::string& what) {
  report(g_last, code, what.c_str());
  return code;
}

inline float mx(float a, float b) { return a > b ? a : b; }
inline int fint(float x) {
  return (x > -2147483904.f && x < 2147483648.f) ? static_cast<int>(x) : static_cast<int>(0x80000000u);
}
inline long long nintl(float x) {
  const float r = std::round(x);  // NINT: half away from zero (lroundf)
  return (r > -9.2233715e18f && r < 9.2233715e18f) ? static_cast<long long>(r) : 0;
}

// CONSTANTS folded to float (prop_params.hpp:27-35): 0.575*GRAV, DERA, RADIUS, DERA*RADIUS.
const float kPiF = static_cast<float>(3.141592653589793);
const float kCg0Fac = 0.575f * 9.806f;
const float kDera = kPiF / 180.f;
const float kRadius = 4.E7f * (1.0f / (2.0f * kPiF));
const float kDeraRadius = kDera * kRadius;

/// gfortran's in-order MINVAL/MAXVAL (prop_xyp3.hpp:64-88).
float minval_io(const float* a, int n) {
  if (n <= 0) return 3.40282347e+38f;
  int i = 0;
  while (i < n && !(a[i] <= __builtin_inff())) ++i;
  if (i == n) return __builtin_nanf("");
  float lim = a[i];
  for (++i; i < n; ++i) lim = (a[i] < lim) ? a[i] : lim;
  return lim;
}
float maxval_io(const float* a, int n) {
  if (n <= 0) return -3.40282347e+38f;
  int i = 0;
  while (i < n && !(a[i] >= -__builtin_inff())) ++i;
  if (i == n) return __builtin_nanf("");
  float lim = a[i];
  for (++i; i < n; ++i) lim = (a[i] > lim) ? a[i] : lim;
  return lim;
}

/// A device copy of one read-only argument, refreshed only when its bytes change.
struct Slot {
  DevBuf dev;
  std::vector<unsigned char> shadow;
  bool valid = false;
  bool changed = true;
  hipError_t put(const void* p, size_t bytes, hipStream_t st) {
    changed = !(valid && shadow.size() == bytes && (bytes == 0 || std::memcmp(shadow.data(), p, bytes) == 0));
    if (!changed) return hipSuccess;
    valid = false;
    hipError_t e = dev.ensure(std::max<size_t>(bytes, 4));
    if (e == hipSuccess && bytes > 0) e = hipMemcpyAsync(dev.p, p, bytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    const auto* b = static_cast<const unsigned char*>(p);
    shadow.assign(b, b + bytes);
    valid = true;
    return hipSuccess;
  }
  template <class T>
  const T* ptr() const {
    return static_cast<const T*>(dev.p);
  }
};

/// Host-built tables and their device copies, rebuilt when the maps change.
struct Derived {
  std::vector<int> cellsea, pos, bnd;
  std::vector<uint64_t> order;
  std::vector<unsigned char> clsx, clsy, act;
  Slot d_cellsea, d_pos, d_bnd, d_order, d_clsx, d_clsy, d_act;
};

struct XyCtx {
  Slot in[24];
  Derived der;
  std::string derived_key;  ///< the map inputs the tables were built from
  Slot comps;
  DevBuf rec, tw, q0, q1, fla, vlcx, vlcy, va;
  hipEvent_t ev[2] = {nullptr, nullptr};
  double ms_gpu = 0;
  long long cell_substeps = 0;
  int substeps = 0;
};

XyCtx& xy() {
  static XyCtx* c = new XyCtx();
  return *c;
}

struct Xy3Args {
  const float* rpar;
  const int* ipar;
  const int* imaps;
  const float* tscal;
  float dtg;
  const int *mapsf, *mapx2, *mapy2, *mapaxy, *mapcxy, *maptrn, *mapsta, *isbpi, *on;
  const float *clats, *sig, *ecos, *esin, *dpdx, *dpdy, *dqdx, *dqdy, *gsqrt, *atrnx, *atrny;
  const float *cg, *cx, *cy, *bbpi0, *bbpin;
  int flbpi, nbi, ncol;
};

/// The checks of xyp3_entry (prop3_shim.cpp:776-793) plus the map properties the kernels rely
/// on: entries inside the grid, each map listing a cell at most once, MAPCXY a permutation.
int check_xyp3(const Xy3Args& a, const char* who) {
  if (a.rpar == nullptr || a.ipar == nullptr || a.imaps == nullptr || a.tscal == nullptr)
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": bad packet");
  const int* ip = a.ipar;
  const int nx = ip[0], ny = ip[1], nsea = ip[2], nk = ip[3], nth = ip[4], nspec = ip[5];
  if (nx < 1 || ny < 1 || nsea < 1 || nk < 1 || nth < 1 || nspec != nk * nth || a.ncol < nsea + 1 ||
      a.nbi < 0 || a.mapsf == nullptr || a.clats == nullptr || a.sig == nullptr || a.ecos == nullptr ||
      a.esin == nullptr || a.dpdx == nullptr || a.dpdy == nullptr || a.dqdx == nullptr || a.dqdy == nullptr ||
      a.gsqrt == nullptr || a.atrnx == nullptr || a.atrny == nullptr || a.mapx2 == nullptr ||
      a.mapy2 == nullptr || a.mapaxy == nullptr || a.mapcxy == nullptr || a.maptrn == nullptr ||
      a.mapsta == nullptr || a.cg == nullptr || a.cx == nullptr || a.cy == nullptr || a.on == nullptr ||
      (a.flbpi != 0 && (a.isbpi == nullptr || a.bbpi0 == nullptr || a.bbpin == nullptr)) ||
      static_cast<long long>(nx) * ny > (1LL << 30))
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": bad sizes or null pointer");
  if (ip[6] == 8) return fail(WW_KOKKOS_ERR_UNSUPPORTED, std::string(who) + ": tripole grid (W3XYP3 stops there)");
  if (ip[6] != -1 && nx < 2)
    return fail(WW_KOKKOS_ERR_UNSUPPORTED, std::string(who) + ": a closed grid needs NX >= 2");
  if (ny < 3)  // W3MAP3 lists no active point on rows 1 and NY; the stencil needs NY >= 3
    return fail(WW_KOKKOS_ERR_UNSUPPORTED, std::string(who) + ": NY < 3");
  const int nxy = nx * ny;
  const int* m = a.imaps;
  if (!(0 <= m[0] && m[0] <= m[1] && m[1] <= m[2] && m[2] <= nxy && 0 <= m[3] && m[3] <= m[4] &&
        m[4] <= m[5] && m[5] <= nxy && 0 <= m[6] && m[6] <= nxy && 0 <= m[7] && m[7] <= nsea))
    return fail(WW_KOKKOS_ERR_TABLE, std::string(who) + ": map counts out of range");
  for (int i = 0; i < nsea; ++i) {
    const int ix = a.mapsf[i], iy = a.mapsf[nsea + i], ixy = a.mapsf[2 * nsea + i];
    if (ix < 1 || ix > nx || iy < 1 || iy > ny || ixy != iy + (ix - 1) * ny)
      return fail(WW_KOKKOS_ERR_TABLE, std::string(who) + ": MAPSF entry outside the grid");
  }
  if (a.flbpi != 0)
    for (int i = 0; i < a.nbi; ++i)
      if (a.isbpi[i] < 1 || a.isbpi[i] > nsea)
        return fail(WW_KOKKOS_ERR_TABLE, std::string(who) + ": ISBPI entry outside the sea points");
  return WW_KOKKOS_OK;
}

/// The tables of Derived from the maps (host).  Returns false when a map repeats a cell or
/// MAPCXY is not a permutation of the sea points.
bool build_derived(const Xy3Args& a, Derived& d) {
  const int nx = a.ipar[0], ny = a.ipar[1], nsea = a.ipar[2];
  const int nxy = nx * ny, nq = ny * (nx + 3), ncent = a.imaps[7];
  d.cellsea.assign(nxy + 1, 0);
  for (int i = 1; i <= nsea; ++i) {
    int& c = d.cellsea[a.mapsf[2 * nsea + i - 1]];
    if (c != 0) return false;
    c = i;
  }
  d.pos.assign(nsea + 1, 0);
  for (int ip = 1; ip <= nsea; ++ip) {
    const int is = a.mapcxy[ip - 1];
    if (is < 1 || is > nsea || d.pos[is] != 0) return false;
    d.pos[is] = ip;
  }
  // rows as W3MAP3 fills them: MAPX2 and MAPAXY rows 2..NY-1, MAPY2 rows 1..NY-1 (the
  // sweeps' stencils then stay inside the plane)
  const auto row = [&](int c) { return (c - 1) % ny + 1; };
  const auto classes = [&](const int* map, int n0, int n1, int n2, int rlo, std::vector<unsigned char>& cls) {
    cls.assign(nxy + 1, 0);
    for (int p = 1; p <= n2; ++p) {
      const int c = map[p - 1];
      if (c < 1 || c > nxy || cls[c] != 0 || row(c) < rlo || row(c) > ny - 1) return false;
      cls[c] = p <= n0 ? 1 : (p <= n1 ? 2 : 3);
    }
    return true;
  };
  if (!classes(a.mapx2, a.imaps[0], a.imaps[1], a.imaps[2], 2, d.clsx)) return false;
  if (!classes(a.mapy2, a.imaps[3], a.imaps[4], a.imaps[5], 1, d.clsy)) return false;
  d.act.assign(nxy + 1, 0);
  for (int p = 1; p <= a.imaps[6]; ++p) {
    const int c = a.mapaxy[p - 1];
    if (c < 1 || c > nxy || d.act[c] != 0 || row(c) < 2 || row(c) > ny - 1) return false;
    d.act[c] = 1;
  }
  d.bnd.assign(nxy + 1, 0);
  if (a.flbpi != 0)
    for (int ibi = 1; ibi <= a.nbi; ++ibi) d.bnd[a.mapsf[2 * nsea + a.isbpi[ibi - 1] - 1]] = ibi;
  // For every target cell t (1-NY .. NY*(NX+2)): the sources s = t - off that scatter (sea
  // points; near-coast ones only with MAPSTA > 0; a MAPTRN source only into its own cell),
  // ordered by their place in MAPCXY, which is the order of the two scatter loops.
  const int dir[9] = {0, 1, -1, ny, -ny, ny + 1, ny - 1, -ny + 1, -ny - 1};
  d.order.assign(nq, ~0ULL);
  std::pair<int, int> lst[9];
  for (int k = 0; k < nq; ++k) {
    const int t = k + 1 - ny;
    int n = 0;
    for (int dd = 0; dd < 9; ++dd) {
      const int s = t - dir[dd];
      if (s < 1 || s > nxy) continue;
      const int is = d.cellsea[s];
      if (is == 0) continue;
      if (d.pos[is] > ncent && a.mapsta[s - 1] <= 0) continue;
      if (a.maptrn[s - 1] != 0 && dd != 0) continue;
      lst[n++] = {d.pos[is], dd};
    }
    std::sort(lst, lst + n);
    uint64_t w = ~0ULL;
    for (int i = n - 1; i >= 0; --i) w = (w << 4) | static_cast<uint64_t>(lst[i].second);
    d.order[k] = w;
  }
  return true;
}

/// W3XYP3's per-component scalars (prop_xyp3.hpp:200-275), float operations in its order.
Comp make_comp(const Xy3Args& a, int isp) {
  const int* ip = a.ipar;
  const int nth = ip[4], nsea = ip[2], flagll = ip[9], flcur = ip[10];
  const float dtcfl = a.rpar[0], dth = a.rpar[1], xfr = a.rpar[2], wdcg = a.rpar[3], wdth = a.rpar[4];
  Comp c{};
  c.isp = isp;
  c.ith = 1 + (isp - 1) % nth;
  c.ik = 1 + (isp - 1) / nth;
  const float cg0 = kCg0Fac / a.sig[1];
  const float cga = kCg0Fac / a.sig[c.ik];
  c.ecosi = a.ecos[c.ith - 1];
  c.esini = a.esin[c.ith - 1];
  const float cgx = c.ecosi * cga, cgy = c.esini * cga;
  float acgx, acgy, cxc, cyc;
  if (flcur != 0) {
    const float cxmin = minval_io(a.cx + 1, nsea), cxmax = maxval_io(a.cx + 1, nsea);
    const float cymin = minval_io(a.cy + 1, nsea), cymax = maxval_io(a.cy + 1, nsea);
    acgx = mx(std::fabs(cgx + cxmin), std::fabs(cgx + cxmax));
    acgy = mx(std::fabs(cgy + cymin), std::fabs(cgy + cymax));
    cxc = mx(std::fabs(cxmin), std::fabs(cxmax));
    cyc = mx(std::fabs(cymin), std::fabs(cymax));
  } else {
    acgx = std::fabs(cgx);
    acgy = std::fabs(cgy);
    cxc = 0.f;
    cyc = 0.f;
  }
  const float t1 = mx(acgx, acgy), t2 = mx(cxc, cyc);
  const float t3 = mx(cg0 * 0.001f, t2);
  const float cgn = mx(t1, t3);
  c.ntloc = 1 + fint(a.dtg / ((dtcfl * cg0) / cgn));
  c.rntloc = static_cast<float>(c.ntloc);
  const float dtloc = a.dtg / c.rntloc;
  c.dtrad = flagll != 0 ? dtloc / kDeraRadius : dtloc;
  float fth = (wdth * dth) * dtloc;
  float fcg = ((wdcg * 0.5f) * (xfr - (1.f / xfr))) * dtloc;
  if (flagll != 0) {
    fth = (fth / kRadius) / kDera;
    fcg = (fcg / kRadius) / kDera;
  }
  fcg = fcg / c.rntloc;
  c.fthx = -(c.esini * fth);
  c.fthy = c.ecosi * fth;
  c.fcgx = c.ecosi * fcg;
  c.fcgy = c.esini * fcg;
  return c;
}

/// One ww_xyp3_batch call with VA on the device (d_va, VA(NSPEC, 0:ncol-1)).
int run_xyp3(const Xy3Args& a, float* d_va, hipStream_t st, const char* who) {
  XyCtx& x = xy();
  const int* ip = a.ipar;
  const int nx = ip[0], ny = ip[1], nsea = ip[2], nk = ip[3], nth = ip[4], nspec = ip[5];
  const int nxy = nx * ny;
  const size_t fz = sizeof(float), iz = sizeof(int);
  std::vector<int> list;
  for (int isp = 1; isp <= nspec; ++isp)
    if (a.on[isp - 1] != 0) list.push_back(isp);
  const int nlist = static_cast<int>(list.size());
  if (nlist == 0) return WW_KOKKOS_OK;
  std::vector<Comp> comps(nlist);
  int maxnt = 0;
  for (int i = 0; i < nlist; ++i) {
    comps[i] = make_comp(a, list[i]);
    if (comps[i].ntloc < 1 || comps[i].ntloc > 100000)
      return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": the CFL sub-step count is not finite");
    maxnt = std::max(maxnt, comps[i].ntloc);
  }
  // the map inputs decide the derived tables
  std::string key;
  const auto add = [&](const void* p, size_t n) { key.append(static_cast<const char*>(p), n); };
  add(ip, 11 * iz);
  add(a.imaps, 8 * iz);
  add(a.mapsf, 3 * nsea * iz);
  add(a.mapx2, a.imaps[2] * iz);
  add(a.mapy2, a.imaps[5] * iz);
  add(a.mapaxy, a.imaps[6] * iz);
  add(a.mapcxy, nsea * iz);
  add(a.maptrn, nxy * iz);
  add(a.mapsta, nxy * iz);
  add(&a.flbpi, iz);
  if (a.flbpi != 0) add(a.isbpi, a.nbi * iz);
  Derived& d = x.der;
  if (key != x.derived_key) {
    if (!build_derived(a, d)) {
      x.derived_key.clear();
      return fail(WW_KOKKOS_ERR_TABLE, std::string(who) + ": a map lists a cell twice or MAPCXY is not a permutation");
    }
    x.derived_key = key;
  }
  hipError_t e = hipSuccess;
  const auto put = [&](Slot& s, const void* p, size_t bytes) {
    if (e == hipSuccess) e = s.put(p, bytes, st);
    return s.dev.p;
  };
  prop::XyGrid g{};
  g.nx = nx;
  g.ny = ny;
  g.nsea = nsea;
  g.nk = nk;
  g.nth = nth;
  g.nspec = nspec;
  g.nxy = nxy;
  g.nq = ny * (nx + 3);
  g.nf = ny * (nx + 1);
  g.cgld = nk + 2;
  g.global = ip[6] != -1;
  g.flcur = ip[10] != 0;
  g.ncent = a.imaps[7];
  g.flbpi = a.flbpi != 0;
  g.dtg = a.dtg;
  g.dsec_bpi0 = g.flbpi ? a.tscal[1] : 0.f;
  g.dsec_bpin = g.flbpi ? a.tscal[2] : 0.f;
  int slot = 0;
  g.mapsf = static_cast<const int*>(put(x.in[slot++], a.mapsf, 3 * nsea * iz));
  g.mapsta = static_cast<const int*>(put(x.in[slot++], a.mapsta, nxy * iz));
  g.maptrn = static_cast<const int*>(put(x.in[slot++], a.maptrn, nxy * iz));
  g.clats = static_cast<const float*>(put(x.in[slot++], a.clats, (nsea + 1) * fz));
  g.dpdx = static_cast<const float*>(put(x.in[slot++], a.dpdx, nxy * fz));
  g.dpdy = static_cast<const float*>(put(x.in[slot++], a.dpdy, nxy * fz));
  g.dqdx = static_cast<const float*>(put(x.in[slot++], a.dqdx, nxy * fz));
  g.dqdy = static_cast<const float*>(put(x.in[slot++], a.dqdy, nxy * fz));
  g.gsqrt = static_cast<const float*>(put(x.in[slot++], a.gsqrt, nxy * fz));
  g.atrnx = static_cast<const float*>(put(x.in[slot++], a.atrnx, 3 * static_cast<size_t>(nxy) * fz));
  g.atrny = static_cast<const float*>(put(x.in[slot++], a.atrny, 3 * static_cast<size_t>(nxy) * fz));
  g.cg = static_cast<const float*>(put(x.in[slot++], a.cg, static_cast<size_t>(nk + 2) * (nsea + 1) * fz));
  g.cx = static_cast<const float*>(put(x.in[slot++], a.cx, (nsea + 1) * fz));
  g.cy = static_cast<const float*>(put(x.in[slot++], a.cy, (nsea + 1) * fz));
  const size_t nbb = g.flbpi ? static_cast<size_t>(nspec) * (a.nbi + 1) : 0;
  g.isbpi = static_cast<const int*>(put(x.in[slot++], a.isbpi, g.flbpi ? a.nbi * iz : 0));
  g.bbpi0 = static_cast<const float*>(put(x.in[slot++], a.bbpi0, nbb * fz));
  g.bbpin = static_cast<const float*>(put(x.in[slot++], a.bbpin, nbb * fz));
  g.cellsea = static_cast<const int*>(put(d.d_cellsea, d.cellsea.data(), d.cellsea.size() * iz));
  g.pos = static_cast<const int*>(put(d.d_pos, d.pos.data(), d.pos.size() * iz));
  g.bnd = static_cast<const int*>(put(d.d_bnd, d.bnd.data(), d.bnd.size() * iz));
  g.order = static_cast<const uint64_t*>(put(d.d_order, d.order.data(), d.order.size() * 8));
  g.clsx = static_cast<const unsigned char*>(put(d.d_clsx, d.clsx.data(), d.clsx.size()));
  g.clsy = static_cast<const unsigned char*>(put(d.d_clsy, d.clsy.data(), d.clsy.size()));
  g.act = static_cast<const unsigned char*>(put(d.d_act, d.act.data(), d.act.size()));
  prop::XyState s{};
  s.nlist = nlist;
  s.comps = static_cast<const Comp*>(put(x.comps, comps.data(), comps.size() * sizeof(Comp)));
  if (e != hipSuccess) return fail(WW_KOKKOS_ERR_RUNTIME, std::string(who) + ": host-to-device copy");
  const size_t nl = static_cast<size_t>(nlist);
  if (x.rec.ensure(nl * (nxy + 1) * sizeof(prop::Rec)) != hipSuccess ||
      x.tw.ensure(nl * (nxy + 1) * sizeof(uint64_t)) != hipSuccess || x.q0.ensure(nl * g.nq * fz) != hipSuccess ||
      x.q1.ensure(nl * g.nq * fz) != hipSuccess || x.fla.ensure(nl * g.nf * fz) != hipSuccess ||
      x.vlcx.ensure(nl * g.nf * fz) != hipSuccess || x.vlcy.ensure(nl * g.nf * fz) != hipSuccess)
    return fail(WW_KOKKOS_ERR_RUNTIME, std::string(who) + ": device allocation");
  s.rec = x.rec.as<prop::Rec>();
  s.tw = x.tw.as<uint64_t>();
  s.q[0] = x.q0.as<float>();
  s.q[1] = x.q1.as<float>();
  s.fla = x.fla.as<float>();
  s.vlcx = x.vlcx.as<float>();
  s.vlcy = x.vlcy.as<float>();
  if (x.ev[0] == nullptr) {
    (void)hipEventCreate(&x.ev[0]);
    (void)hipEventCreate(&x.ev[1]);
  }
  (void)hipEventRecord(x.ev[0], st);
  e = prop::xyp3_gather(g, s, d_va, st);
  if (e == hipSuccess) e = prop::xyp3_records(g, s, st);
  const int dttst = fint(a.tscal[0]);
  bool yfirst = (nintl(static_cast<float>(dttst) / a.dtg) % 2) == 0;
  long long cells = 0;
  for (int it = 1; it <= maxnt && e == hipSuccess; ++it) {
    e = prop::xyp3_substep(g, s, it, yfirst, ip[7] != 0, ip[8] != 0, st);
    yfirst = !yfirst;
    for (const Comp& c : comps) cells += c.ntloc >= it ? nxy : 0;
  }
  if (e == hipSuccess) e = prop::xyp3_store(g, s, d_va, st);
  (void)hipEventRecord(x.ev[1], st);
  if (e == hipSuccess) e = hipEventSynchronize(x.ev[1]);
  if (e != hipSuccess) return fail(WW_KOKKOS_ERR_KERNEL, std::string(who) + ": " + hipGetErrorString(e));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, x.ev[0], x.ev[1]);
  x.ms_gpu = ms;
  x.cell_substeps = cells;
  x.substeps = maxnt;
  return WW_KOKKOS_OK;
}

// ---------------------------------------------------------------------------------------
// W3KTP3
// ---------------------------------------------------------------------------------------
struct KtCtx {
  Slot in[16];
  DevBuf va, cth, ck;
  hipEvent_t ev[2] = {nullptr, nullptr};
  double ms_gpu = 0;
};
KtCtx& kt() {
  static KtCtx* c = new KtCtx();
  return *c;
}

struct Kt3Args {
  const float* rpar;
  const int* ipar;
  const float *sig, *dsip, *ecos, *esin, *ec2, *esc, *es2, *cg, *wn, *pscal;
  const int *mapwn, *mapth2, *mapwn2, *jsea, *isea;
  int itstep, ntimes, nsea, npts, ncol;
  float facth, fack;
};

/// The checks of ktp3_entry (prop3_shim.cpp:843-867) plus the map ranges the kernel indexes.
int check_ktp3(const Kt3Args& a, prop::KtGrid& g, const char* who) {
  if (a.rpar == nullptr || a.ipar == nullptr) return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": bad packet");
  g.dmin = a.rpar[0];
  g.ctmax = a.rpar[1];
  g.fachfa = a.rpar[2];
  g.nk = a.ipar[0];
  g.nth = a.ipar[1];
  g.nspec = a.ipar[2];
  g.nk2 = a.ipar[3];
  g.flcth = a.ipar[4];
  g.flck = a.ipar[5];
  g.flcur = a.ipar[6];
  if (g.nk < 1 || g.nth < 1 || g.nspec != g.nk * g.nth || g.nk2 != g.nk + 2 || a.nsea < 1 || a.npts < 0 ||
      a.npts > a.nsea || a.ntimes < 0 || a.ncol < 2 || a.sig == nullptr || a.dsip == nullptr ||
      a.ecos == nullptr || a.esin == nullptr || a.ec2 == nullptr || a.esc == nullptr || a.es2 == nullptr ||
      a.mapwn == nullptr || a.mapth2 == nullptr || a.mapwn2 == nullptr || a.cg == nullptr || a.wn == nullptr ||
      (a.npts > 0 && (a.jsea == nullptr || a.isea == nullptr || a.pscal == nullptr)))
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": bad sizes or null pointer");
  for (int j = 0; j < a.npts; ++j)
    if (a.jsea[j] < 1 || a.jsea[j] >= a.ncol || a.isea[j] < 1 || a.isea[j] > a.nsea)
      return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": a point outside VA or the sea");
  if (prop::ktp3_lds_bytes(g.nk, g.nth) > 64 * 1024)
    return fail(WW_KOKKOS_ERR_UNSUPPORTED, std::string(who) + ": NK x NTH too large for one workgroup's LDS");
  // the rotated-grid maps must stay inside the work spectrum
  const int nk2 = g.nk2, lo = 1, hi = nk2 * g.nth;
  for (int i = 0; i < g.nspec; ++i) {
    const int m2 = a.mapth2[i], mw = a.mapwn[i];
    if (m2 < lo || m2 > hi || mw < 1 || mw > g.nk)
      return fail(WW_KOKKOS_ERR_TABLE, std::string(who) + ": MAPTH2/MAPWN entry outside the spectrum");
  }
  for (int i = 0; i < g.nspec + g.nth; ++i)
    if (a.mapwn2[i] < lo || a.mapwn2[i] > hi)
      return fail(WW_KOKKOS_ERR_TABLE, std::string(who) + ": MAPWN2 entry outside the spectrum");
  return WW_KOKKOS_OK;
}

int run_ktp3(const Kt3Args& a, const prop::KtGrid& g0, float* d_va, float* d_cth, float* d_ck, hipStream_t st,
             const char* who) {
  KtCtx& x = kt();
  prop::KtGrid g = g0;
  const size_t fz = sizeof(float), iz = sizeof(int);
  const size_t ntab = static_cast<size_t>(g.nspec + g.nth), nk2 = g.nk2;
  hipError_t e = hipSuccess;
  int slot = 0;
  const auto put = [&](const void* p, size_t bytes) {
    Slot& s = x.in[slot++];
    if (e == hipSuccess) e = s.put(p, bytes, st);
    return s.dev.p;
  };
  g.sig = static_cast<const float*>(put(a.sig, nk2 * fz));
  g.dsip = static_cast<const float*>(put(a.dsip, nk2 * fz));
  g.ecos = static_cast<const float*>(put(a.ecos, ntab * fz));
  g.esin = static_cast<const float*>(put(a.esin, ntab * fz));
  g.ec2 = static_cast<const float*>(put(a.ec2, ntab * fz));
  g.esc = static_cast<const float*>(put(a.esc, ntab * fz));
  g.es2 = static_cast<const float*>(put(a.es2, ntab * fz));
  g.mapwn = static_cast<const int*>(put(a.mapwn, ntab * iz));
  g.mapth2 = static_cast<const int*>(put(a.mapth2, nk2 * g.nth * iz));
  g.mapwn2 = static_cast<const int*>(put(a.mapwn2, ntab * iz));
  const size_t ncell = nk2 * (static_cast<size_t>(a.nsea) + 1);
  const float* cg = static_cast<const float*>(put(a.cg, ncell * fz));
  const float* wn = static_cast<const float*>(put(a.wn, ncell * fz));
  const int* js = static_cast<const int*>(put(a.jsea, a.npts * iz));
  const int* is = static_cast<const int*>(put(a.isea, a.npts * iz));
  const float* ps = static_cast<const float*>(put(a.pscal, 10 * static_cast<size_t>(a.npts) * fz));
  if (e != hipSuccess) return fail(WW_KOKKOS_ERR_RUNTIME, std::string(who) + ": host-to-device copy");
  if (x.ev[0] == nullptr) {
    (void)hipEventCreate(&x.ev[0]);
    (void)hipEventCreate(&x.ev[1]);
  }
  (void)hipEventRecord(x.ev[0], st);
  e = prop::ktp3_batch(g, a.itstep, a.ntimes, a.facth, a.fack, cg, wn, a.npts, js, is, ps, d_va, d_cth, d_ck, st);
  (void)hipEventRecord(x.ev[1], st);
  if (e == hipSuccess) e = hipEventSynchronize(x.ev[1]);
  if (e != hipSuccess) return fail(WW_KOKKOS_ERR_KERNEL, std::string(who) + ": " + hipGetErrorString(e));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, x.ev[0], x.ev[1]);
  x.ms_gpu = ms;
  return WW_KOKKOS_OK;
}

}  // namespace
}  // namespace prop3h
}  // namespace ww3gpu

using namespace ww3gpu;
using namespace ww3gpu::prop3h;

extern "C" {

int ww_xyp3_batch(const float* rpar, int nrpar, const int* ipar, int nipar, const int* imaps, int nimaps,
                  const float* tscal, int ntscal, float dtg, const int* mapsf, const float* clats,
                  const float* sig, const float* ecos, const float* esin, const float* dpdx, const float* dpdy,
                  const float* dqdx, const float* dqdy, const float* gsqrt, const float* atrnx,
                  const float* atrny, const int* mapx2, const int* mapy2, const int* mapaxy, const int* mapcxy,
                  const int* maptrn, const int* mapsta, const float* cg, const float* cx, const float* cy,
                  int flbpi, int nbi, const int* isbpi, const float* bbpi0, const float* bbpin, const int* on,
                  int ncol, float* va) {
  const char* who = "ww_xyp3_batch";
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!runtime().ready) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, std::string(who) + " before ww_kokkos_init");
  if (nrpar != WW_XYP3_NRPAR || nipar != WW_XYP3_NIPAR || nimaps != WW_XYP3_NIMAPS || ntscal != WW_XYP3_NTSCAL)
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": bad packet");
  const Xy3Args a{rpar, ipar, imaps, tscal, dtg, mapsf, mapx2, mapy2, mapaxy, mapcxy, maptrn, mapsta, isbpi, on,
                  clats, sig, ecos, esin, dpdx, dpdy, dqdx, dqdy, gsqrt, atrnx, atrny, cg, cx, cy, bbpi0, bbpin,
                  flbpi, nbi, ncol};
  int rc = check_xyp3(a, who);
  if (rc != WW_KOKKOS_OK) return rc;
  if (va == nullptr) return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": bad sizes or null pointer");
  const size_t bytes = static_cast<size_t>(ipar[5]) * ncol * sizeof(float);
  XyCtx& x = xy();
  hipStream_t st = runtime().stream;
  if (x.va.ensure(bytes) != hipSuccess || hipMemcpyAsync(x.va.p, va, bytes, hipMemcpyHostToDevice, st) != hipSuccess)
    return fail(WW_KOKKOS_ERR_RUNTIME, std::string(who) + ": VA host-to-device copy");
  rc = run_xyp3(a, x.va.as<float>(), st, who);
  if (rc != WW_KOKKOS_OK) return rc;
  if (hipMemcpyAsync(va, x.va.p, bytes, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    return fail(WW_KOKKOS_ERR_RUNTIME, std::string(who) + ": VA device-to-host copy");
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

int ww3gpu_xyp3_batch_dev(const float* rpar, int nrpar, const int* ipar, int nipar, const int* imaps, int nimaps,
                          const float* tscal, int ntscal, float dtg, const int* mapsf, const float* clats,
                          const float* sig, const float* ecos, const float* esin, const float* dpdx,
                          const float* dpdy, const float* dqdx, const float* dqdy, const float* gsqrt,
                          const float* atrnx, const float* atrny, const int* mapx2, const int* mapy2,
                          const int* mapaxy, const int* mapcxy, const int* maptrn, const int* mapsta,
                          const float* cg, const float* cx, const float* cy, int flbpi, int nbi, const int* isbpi,
                          const float* bbpi0, const float* bbpin, const int* on, int ncol, float* d_va,
                          void* stream) {
  const char* who = "ww3gpu_xyp3_batch_dev";
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!runtime().ready) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, std::string(who) + " before ww_kokkos_init");
  if (nrpar != WW_XYP3_NRPAR || nipar != WW_XYP3_NIPAR || nimaps != WW_XYP3_NIMAPS || ntscal != WW_XYP3_NTSCAL)
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": bad packet");
  const Xy3Args a{rpar, ipar, imaps, tscal, dtg, mapsf, mapx2, mapy2, mapaxy, mapcxy, maptrn, mapsta, isbpi, on,
                  clats, sig, ecos, esin, dpdx, dpdy, dqdx, dqdy, gsqrt, atrnx, atrny, cg, cx, cy, bbpi0, bbpin,
                  flbpi, nbi, ncol};
  int rc = check_xyp3(a, who);
  if (rc != WW_KOKKOS_OK) return rc;
  if (d_va == nullptr) return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": null VA");
  hipStream_t st = stream != nullptr ? static_cast<hipStream_t>(stream) : runtime().stream;
  rc = run_xyp3(a, d_va, st, who);
  if (rc == WW_KOKKOS_OK) report(g_last, WW_KOKKOS_OK, "");
  return rc;
}

int ww_xyp3_batch_va(int igrid, const float* rpar, int nrpar, const int* ipar, int nipar, const int* imaps,
                     int nimaps, const float* tscal, int ntscal, float dtg, const int* mapsf, const float* clats,
                     const float* sig, const float* ecos, const float* esin, const float* dpdx,
                     const float* dpdy, const float* dqdx, const float* dqdy, const float* gsqrt,
                     const float* atrnx, const float* atrny, const int* mapx2, const int* mapy2,
                     const int* mapaxy, const int* mapcxy, const int* maptrn, const int* mapsta, const float* cg,
                     const float* cx, const float* cy, int flbpi, int nbi, const int* isbpi, const float* bbpi0,
                     const float* bbpin, const int* on, float* va, int nspec, int nsealm) {
  const char* who = "ww_xyp3_batch_va";
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!runtime().ready) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, std::string(who) + " before ww_kokkos_init");
  if (nrpar != WW_XYP3_NRPAR || nipar != WW_XYP3_NIPAR || nimaps != WW_XYP3_NIMAPS || ntscal != WW_XYP3_NTSCAL)
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": bad packet");
  const Xy3Args a{rpar, ipar, imaps, tscal, dtg, mapsf, mapx2, mapy2, mapaxy, mapcxy, maptrn, mapsta, isbpi, on,
                  clats, sig, ecos, esin, dpdx, dpdy, dqdx, dqdy, gsqrt, atrnx, atrny, cg, cx, cy, bbpi0, bbpin,
                  flbpi, nbi, nsealm + 1};
  int rc = check_xyp3(a, who);
  if (rc != WW_KOKKOS_OK) return rc;
  if (nspec != ipar[5]) return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": NSPEC differs from the packet");
  float* d_va = nullptr;
  if ((rc = resident::device_va(igrid, va, nspec, nsealm, true, &d_va, who)) != WW_KOKKOS_OK) return rc;
  rc = run_xyp3(a, d_va, runtime().stream, who);
  if (rc == WW_KOKKOS_OK) report(g_last, WW_KOKKOS_OK, "");
  return rc;
}

int ww_xyp3_batch_mpi(int igrid, const float* rpar, int nrpar, const int* ipar, int nipar, const int* imaps,
                      int nimaps, const float* tscal, int ntscal, float dtg, const int* mapsf, const float* clats,
                      const float* sig, const float* ecos, const float* esin, const float* dpdx,
                      const float* dpdy, const float* dqdx, const float* dqdy, const float* gsqrt,
                      const float* atrnx, const float* atrny, const int* mapx2, const int* mapy2,
                      const int* mapaxy, const int* mapcxy, const int* maptrn, const int* mapsta, const float* cg,
                      const float* cx, const float* cy, int flbpi, int nbi, const int* isbpi, const float* bbpi0,
                      const float* bbpin, const int* on, float* va, int nspec, int nsealm, int naproc, int iaproc,
                      const int* poff, const float* rbuf, float* sbuf, int nbuf) {
  return fail(WW_KOKKOS_ERR_UNSUPPORTED, "ww_xyp3_batch_mpi: one GPU, no MPI exchange");
}

int ww_xyp3_mpi_pack(int igrid, float* va, int nspec, int nsealm, const int* isps, int nisp, float* buf) {
  return fail(WW_KOKKOS_ERR_UNSUPPORTED, "ww_xyp3_mpi_pack: one GPU, no MPI exchange");
}

int ww_xyp3_mpi_unpack(int igrid, float* va, int nspec, int nsealm, const int* isps, int nisp, const float* buf) {
  return fail(WW_KOKKOS_ERR_UNSUPPORTED, "ww_xyp3_mpi_unpack: one GPU, no MPI exchange");
}

int ww_ktp3_batch(const float* rpar, int nrpar, const int* ipar, int nipar, const float* sig, const float* dsip,
                  const float* ecos, const float* esin, const float* ec2, const float* esc, const float* es2,
                  const int* mapwn, const int* mapth2, const int* mapwn2, int itstep, int ntimes, float facth,
                  float fack, int nsea, const float* cg, const float* wn, int npts, const int* jsea,
                  const int* isea, const float* pscal, int ncol, float* va, float* cflthmax, float* cflkmax) {
  const char* who = "ww_ktp3_batch";
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!runtime().ready) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, std::string(who) + " before ww_kokkos_init");
  if (nrpar != WW_KTP3_NRPAR || nipar != WW_KTP3_NIPAR) return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": bad packet");
  const Kt3Args a{rpar, ipar, sig, dsip, ecos, esin, ec2, esc, es2, cg, wn, pscal, mapwn, mapth2, mapwn2,
                  jsea, isea, itstep, ntimes, nsea, npts, ncol, facth, fack};
  prop::KtGrid g{};
  int rc = check_ktp3(a, g, who);
  if (rc != WW_KOKKOS_OK) return rc;
  if (va == nullptr || cflthmax == nullptr || cflkmax == nullptr)
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": bad sizes or null pointer");
  if (ntimes == 0 || npts == 0) {
    report(g_last, WW_KOKKOS_OK, "");
    return WW_KOKKOS_OK;
  }
  KtCtx& x = kt();
  hipStream_t st = runtime().stream;
  const size_t bva = static_cast<size_t>(g.nspec) * ncol * sizeof(float), bc = static_cast<size_t>(ncol - 1) * sizeof(float);
  if (x.va.ensure(bva) != hipSuccess || x.cth.ensure(bc) != hipSuccess || x.ck.ensure(bc) != hipSuccess ||
      hipMemcpyAsync(x.va.p, va, bva, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(x.cth.p, cflthmax, bc, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(x.ck.p, cflkmax, bc, hipMemcpyHostToDevice, st) != hipSuccess)
    return fail(WW_KOKKOS_ERR_RUNTIME, std::string(who) + ": host-to-device copy");
  rc = run_ktp3(a, g, x.va.as<float>(), x.cth.as<float>(), x.ck.as<float>(), st, who);
  if (rc != WW_KOKKOS_OK) return rc;
  if (hipMemcpyAsync(va, x.va.p, bva, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(cflthmax, x.cth.p, bc, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(cflkmax, x.ck.p, bc, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(WW_KOKKOS_ERR_RUNTIME, std::string(who) + ": device-to-host copy");
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

int ww3gpu_ktp3_batch_dev(const float* rpar, int nrpar, const int* ipar, int nipar, const float* sig,
                          const float* dsip, const float* ecos, const float* esin, const float* ec2,
                          const float* esc, const float* es2, const int* mapwn, const int* mapth2,
                          const int* mapwn2, int itstep, int ntimes, float facth, float fack, int nsea,
                          const float* cg, const float* wn, int npts, const int* jsea, const int* isea,
                          const float* pscal, int ncol, float* d_va, float* d_cflthmax, float* d_cflkmax,
                          void* stream) {
  const char* who = "ww3gpu_ktp3_batch_dev";
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!runtime().ready) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, std::string(who) + " before ww_kokkos_init");
  if (nrpar != WW_KTP3_NRPAR || nipar != WW_KTP3_NIPAR) return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": bad packet");
  const Kt3Args a{rpar, ipar, sig, dsip, ecos, esin, ec2, esc, es2, cg, wn, pscal, mapwn, mapth2, mapwn2,
                  jsea, isea, itstep, ntimes, nsea, npts, ncol, facth, fack};
  prop::KtGrid g{};
  int rc = check_ktp3(a, g, who);
  if (rc != WW_KOKKOS_OK) return rc;
  if (d_va == nullptr || d_cflthmax == nullptr || d_cflkmax == nullptr)
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": null device array");
  if (ntimes == 0 || npts == 0) return WW_KOKKOS_OK;
  hipStream_t st = stream != nullptr ? static_cast<hipStream_t>(stream) : runtime().stream;
  rc = run_ktp3(a, g, d_va, d_cflthmax, d_cflkmax, st, who);
  if (rc == WW_KOKKOS_OK) report(g_last, WW_KOKKOS_OK, "");
  return rc;
}

int ww_ktp3_batch_va(int igrid, const float* rpar, int nrpar, const int* ipar, int nipar, const float* sig,
                     const float* dsip, const float* ecos, const float* esin, const float* ec2, const float* esc,
                     const float* es2, const int* mapwn, const int* mapth2, const int* mapwn2, int itstep,
                     int ntimes, float facth, float fack, int nsea, const float* cg, const float* wn, int npts,
                     const int* jsea, const int* isea, const float* pscal, float* va, int nspec, int nsealm,
                     float* cflthmax, float* cflkmax) {
  const char* who = "ww_ktp3_batch_va";
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!runtime().ready) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, std::string(who) + " before ww_kokkos_init");
  if (nrpar != WW_KTP3_NRPAR || nipar != WW_KTP3_NIPAR) return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": bad packet");
  const Kt3Args a{rpar, ipar, sig, dsip, ecos, esin, ec2, esc, es2, cg, wn, pscal, mapwn, mapth2, mapwn2,
                  jsea, isea, itstep, ntimes, nsea, npts, nsealm + 1, facth, fack};
  prop::KtGrid g{};
  int rc = check_ktp3(a, g, who);
  if (rc != WW_KOKKOS_OK) return rc;
  if (nspec != g.nspec || cflthmax == nullptr || cflkmax == nullptr)
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, std::string(who) + ": NSPEC differs from the packet, or a null array");
  if (ntimes == 0 || npts == 0) {
    report(g_last, WW_KOKKOS_OK, "");
    return WW_KOKKOS_OK;
  }
  float* d_va = nullptr;
  if ((rc = resident::device_va(igrid, va, nspec, nsealm, true, &d_va, who)) != WW_KOKKOS_OK) return rc;
  KtCtx& x = kt();
  hipStream_t st = runtime().stream;
  const size_t bc = static_cast<size_t>(nsealm) * sizeof(float);
  if (x.cth.ensure(bc) != hipSuccess || x.ck.ensure(bc) != hipSuccess ||
      hipMemcpyAsync(x.cth.p, cflthmax, bc, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(x.ck.p, cflkmax, bc, hipMemcpyHostToDevice, st) != hipSuccess)
    return fail(WW_KOKKOS_ERR_RUNTIME, std::string(who) + ": host-to-device copy");
  rc = run_ktp3(a, g, d_va, x.cth.as<float>(), x.ck.as<float>(), st, who);
  if (rc != WW_KOKKOS_OK) return rc;
  if (hipMemcpyAsync(cflthmax, x.cth.p, bc, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(cflkmax, x.ck.p, bc, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(WW_KOKKOS_ERR_RUNTIME, std::string(who) + ": device-to-host copy");
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

int ww3gpu_prop3_last_error(void) { return g_last.load(std::memory_order_relaxed); }

/// GPU milliseconds of the last W3XYP3 call (gather to store) and its cell-component
/// sub-steps (NX*NY per component per CFL sub-step) and the most sub-steps of a component; of
/// the last W3KTP3 call, its GPU milliseconds.
int ww3gpu_prop3_stats(double* out, int n) {
  if (out == nullptr || n < 4) return WW_KOKKOS_ERR_BAD_SHAPE;
  out[0] = xy().ms_gpu;
  out[1] = static_cast<double>(xy().cell_substeps);
  out[2] = xy().substeps;
  out[3] = kt().ms_gpu;
  return WW_KOKKOS_OK;
}

}  // extern "C"

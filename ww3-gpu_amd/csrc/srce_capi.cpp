// ww3-gpu_amd/csrc/srce_capi.cpp
// The ST4 and W3SRCE entry points of the C ABI: per-grid parameter packets and tables,
// the host side of INSIN4's saturation and cumulative tables, and the host loop that
// drives one time step's sub-steps on the device.
//
// Reference behaviour kept (kokkos/src/fortran_iface/st4_shim.cpp:178-209, 212-248,
// 480-580; srce_shim.cpp:679-760, 866-930): a grid is configured by init, array
// registrations (copied, LBOUND and type checked) and commit (every array the reference
// checks, with the same extents); module arrays (FWTABLE and the stress tables) and DELAB
// are one copy per process, the last registration wins; ww_srce_init needs the grid's DIA
// and ST4 contexts with the same NK and NTH, and ww_srce_batch checks them again at every
// call; validation failures write no output; npts == 0 is a no-op.
//
// The sub-step loop.  Every point of a batch integrates to DTG in its own number of
// sub-steps.  k_init runs W3SRCE's section 1 and the first sub-step's input terms (W3SLN1,
// W3SIN4, KDMEAN); then one device sub-step is the DIA over the points still integrating and
// k_step (W3SDS4, W3SDB1, W3SBT1, the time-step limit, the update, W3SPR4, the tail, the next
// sub-step's W3SIN4), which lists and counts the points that are not done.  The host reads
// that count (4 bytes), sizes the next launches by it and stops at zero; k_finish runs
// section 3.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "runtime.hpp"
#include "snl1.hpp"
#include "srce.hpp"
#include "resident.hpp"

namespace ww3gpu {
namespace st4h {
namespace {

// W3SRC4MD / W3GDATMD / CONSTANTS extents the commit checks (reference st4_params.hpp:39-52).
constexpr int kSizeFwTable = 300;
constexpr int kItauMax = 200, kJuMax = 200;
constexpr int kIustar = 100, kIalpha = 200, kIlevTail = 50;
constexpr int kNkhs = 2000, kNkd = 1300, kNdtab = 2000;

struct Spec {
  const char* name;
  bool is_int;
  int rank;
  int lbound;
  bool module;
};

// One row per WwSt4ArrayId: the registration contract of reference st4_arrays.hpp:64-84.
constexpr Spec kSpecs[WW_ST4_NARRAYS] = {
    {"SIG", false, 1, 0, false},        {"SIG2", false, 1, 1, false},
    {"DSIP", false, 1, 0, false},       {"DDEN", false, 1, 1, false},
    {"DDEN2", false, 1, 1, false},      {"ESIN", false, 1, 1, false},
    {"ECOS", false, 1, 1, false},       {"TH", false, 1, 1, false},
    {"SATWEIGHTS", false, 2, 1, false}, {"CUMULW", false, 2, 1, false},
    {"DCKI", false, 2, 1, false},       {"QBI", false, 2, 1, false},
    {"SATINDICES", true, 2, 1, false},  {"IKTAB", true, 2, 1, false},
    {"TAUT", false, 2, 0, true},        {"TAUHFT", false, 2, 0, true},
    {"TAUHFT2", false, 3, 0, true},     {"FWTABLE", false, 1, 0, true},
};

struct Arr {
  bool set = false;
  int rank = 0;
  size_t ext[3] = {0, 0, 0};
  std::vector<float> f;
  std::vector<int> i;
};

struct Grid {
  bool configured = false;
  bool ready = false;
  float r[WW_ST4_NRPAR] = {};
  int ip[WW_ST4_NIPAR] = {};
  Arr a[WW_ST4_NARRAYS];
  DevBuf tables;
  srce::St4Dev dev{};  ///< scalars and device pointers; fwtable and delab set per call
};

struct Module {
  Arr a[WW_ST4_NARRAYS];
  DevBuf fwtable;
  float delab = 0.f;
};

std::map<int, Grid>& grids() {
  static auto* m = new std::map<int, Grid>();
  return *m;
}

Module& module() {
  static auto* m = new Module();
  return *m;
}

std::atomic<int> g_last{WW_KOKKOS_OK};
bool g_hooked = false;

void release_all() {
  for (auto& kv : grids()) kv.second.tables.release();
  grids().clear();
  module().fwtable.release();
}

void hook() {
  if (!g_hooked) {
    on_finalize(release_all);
    g_hooked = true;
  }
}

int fnint(float x) {  // NINT of a REAL: round half away, 0 when out of range
  const float r = std::round(x);
  return (r > -9.2233715e18f && r < 9.2233715e18f) ? static_cast<int>(static_cast<long long>(r)) : 0;
}

template <class T>
int set_array(int igrid, int id, const T* data, int rank, const int* lbound, const int* extent) {
  constexpr bool is_int = std::is_same<T, int>::value;
  const auto it = grids().find(igrid);
  if (it == grids().end() || !it->second.configured) {
    report(g_last, WW_KOKKOS_ERR_NOT_INITIALISED, "ww_st4_set_* before ww_st4_init for this grid");
    return WW_KOKKOS_ERR_NOT_INITIALISED;
  }
  if (id < 0 || id >= WW_ST4_NARRAYS || data == nullptr || lbound == nullptr || extent == nullptr) {
    report(g_last, WW_KOKKOS_ERR_TABLE, "ww_st4_set_*: bad array id or null pointer");
    return WW_KOKKOS_ERR_TABLE;
  }
  const Spec& sp = kSpecs[id];
  if (sp.is_int != is_int || sp.rank != rank) {
    report(g_last, WW_KOKKOS_ERR_TABLE, (std::string("ww_st4_set_*: type or rank of ") + sp.name).c_str());
    return WW_KOKKOS_ERR_TABLE;
  }
  size_t n = 1;
  for (int r = 0; r < rank; ++r) {
    if (lbound[r] != sp.lbound || extent[r] <= 0) {
      report(g_last, WW_KOKKOS_ERR_TABLE,
             (std::string("ww_st4_set_*: LBOUND or extent of ") + sp.name).c_str());
      return WW_KOKKOS_ERR_TABLE;
    }
    n *= static_cast<size_t>(extent[r]);
  }
  Arr& dst = sp.module ? module().a[id] : it->second.a[id];
  dst.set = true;
  dst.rank = rank;
  for (int r = 0; r < 3; ++r) dst.ext[r] = r < rank ? static_cast<size_t>(extent[r]) : 0;
  if (is_int) {
    dst.i.assign(reinterpret_cast<const int*>(data), reinterpret_cast<const int*>(data) + n);
    dst.f.clear();
  } else {
    dst.f.assign(reinterpret_cast<const float*>(data), reinterpret_cast<const float*>(data) + n);
    dst.i.clear();
  }
  it->second.ready = false;
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

/// The extents the commit requires (reference st4_shim.cpp:212-248): at least the grid's
/// for the spectral arrays, exactly the allocation's for the tables.
bool check_arrays(const Grid& g, std::string& why) {
  const size_t nk = g.ip[WW_ST4_I_NK], nth = g.ip[WW_ST4_I_NTH], nspec = g.ip[WW_ST4_I_NSPEC];
  const size_t mk = g.ip[WW_ST4_I_MK], mth = g.ip[WW_ST4_I_MTH], mspec = g.ip[WW_ST4_I_MSPEC];
  const size_t nsat = 2 * (mth / 2 - 1) + 1;
  const Module& m = module();
  auto need = [&](const Arr& v, int id, size_t e0, size_t e1, size_t e2, bool exact) {
    if (!v.set) {
      why = std::string(kSpecs[id].name) + " not registered";
      return false;
    }
    const size_t e[3] = {e0, e1, e2};
    for (int r = 0; r < v.rank; ++r)
      if (exact ? v.ext[r] != e[r] : v.ext[r] < e[r]) {
        why = std::string(kSpecs[id].name) + " has the wrong shape";
        return false;
      }
    return true;
  };
  bool ok = need(g.a[WW_ST4_A_SIG], WW_ST4_A_SIG, nk + 2, 0, 0, false) &&
            need(g.a[WW_ST4_A_SIG2], WW_ST4_A_SIG2, nspec, 0, 0, false) &&
            need(g.a[WW_ST4_A_DSIP], WW_ST4_A_DSIP, nk + 2, 0, 0, false) &&
            need(g.a[WW_ST4_A_DDEN], WW_ST4_A_DDEN, nk, 0, 0, false) &&
            need(g.a[WW_ST4_A_DDEN2], WW_ST4_A_DDEN2, nspec, 0, 0, false) &&
            need(g.a[WW_ST4_A_ESIN], WW_ST4_A_ESIN, nspec + nth, 0, 0, false) &&
            need(g.a[WW_ST4_A_ECOS], WW_ST4_A_ECOS, nspec + nth, 0, 0, false) &&
            need(g.a[WW_ST4_A_TH], WW_ST4_A_TH, nth, 0, 0, false) &&
            need(g.a[WW_ST4_A_SATWEIGHTS], WW_ST4_A_SATWEIGHTS, nsat, mth, 0, true) &&
            need(g.a[WW_ST4_A_SATINDICES], WW_ST4_A_SATINDICES, nsat, mth, 0, true) &&
            need(g.a[WW_ST4_A_CUMULW], WW_ST4_A_CUMULW, mspec, mspec, 0, true) &&
            need(g.a[WW_ST4_A_DCKI], WW_ST4_A_DCKI, kNkhs, kNkd, 0, true) &&
            need(g.a[WW_ST4_A_QBI], WW_ST4_A_QBI, kNkhs, kNkd, 0, true) &&
            need(g.a[WW_ST4_A_IKTAB], WW_ST4_A_IKTAB, mk, kNdtab, 0, true) &&
            need(m.a[WW_ST4_A_FWTABLE], WW_ST4_A_FWTABLE, kSizeFwTable + 1, 0, 0, true);
  if (ok && g.r[WW_ST4_R_SINTAILPAR] > 0.5f) {
    ok = need(m.a[WW_ST4_A_TAUT], WW_ST4_A_TAUT, kItauMax + 1, kJuMax + 1, 0, true) &&
         need(m.a[WW_ST4_A_TAUHFT], WW_ST4_A_TAUHFT, kIustar + 1, kIalpha + 1, 0, true);
    if (ok && g.r[WW_ST4_R_TTAUWSHELTER] > 0.f)
      ok = need(m.a[WW_ST4_A_TAUHFT2], WW_ST4_A_TAUHFT2, kIustar + 1, kIalpha + 1, kIlevTail + 1, true);
  }
  return ok;
}

}  // namespace

/// INSIN4's SATWEIGHTS/SATINDICES and CUMULW (reference st4_insin4.cpp:110-138, 239-262),
/// host libm: COS and REAL**REAL of the Fortran are cosf and powf, SQRT is sqrtf.
/// satw/sati: (nsat, mth), nsat = 2*(MTH/2-1)+1, column-major; cumulw: (mspec, mspec).
void insin4_tables(const float* r, const int* ip, const float* sig0, const float* dsip0,
                   const float* ecos, const float* th, float* satw, int* sati, float* cumulw) {
  const int nk = ip[WW_ST4_I_NK], nth = ip[WW_ST4_I_NTH];
  const int mth = ip[WW_ST4_I_MTH], mspec = ip[WW_ST4_I_MSPEC];
  const int nsat = 2 * (mth / 2 - 1) + 1;
  const float dth = r[WW_ST4_R_DTH];
  const float pi = static_cast<float>(3.141592653589793);
  const float rade = 180.f / pi;
  const size_t nsat_all = static_cast<size_t>(nsat) * mth;
  for (size_t i = 0; i < nsat_all; ++i) sati[i] = 1;
  if (r[WW_ST4_R_SSDSDTH] < 180.f) {
    for (size_t i = 0; i < nsat_all; ++i) satw[i] = 0.f;
    int sdsnth = fnint(r[WW_ST4_R_SSDSDTH] / (dth * rade));
    if (sdsnth > nth / 2 - 1) sdsnth = nth / 2 - 1;
    for (int ith = 1; ith <= nth; ++ith)
      for (int ii = ith - sdsnth; ii <= ith + sdsnth; ++ii) {
        const int jj = ii < 1 ? ii + nth : (ii > nth ? ii - nth : ii);
        const size_t at = static_cast<size_t>(ii - (ith - sdsnth)) + static_cast<size_t>(nsat) * (ith - 1);
        sati[at] = jj;
        satw[at] = std::pow(std::cos(th[ith - 1] - th[jj - 1]), r[WW_ST4_R_SSDSCOS]);
      }
  } else {
    for (size_t i = 0; i < nsat_all; ++i) satw[i] = 1.f;
  }
  std::memset(cumulw, 0, sizeof(float) * static_cast<size_t>(mspec) * mspec);
  if (r[WW_ST4_R_SSDSC + 2] < 0.f) {
    const float grav = 9.806f;
    const int dik = fnint(r[WW_ST4_R_SSDSBRF1] / (r[WW_ST4_R_XFR] - 1.f));
    for (int ik = 1; ik <= nk; ++ik) {
      const float c = grav / sig0[ik];
      for (int ith = 1; ith <= nth; ++ith) {
        const size_t is = static_cast<size_t>(ith + (ik - 1) * nth);
        for (int ik2 = 1; ik2 <= ik - dik; ++ik2) {
          const float c2 = grav / sig0[ik2];
          for (int ith2 = 1; ith2 <= nth; ++ith2) {
            const size_t is2 = static_cast<size_t>(ith2 + (ik2 - 1) * nth);
            const float cw = ecos[std::abs(ith2 - ith)];
            cumulw[(is2 - 1) + static_cast<size_t>(mspec) * (is - 1)] =
                std::sqrt(c * c + c2 * c2 - 2.f * c * c2 * cw) * dsip0[ik2] / (0.5f * c2) * dth;
          }
        }
      }
    }
  }
}

namespace {

/// Device copies of the grid arrays the kernels read, and the St4Dev that points at them.
int upload(Grid& g) {
  const int nk = g.ip[WW_ST4_I_NK], nth = g.ip[WW_ST4_I_NTH], nspec = g.ip[WW_ST4_I_NSPEC];
  const int mspec = g.ip[WW_ST4_I_MSPEC];
  const Arr& sw = g.a[WW_ST4_A_SATWEIGHTS];
  const int nsat = static_cast<int>(sw.ext[0]);
  // Layout of the block, in 4-byte words, each array 64-word aligned.
  size_t off = 0;
  auto take = [&](size_t n) {
    const size_t o = off;
    off += (n + 63) / 64 * 64;
    return o;
  };
  const size_t o_sig = take(nk + 2), o_sig2 = take(nspec), o_dsip = take(nk + 2), o_dden = take(nk);
  const size_t o_dden2 = take(nspec), o_esin = take(nspec + nth), o_ecos = take(nspec + nth);
  const size_t o_satw = take(static_cast<size_t>(nsat) * nth);
  const size_t o_sati = take(static_cast<size_t>(nsat) * nth);
  const size_t o_cum = take(static_cast<size_t>(nk) * nk * nth);
  const int ms = srce::kMirStride(nth);
  const size_t o_mir = take(static_cast<size_t>(nk) * nk * ms);
  std::vector<float> h(off, 0.f);
  auto put = [&](size_t o, const Arr& a, size_t n) { std::memcpy(&h[o], a.f.data(), n * sizeof(float)); };
  put(o_sig, g.a[WW_ST4_A_SIG], nk + 2);
  put(o_sig2, g.a[WW_ST4_A_SIG2], nspec);
  put(o_dsip, g.a[WW_ST4_A_DSIP], nk + 2);
  put(o_dden, g.a[WW_ST4_A_DDEN], nk);
  put(o_dden2, g.a[WW_ST4_A_DDEN2], nspec);
  put(o_esin, g.a[WW_ST4_A_ESIN], nspec + nth);
  put(o_ecos, g.a[WW_ST4_A_ECOS], nspec + nth);
  put(o_satw, sw, static_cast<size_t>(nsat) * nth);
  std::memcpy(&h[o_sati], g.a[WW_ST4_A_SATINDICES].i.data(), sizeof(int) * static_cast<size_t>(nsat) * nth);
  // CUMULW(IS2, IS) for IS = (IK, ITH), IS2 = (IK2, ITH2) depends on IK, IK2 and |ITH2 - ITH|
  // only (INSIN4's construction): keep one row per (IK, IK2), read from ITH = 1, and check
  // every registered entry against it.
  const std::vector<float>& cw = g.a[WW_ST4_A_CUMULW].f;
  const auto at = [&](int is, int is2) { return cw[static_cast<size_t>(is2) + static_cast<size_t>(mspec) * is]; };
  for (int ik = 0; ik < nk; ++ik)
    for (int ik2 = 0; ik2 < nk; ++ik2)
      for (int d = 0; d < nth; ++d)
        h[o_cum + (static_cast<size_t>(ik) * nk + ik2) * nth + d] = at(ik * nth, ik2 * nth + d);
  for (int is = 0; is < nspec; ++is)
    for (int is2 = 0; is2 < nspec; ++is2) {
      const int ik = is / nth, ith = is % nth, ik2 = is2 / nth, ith2 = is2 % nth;
      const float v = at(is, is2);
      const float k = h[o_cum + (static_cast<size_t>(ik) * nk + ik2) * nth + std::abs(ith2 - ith)];
      if (std::memcmp(&v, &k, sizeof v) != 0) return WW_KOKKOS_ERR_UNSUPPORTED;
    }
  for (size_t r = 0; r < static_cast<size_t>(nk) * nk; ++r)
    for (int m = 0; m < 2 * nth - 1; ++m)
      h[o_mir + r * ms + m] = h[o_cum + r * nth + std::abs(m - (nth - 1))];
  if (g.tables.ensure(off * sizeof(float)) != hipSuccess ||
      hipMemcpy(g.tables.p, h.data(), off * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return WW_KOKKOS_ERR_RUNTIME;
  // The module's FWTABLE: one device copy, replaced at every commit.
  Module& m = module();
  const Arr& fw = m.a[WW_ST4_A_FWTABLE];
  if (m.fwtable.ensure(fw.f.size() * sizeof(float)) != hipSuccess ||
      hipMemcpy(m.fwtable.p, fw.f.data(), fw.f.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return WW_KOKKOS_ERR_RUNTIME;

  srce::St4Dev& d = g.dev;
  const float* r = g.r;
  d = srce::St4Dev{};
  d.nk = nk;
  d.nth = nth;
  d.nspec = nspec;
  d.stab3 = g.ip[WW_ST4_I_STAB3] != 0;
  d.ig1 = g.ip[WW_ST4_I_IG1] != 0;
  d.nsat = nsat;
  d.dth = r[WW_ST4_R_DTH];
  d.xfr = r[WW_ST4_R_XFR];
  d.fte = r[WW_ST4_R_FTE];
  d.ftf = r[WW_ST4_R_FTF];
  d.aaircmin = r[WW_ST4_R_AAIRCMIN];
  d.aairgb = r[WW_ST4_R_AAIRGB];
  d.igpars5 = r[WW_ST4_R_IGPARS5];
  d.wwnmeanptail = r[WW_ST4_R_WWNMEANPTAIL];
  d.sstxftftail = r[WW_ST4_R_SSTXFTFTAIL];
  d.aalpha = r[WW_ST4_R_AALPHA];
  d.bbeta = r[WW_ST4_R_BBETA];
  d.zz0max = r[WW_ST4_R_ZZ0MAX];
  d.zz0rat = r[WW_ST4_R_ZZ0RAT];
  d.zzalp = r[WW_ST4_R_ZZALP];
  d.ssinthp = r[WW_ST4_R_SSINTHP];
  d.ttauwshelter = r[WW_ST4_R_TTAUWSHELTER];
  for (int j = 0; j < 7; ++j) d.sswellf[j] = r[WW_ST4_R_SSWELLF + j];
  for (int j = 0; j < 21; ++j) d.ssdsc[j] = r[WW_ST4_R_SSDSC + j];
  d.ssdsbr = r[WW_ST4_R_SSDSBR];
  for (int j = 0; j < 5; ++j) d.sintailpar[j] = r[WW_ST4_R_SINTAILPAR + j];
  d.ssdsp = r[WW_ST4_R_SSDSP];
  d.wwnmeanp = r[WW_ST4_R_WWNMEANP];
  d.sstxftwn = r[WW_ST4_R_SSTXFTWN];
  d.ssdsbrf1 = r[WW_ST4_R_SSDSBRF1];
  for (int j = 0; j < 10; ++j) d.capchnk[j] = r[WW_ST4_R_CAPCHNK + j];
  d.zzwnd = r[WW_ST4_R_ZZWND];
  d.ssdsdth = r[WW_ST4_R_SSDSDTH];
  for (int j = 0; j < 5; ++j) d.ssdsbm[j] = r[WW_ST4_R_SSDSBM + j];
  const float* b = g.tables.as<float>();
  d.sig = b + o_sig;
  d.sig2 = b + o_sig2;
  d.dsip = b + o_dsip;
  d.dden = b + o_dden;
  d.dden2 = b + o_dden2;
  d.esin = b + o_esin;
  d.ecos = b + o_ecos;
  d.satw = b + o_satw;
  d.sati = reinterpret_cast<const int*>(b + o_sati);
  d.cumulk = b + o_cum;
  d.cumulm = b + o_mir;
  d.fwtable = nullptr;
  return WW_KOKKOS_OK;
}

/// The configurations the kernels translate (others are refused at commit).
bool supported(const Grid& g, std::string& why) {
  if (g.r[WW_ST4_R_SINTAILPAR] > 0.5f) {
    why = "SINTAILPAR(1) > 0.5 (tabulated u* and tail stress) is not supported";
    return false;
  }
  if (fnint(g.r[WW_ST4_R_SSDSC]) != 1) {
    why = "SDSBCHOICE = NINT(SSDSC(1)) other than 1 is not supported";
    return false;
  }
  // The friction-factor lookup indexes FWTABLE with (LOG10(...) + 1) / DELAB.
  if (g.r[WW_ST4_R_SSWELLF + 1] != 0.f && !(module().delab > 0.f)) {
    why = "SWELLF(2) /= 0 needs DELAB > 0 for the FWTABLE lookup";
    return false;
  }
  return true;
}

}  // namespace

/// Grid igrid's committed ST4 state for W3SRCE, with the module's FWTABLE and DELAB as of
/// this call.  False when the grid is not committed.
bool grid_state(int igrid, srce::St4Dev& out) {
  if (!runtime().ready) return false;
  const auto it = grids().find(igrid);
  if (it == grids().end() || !it->second.ready) return false;
  out = it->second.dev;
  out.fwtable = module().fwtable.as<float>();
  out.delab = module().delab;
  return true;
}

int last_status() { return g_last.load(std::memory_order_relaxed); }

/// One host array of a per-routine call: copied to the device before the kernel, and back
/// after it when `out` (every OUT argument is copied in too: ww_kokkos_c.hpp:149-152).
struct HostArr {
  void* h;
  size_t bytes;
  bool out;
  void* d = nullptr;
};

/// ww_spr4 / ww_sin4 / ww_sds4: stage the arrays, let `bind` point the call at the device
/// copies, run the routine, copy the outputs back.
template <class Bind>
void routine_with(int igrid, int npts, srce::St4Call& io, std::vector<HostArr>& arrs, const Bind& bind,
                  const char* who) {
  srce::St4Dev q;
  if (!grid_state(igrid, q)) {
    report(g_last, WW_KOKKOS_ERR_NOT_INITIALISED, (std::string(who) + ": grid not committed").c_str());
    return;
  }
  if (npts < 0) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, (std::string(who) + ": npts < 0").c_str());
    return;
  }
  for (const HostArr& x : arrs)
    if (npts > 0 && x.h == nullptr) {
      report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, (std::string(who) + ": a null array").c_str());
      return;
    }
  if (npts == 0) {
    report(g_last, WW_KOKKOS_OK, "");
    return;
  }
  const size_t n = static_cast<size_t>(npts), nspec = q.nspec, nk = q.nk;
  const auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  size_t total = 0;
  for (const HostArr& x : arrs) total += al(x.bytes);
  const size_t sc_pin = al(n * srce::N_PIN * 4), sc_l8 = al(n * nspec), sc_rnw = al(n * nspec * 4),
               sc_k = al(n * nk * 4);
  total += sc_pin + sc_l8 + sc_rnw + 2 * sc_k;
  static DevBuf* buf = new DevBuf();
  hipStream_t st = runtime().stream;
  if (buf->ensure(total) != hipSuccess) {
    report(g_last, WW_KOKKOS_ERR_RUNTIME, (std::string(who) + ": device allocation").c_str());
    return;
  }
  char* cur = static_cast<char*>(buf->p);
  for (HostArr& x : arrs) {
    x.d = cur;
    if (hipMemcpyAsync(cur, x.h, x.bytes, hipMemcpyHostToDevice, st) != hipSuccess) {
      report(g_last, WW_KOKKOS_ERR_RUNTIME, (std::string(who) + ": host-to-device copy").c_str());
      return;
    }
    cur += al(x.bytes);
  }
  io.pin = reinterpret_cast<float*>(cur);
  cur += sc_pin;
  io.llws8 = reinterpret_cast<unsigned char*>(cur);
  cur += sc_l8;
  io.rnw = reinterpret_cast<float*>(cur);
  cur += sc_rnw;
  io.mft = reinterpret_cast<float*>(cur);
  cur += sc_k;
  io.kband = reinterpret_cast<float*>(cur);
  io.npts = npts;
  bind();
  hipError_t e = srce::k_routine(io, q, st);
  if (e == hipSuccess)
    for (const HostArr& x : arrs)
      if (x.out && e == hipSuccess) e = hipMemcpyAsync(x.h, x.d, x.bytes, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    report(g_last, WW_KOKKOS_ERR_KERNEL, (std::string(who) + ": " + hipGetErrorString(e)).c_str());
    return;
  }
  report(g_last, WW_KOKKOS_OK, "");
}

}  // namespace st4h

// ======================================================================================
// W3SRCE
// ======================================================================================
namespace srceh {
namespace {

using srce::SrceDev;
using srce::St4Dev;
using srce::State;
using srce::StepFlags;

/// Timer ids of ww3gpu_srce_timers: init, DIA, sub-step, finish.
enum Timer : int { T_INIT = 0, T_DIA, T_STEP, T_FINISH, kTimers };

struct Grid {
  bool ready = false;
  float r[WW_SRCE_NRPAR] = {};
  int ip[WW_SRCE_NIPAR] = {};
  SrceDev dev{};
  DevBuf tables;  ///< sig, dden, ecos, esin, mapwn
  DevBuf state;   ///< the resident state of the last batch
  DevBuf stage;   ///< device copies of a host-array call's arrays
  DevBuf cols;    ///< the resident-VA columns of a ww_srce_batch_va call
  int* counter = nullptr;  ///< pinned: the host's copy of one sub-step's active count
  int substeps = -1;
  long point_substeps = -1;  ///< sum over the last call's sub-steps of the points integrated
  // Per-kernel timing (ww3gpu_srce_timing): event pairs around each launch on the call's
  // stream, resolved when the timers are read.
  bool timing = false;
  std::vector<hipEvent_t> events;  ///< pool, reused call after call
  size_t used = 0;                 ///< events of the pool recorded since the last resolve
  std::vector<int> pending;        ///< timer id of each recorded pair
  double ms[kTimers] = {};
  long launches[kTimers] = {};
};

std::map<int, Grid>& grids() {
  static auto* m = new std::map<int, Grid>();
  return *m;
}

std::atomic<int> g_last{WW_KOKKOS_OK};
bool g_hooked = false;

void release_grid(Grid& g) {
  for (hipEvent_t e : g.events) (void)hipEventDestroy(e);
  g.events.clear();
  g.used = 0;
  g.pending.clear();
  g.tables.release();
  g.state.release();
  g.stage.release();
  if (g.counter != nullptr) (void)hipHostFree(g.counter);
  g.counter = nullptr;
}

void release_all() {
  for (auto& kv : grids()) release_grid(kv.second);
  grids().clear();
}

Grid* find(int igrid) {
  if (!runtime().ready) return nullptr;
  const auto it = grids().find(igrid);
  if (it == grids().end() || !it->second.ready) return nullptr;
  return &it->second;
}

size_t al64(size_t n) { return (n + 63) / 64 * 64; }

/// Record the start (begin = true) or the end of a timed launch on `st`.
void mark(Grid& g, int id, bool begin, hipStream_t st) {
  if (!g.timing) return;
  if (g.used == g.events.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) {
      g.timing = false;
      return;
    }
    g.events.push_back(e);
  }
  (void)hipEventRecord(g.events[g.used++], st);
  if (begin) g.pending.push_back(id);
}

/// Wait for the recorded events and add their intervals to the timers.
void resolve_timers(Grid& g) {
  if (g.used == 0) return;
  (void)hipEventSynchronize(g.events[g.used - 1]);
  for (size_t k = 0; k < g.pending.size() && 2 * k + 1 < g.used; ++k) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g.events[2 * k], g.events[2 * k + 1]) == hipSuccess) {
      g.ms[g.pending[k]] += ms;
      g.launches[g.pending[k]] += 1;
    }
  }
  g.used = 0;
  g.pending.clear();
}

/// The context a compute call needs: W3SRCE grid, ST4 state, DIA view, consistent shapes.
int resolve(int igrid, const char* who, Grid*& g, St4Dev& q, snl1::GridView& dv) {
  g = find(igrid);
  if (g == nullptr) {
    report(g_last, WW_KOKKOS_ERR_NOT_INITIALISED, (std::string(who) + " before ww_srce_init for this grid").c_str());
    return WW_KOKKOS_ERR_NOT_INITIALISED;
  }
  if (!st4h::grid_state(igrid, q) || !snl1::grid_view(igrid, dv)) {
    report(g_last, WW_KOKKOS_ERR_NOT_INITIALISED,
           (std::string(who) + ": the DIA or ST4 context of this grid is gone").c_str());
    return WW_KOKKOS_ERR_NOT_INITIALISED;
  }
  if (dv.g->nk != g->dev.nk || dv.g->nth != g->dev.nth || q.nk != g->dev.nk || q.nth != g->dev.nth) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, (std::string(who) + ": NK/NTH differ from the DIA or ST4 context").c_str());
    return WW_KOKKOS_ERR_BAD_SHAPE;
  }
  return WW_KOKKOS_OK;
}

/// One call's sub-step loop on device arrays (point-major, the C ABI's shapes).
int run(Grid& g, const St4Dev& q, const snl1::GridView& dv, const StepFlags& f, int npts, float* spec,
        const float* wn1, const float* cg1, const float* pin, float* psc, float* whitecap,
        float* taubbl, float* tauice, int* shave, hipStream_t st, const char* who) {
  const int nk = g.dev.nk, nth = g.dev.nth;
  if (g.state.ensure(srce::state_bytes(nk, nth, npts)) != hipSuccess) {
    report(g_last, WW_KOKKOS_ERR_RUNTIME, (std::string(who) + ": device allocation").c_str());
    return WW_KOKKOS_ERR_RUNTIME;
  }
  if (g.counter == nullptr && hipHostMalloc(reinterpret_cast<void**>(&g.counter), sizeof(int)) != hipSuccess) {
    g.counter = nullptr;
    report(g_last, WW_KOKKOS_ERR_RUNTIME, (std::string(who) + ": pinned allocation").c_str());
    return WW_KOKKOS_ERR_RUNTIME;
  }
  State s = srce::carve_state(g.state.p, nk, nth, npts);
  s.spec = spec;
  s.wn1 = wn1;
  s.cg1 = cg1;
  s.pin = pin;
  s.psc = psc;
  s.whitecap = whitecap;
  s.taubbl = taubbl;
  s.tauice = tauice;
  s.shave = shave;
  auto fail = [&](hipError_t e, const char* what) {
    std::string m = std::string(who) + ": " + what + ": " + hipGetErrorString(e);
    report(g_last, WW_KOKKOS_ERR_KERNEL, m.c_str());
    return WW_KOKKOS_ERR_KERNEL;
  };
  resolve_timers(g);
  mark(g, T_INIT, true, st);
  hipError_t e = srce::k_init(s, q, g.dev, f, st);
  mark(g, T_INIT, false, st);
  if (e != hipSuccess) return fail(e, "init");
  int iter = 0, n = npts;
  long pts = 0;
  for (;; ++iter) {
    if (iter == srce::kMaxIter) {
      report(g_last, WW_KOKKOS_ERR_KERNEL,
             (std::string(who) + ": the sub-step loop did not reach DTG within its step limit").c_str());
      return WW_KOKKOS_ERR_KERNEL;
    }
    pts += n;
    // The DIA of the points still integrating (the list the last sub-step left), dealt to
    // the 8 XCDs in contiguous runs of list slots.
    const snl1::Layout L{1, dv.g->nspec, 1, nk, nullptr, (n + 7) / 8, s.list[iter & 1]};
    mark(g, T_DIA, true, st);
    e = snl1::launch(*dv.g, dv.af11_row, dv.sig, n, s.spec, s.cg1, s.kdmean, s.vsnl, s.vdnl, st,
                     dv.variant, L);
    mark(g, T_DIA, false, st);
    if (e != hipSuccess) return fail(e, "sub-step (DIA)");
    mark(g, T_STEP, true, st);
    e = srce::k_step(s, q, g.dev, f, iter, n, st);
    mark(g, T_STEP, false, st);
    if (e != hipSuccess) return fail(e, "sub-step");
    e = hipMemcpyAsync(g.counter, s.count + iter, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(e, "sub-step count");
    n = *g.counter;
    if (n == 0) break;
  }
  g.substeps = iter + 1;
  g.point_substeps = pts;
  mark(g, T_FINISH, true, st);
  e = srce::k_finish(s, q, g.dev, f, st);
  mark(g, T_FINISH, false, st);
  if (e != hipSuccess) return fail(e, "finish");
  return WW_KOKKOS_OK;
}

bool args_ok(int npts, const void* spec, const void* wn1, const void* cg1, const void* pin, const void* psc,
             const void* whitecap, const void* taubbl, const void* tauice, const void* shave) {
  return npts >= 0 && (npts == 0 || (spec != nullptr && wn1 != nullptr && cg1 != nullptr && pin != nullptr &&
                                     psc != nullptr && whitecap != nullptr && taubbl != nullptr &&
                                     tauice != nullptr && shave != nullptr));
}

int check_step(const St4Dev& q, int it, const char* who) {
  if (it != 0 && !(q.sintailpar[3] > 0.5f)) {
    report(g_last, WW_KOKKOS_ERR_UNSUPPORTED,
           (std::string(who) + ": IT > 0 with TAUWBUG off reads LLWS before it is set").c_str());
    return WW_KOKKOS_ERR_UNSUPPORTED;
  }
  return WW_KOKKOS_OK;
}

/// ww_srce_batch with host arrays; with d_va, the spectra are instead the columns jsea of the
/// resident VA on the device (ww_srce_batch_va), gathered before and scattered after.
int batch_host(int igrid, const StepFlags& f, int npts, float* spec, const float* wn1, const float* cg1,
               const float* pin, float* psc, float* whitecap, float* taubbl, float* tauice, int* shave,
               const char* who, float* d_va = nullptr, const int* jsea = nullptr) {
  Grid* g = nullptr;
  St4Dev q;
  snl1::GridView dv;
  int rc = resolve(igrid, who, g, q, dv);
  if (rc != WW_KOKKOS_OK) return rc;
  if (!args_ok(npts, d_va != nullptr ? d_va : spec, wn1, cg1, pin, psc, whitecap, taubbl, tauice, shave)) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, (std::string(who) + ": npts < 0 or a null array").c_str());
    return WW_KOKKOS_ERR_BAD_SHAPE;
  }
  if ((rc = check_step(q, f.it, who)) != WW_KOKKOS_OK) return rc;
  if (npts == 0) {
    report(g_last, WW_KOKKOS_OK, "");
    return WW_KOKKOS_OK;
  }
  const size_t n = static_cast<size_t>(npts);
  const size_t nspec = g->dev.nspec, nk = g->dev.nk;
  const size_t w_spec = al64(nspec * n), w_k = al64(nk * n), w_pin = al64(WW_SRCE_NPIN * n);
  const size_t w_psc = al64(WW_SRCE_NPSC * n), w_4 = al64(4 * n), w_2 = al64(2 * n), w_1 = al64(n);
  const size_t words = w_spec + 2 * w_k + w_pin + w_psc + w_4 + 2 * w_2 + w_1;
  if (g->stage.ensure(words * sizeof(float)) != hipSuccess) {
    report(g_last, WW_KOKKOS_ERR_RUNTIME, (std::string(who) + ": device allocation").c_str());
    return WW_KOKKOS_ERR_RUNTIME;
  }
  float* b = g->stage.as<float>();
  float* d_spec = b;
  float* d_wn1 = d_spec + w_spec;
  float* d_cg1 = d_wn1 + w_k;
  float* d_pin = d_cg1 + w_k;
  float* d_psc = d_pin + w_pin;
  float* d_wc = d_psc + w_psc;
  float* d_tb = d_wc + w_4;
  float* d_ti = d_tb + w_2;
  int* d_sh = reinterpret_cast<int*>(d_ti + w_2);
  hipStream_t st = runtime().stream;
  const auto h2d = [&](void* d, const void* h, size_t bytes) {
    return hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st) == hipSuccess;
  };
  if (d_va != nullptr) {
    if (g->cols.ensure(n * sizeof(int)) != hipSuccess || !h2d(g->cols.p, jsea, n * sizeof(int)) ||
        resident::gather_cols(d_va, static_cast<int>(nspec), g->cols.as<int>(), npts, d_spec, st) != hipSuccess) {
      report(g_last, WW_KOKKOS_ERR_RUNTIME, (std::string(who) + ": gathering the resident columns").c_str());
      return WW_KOKKOS_ERR_RUNTIME;
    }
  }
  if ((d_va == nullptr && !h2d(d_spec, spec, nspec * n * 4)) || !h2d(d_wn1, wn1, nk * n * 4) ||
      !h2d(d_cg1, cg1, nk * n * 4) ||
      !h2d(d_pin, pin, WW_SRCE_NPIN * n * 4) || !h2d(d_psc, psc, WW_SRCE_NPSC * n * 4) ||
      !h2d(d_wc, whitecap, 4 * n * 4) || !h2d(d_tb, taubbl, 2 * n * 4) || !h2d(d_ti, tauice, 2 * n * 4)) {
    report(g_last, WW_KOKKOS_ERR_RUNTIME, (std::string(who) + ": host-to-device copy").c_str());
    return WW_KOKKOS_ERR_RUNTIME;
  }
  rc = run(*g, q, dv, f, npts, d_spec, d_wn1, d_cg1, d_pin, d_psc, d_wc, d_tb, d_ti, d_sh, st, who);
  if (rc != WW_KOKKOS_OK) {
    (void)hipStreamSynchronize(st);
    return rc;
  }
  const auto d2h = [&](void* h, const void* d, size_t bytes) {
    return hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, st) == hipSuccess;
  };
  const bool spec_back =
      d_va != nullptr
          ? resident::scatter_cols(d_va, static_cast<int>(nspec), g->cols.as<int>(), npts, d_spec, st) == hipSuccess
          : d2h(spec, d_spec, nspec * n * 4);
  if (!spec_back || !d2h(psc, d_psc, WW_SRCE_NPSC * n * 4) ||
      !d2h(whitecap, d_wc, 4 * n * 4) || !d2h(taubbl, d_tb, 2 * n * 4) || !d2h(tauice, d_ti, 2 * n * 4) ||
      !d2h(shave, d_sh, n * 4) || hipStreamSynchronize(st) != hipSuccess) {
    report(g_last, WW_KOKKOS_ERR_RUNTIME, (std::string(who) + ": device-to-host copy").c_str());
    return WW_KOKKOS_ERR_RUNTIME;
  }
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

}  // namespace
}  // namespace srceh
}  // namespace ww3gpu

using namespace ww3gpu;

extern "C" {

// ------------------------------------------------------------------------------------
// ST4
// ------------------------------------------------------------------------------------
int ww_st4_init(int igrid, const float* rpar, int nrpar, const int* ipar, int nipar) {
  using namespace ww3gpu::st4h;
  grids().erase(igrid);
  if (!runtime().ready) {
    report(g_last, WW_KOKKOS_ERR_NOT_INITIALISED, "ww_st4_init before ww_kokkos_init");
    return WW_KOKKOS_ERR_NOT_INITIALISED;
  }
  hook();
  if (nrpar != WW_ST4_NRPAR || nipar != WW_ST4_NIPAR || rpar == nullptr || ipar == nullptr) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww_st4_init: RPAR/IPAR length differs from the ST4 packet");
    return WW_KOKKOS_ERR_BAD_SHAPE;
  }
  if (ipar[WW_ST4_I_FLX5] != 0) {
    report(g_last, WW_KOKKOS_ERR_UNSUPPORTED, "ww_st4_init: W3_FLX5 is not supported");
    return WW_KOKKOS_ERR_UNSUPPORTED;
  }
  const int nk = ipar[WW_ST4_I_NK], nth = ipar[WW_ST4_I_NTH], nspec = ipar[WW_ST4_I_NSPEC];
  if (nk <= 0 || nth <= 0 || nspec != nk * nth || ipar[WW_ST4_I_MK] < nk || ipar[WW_ST4_I_MTH] < nth ||
      ipar[WW_ST4_I_MSPEC] < nspec) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww_st4_init: NK, NTH, NSPEC, MK, MTH, MSPEC inconsistent");
    return WW_KOKKOS_ERR_BAD_SHAPE;
  }
  Grid& g = grids()[igrid];
  std::memcpy(g.r, rpar, sizeof g.r);
  std::memcpy(g.ip, ipar, sizeof g.ip);
  g.configured = true;
  module().delab = rpar[WW_ST4_R_DELAB];  // module scalar: the last grid to initialise wins
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

int ww_st4_set_real(int igrid, int id, const float* data, int rank, const int* lbound, const int* extent) {
  return ww3gpu::st4h::set_array<float>(igrid, id, data, rank, lbound, extent);
}

int ww_st4_set_int(int igrid, int id, const int* data, int rank, const int* lbound, const int* extent) {
  return ww3gpu::st4h::set_array<int>(igrid, id, data, rank, lbound, extent);
}

int ww_st4_commit(int igrid) {
  using namespace ww3gpu::st4h;
  const auto it = grids().find(igrid);
  if (it == grids().end() || !it->second.configured || !runtime().ready) {
    report(g_last, WW_KOKKOS_ERR_NOT_INITIALISED, "ww_st4_commit before ww_st4_init for this grid");
    return WW_KOKKOS_ERR_NOT_INITIALISED;
  }
  Grid& g = it->second;
  std::string why;
  if (!check_arrays(g, why)) {
    report(g_last, WW_KOKKOS_ERR_TABLE, ("ww_st4_commit: " + why).c_str());
    return WW_KOKKOS_ERR_TABLE;
  }
  if (!supported(g, why)) {
    report(g_last, WW_KOKKOS_ERR_UNSUPPORTED, ("ww_st4_commit: " + why).c_str());
    return WW_KOKKOS_ERR_UNSUPPORTED;
  }
  // The saturation integral gathers A(IS0 + SATINDICES(J,ITH)): every entry must name a
  // direction of the grid, or the kernel would read outside the point's spectrum.
  for (const int v : g.a[WW_ST4_A_SATINDICES].i)
    if (v < 1 || v > g.ip[WW_ST4_I_NTH]) {
      report(g_last, WW_KOKKOS_ERR_TABLE, "ww_st4_commit: SATINDICES outside 1..NTH");
      return WW_KOKKOS_ERR_TABLE;
    }
  if (ww_kokkos_flag("WW_KOKKOS_ST4_TABLES")) {
    Arr& sw = g.a[WW_ST4_A_SATWEIGHTS];
    Arr& si = g.a[WW_ST4_A_SATINDICES];
    Arr& cw = g.a[WW_ST4_A_CUMULW];
    std::vector<float> w2(sw.f.size()), c2(cw.f.size());
    std::vector<int> i2(si.i.size());
    insin4_tables(g.r, g.ip, g.a[WW_ST4_A_SIG].f.data(), g.a[WW_ST4_A_DSIP].f.data(),
                  g.a[WW_ST4_A_ECOS].f.data(), g.a[WW_ST4_A_TH].f.data(), w2.data(), i2.data(), c2.data());
    long long diff = 0;
    for (size_t i = 0; i < w2.size(); ++i) diff += std::memcmp(&w2[i], &sw.f[i], 4) != 0;
    for (size_t i = 0; i < i2.size(); ++i) diff += i2[i] != si.i[i];
    for (size_t i = 0; i < c2.size(); ++i) diff += std::memcmp(&c2[i], &cw.f[i], 4) != 0;
    std::fprintf(stderr, "ww3gpu: ww_st4_commit grid %d: INSIN4 tables rebuilt, %lld values differ\n", igrid,
                 diff);
    sw.f.swap(w2);
    si.i.swap(i2);
    cw.f.swap(c2);
  }
  const int rc = upload(g);
  if (rc != WW_KOKKOS_OK) {
    report(g_last, rc, rc == WW_KOKKOS_ERR_UNSUPPORTED
                           ? "ww_st4_commit: CUMULW is not a function of (IK, IK2, |ITH2-ITH|)"
                           : "ww_st4_commit: device copy");
    return rc;
  }
  g.ready = true;
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

int ww_st4_enabled(int which) {
  switch (which) {
    case 1: return ww_kokkos_flag("WW_KOKKOS_SDS4") != 0 ? 1 : 0;
    case 2: return ww_kokkos_flag("WW_KOKKOS_SIN4") != 0 ? 1 : 0;
    case 3: return ww_kokkos_flag("WW_KOKKOS_SPR4") != 0 ? 1 : 0;
    case 4: return ww_kokkos_flag("WW_KOKKOS_ST4_TABLES") != 0 ? 1 : 0;
    default: return 0;
  }
}

int ww_st4_last_error(void) { return ww3gpu::st4h::last_status(); }

void ww_sds4(int igrid, int npts, const float* a, const float* k, const float* cg, const float* ustar,
             const float* usdir, const float* depth, const float* dair, const float* dlwmean, int flogrd57,
             int flogrd58, float* srhs, float* ddiag, float* brlambda, float* whitecap) {
  using ww3gpu::st4h::HostArr;
  srce::St4Dev q;
  const size_t n = npts > 0 ? static_cast<size_t>(npts) : 0;
  const size_t nspec = ww3gpu::st4h::grid_state(igrid, q) ? q.nspec : 0, nk = nspec ? q.nk : 0;
  const auto in = [&](const float* p, size_t m) { return HostArr{const_cast<float*>(p), m * 4, false}; };
  const auto io = [&](float* p, size_t m) { return HostArr{p, m * 4, true}; };
  std::vector<HostArr> v{in(a, n * nspec), in(k, n * nk), in(cg, n * nk), in(ustar, n), in(usdir, n),
                         in(depth, n), in(dair, n), in(dlwmean, n), io(srhs, n * nspec), io(ddiag, n * nspec),
                         io(brlambda, n * nspec), io(whitecap, n * 4)};
  srce::St4Call c{};
  c.which = 2;
  c.flogrd57 = flogrd57 != 0;
  c.flogrd58 = flogrd58 != 0;
  const auto run = [&]() {
    c.a = static_cast<const float*>(v[0].d);
    c.k = static_cast<const float*>(v[1].d);
    c.cg = static_cast<const float*>(v[2].d);
    for (int i = 0; i < 4; ++i) c.in[i] = static_cast<const float*>(v[3 + i].d);
    c.s = static_cast<float*>(v[8].d);
    c.d = static_cast<float*>(v[9].d);
    c.brlambda = static_cast<float*>(v[10].d);
    c.whitecap = static_cast<float*>(v[11].d);
  };
  ww3gpu::st4h::routine_with(igrid, npts, c, v, run, "ww_sds4");
}

void ww_sin4(int igrid, int npts, const float* a, const float* cg, const float* k, const float* u,
             const float* ustar, const float* drat, const float* as, const float* usdir, const float* z0,
             const float* cd, float* tauwx, float* tauwy, float* tauwnx, float* tauwny, float* s, float* d,
             int* llws) {
  using ww3gpu::st4h::HostArr;
  srce::St4Dev q;
  const size_t n = npts > 0 ? static_cast<size_t>(npts) : 0;
  const size_t nspec = ww3gpu::st4h::grid_state(igrid, q) ? q.nspec : 0, nk = nspec ? q.nk : 0;
  const auto in = [&](const void* p, size_t m) { return HostArr{const_cast<void*>(p), m * 4, false}; };
  const auto io = [&](void* p, size_t m) { return HostArr{p, m * 4, true}; };
  std::vector<HostArr> v{in(a, n * nspec), in(cg, n * nk), in(k, n * nspec), in(u, n), in(ustar, n),
                         in(drat, n), in(as, n), in(usdir, n), in(z0, n), in(cd, n), io(tauwx, n), io(tauwy, n),
                         io(tauwnx, n), io(tauwny, n), io(s, n * nspec), io(d, n * nspec), io(llws, n * nspec)};
  srce::St4Call c{};
  c.which = 1;
  const auto run = [&]() {
    c.a = static_cast<const float*>(v[0].d);
    c.cg = static_cast<const float*>(v[1].d);
    c.k = static_cast<const float*>(v[2].d);
    for (int i = 0; i < 6; ++i) c.in[i] = static_cast<const float*>(v[3 + i].d);
    for (int i = 0; i < 4; ++i) c.out[i] = static_cast<float*>(v[10 + i].d);
    c.s = static_cast<float*>(v[14].d);
    c.d = static_cast<float*>(v[15].d);
    c.llws = static_cast<int*>(v[16].d);
  };
  ww3gpu::st4h::routine_with(igrid, npts, c, v, run, "ww_sin4");
}

void ww_spr4(int igrid, int npts, const float* a, const float* cg, const float* wn, const float* u,
             const float* udir, const float* tauwx, const float* tauwy, const int* llws, float* ustar,
             float* usdir, float* emean, float* fmean, float* fmean1, float* wnmean, float* amax, float* cd,
             float* z0, float* charn, float* fmeanws, float* dlwmean) {
  using ww3gpu::st4h::HostArr;
  srce::St4Dev q;
  const size_t n = npts > 0 ? static_cast<size_t>(npts) : 0;
  const size_t nspec = ww3gpu::st4h::grid_state(igrid, q) ? q.nspec : 0, nk = nspec ? q.nk : 0;
  const auto in = [&](const void* p, size_t m) { return HostArr{const_cast<void*>(p), m * 4, false}; };
  const auto io = [&](void* p, size_t m) { return HostArr{p, m * 4, true}; };
  std::vector<HostArr> v{in(a, n * nspec), in(cg, n * nk), in(wn, n * nk), in(u, n), in(udir, n),
                         in(tauwx, n), in(tauwy, n), in(llws, n * nspec), io(ustar, n), io(usdir, n),
                         io(emean, n), io(fmean, n), io(fmean1, n), io(wnmean, n), io(amax, n), io(cd, n),
                         io(z0, n), io(charn, n), io(fmeanws, n), io(dlwmean, n)};
  srce::St4Call c{};
  c.which = 0;
  const auto run = [&]() {
    c.a = static_cast<const float*>(v[0].d);
    c.cg = static_cast<const float*>(v[1].d);
    c.k = static_cast<const float*>(v[2].d);
    for (int i = 0; i < 4; ++i) c.in[i] = static_cast<const float*>(v[3 + i].d);
    c.llws = static_cast<int*>(v[7].d);
    for (int i = 0; i < 12; ++i) c.out[i] = static_cast<float*>(v[8 + i].d);
  };
  ww3gpu::st4h::routine_with(igrid, npts, c, v, run, "ww_spr4");
}

int ww3gpu_insin4_tables(const float* rpar, int nrpar, const int* ipar, int nipar, const float* sig0,
                         const float* dsip0, const float* ecos, const float* th, float* satw, int* sati,
                         float* cumulw) {
  if (nrpar != WW_ST4_NRPAR || nipar != WW_ST4_NIPAR || rpar == nullptr || ipar == nullptr || sig0 == nullptr ||
      dsip0 == nullptr || ecos == nullptr || th == nullptr || satw == nullptr || sati == nullptr ||
      cumulw == nullptr)
    return WW_KOKKOS_ERR_BAD_SHAPE;
  const int nk = ipar[WW_ST4_I_NK], nth = ipar[WW_ST4_I_NTH];
  if (nk <= 0 || nth <= 0 || ipar[WW_ST4_I_MTH] < nth || ipar[WW_ST4_I_MSPEC] < nk * nth)
    return WW_KOKKOS_ERR_BAD_SHAPE;
  ww3gpu::st4h::insin4_tables(rpar, ipar, sig0, dsip0, ecos, th, satw, sati, cumulw);
  return WW_KOKKOS_OK;
}

// ------------------------------------------------------------------------------------
// W3SRCE
// ------------------------------------------------------------------------------------
int ww_srce_init(int igrid, const float* rpar, int nrpar, const int* ipar, int nipar, const float* sig,
                 const float* dden, const float* ecos, const float* esin, const int* mapwn) {
  using namespace ww3gpu::srceh;
  if (grids().count(igrid)) {
    release_grid(grids()[igrid]);
    grids().erase(igrid);
  }
  if (!runtime().ready) {
    report(g_last, WW_KOKKOS_ERR_NOT_INITIALISED, "ww_srce_init before ww_kokkos_init");
    return WW_KOKKOS_ERR_NOT_INITIALISED;
  }
  if (!g_hooked) {
    on_finalize(release_all);
    g_hooked = true;
  }
  if (nrpar != WW_SRCE_NRPAR || nipar != WW_SRCE_NIPAR || rpar == nullptr || ipar == nullptr || sig == nullptr ||
      dden == nullptr || ecos == nullptr || esin == nullptr || mapwn == nullptr) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww_srce_init: RPAR/IPAR length differs from the W3SRCE packet");
    return WW_KOKKOS_ERR_BAD_SHAPE;
  }
  if (ipar[WW_SRCE_I_LAB_SWITCH] != 1) {
    report(g_last, WW_KOKKOS_ERR_UNSUPPORTED, "ww_srce_init: only the lab switch's source terms are translated");
    return WW_KOKKOS_ERR_UNSUPPORTED;
  }
  if (ipar[WW_SRCE_I_SEED] != 0 || ipar[WW_SRCE_I_FLD2] != 0) {
    report(g_last, WW_KOKKOS_ERR_UNSUPPORTED, "ww_srce_init: SEED and FLD2 are not supported");
    return WW_KOKKOS_ERR_UNSUPPORTED;
  }
  const int nk = ipar[WW_SRCE_I_NK], nth = ipar[WW_SRCE_I_NTH], nspec = ipar[WW_SRCE_I_NSPEC];
  if (nk <= 0 || nth <= 0 || nspec != nk * nth) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww_srce_init: nspec != nk*nth");
    return WW_KOKKOS_ERR_BAD_SHAPE;
  }
  srce::St4Dev q;
  snl1::GridView dv;
  if (!snl1::grid_view(igrid, dv) || !st4h::grid_state(igrid, q)) {
    report(g_last, WW_KOKKOS_ERR_NOT_INITIALISED, "ww_srce_init before ww_snl1_init and ww_st4_commit for this grid");
    return WW_KOKKOS_ERR_NOT_INITIALISED;
  }
  if (dv.g->nk != nk || dv.g->nth != nth || q.nk != nk || q.nth != nth) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww_srce_init: NK/NTH differ from the DIA or ST4 context");
    return WW_KOKKOS_ERR_BAD_SHAPE;
  }
  for (int is = 0; is < nspec; ++is)
    if (mapwn[is] < 1 || mapwn[is] > nk) {
      report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww_srce_init: MAPWN outside 1..NK");
      return WW_KOKKOS_ERR_BAD_SHAPE;
    }
  if (srce::group_lds_bytes(nk, nth) > 64u * 1024u) {
    report(g_last, WW_KOKKOS_ERR_UNSUPPORTED, "ww_srce_init: spectrum too large for the kernels' LDS");
    return WW_KOKKOS_ERR_UNSUPPORTED;
  }
  Grid& g = grids()[igrid];
  std::memcpy(g.r, rpar, sizeof g.r);
  std::memcpy(g.ip, ipar, sizeof g.ip);
  const size_t o_sig = 0, o_dden = al64(nk), o_ecos = o_dden + al64(nk), o_esin = o_ecos + al64(nspec);
  const size_t o_map = o_esin + al64(nspec), words = o_map + al64(nspec);
  std::vector<float> h(words, 0.f);
  std::memcpy(&h[o_sig], sig, sizeof(float) * nk);
  std::memcpy(&h[o_dden], dden, sizeof(float) * nk);
  std::memcpy(&h[o_ecos], ecos, sizeof(float) * nspec);
  std::memcpy(&h[o_esin], esin, sizeof(float) * nspec);
  std::memcpy(&h[o_map], mapwn, sizeof(int) * nspec);
  if (g.tables.ensure(words * sizeof(float)) != hipSuccess ||
      hipMemcpy(g.tables.p, h.data(), words * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    release_grid(g);
    grids().erase(igrid);
    report(g_last, WW_KOKKOS_ERR_RUNTIME, "ww_srce_init: device copy");
    return WW_KOKKOS_ERR_RUNTIME;
  }
  srce::SrceDev& d = g.dev;
  const float* r = rpar;
  d.nk = nk;
  d.nth = nth;
  d.nspec = nspec;
  d.ic_numerics = ipar[WW_SRCE_I_IC_NUMERICS] != 0;
  d.fdonly = ipar[WW_SRCE_I_FDONLY] != 0;
  d.ln1 = ipar[WW_SRCE_I_LN1] != 0;
  d.dmin = r[WW_SRCE_R_DMIN];
  d.dtmax = r[WW_SRCE_R_DTMAX];
  d.dtmin = r[WW_SRCE_R_DTMIN];
  d.facti1 = r[WW_SRCE_R_FACTI1];
  d.facti2 = r[WW_SRCE_R_FACTI2];
  d.fachfa = r[WW_SRCE_R_FACHFA];
  d.facp = r[WW_SRCE_R_FACP];
  d.xflt = r[WW_SRCE_R_XFLT];
  d.xrel = r[WW_SRCE_R_XREL];
  d.fhmax = r[WW_SRCE_R_FHMAX];
  d.ffxfm = r[WW_SRCE_R_FFXFM];
  d.ffxpm = r[WW_SRCE_R_FFXPM];
  d.ffxfa = r[WW_SRCE_R_FFXFA];
  for (int j = 0; j < 4; ++j) d.icescales[j] = r[WW_SRCE_R_ICESCALES + j];
  d.slnc1 = r[WW_SRCE_R_SLNC1];
  d.fspm = r[WW_SRCE_R_FSPM];
  d.fshf = r[WW_SRCE_R_FSHF];
  d.sbtc1 = r[WW_SRCE_R_SBTC1];
  d.sdbc1 = r[WW_SRCE_R_SDBC1];
  d.sdbc2 = r[WW_SRCE_R_SDBC2];
  const float* b = g.tables.as<float>();
  d.sig = b + o_sig;
  d.dden = b + o_dden;
  d.ecos = b + o_ecos;
  d.esin = b + o_esin;
  d.mapwn = reinterpret_cast<const int*>(b + o_map);
  g.ready = true;
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

void ww_srce_batch(int igrid, int it, int inflags24, int flogrd57, int flogrd58, float dtg, int npts,
                   float* spec, const float* wn1, const float* cg1, const float* pin, float* psc,
                   float* whitecap, float* taubbl, float* tauice, int* shave) {
  const srce::StepFlags f{it, inflags24 != 0, flogrd57 != 0, flogrd58 != 0, dtg};
  (void)ww3gpu::srceh::batch_host(igrid, f, npts, spec, wn1, cg1, pin, psc, whitecap, taubbl, tauice, shave,
                                  "ww_srce_batch");
}

void ww_srce_batch_va(int igrid, int it, int inflags24, int flogrd57, int flogrd58, float dtg, float* va,
                      int nspec, int nsealm, int npts, const int* jsea, const float* wn1, const float* cg1,
                      const float* pin, float* psc, float* whitecap, float* taubbl, float* tauice, int* shave) {
  using namespace ww3gpu::srceh;
  const char* who = "ww_srce_batch_va";
  Grid* g = nullptr;
  srce::St4Dev q;
  snl1::GridView dv;
  if (resolve(igrid, who, g, q, dv) != WW_KOKKOS_OK) return;
  if (nspec != q.nspec || npts < 0 || npts > nsealm || (npts > 0 && jsea == nullptr)) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww_srce_batch_va: NSPEC differs from the grid, or a bad point list");
    return;
  }
  std::vector<char> seen(static_cast<size_t>(nsealm) + 1, 0);
  for (int i = 0; i < npts; ++i) {
    if (jsea[i] < 1 || jsea[i] > nsealm || seen[jsea[i]]) {
      report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww_srce_batch_va: a JSEA outside 1..NSEALM or listed twice");
      return;
    }
    seen[jsea[i]] = 1;
  }
  float* d_va = nullptr;
  const int rc = ww3gpu::resident::device_va(igrid, va, nspec, nsealm, true, &d_va, who);
  if (rc != WW_KOKKOS_OK) {
    report(g_last, rc, "ww_srce_batch_va: resident VA");
    return;
  }
  const srce::StepFlags f{it, inflags24 != 0, flogrd57 != 0, flogrd58 != 0, dtg};
  (void)batch_host(igrid, f, npts, nullptr, wn1, cg1, pin, psc, whitecap, taubbl, tauice, shave, who, d_va, jsea);
}

void ww_srce(int igrid, int it, int inflags24, int flogrd57, int flogrd58, float* spec, const float* wn1,
             const float* cg1, float d_inp, float u10abs, float u10dir, float as, float ice, float dtg,
             float dair, float coef, float* ustar, float* ustdir, float* tauwx, float* tauwy, float* dtdyn,
             float* fcut, float* tauox, float* tauoy, float* tauwix, float* tauwiy, float* tauwnx,
             float* tauwny, float* phiaw, float* charn, float* tws, float* phioc, float* whitecap,
             float* phibbl, float* taubbl, float* tauice, float* phice, float* tauocx, float* tauocy,
             float* wnmean, float* fpi, int* shave) {
  float* sc[WW_SRCE_NPSC] = {ustar, ustdir, tauwx,  tauwy,  dtdyn, fcut,   tauox,  tauoy,
                             tauwix, tauwiy, tauwnx, tauwny, phiaw, charn,  tws,    phioc,
                             phibbl, phice,  tauocx, tauocy, wnmean, fpi};
  for (float* p : sc)
    if (p == nullptr) {
      ww3gpu::report(ww3gpu::srceh::g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww_srce: a null argument");
      return;
    }
  const float pin[WW_SRCE_NPIN] = {d_inp, u10abs, u10dir, as, ice, dair, coef};
  float psc[WW_SRCE_NPSC];
  for (int j = 0; j < WW_SRCE_NPSC; ++j) psc[j] = *sc[j];
  const srce::StepFlags f{it, inflags24 != 0, flogrd57 != 0, flogrd58 != 0, dtg};
  int sh = shave != nullptr ? *shave : 0;
  const int rc = ww3gpu::srceh::batch_host(igrid, f, 1, spec, wn1, cg1, pin, psc, whitecap, taubbl, tauice,
                                           shave != nullptr ? &sh : nullptr, "ww_srce");
  if (rc != WW_KOKKOS_OK) return;
  for (int j = 0; j < WW_SRCE_NPSC; ++j) *sc[j] = psc[j];
  *shave = sh;
}

int ww_srce_enabled(void) { return ww_kokkos_flag("WW_KOKKOS_SRCE") != 0 ? 1 : 0; }

int ww_srce_last_error(void) { return ww3gpu::srceh::g_last.load(std::memory_order_relaxed); }

int ww3gpu_srce_batch_dev(int igrid, int it, int inflags24, int flogrd57, int flogrd58, float dtg, int npts,
                          float* spec, const float* wn1, const float* cg1, const float* pin, float* psc,
                          float* whitecap, float* taubbl, float* tauice, int* shave, void* stream) {
  using namespace ww3gpu::srceh;
  const char* who = "ww3gpu_srce_batch_dev";
  Grid* g = nullptr;
  srce::St4Dev q;
  snl1::GridView dv;
  int rc = resolve(igrid, who, g, q, dv);
  if (rc != WW_KOKKOS_OK) return rc;
  if (!args_ok(npts, spec, wn1, cg1, pin, psc, whitecap, taubbl, tauice, shave)) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww3gpu_srce_batch_dev: npts < 0 or a null array");
    return WW_KOKKOS_ERR_BAD_SHAPE;
  }
  const srce::StepFlags f{it, inflags24 != 0, flogrd57 != 0, flogrd58 != 0, dtg};
  if ((rc = check_step(q, it, who)) != WW_KOKKOS_OK) return rc;
  if (npts == 0) {
    report(g_last, WW_KOKKOS_OK, "");
    return WW_KOKKOS_OK;
  }
  rc = run(*g, q, dv, f, npts, spec, wn1, cg1, pin, psc, whitecap, taubbl, tauice, shave,
           static_cast<hipStream_t>(stream), who);
  if (rc == WW_KOKKOS_OK) report(g_last, WW_KOKKOS_OK, "");
  return rc;
}

int ww3gpu_srce_timing(int igrid, int enable) {
  ww3gpu::srceh::Grid* g = ww3gpu::srceh::find(igrid);
  if (g == nullptr) return WW_KOKKOS_ERR_NOT_INITIALISED;
  ww3gpu::srceh::resolve_timers(*g);
  g->timing = enable != 0;
  for (int k = 0; k < ww3gpu::srceh::kTimers; ++k) {
    g->ms[k] = 0.0;
    g->launches[k] = 0;
  }
  return WW_KOKKOS_OK;
}

int ww3gpu_srce_timers(int igrid, double* ms, long* launches, int n) {
  ww3gpu::srceh::Grid* g = ww3gpu::srceh::find(igrid);
  if (g == nullptr) return WW_KOKKOS_ERR_NOT_INITIALISED;
  if (n < 0 || (n > 0 && (ms == nullptr || launches == nullptr))) return WW_KOKKOS_ERR_BAD_SHAPE;
  ww3gpu::srceh::resolve_timers(*g);
  for (int k = 0; k < n && k < ww3gpu::srceh::kTimers; ++k) {
    ms[k] = g->ms[k];
    launches[k] = g->launches[k];
  }
  return WW_KOKKOS_OK;
}

int ww3gpu_srce_substeps(int igrid) {
  ww3gpu::srceh::Grid* g = ww3gpu::srceh::find(igrid);
  return g == nullptr ? -1 : g->substeps;
}

long ww3gpu_srce_point_substeps(int igrid) {
  ww3gpu::srceh::Grid* g = ww3gpu::srceh::find(igrid);
  return g == nullptr ? -1 : g->point_substeps;
}

int ww3gpu_srce_group_lanes(int lanes) {
  if (lanes != 0 && lanes != 16 && lanes != 32 && lanes != 64) return WW_KOKKOS_ERR_BAD_SHAPE;
  ww3gpu::srce::set_group_lanes(lanes);
  return WW_KOKKOS_OK;
}

int ww3gpu_srce_width_thresholds(int half_below, int wide_below) {
  if (half_below == 0 && wide_below == 0) {
    half_below = ww3gpu::srce::kHalfBelow;
    wide_below = ww3gpu::srce::kWideBelow;
  }
  if (half_below < 0 || wide_below < 0 || wide_below > half_below) return WW_KOKKOS_ERR_BAD_SHAPE;
  ww3gpu::srce::set_width_thresholds(half_below, wide_below);
  return WW_KOKKOS_OK;
}

}  // extern "C"

// ww3-gpu_amd/csrc/snl1_capi.cpp
// The W3SNL1 entry points of the C ABI: per-grid contexts, validation, copies.
//
// Reference behaviour (kokkos/src/fortran_iface/snl1_shim.cpp:254-419): one context per
// igrid in a never-destroyed map; init erases the grid first so a failure leaves it
// unconfigured; npts == 0 is a legal no-op; a null pointer or negative npts is BAD_SHAPE
// and writes nothing; ww_snl1 copies in, runs, copies out and returns with the results
// in the caller's arrays.
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "runtime.hpp"
#include "snl1.hpp"

namespace ww3gpu {
namespace snl1 {
namespace {

struct Ctx {
  Tables t;
  DevBuf sig, af11;
  DevBuf a, cg, kd, s, d;  // staging for the host-array entry point
  int variant = 0;         // 0: specialised kernel when available, 1: general kernel
  bool ready = false;
};

std::map<int, Ctx>& contexts() {
  static auto* m = new std::map<int, Ctx>();
  return *m;
}

std::atomic<int> g_last{WW_KOKKOS_OK};
bool g_hooked = false;

void release_all() {
  for (auto& kv : contexts()) {
    Ctx& c = kv.second;
    c.sig.release();
    c.af11.release();
    c.a.release();
    c.cg.release();
    c.kd.release();
    c.s.release();
    c.d.release();
  }
  contexts().clear();
}

Ctx* find(int igrid) {
  if (!runtime().ready) return nullptr;
  const auto it = contexts().find(igrid);
  if (it == contexts().end() || !it->second.ready) return nullptr;
  return &it->second;
}

int upload_tables(Ctx& c) {
  const size_t nsig = c.t.sig.size() * sizeof(float);
  const size_t naf = c.t.af11_row.size() * sizeof(float);
  if (c.sig.ensure(nsig) != hipSuccess || c.af11.ensure(naf) != hipSuccess) return WW_KOKKOS_ERR_RUNTIME;
  if (hipMemcpy(c.sig.p, c.t.sig.data(), nsig, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c.af11.p, c.t.af11_row.data(), naf, hipMemcpyHostToDevice) != hipSuccess)
    return WW_KOKKOS_ERR_RUNTIME;
  return WW_KOKKOS_OK;
}

int run(Ctx& c, int npts, const float* a, const float* cg, const float* kd, float* s, float* d,
        hipStream_t stream) {
  const hipError_t e =
      launch(c.t.g, c.af11.as<float>(), c.sig.as<float>(), npts, a, cg, kd, s, d, stream, c.variant,
             point_major(c.t.g.nk, c.t.g.nspec));
  if (e != hipSuccess) {
    std::fprintf(stderr, "ww3gpu: DIA launch: %s\n", hipGetErrorString(e));
    return WW_KOKKOS_ERR_KERNEL;
  }
  return WW_KOKKOS_OK;
}

}  // namespace

bool grid_view(int igrid, GridView& v) {
  Ctx* c = find(igrid);
  if (c == nullptr) return false;
  v.g = &c->t.g;
  v.af11_row = c->af11.as<float>();
  v.sig = c->sig.as<float>();
  v.variant = c->variant;
  return true;
}

/// The configured grid's DIA set-up (SIG, CONS parameters, tables), or nullptr.
const Tables* tables_of(int igrid) {
  const Ctx* c = find(igrid);
  return c == nullptr ? nullptr : &c->t;
}

}  // namespace snl1
}  // namespace ww3gpu

namespace ww3gpu {
namespace snl1 {
namespace {
void export_tables(const Tables& t, int* dims9, float* scal19, int* ipim, int* ic, float* af11) {
  const Geom& g = t.g;
  if (dims9 != nullptr) {
    const int v[9] = {g.nk, g.nfrhgh, g.nfrchg, g.nspecx, g.nfrhgh * g.nth,
                      t.ifr_off[0], t.ifr_off[1], t.ifr_off[2], t.ifr_off[3]};
    std::memcpy(dims9, v, sizeof v);
  }
  if (scal19 != nullptr) std::memcpy(scal19, t.scal, sizeof t.scal);
  if (ipim != nullptr) std::memcpy(ipim, t.ipim.data(), t.ipim.size() * sizeof(int));
  if (ic != nullptr) std::memcpy(ic, t.ic.data(), t.ic.size() * sizeof(int));
  if (af11 != nullptr)
    for (int r = 0; r < g.nfrchg; ++r)
      for (int j = 0; j < g.nth; ++j) af11[r * g.nth + j] = t.af11_row[r];
}
}  // namespace
}  // namespace snl1
}  // namespace ww3gpu

using namespace ww3gpu;
using namespace ww3gpu::snl1;

extern "C" {

int ww_snl1_init(int igrid, int nk, int nth, float xfr, float dth, float lam, float snlc1,
                 float kdcon, float kdmn, float snls1, float snls2, float snls3, float fachfe,
                 const float* sig) {
  contexts().erase(igrid);
  if (!runtime().ready) {
    report(g_last, WW_KOKKOS_ERR_NOT_INITIALISED, "ww_snl1_init before ww_kokkos_init");
    return WW_KOKKOS_ERR_NOT_INITIALISED;
  }
  if (!g_hooked) {
    on_finalize(release_all);
    g_hooked = true;
  }
  if (nk <= 0 || nth <= 0 || sig == nullptr) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww_snl1_init: nk and nth must be > 0 and sig non-null");
    return WW_KOKKOS_ERR_BAD_SHAPE;
  }
  const Config cfg{nk, nth, xfr, dth, lam, snlc1, kdcon, kdmn, snls1, snls2, snls3, fachfe};
  Ctx c;
  if (!build_tables(cfg, sig, c.t)) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww_snl1_init: INSNL1 rejected the configuration");
    return WW_KOKKOS_ERR_BAD_SHAPE;
  }
  if (!c.t.structured) {
    report(g_last, WW_KOKKOS_ERR_UNSUPPORTED,
           "ww_snl1_init: quadruplet tables not of the row/shift form the kernel needs");
    return WW_KOKKOS_ERR_UNSUPPORTED;
  }
  if (lds_bytes(c.t.g) > 160u * 1024u || c.t.g.nspecx > 16 * 256) {
    report(g_last, WW_KOKKOS_ERR_UNSUPPORTED, "ww_snl1_init: spectrum too large for one workgroup");
    return WW_KOKKOS_ERR_UNSUPPORTED;
  }
  Ctx& slot = contexts()[igrid];
  slot.t = std::move(c.t);
  const int rc = upload_tables(slot);
  if (rc != WW_KOKKOS_OK) {
    slot.sig.release();
    slot.af11.release();
    contexts().erase(igrid);
    report(g_last, rc, "ww_snl1_init: device allocation");
    return rc;
  }
  slot.ready = true;
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

void ww_snl1(int igrid, int npts, const float* a, const float* cg, const float* kdmean, float* s,
             float* d) {
  Ctx* c = find(igrid);
  if (c == nullptr) {
    report(g_last, WW_KOKKOS_ERR_NOT_INITIALISED, "ww_snl1 before ww_snl1_init for this grid");
    return;
  }
  if (npts < 0) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww_snl1: npts is negative");
    return;
  }
  if (npts == 0) {
    report(g_last, WW_KOKKOS_OK, "");
    return;
  }
  if (a == nullptr || cg == nullptr || kdmean == nullptr || s == nullptr || d == nullptr) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww_snl1: null array pointer");
    return;
  }
  const Geom& g = c->t.g;
  const size_t nsp = static_cast<size_t>(g.nspec) * npts * sizeof(float);
  const size_t ncg = static_cast<size_t>(g.nk) * npts * sizeof(float);
  const size_t nkd = static_cast<size_t>(npts) * sizeof(float);
  if (c->a.ensure(nsp) != hipSuccess || c->cg.ensure(ncg) != hipSuccess ||
      c->kd.ensure(nkd) != hipSuccess || c->s.ensure(nsp) != hipSuccess ||
      c->d.ensure(nsp) != hipSuccess) {
    report(g_last, WW_KOKKOS_ERR_RUNTIME, "ww_snl1: device allocation");
    return;
  }
  hipStream_t st = runtime().stream;
  if (hipMemcpyAsync(c->a.p, a, nsp, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(c->cg.p, cg, ncg, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(c->kd.p, kdmean, nkd, hipMemcpyHostToDevice, st) != hipSuccess) {
    report(g_last, WW_KOKKOS_ERR_RUNTIME, "ww_snl1: host-to-device copy");
    return;
  }
  int rc = run(*c, npts, c->a.as<float>(), c->cg.as<float>(), c->kd.as<float>(), c->s.as<float>(),
               c->d.as<float>(), st);
  if (rc != WW_KOKKOS_OK) {
    report(g_last, rc, "ww_snl1: kernel");
    return;
  }
  if (hipMemcpyAsync(s, c->s.p, nsp, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(d, c->d.p, nsp, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    report(g_last, WW_KOKKOS_ERR_RUNTIME, "ww_snl1: device-to-host copy");
    return;
  }
  report(g_last, WW_KOKKOS_OK, "");
}

int ww_snl1_enabled(void) { return ww_kokkos_flag("WW_KOKKOS_SNL1") != 0 ? 1 : 0; }

int ww_snl1_last_error(void) { return g_last.load(std::memory_order_relaxed); }

int ww3gpu_snl1_dev(int igrid, int npts, const float* a, const float* cg, const float* kdmean,
                    float* s, float* d, void* stream) {
  Ctx* c = find(igrid);
  if (c == nullptr) {
    report(g_last, WW_KOKKOS_ERR_NOT_INITIALISED, "ww3gpu_snl1_dev before ww_snl1_init for this grid");
    return WW_KOKKOS_ERR_NOT_INITIALISED;
  }
  if (npts < 0) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww3gpu_snl1_dev: npts is negative");
    return WW_KOKKOS_ERR_BAD_SHAPE;
  }
  if (npts == 0) {
    report(g_last, WW_KOKKOS_OK, "");
    return WW_KOKKOS_OK;
  }
  if (a == nullptr || cg == nullptr || kdmean == nullptr || s == nullptr || d == nullptr) {
    report(g_last, WW_KOKKOS_ERR_BAD_SHAPE, "ww3gpu_snl1_dev: null array pointer");
    return WW_KOKKOS_ERR_BAD_SHAPE;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);  // NULL is the default stream
  const int rc = run(*c, npts, a, cg, kdmean, s, d, st);
  report(g_last, rc, "ww3gpu_snl1_dev: kernel");
  return rc;
}


int ww3gpu_snl1_tables(int igrid, int* dims9, float* scal19, int* ipim, int* ic, float* af11) {
  Ctx* c = find(igrid);
  if (c == nullptr) return WW_KOKKOS_ERR_NOT_INITIALISED;
  export_tables(c->t, dims9, scal19, ipim, ic, af11);
  return WW_KOKKOS_OK;
}

int ww3gpu_insnl1(int nk, int nth, float xfr, float dth, float lam, const float* sig,
                  int* dims9, float* scal19, int* ipim, int* ic, float* af11, int* structured) {
  // Only the geometry inputs matter to the tables; the CONS parameters are placeholders.
  const Config cfg{nk, nth, xfr, dth, lam, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  Tables t;
  if (!build_tables(cfg, sig, t)) return WW_KOKKOS_ERR_BAD_SHAPE;
  export_tables(t, dims9, scal19, ipim, ic, af11);
  if (structured != nullptr) *structured = t.structured ? 1 : 0;
  return WW_KOKKOS_OK;
}

int ww3gpu_snl1_set_af11(int igrid, const float* af11, int n) {
  Ctx* c = find(igrid);
  if (c == nullptr) return WW_KOKKOS_ERR_NOT_INITIALISED;
  const Geom& g = c->t.g;
  if (af11 == nullptr || n != g.nspecx) return WW_KOKKOS_ERR_BAD_SHAPE;
  std::vector<float> rows(static_cast<size_t>(g.nfrchg));
  for (int r = 0; r < g.nfrchg; ++r) {
    rows[r] = af11[r * g.nth];
    for (int j = 1; j < g.nth; ++j)
      if (std::memcmp(&af11[r * g.nth + j], &rows[r], sizeof(float)) != 0) return WW_KOKKOS_ERR_TABLE;
  }
  c->t.af11_row = rows;
  if (hipMemcpy(c->af11.p, rows.data(), rows.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return WW_KOKKOS_ERR_RUNTIME;
  return WW_KOKKOS_OK;
}

int ww3gpu_snl1_set_variant(int igrid, int variant) {
  Ctx* c = find(igrid);
  if (c == nullptr) return WW_KOKKOS_ERR_NOT_INITIALISED;
  if (variant < 0 || variant > 1) return WW_KOKKOS_ERR_BAD_SHAPE;
  c->variant = variant;
  return WW_KOKKOS_OK;
}

int ww3gpu_snl1_kernel_variant(int igrid) {
  Ctx* c = find(igrid);
  if (c == nullptr) return -1;
  return (c->variant == 0 && has_specialized(c->t.g)) ? 0 : 1;
}

long ww3gpu_snl1_lds_bytes(int igrid) {
  Ctx* c = find(igrid);
  if (c == nullptr) return -1;
  return static_cast<long>(lds_bytes(c->t.g));
}

}  // extern "C"

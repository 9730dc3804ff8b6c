// ww3-gpu_amd/csrc/snlsur_capi.cpp
// The W3SNL1 surrogate behind the C ABI: per-grid networks, training on DIA outputs,
// inference with the conservation check and the exact-DIA fallback.  None of these entry
// points has a counterpart in the reference; they sit beside ww_snl1 and need the grid's
// ww_snl1_init first (the normalisation uses its SIG and CONS parameters, the fallback its
// tables).
//
// Network: input NSPEC + 1 features (padded to a multiple of 128), three GELU layers of 512,
// a linear output of NSPEC (padded likewise).  fp32 master parameters and Adam moments,
// bf16 copies W [out][in] and W^T [in][out] for the forward and backward GEMMs.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <random>
#include <vector>

#include "runtime.hpp"
#include "snl1.hpp"
#include "snlsur.hpp"

namespace ww3gpu {
namespace snl1 {
const Tables* tables_of(int igrid);
}  // namespace snl1

namespace sur {
namespace {

constexpr int kLayers = 4;
constexpr int kInferChunk = 1 << 20;  // points per inference pass (work buffers ~8.5 KB per point)

struct Net {
  Norm norm{};
  int rows[kLayers], cols[kLayers];
  size_t woff[kLayers], boff[kLayers];  // offsets (floats) into the parameter vector
  size_t bfoff[kLayers];                 // offsets (bf16 elements) into wb / wt
  size_t nparam = 0, nbf = 0;
  DevBuf p, g, m, v, wb, wt, sig, wsig;
  int adam_t = 0;
  // work buffers
  DevBuf x, xt, y, z[3], h[3], ht[3], o, go, got, dz[2], dzt[2], loss, scale, shift, flag, list, nlist;
  DevBuf fa, fcg, fkd, fs, fd;
  DevBuf data_a, data_cg, data_kd, data_s, data_d;  // a training set (released after the fit)
  bool ready = false;
};

std::map<int, Net>& nets() {
  static auto* m = new std::map<int, Net>();
  return *m;
}

std::atomic<int> g_last{WW_KOKKOS_OK};
bool g_hooked = false;

void release(Net& n) {
  DevBuf* all[] = {&n.p, &n.g, &n.m, &n.v, &n.wb, &n.wt, &n.sig, &n.wsig, &n.x, &n.xt, &n.y, &n.o, &n.go,
                   &n.got, &n.loss, &n.scale, &n.shift, &n.flag, &n.list, &n.nlist, &n.fa, &n.fcg, &n.fkd,
                   &n.fs, &n.fd, &n.data_a, &n.data_cg, &n.data_kd, &n.data_s, &n.data_d};
  for (DevBuf* b : all) b->release();
  for (int i = 0; i < 3; ++i) {
    n.z[i].release();
    n.h[i].release();
    n.ht[i].release();
  }
  for (int i = 0; i < 2; ++i) {
    n.dz[i].release();
    n.dzt[i].release();
  }
}

Net* find(int igrid) {
  if (!runtime().ready) return nullptr;
  auto it = nets().find(igrid);
  return it == nets().end() || !it->second.ready ? nullptr : &it->second;
}

int fail(int code, const char* what) {
  report(g_last, code, what);
  return code;
}

#define SUR_TRY(expr, what)                                      \
  do {                                                           \
    if ((expr) != hipSuccess) return fail(WW_KOKKOS_ERR_KERNEL, what); \
  } while (0)
#define SUR_ALLOC(buf, bytes)                                                     \
  do {                                                                            \
    if ((buf).ensure(bytes) != hipSuccess) return fail(WW_KOKKOS_ERR_RUNTIME, "surrogate: out of device memory"); \
  } while (0)

float* P(Net& n, size_t off) { return n.p.as<float>() + off; }
float* G(Net& n, size_t off) { return n.g.as<float>() + off; }

/// Work buffers for a batch of mpad rows; transposed copies only when training.
int ensure_work(Net& n, int mpad, bool train) {
  const size_t M = static_cast<size_t>(mpad);
  const Norm& q = n.norm;
  SUR_ALLOC(n.x, M * q.kin * 2);
  SUR_ALLOC(n.o, M * q.nout * 4);
  SUR_ALLOC(n.scale, M * 4);
  SUR_ALLOC(n.shift, M * 4);
  SUR_ALLOC(n.flag, M * 4);
  SUR_ALLOC(n.list, M * 4);
  SUR_ALLOC(n.nlist, 256);
  for (int i = 0; i < 3; ++i) SUR_ALLOC(n.h[i], M * kHidden * 2);
  if (train) {
    SUR_ALLOC(n.xt, M * q.kin * 2);
    SUR_ALLOC(n.y, M * q.nout * 4);
    SUR_ALLOC(n.go, M * q.nout * 2);
    SUR_ALLOC(n.got, M * q.nout * 2);
    SUR_ALLOC(n.loss, M * 4 + 64);
    for (int i = 0; i < 3; ++i) {
      SUR_ALLOC(n.z[i], M * kHidden * 4);
      SUR_ALLOC(n.ht[i], M * kHidden * 2);
    }
    for (int i = 0; i < 2; ++i) {
      SUR_ALLOC(n.dz[i], M * kHidden * 2);
      SUR_ALLOC(n.dzt[i], M * kHidden * 2);
    }
  }
  return WW_KOKKOS_OK;
}

uint16_t* WB(Net& n, int l) { return n.wb.as<uint16_t>() + n.bfoff[l]; }
uint16_t* WT(Net& n, int l) { return n.wt.as<uint16_t>() + n.bfoff[l]; }

/// Forward pass over mpad rows of n.x.  Training keeps pre-activations and transposes.
int forward(Net& n, int mpad, bool train, hipStream_t st) {
  const uint16_t* in = n.x.as<uint16_t>();
  int ld = n.norm.kin;
  for (int l = 0; l < 3; ++l) {
    Epi e{};
    e.mode = kGeluFwd;
    e.bias = P(n, n.boff[l]);
    e.hb = n.h[l].as<uint16_t>();
    e.ldh = kHidden;
    if (train) {
      e.zf = n.z[l].as<float>();
      e.ldz = kHidden;
      e.ht = n.ht[l].as<uint16_t>();
      e.ldt = mpad;
    }
    SUR_TRY(gemm_nt(mpad, kHidden, ld, in, ld, WB(n, l), ld, e, st), "surrogate: hidden layer");
    in = n.h[l].as<uint16_t>();
    ld = kHidden;
  }
  Epi e{};
  e.mode = kLinearFwd;
  e.bias = P(n, n.boff[3]);
  if (train) {  // the loss reads fp32; inference keeps the output in bf16
    e.of = n.o.as<float>();
    e.ldo = n.norm.nout;
  } else {
    e.hb = n.o.as<uint16_t>();
    e.ldh = n.norm.nout;
  }
  SUR_TRY(gemm_nt(mpad, n.norm.nout, kHidden, in, kHidden, WB(n, 3), kHidden, e, st), "surrogate: output layer");
  return WW_KOKKOS_OK;
}

/// Loss and gradients for a batch of B rows whose features/targets are in n.x, n.xt, n.y.
int backward(Net& n, int B, hipStream_t st) {
  const int nout = n.norm.nout;
  SUR_TRY(loss_grad(B, nout, n.o.as<float>(), n.y.as<float>(), n.loss.as<float>(), n.go.as<uint16_t>(),
                    n.got.as<uint16_t>(), st), "surrogate: loss");
  // Output layer: dW4 = dO^T H3, db4 = column sums of dO, dZ3 = (dO W4) * gelu'(Z3).
  const uint16_t* gt = n.got.as<uint16_t>();  // [out][B] gradient of this layer's output
  const uint16_t* gr = n.go.as<uint16_t>();   // [B][out]
  int out = nout;
  for (int l = 3; l >= 0; --l) {
    const int in = n.cols[l];
    const uint16_t* act_t = l > 0 ? n.ht[l - 1].as<uint16_t>() : n.xt.as<uint16_t>();  // [in][B]
    Epi e{};
    e.mode = kPlain;
    e.of = G(n, n.woff[l]);
    e.ldo = in;
    SUR_TRY(gemm_nt(out, in, B, gt, B, act_t, B, e, st), "surrogate: weight gradient");
    SUR_TRY(column_sums(out, B, gt, G(n, n.boff[l]), st), "surrogate: bias gradient");
    if (l == 0) break;
    Epi d{};
    d.mode = kGeluBwd;
    d.zf = n.z[l - 1].as<float>();
    d.ldz = kHidden;
    d.hb = n.dz[l & 1].as<uint16_t>();
    d.ldh = kHidden;
    d.ht = n.dzt[l & 1].as<uint16_t>();
    d.ldt = B;
    SUR_TRY(gemm_nt(B, kHidden, out, gr, out, WT(n, l), out, d, st), "surrogate: input gradient");
    gt = n.dzt[l & 1].as<uint16_t>();
    gr = n.dz[l & 1].as<uint16_t>();
    out = kHidden;
  }
  return WW_KOKKOS_OK;
}

int adam_step(Net& n, float lr, hipStream_t st) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  ++n.adam_t;
  const float bc1 = 1.0f - std::pow(b1, static_cast<float>(n.adam_t));
  const float bc2 = 1.0f - std::pow(b2, static_cast<float>(n.adam_t));
  for (int l = 0; l < kLayers; ++l) {
    SUR_TRY(adam(n.rows[l], n.cols[l], P(n, n.woff[l]), G(n, n.woff[l]), n.m.as<float>() + n.woff[l],
                 n.v.as<float>() + n.woff[l], WB(n, l), WT(n, l), lr, b1, b2, eps, bc1, bc2, st), "surrogate: adam");
    SUR_TRY(adam(1, n.rows[l], P(n, n.boff[l]), G(n, n.boff[l]), n.m.as<float>() + n.boff[l],
                 n.v.as<float>() + n.boff[l], nullptr, nullptr, lr, b1, b2, eps, bc1, bc2, st), "surrogate: adam");
  }
  return WW_KOKKOS_OK;
}

int refresh_bf16(Net& n, hipStream_t st) {
  for (int l = 0; l < kLayers; ++l)
    SUR_TRY(to_bf16(n.rows[l], n.cols[l], P(n, n.woff[l]), WB(n, l), WT(n, l), st), "surrogate: bf16 copies");
  return WW_KOKKOS_OK;
}

/// One training step on B points (B a multiple of 128) of device arrays.
int train_batch(Net& n, int B, const float* a, const float* cg, const float* kd, const float* s, float lr,
                hipStream_t st) {
  int rc = ensure_work(n, B, true);
  if (rc != WW_KOKKOS_OK) return rc;
  SUR_TRY(prep(n.norm, B, B, a, cg, kd, s, n.x.as<uint16_t>(), n.xt.as<uint16_t>(), B, n.y.as<float>(),
               n.scale.as<float>(), n.shift.as<int>(), n.flag.as<int>(), st), "surrogate: features");
  if ((rc = forward(n, B, true, st)) != WW_KOKKOS_OK) return rc;
  if ((rc = backward(n, B, st)) != WW_KOKKOS_OK) return rc;
  return adam_step(n, lr, st);
}

}  // namespace
}  // namespace sur
}  // namespace ww3gpu

using namespace ww3gpu;
using namespace ww3gpu::sur;

extern "C" {

int ww3gpu_snlsur_init(int igrid, unsigned long long seed) {
  const snl1::Tables* t = snl1::tables_of(igrid);
  if (t == nullptr) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, "ww3gpu_snlsur_init before ww_snl1_init for this grid");
  if (!g_hooked) {
    on_finalize([] {
      for (auto& kv : nets()) release(kv.second);
      nets().clear();
    });
    g_hooked = true;
  }
  nets().erase(igrid);
  Net& n = nets()[igrid];
  const snl1::Geom& g = t->g;
  Norm& q = n.norm;
  q.nk = g.nk;
  q.nth = g.nth;
  q.nspec = g.nspec;
  q.kin = round_up(g.nspec + 1, kTile);
  q.nout = round_up(g.nspec, kTile);
  q.tpiinv = g.tpiinv;
  q.snlc1 = g.snlc1;
  q.kdcon = g.kdcon;
  q.kdmn = g.kdmn;
  q.snls1 = g.snls1;
  q.snls2 = g.snls2;
  q.snls3 = g.snls3;
  const int in[kLayers] = {q.kin, kHidden, kHidden, kHidden};
  const int out[kLayers] = {kHidden, kHidden, kHidden, q.nout};
  size_t off = 0, bo = 0;
  for (int l = 0; l < kLayers; ++l) {
    n.rows[l] = out[l];
    n.cols[l] = in[l];
    n.woff[l] = off;
    off += static_cast<size_t>(out[l]) * in[l];
    n.boff[l] = off;
    off += out[l];
    n.bfoff[l] = bo;
    bo += static_cast<size_t>(out[l]) * in[l];
  }
  n.nparam = off;
  n.nbf = bo;
  // Uniform initialisation with variance 2 / fan_in (the output layer 100 times smaller), biases 0.
  std::vector<float> host(n.nparam, 0.0f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < kLayers; ++l) {
    const double lim = std::sqrt(6.0 / in[l]) * (l == 3 ? 0.01 : 1.0);
    std::uniform_real_distribution<double> u(-lim, lim);
    for (size_t i = 0; i < static_cast<size_t>(out[l]) * in[l]; ++i) host[n.woff[l] + i] = static_cast<float>(u(rng));
  }
  // Energy weights sig_k dsig_k with dsig_k = sig_k (XFR - 1/XFR) / 2.
  std::vector<float> wsig(g.nk);
  const double xfr = g.nk > 1 ? static_cast<double>(t->sig[1]) / t->sig[0] : 1.1;
  for (int k = 0; k < g.nk; ++k)
    wsig[k] = static_cast<float>(static_cast<double>(t->sig[k]) * t->sig[k] * 0.5 * (xfr - 1.0 / xfr));
  const size_t pb = n.nparam * sizeof(float);
  SUR_ALLOC(n.p, pb);
  SUR_ALLOC(n.g, pb);
  SUR_ALLOC(n.m, pb);
  SUR_ALLOC(n.v, pb);
  SUR_ALLOC(n.wb, n.nbf * 2);
  SUR_ALLOC(n.wt, n.nbf * 2);
  SUR_ALLOC(n.sig, g.nk * sizeof(float));
  SUR_ALLOC(n.wsig, g.nk * sizeof(float));
  hipStream_t st = runtime().stream;
  if (hipMemcpy(n.p.p, host.data(), pb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(n.g.p, 0, pb) != hipSuccess || hipMemset(n.m.p, 0, pb) != hipSuccess ||
      hipMemset(n.v.p, 0, pb) != hipSuccess ||
      hipMemcpy(n.sig.p, t->sig.data(), g.nk * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(n.wsig.p, wsig.data(), g.nk * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return fail(WW_KOKKOS_ERR_RUNTIME, "ww3gpu_snlsur_init: upload");
  q.sig = n.sig.as<float>();
  q.wsig = n.wsig.as<float>();
  int rc = refresh_bf16(n, st);
  if (rc != WW_KOKKOS_OK) return rc;
  if (hipStreamSynchronize(st) != hipSuccess) return fail(WW_KOKKOS_ERR_KERNEL, "ww3gpu_snlsur_init: sync");
  n.ready = true;
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

long long ww3gpu_snlsur_nparams(int igrid, int* dims4) {
  Net* n = find(igrid);
  if (n == nullptr) return -WW_KOKKOS_ERR_NOT_INITIALISED;
  if (dims4 != nullptr) {
    dims4[0] = n->norm.kin;
    dims4[1] = kHidden;
    dims4[2] = n->norm.nout;
    dims4[3] = n->norm.nspec;
  }
  return static_cast<long long>(n->nparam);
}

int ww3gpu_snlsur_params(int igrid, int set, float* host, long long count) {
  Net* n = find(igrid);
  if (n == nullptr) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, "ww3gpu_snlsur_params: no surrogate for this grid");
  if (host == nullptr || count != static_cast<long long>(n->nparam))
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, "ww3gpu_snlsur_params: count differs from the parameter count");
  const size_t b = n->nparam * sizeof(float);
  hipStream_t st = runtime().stream;
  if (set) {
    if (hipMemcpyAsync(n->p.p, host, b, hipMemcpyHostToDevice, st) != hipSuccess)
      return fail(WW_KOKKOS_ERR_RUNTIME, "ww3gpu_snlsur_params: upload");
    const int rc = refresh_bf16(*n, st);
    if (rc != WW_KOKKOS_OK) return rc;
  } else if (hipMemcpyAsync(host, n->p.p, b, hipMemcpyDeviceToHost, st) != hipSuccess) {
    return fail(WW_KOKKOS_ERR_RUNTIME, "ww3gpu_snlsur_params: download");
  }
  if (hipStreamSynchronize(st) != hipSuccess) return fail(WW_KOKKOS_ERR_KERNEL, "ww3gpu_snlsur_params: sync");
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

int ww3gpu_snlsur_grads(int igrid, int B, const float* x_host, const float* y_host, float* loss_host,
                        float* grad_host) {
  Net* n = find(igrid);
  if (n == nullptr) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, "ww3gpu_snlsur_grads: no surrogate for this grid");
  if (B <= 0 || B % kTile || x_host == nullptr || y_host == nullptr)
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, "ww3gpu_snlsur_grads: B must be a positive multiple of 128");
  int rc = ensure_work(*n, B, true);
  if (rc != WW_KOKKOS_OK) return rc;
  const Norm& q = n->norm;
  hipStream_t st = runtime().stream;
  DevBuf xf;
  if (xf.ensure(static_cast<size_t>(B) * q.kin * 4) != hipSuccess) return fail(WW_KOKKOS_ERR_RUNTIME, "out of memory");
  if (hipMemcpyAsync(xf.p, x_host, static_cast<size_t>(B) * q.kin * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(n->y.p, y_host, static_cast<size_t>(B) * q.nout * 4, hipMemcpyHostToDevice, st) != hipSuccess) {
    xf.release();
    return fail(WW_KOKKOS_ERR_RUNTIME, "ww3gpu_snlsur_grads: upload");
  }
  rc = to_bf16(B, q.kin, xf.as<float>(), n->x.as<uint16_t>(), n->xt.as<uint16_t>(), st) == hipSuccess
           ? forward(*n, B, true, st)
           : fail(WW_KOKKOS_ERR_KERNEL, "ww3gpu_snlsur_grads: features");
  if (rc == WW_KOKKOS_OK) rc = backward(*n, B, st);
  if (rc == WW_KOKKOS_OK &&
      ((loss_host != nullptr &&
        hipMemcpyAsync(loss_host, n->loss.p, static_cast<size_t>(B) * 4, hipMemcpyDeviceToHost, st) != hipSuccess) ||
       (grad_host != nullptr &&
        hipMemcpyAsync(grad_host, n->g.p, n->nparam * 4, hipMemcpyDeviceToHost, st) != hipSuccess) ||
       hipStreamSynchronize(st) != hipSuccess))
    rc = fail(WW_KOKKOS_ERR_RUNTIME, "ww3gpu_snlsur_grads: download");
  (void)hipStreamSynchronize(st);
  xf.release();
  if (rc == WW_KOKKOS_OK) report(g_last, WW_KOKKOS_OK, "");
  return rc;
}

int ww3gpu_snlsur_forward(int igrid, int M, const float* x_host, float* o_host) {
  Net* n = find(igrid);
  if (n == nullptr) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, "ww3gpu_snlsur_forward: no surrogate for this grid");
  if (M <= 0 || M % kTile || x_host == nullptr || o_host == nullptr)
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, "ww3gpu_snlsur_forward: M must be a positive multiple of 128");
  int rc = ensure_work(*n, M, true);  // the training form keeps the output in fp32
  if (rc != WW_KOKKOS_OK) return rc;
  const Norm& q = n->norm;
  hipStream_t st = runtime().stream;
  DevBuf xf;
  if (xf.ensure(static_cast<size_t>(M) * q.kin * 4) != hipSuccess) return fail(WW_KOKKOS_ERR_RUNTIME, "out of memory");
  if (hipMemcpyAsync(xf.p, x_host, static_cast<size_t>(M) * q.kin * 4, hipMemcpyHostToDevice, st) != hipSuccess) {
    xf.release();
    return fail(WW_KOKKOS_ERR_RUNTIME, "ww3gpu_snlsur_forward: upload");
  }
  rc = to_bf16(M, q.kin, xf.as<float>(), n->x.as<uint16_t>(), nullptr, st) == hipSuccess
           ? forward(*n, M, true, st)
           : fail(WW_KOKKOS_ERR_KERNEL, "ww3gpu_snlsur_forward: features");
  if (rc == WW_KOKKOS_OK &&
      (hipMemcpyAsync(o_host, n->o.p, static_cast<size_t>(M) * q.nout * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
       hipStreamSynchronize(st) != hipSuccess))
    rc = fail(WW_KOKKOS_ERR_RUNTIME, "ww3gpu_snlsur_forward: download");
  (void)hipStreamSynchronize(st);
  xf.release();
  if (rc == WW_KOKKOS_OK) report(g_last, WW_KOKKOS_OK, "");
  return rc;
}

int ww3gpu_snlsur_synth_dev(int igrid, int npts, unsigned long long seed, float bimodal, float* a, float* cg,
                            float* kd, void* stream) {
  Net* n = find(igrid);
  if (n == nullptr) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, "ww3gpu_snlsur_synth_dev: no surrogate for this grid");
  if (npts <= 0 || a == nullptr || cg == nullptr || kd == nullptr)
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, "ww3gpu_snlsur_synth_dev: bad arguments");
  if (synth(n->norm.nk, n->norm.nth, n->norm.sig, npts, seed, bimodal, a, cg, kd,
            static_cast<hipStream_t>(stream)) != hipSuccess)
    return fail(WW_KOKKOS_ERR_KERNEL, "ww3gpu_snlsur_synth_dev: launch");
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

int ww3gpu_snlsur_train_dev(int igrid, int B, const float* a, const float* cg, const float* kd, const float* s,
                            float lr, float* mean_loss, void* stream) {
  Net* n = find(igrid);
  if (n == nullptr) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, "ww3gpu_snlsur_train_dev: no surrogate for this grid");
  if (B <= 0 || B % kTile || a == nullptr || cg == nullptr || kd == nullptr || s == nullptr)
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, "ww3gpu_snlsur_train_dev: B must be a positive multiple of 128");
  hipStream_t st = static_cast<hipStream_t>(stream);
  int rc = train_batch(*n, B, a, cg, kd, s, lr, st);
  if (rc != WW_KOKKOS_OK) return rc;
  if (mean_loss != nullptr) {
    std::vector<float> l(B);
    if (hipMemcpyAsync(l.data(), n->loss.p, static_cast<size_t>(B) * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(WW_KOKKOS_ERR_RUNTIME, "ww3gpu_snlsur_train_dev: loss");
    double acc = 0.0;
    for (float v : l) acc += v;
    *mean_loss = static_cast<float>(acc / B);
  }
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

int ww3gpu_snlsur_fit(int igrid, int nsamples, int batch, int epochs, float lr, unsigned long long seed,
                      float bimodal, float* epoch_loss, void* stream) {
  Net* n = find(igrid);
  if (n == nullptr) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, "ww3gpu_snlsur_fit: no surrogate for this grid");
  if (batch <= 0 || batch % kTile || nsamples < batch || epochs <= 0)
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, "ww3gpu_snlsur_fit: batch must be a multiple of 128 and <= nsamples");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Norm& q = n->norm;
  const size_t N = static_cast<size_t>(nsamples);
  SUR_ALLOC(n->data_a, N * q.nspec * 4);
  SUR_ALLOC(n->data_s, N * q.nspec * 4);
  SUR_ALLOC(n->data_d, N * q.nspec * 4);
  SUR_ALLOC(n->data_cg, N * q.nk * 4);
  SUR_ALLOC(n->data_kd, N * 4);
  float* A = n->data_a.as<float>();
  float* S = n->data_s.as<float>();
  float* CG = n->data_cg.as<float>();
  float* KD = n->data_kd.as<float>();
  if (synth(q.nk, q.nth, q.sig, nsamples, seed, bimodal, A, CG, KD, st) != hipSuccess)
    return fail(WW_KOKKOS_ERR_KERNEL, "ww3gpu_snlsur_fit: spectra");
  // Targets from the exact DIA of this library.
  if (ww3gpu_snl1_dev(igrid, nsamples, A, CG, KD, S, n->data_d.as<float>(), st) != WW_KOKKOS_OK)
    return fail(WW_KOKKOS_ERR_KERNEL, "ww3gpu_snlsur_fit: DIA targets");
  const int nb = nsamples / batch;
  const long total = static_cast<long>(nb) * epochs;
  std::mt19937_64 rng(seed ^ 0x5bd1e995ull);
  std::vector<int> order(nb);
  long step = 0;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int b = 0; b < nb; ++b) order[b] = b;
    std::shuffle(order.begin(), order.end(), rng);
    double acc = 0.0;
    for (int b = 0; b < nb; ++b, ++step) {
      // Cosine decay from lr to lr / 100.
      const double frac = static_cast<double>(step) / static_cast<double>(total);
      const float lrs = static_cast<float>(lr * (0.01 + 0.99 * 0.5 * (1.0 + std::cos(3.141592653589793 * frac))));
      const size_t o = static_cast<size_t>(order[b]) * batch;
      const int rc = train_batch(*n, batch, A + o * q.nspec, CG + o * q.nk, KD + o, S + o * q.nspec, lrs, st);
      if (rc != WW_KOKKOS_OK) return rc;
      if (epoch_loss != nullptr && b == nb - 1) {
        std::vector<float> l(batch);
        if (hipMemcpyAsync(l.data(), n->loss.p, static_cast<size_t>(batch) * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
          return fail(WW_KOKKOS_ERR_RUNTIME, "ww3gpu_snlsur_fit: loss");
        for (float v : l) acc += v;
        epoch_loss[ep] = static_cast<float>(acc / batch);
      }
    }
  }
  if (hipStreamSynchronize(st) != hipSuccess) return fail(WW_KOKKOS_ERR_KERNEL, "ww3gpu_snlsur_fit: sync");
  n->data_a.release();
  n->data_s.release();
  n->data_d.release();
  n->data_cg.release();
  n->data_kd.release();
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

int ww3gpu_snlsur_dev(int igrid, int npts, const float* a, const float* cg, const float* kd, float* s,
                      int* flags, int* nfallback, void* stream) {
  Net* n = find(igrid);
  if (n == nullptr) return fail(WW_KOKKOS_ERR_NOT_INITIALISED, "ww3gpu_snlsur_dev: no surrogate for this grid");
  if (npts < 0) return fail(WW_KOKKOS_ERR_BAD_SHAPE, "ww3gpu_snlsur_dev: npts is negative");
  if (nfallback != nullptr) *nfallback = 0;
  if (npts == 0) {
    report(g_last, WW_KOKKOS_OK, "");
    return WW_KOKKOS_OK;
  }
  if (a == nullptr || cg == nullptr || kd == nullptr || s == nullptr)
    return fail(WW_KOKKOS_ERR_BAD_SHAPE, "ww3gpu_snlsur_dev: null array pointer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Norm& q = n->norm;
  int total_fb = 0;
  for (int p0 = 0; p0 < npts; p0 += kInferChunk) {
    const int np = std::min(kInferChunk, npts - p0);
    const int mpad = round_up(np, kTile);
    int rc = ensure_work(*n, mpad, false);
    if (rc != WW_KOKKOS_OK) return rc;
    const float* ac = a + static_cast<size_t>(p0) * q.nspec;
    const float* cgc = cg + static_cast<size_t>(p0) * q.nk;
    const float* kdc = kd + p0;
    float* sc = s + static_cast<size_t>(p0) * q.nspec;
    SUR_TRY(prep(q, np, mpad, ac, cgc, kdc, nullptr, n->x.as<uint16_t>(), nullptr, 0, nullptr, n->scale.as<float>(),
                 n->shift.as<int>(), n->flag.as<int>(), st), "ww3gpu_snlsur_dev: features");
    if ((rc = forward(*n, mpad, false, st)) != WW_KOKKOS_OK) return rc;
    SUR_TRY(hipMemsetAsync(n->nlist.p, 0, sizeof(int), st), "ww3gpu_snlsur_dev: clear");
    SUR_TRY(post(q, np, n->o.as<uint16_t>(), cgc, n->scale.as<float>(), n->shift.as<int>(), n->flag.as<int>(), sc,
                 n->list.as<int>(), n->nlist.as<int>(), st), "ww3gpu_snlsur_dev: output");
    if (flags != nullptr)
      SUR_TRY(hipMemcpyAsync(flags + p0, n->flag.p, static_cast<size_t>(np) * 4, hipMemcpyDeviceToDevice, st),
              "ww3gpu_snlsur_dev: flags");
    int nfb = 0;
    SUR_TRY(hipMemcpyAsync(&nfb, n->nlist.p, sizeof(int), hipMemcpyDeviceToHost, st), "ww3gpu_snlsur_dev: count");
    SUR_TRY(hipStreamSynchronize(st), "ww3gpu_snlsur_dev: sync");
    if (nfb > 0) {
      const size_t F = static_cast<size_t>(nfb);
      SUR_ALLOC(n->fa, F * q.nspec * 4);
      SUR_ALLOC(n->fs, F * q.nspec * 4);
      SUR_ALLOC(n->fd, F * q.nspec * 4);
      SUR_ALLOC(n->fcg, F * q.nk * 4);
      SUR_ALLOC(n->fkd, F * 4);
      SUR_TRY(gather(q.nspec, q.nk, n->list.as<int>(), nfb, ac, cgc, kdc, n->fa.as<float>(), n->fcg.as<float>(),
                     n->fkd.as<float>(), st), "ww3gpu_snlsur_dev: gather");
      if (ww3gpu_snl1_dev(igrid, nfb, n->fa.as<float>(), n->fcg.as<float>(), n->fkd.as<float>(), n->fs.as<float>(),
                          n->fd.as<float>(), st) != WW_KOKKOS_OK)
        return fail(WW_KOKKOS_ERR_KERNEL, "ww3gpu_snlsur_dev: DIA fallback");
      SUR_TRY(scatter(q.nspec, n->list.as<int>(), nfb, n->fs.as<float>(), sc, st), "ww3gpu_snlsur_dev: scatter");
    }
    total_fb += nfb;
  }
  if (nfallback != nullptr) *nfallback = total_fb;
  report(g_last, WW_KOKKOS_OK, "");
  return WW_KOKKOS_OK;
}

int ww3gpu_snlsur_last_error(void) { return g_last.load(std::memory_order_relaxed); }

int ww3gpu_sur_gemm_dev(int M, int N, int K, const void* a, int lda, const void* b, int ldb, float* c, int ldc,
                        void* stream) {
  if (a == nullptr || b == nullptr || c == nullptr || lda < K || ldb < K || ldc < N || lda % 8 || ldb % 8)
    return WW_KOKKOS_ERR_BAD_SHAPE;
  Epi e{};
  e.mode = kPlain;
  e.of = c;
  e.ldo = ldc;
  if (gemm_nt(M, N, K, static_cast<const uint16_t*>(a), lda, static_cast<const uint16_t*>(b), ldb, e,
              static_cast<hipStream_t>(stream)) != hipSuccess)
    return WW_KOKKOS_ERR_BAD_SHAPE;
  return WW_KOKKOS_OK;
}

}  // extern "C"

// ww3-gpu_amd/csrc/snlsur_kernels.hip
// Kernels of the W3SNL1 surrogate: a bf16 GEMM on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16) with the layer epilogues fused, the normalisation around the
// network, the training loss, Adam, and a seeded spectrum generator for training data.
#include <algorithm>
#include <cmath>

#include "snlsur.hpp"

namespace ww3gpu {
namespace sur {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);                     // round to nearest even
  return static_cast<uint16_t>(u >> 16);
}

// GELU in its tanh form, 0.5 z (1 + tanh(c (z + 0.044715 z^3))) with c = sqrt(2/pi),
// written as z * sigmoid(2 c (z + 0.044715 z^3)): one exponential and one reciprocal.
__device__ __forceinline__ float gelu(float z) {
  const float u = 1.5957691f * (z + 0.044715f * z * z * z);
  return z * __frcp_rn(1.0f + __expf(-u));
}

__device__ __forceinline__ float gelu_d(float z) {
  const float u = 1.5957691f * (z + 0.044715f * z * z * z);
  const float sg = __frcp_rn(1.0f + __expf(-u));
  return sg + z * sg * (1.0f - sg) * 1.5957691f * (1.0f + 0.134145f * z * z);
}

__device__ __forceinline__ float wave_sum(float v) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// ---------------------------------------------------------------------------------------
// GEMM C = A B^T.  A workgroup of four wavefronts computes a 128 x 128 tile of C; each
// wavefront a 64 x 64 quarter as 4 x 4 MFMA tiles of 16 x 16.  The reduction runs in steps
// of 64: the next step's A and B tiles (128 x 64 bf16 each) are loaded into registers while
// the MFMAs consume the current ones from LDS, then stored into the other LDS stage.  LDS
// rows are 144 bytes (128 + 16), so the 16-byte fragment reads of 16 consecutive rows fall
// into distinct banks.
// MFMA 16x16x32 operand layout: lane l holds row (l & 15) of A (and column (l & 15) of B)
// with k = 8 (l >> 4) .. 8 (l >> 4) + 7; the result in lane l is column (l & 15), rows
// 4 (l >> 4) .. 4 (l >> 4) + 3.
constexpr int kRowB = kTileK * 2 + 16;
constexpr int kStageB = 2 * kTile * kRowB;

// Persistent tiles: a grid of G workgroups (two per CU) walks the T output tiles, tile
// index t = blockIdx.x, + G, + 2G, ...  While a tile's epilogue runs, the first A and B
// slabs of the next tile are already in flight.
// Tile order: workgroups are dealt to the eight XCDs in turn (workgroup g to XCD g % 8), and
// each XCD has its own L2.  Tile (t % 8) * (T / 8) + t / 8 (n fastest within an m row)
// gives every XCD a contiguous run of tiles (G is a multiple of 8), so the N / 128 tiles
// that read the same rows of A run on one XCD, close together in time, and share them
// through its L2.
__device__ __forceinline__ int tile_of(int t, int T) { return T % 8 == 0 ? (t % 8) * (T / 8) + t / 8 : t; }

template <int MODE>
__global__ __launch_bounds__(256) void gemm_nt_kernel(int K, int ntn, int T, const uint16_t* __restrict__ A, int lda,
                                                      const uint16_t* __restrict__ B, int ldb, Epi e) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[2 * kStageB];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int G = static_cast<int>(gridDim.x);
  int t = static_cast<int>(blockIdx.x);
  if (t >= T) return;
  const int lr = tid >> 3, lc = tid & 7;
  int m0 = (tile_of(t, T) / ntn) * kTile, n0 = (tile_of(t, T) % ntn) * kTile;
  const uint16_t* ga = A + static_cast<size_t>(m0 + lr) * lda + lc * 8;
  const uint16_t* gb = B + static_cast<size_t>(n0 + lr) * ldb + lc * 8;

  u32x4 ra[4], rb[4];
  // Four rows 32 apart per thread, written out (a loop here would leave the arrays in
  // scratch memory).
#define SUR_LD(r, k0)                                                                         \
  ra[r] = *reinterpret_cast<const u32x4*>(ga + static_cast<size_t>(32 * (r)) * lda + (k0));   \
  rb[r] = *reinterpret_cast<const u32x4*>(gb + static_cast<size_t>(32 * (r)) * ldb + (k0));
#define SUR_GLOAD(k0) { SUR_LD(0, k0) SUR_LD(1, k0) SUR_LD(2, k0) SUR_LD(3, k0) }
#define SUR_ST(r, sa_, sb_)                                                                   \
  *reinterpret_cast<u32x4*>((sa_) + (lr + 32 * (r)) * kRowB + lc * 16) = ra[r];               \
  *reinterpret_cast<u32x4*>((sb_) + (lr + 32 * (r)) * kRowB + lc * 16) = rb[r];
#define SUR_SSTORE(stage)                                                                     \
  {                                                                                           \
    unsigned char* sa_ = lds + (stage) * kStageB;                                             \
    unsigned char* sb_ = sa_ + kTile * kRowB;                                                 \
    SUR_ST(0, sa_, sb_) SUR_ST(1, sa_, sb_) SUR_ST(2, sa_, sb_) SUR_ST(3, sa_, sb_)           \
  }

  const int nkt = K / kTileK;
  SUR_GLOAD(0)
  SUR_SSTORE(0)
  __syncthreads();
  while (true) {
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int kt = 0; kt < nkt; ++kt) {
      const int cur = kt & 1;
      if (kt + 1 < nkt) {
        SUR_GLOAD((kt + 1) * kTileK)
      }
      const unsigned char* sa = lds + cur * kStageB;
      const unsigned char* sb = sa + kTile * kRowB;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        bf16x8 fa[4], fb[4];
        const int koff = s * 64 + (lane >> 4) * 16;
#pragma unroll
        for (int i = 0; i < 4; ++i)
          fa[i] = *reinterpret_cast<const bf16x8*>(sa + (wm * 64 + i * 16 + (lane & 15)) * kRowB + koff);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          fb[j] = *reinterpret_cast<const bf16x8*>(sb + (wn * 64 + j * 16 + (lane & 15)) * kRowB + koff);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
      }
      if (kt + 1 < nkt) SUR_SSTORE(cur ^ 1)
      __syncthreads();
    }

    // Next tile's first slab into registers before this tile's epilogue.
    const int tn = t + G;
    const bool more = tn < T;
    const int em0 = m0, en0 = n0;
    if (more) {
      m0 = (tile_of(tn, T) / ntn) * kTile;
      n0 = (tile_of(tn, T) % ntn) * kTile;
      ga = A + static_cast<size_t>(m0 + lr) * lda + lc * 8;
      gb = B + static_cast<size_t>(n0 + lr) * ldb + lc * 8;
      SUR_GLOAD(0)
    }

#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int n = en0 + wn * 64 + j * 16 + (lane & 15);
        const int mb = em0 + wm * 64 + i * 16 + 4 * (lane >> 4);
        float bn = 0.f;
        if constexpr (MODE == kGeluFwd || MODE == kLinearFwd) bn = e.bias[n];
        uint16_t hv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float c = acc[i][j][r];
          const size_t m = static_cast<size_t>(mb + r);
          if constexpr (MODE == kGeluFwd) {
            const float z = c + bn;
            if (e.zf != nullptr) e.zf[m * e.ldz + n] = z;
            hv[r] = to_bf(gelu(z));
            e.hb[m * e.ldh + n] = hv[r];
          } else if constexpr (MODE == kLinearFwd) {
            if (e.of != nullptr)
              e.of[m * e.ldo + n] = c + bn;
            else
              e.hb[m * e.ldh + n] = to_bf(c + bn);
          } else if constexpr (MODE == kPlain) {
            e.of[m * e.ldo + n] = c;
          } else {
            hv[r] = to_bf(c * gelu_d(e.zf[m * e.ldz + n]));
            e.hb[m * e.ldh + n] = hv[r];
          }
        }
        if constexpr (MODE == kGeluFwd || MODE == kGeluBwd) {
          if (e.ht != nullptr) {
            uint2 pk;
            pk.x = static_cast<uint32_t>(hv[0]) | (static_cast<uint32_t>(hv[1]) << 16);
            pk.y = static_cast<uint32_t>(hv[2]) | (static_cast<uint32_t>(hv[3]) << 16);
            *reinterpret_cast<uint2*>(e.ht + static_cast<size_t>(n) * e.ldt + mb) = pk;
          }
        }
      }
    }
    if (!more) break;
    // Every wave left the k loop through its final barrier, so no wave still reads LDS.
    SUR_SSTORE(0)
    __syncthreads();
    t = tn;
  }
#undef SUR_LD
#undef SUR_GLOAD
#undef SUR_ST
#undef SUR_SSTORE
}

// ---------------------------------------------------------------------------------------
// Normalisation around the network: one wavefront per point, four points per workgroup.
// A point's spectrum is read from HBM once, as UE = A / CONX, into LDS; the rotated
// features, the scale and (training) the targets are formed from there.  The divisions by
// CONX are multiplications by one reciprocal per frequency: the surrogate's inputs need no
// bit parity with W3SNL1.
constexpr int kMaxSpec = 2048;  // NSPEC limit of the normalisation kernels
constexpr int kMaxK = 64;       // NK limit

// i / nth for 0 <= i < 2048 and nth <= 64 without an integer division: (i + 1/2) / nth sits
// at least 1/(2 nth) >= 2^-7 from an integer, and the float product errs by at most
// 2048 * 2^-23 = 2^-12, so truncation gives the exact quotient.
__device__ __forceinline__ int div_nth(int i, float inv_nth) {
  return static_cast<int>((static_cast<float>(i) + 0.5f) * inv_nth);
}

__device__ __forceinline__ float conx_of(const Norm& nm, const float* cgr, int k) {
  return nm.tpiinv / nm.sig[k] * cgr[k];
}

__global__ __launch_bounds__(256) void prep_kernel(Norm nm, int npts, const float* __restrict__ a,
                                                   const float* __restrict__ cg, const float* __restrict__ kd,
                                                   const float* __restrict__ s_ref, uint16_t* __restrict__ x,
                                                   uint16_t* __restrict__ xt, int ldt, float* __restrict__ y,
                                                   float* scale, int* shift, int* flag) {
  extern __shared__ float rows[];  // 4 x NSPEC floats: each wavefront's UE (sized at launch)
  __shared__ float icx[4][kMaxK];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p = blockIdx.x * 4 + wv;
  const bool live = p < npts;
  const int nspec = nm.nspec, nth = nm.nth, kin = nm.kin, nout = nm.nout;
  const float inv_nth = 1.0f / static_cast<float>(nth);
  uint16_t* xr = x + static_cast<size_t>(p) * kin;
  float* yr = y != nullptr ? y + static_cast<size_t>(p) * nout : nullptr;
  const float* ar = a + static_cast<size_t>(p) * nspec;
  const float* cr = cg + static_cast<size_t>(p) * nm.nk;
  float* ue = rows + wv * nspec;
  if (live && lane < nm.nk) icx[wv][lane] = 1.0f / conx_of(nm, cr, lane);
  __syncthreads();
  float best = -1.0f;
  int bi = 0x7fffffff;
  bool bad = !live;
  if (live) {
#pragma unroll 6
    for (int i = lane; i < nspec; i += 64) {
      const float u = ar[i] * icx[wv][div_nth(i, inv_nth)];
      ue[i] = u;
      if (!(u >= 0.0f) || isinf(u)) bad = true;
      if (u > best) {
        best = u;
        bi = i;
      }
    }
  }
  __syncthreads();
  for (int off = 32; off >= 1; off >>= 1) {
    const float ob = __shfl_xor(best, off);
    const int oi = __shfl_xor(bi, off);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
  const float kdp = live ? kd[p] : 1.0f;
  bad = __any(bad) || !(best > 0.0f) || !(kdp > 0.0f) || isinf(kdp);
  const int jp = bad ? 0 : bi % nth;
  const float inv_m = bad ? 0.0f : 1.0f / best;
  // Features in the rotated order, and q = sum f_k^11 u^3 in that order, so that a rotated
  // input gives the same q bit for bit.
  float q = 0.0f;
#pragma unroll 6
  for (int i = lane; i < nspec; i += 64) {
    const int k = div_nth(i, inv_nth), t = i - k * nth;
    int ts = t + jp;
    if (ts >= nth) ts -= nth;
    const float u = bad ? 0.0f : ue[k * nth + ts] * inv_m;
    const float f = nm.sig[k] * nm.tpiinv, f2 = f * f, f4 = f2 * f2, f8 = f4 * f4;
    q += f8 * f2 * f * (u * u * u);
    const uint16_t h = to_bf(u);
    xr[i] = h;
    if (xt != nullptr) xt[static_cast<size_t>(i) * ldt + p] = h;
  }
  for (int i = nspec + lane; i < kin; i += 64) {
    const uint16_t h = to_bf(i == nspec && !bad ? 0.25f * logf(kdp) : 0.0f);
    xr[i] = h;
    if (xt != nullptr) xt[static_cast<size_t>(i) * ldt + p] = h;
  }
  q = wave_sum(q);
  float sc = 0.0f;
  if (!bad) {
    const float xx = fmaxf(nm.kdcon * kdp, nm.kdmn);
    const float cons =
        nm.snlc1 * (1.0f + nm.snls1 / xx * (1.0f - nm.snls2 * xx) * expf(fmaxf(-1.e15f, nm.snls3 * xx)));
    sc = cons * (best * best * best) * (kScaleC0 * q);
    if (!(sc > 0.0f) || isinf(sc)) bad = true;
  }
  if (yr != nullptr) {
    const float inv_sc = bad ? 0.0f : 1.0f / sc;
    for (int i = lane; i < nout; i += 64) {
      float v = 0.0f;
      if (i < nspec && !bad) {
        const int k = div_nth(i, inv_nth), t = i - k * nth;
        int ts = t + jp;
        if (ts >= nth) ts -= nth;
        v = s_ref[static_cast<size_t>(p) * nspec + k * nth + ts] * icx[wv][k] * inv_sc;
      }
      yr[i] = v;
    }
  }
  if (lane == 0) {
    scale[p] = bad ? 0.0f : sc;
    shift[p] = jp;
    flag[p] = bad ? 1 : 0;
  }
}

__global__ __launch_bounds__(256) void post_kernel(Norm nm, int npts, const uint16_t* __restrict__ o,
                                                   const float* __restrict__ cg, const float* __restrict__ scale,
                                                   const int* __restrict__ shift, int* flag, float* __restrict__ s,
                                                   int* list, int* nlist) {
  __shared__ float cx[4][kMaxK];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p = blockIdx.x * 4 + wv;
  const bool live = p < npts;
  const float* cr = cg + static_cast<size_t>(p) * nm.nk;
  const float sc = live ? scale[p] : 0.0f;
  if (live && lane < nm.nk) cx[wv][lane] = conx_of(nm, cr, lane) * sc;
  __syncthreads();
  if (!live) return;
  const int nspec = nm.nspec, nth = nm.nth;
  const float inv_nth = 1.0f / static_cast<float>(nth);
  const int jp = shift[p];
  const uint16_t* orow = o + static_cast<size_t>(p) * nm.nout;
  float* sr = s + static_cast<size_t>(p) * nspec;
  float er = 0.0f, ea = 0.0f;
  bool nonfinite = false;
#pragma unroll 6
  for (int i = lane; i < nspec; i += 64) {
    const int k = div_nth(i, inv_nth), t = i - k * nth;
    int tc = t - jp;
    if (tc < 0) tc += nth;
    const float v = __uint_as_float(static_cast<uint32_t>(orow[k * nth + tc]) << 16) * cx[wv][k];
    sr[i] = v;
    const float wk = nm.wsig[k];
    er += v * wk;
    ea += fabsf(v) * wk;
    if (!isfinite(v)) nonfinite = true;
  }
  er = wave_sum(er);
  ea = wave_sum(ea);
  const bool fail = flag[p] != 0 || __any(nonfinite) || !(fabsf(er) <= kEnergyTol * ea);
  if (lane == 0) {
    flag[p] = fail ? 1 : 0;
    if (fail) list[atomicAdd(nlist, 1)] = p;
  }
}

__global__ __launch_bounds__(256) void loss_kernel(int B, int nout, const float* __restrict__ o,
                                                   const float* __restrict__ y, float* loss,
                                                   uint16_t* __restrict__ go, uint16_t* __restrict__ got) {
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= B) return;
  const float* orow = o + static_cast<size_t>(p) * nout;
  const float* yrow = y + static_cast<size_t>(p) * nout;
  float yy = 0.0f, ee = 0.0f;
  for (int i = lane; i < nout; i += 64) {
    const float d = orow[i] - yrow[i];
    yy += yrow[i] * yrow[i];
    ee += d * d;
  }
  yy = wave_sum(yy);
  ee = wave_sum(ee);
  const float wgt = yy > 0.0f ? 1.0f / yy : 0.0f;
  if (lane == 0) loss[p] = ee * wgt;
  const float gs = 2.0f * wgt / static_cast<float>(B);
  for (int i = lane; i < nout; i += 64) {
    const uint16_t h = to_bf((orow[i] - yrow[i]) * gs);
    go[static_cast<size_t>(p) * nout + i] = h;
    got[static_cast<size_t>(i) * B + p] = h;
  }
}

__global__ __launch_bounds__(64) void column_sum_kernel(int M, const uint16_t* __restrict__ gt, float* db) {
  const int n = blockIdx.x, lane = threadIdx.x;
  const uint16_t* row = gt + static_cast<size_t>(n) * M;
  float acc = 0.0f;
  for (int m = lane; m < M; m += 64) acc += __uint_as_float(static_cast<uint32_t>(row[m]) << 16);
  acc = wave_sum(acc);
  if (lane == 0) db[n] = acc;
}

__global__ __launch_bounds__(256) void adam_kernel(int rows, int cols, float* w, const float* g, float* m,
                                                   float* v, uint16_t* wb, uint16_t* wt, float lr, float b1,
                                                   float b2, float eps, float bc1, float bc2, int update) {
  const size_t n = static_cast<size_t>(rows) * cols;
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256) {
    float wi = w[i];
    if (update) {
      const float gi = g[i];
      const float mi = b1 * m[i] + (1.0f - b1) * gi;
      const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
      m[i] = mi;
      v[i] = vi;
      wi -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
      w[i] = wi;
    }
    const uint16_t h = to_bf(wi);
    if (wb != nullptr) wb[i] = h;
    if (wt != nullptr) {
      const size_t r = i / cols, c = i - r * cols;
      wt[c * rows + r] = h;
    }
  }
}

__global__ __launch_bounds__(256) void gather_kernel(int nspec, int nk, const int* list, const float* a,
                                                     const float* cg, const float* kd, float* ap, float* cgp,
                                                     float* kdp) {
  const int q = blockIdx.x;
  const size_t p = static_cast<size_t>(list[q]);
  for (int i = threadIdx.x; i < nspec; i += 256) ap[static_cast<size_t>(q) * nspec + i] = a[p * nspec + i];
  for (int i = threadIdx.x; i < nk; i += 256) cgp[static_cast<size_t>(q) * nk + i] = cg[p * nk + i];
  if (threadIdx.x == 0) kdp[q] = kd[p];
}

__global__ __launch_bounds__(256) void scatter_kernel(int nspec, const int* list, const float* sp, float* s) {
  const int q = blockIdx.x;
  const size_t p = static_cast<size_t>(list[q]);
  for (int i = threadIdx.x; i < nspec; i += 256) s[p * nspec + i] = sp[static_cast<size_t>(q) * nspec + i];
}

// ---------------------------------------------------------------------------------------
// Training spectra.  Counter-based uniforms (SplitMix64 finaliser of seed, point, draw), so
// a point's sea state does not depend on the launch shape.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double uni(unsigned long long seed, long long p, int j) {
  return static_cast<double>(mix64(seed ^ mix64(static_cast<unsigned long long>(p) * 64ull + j)) >> 11) *
         0x1.0p-53;
}

constexpr int kSynthMax = 64;  // NK and NTH limit of the generator

__global__ __launch_bounds__(256) void synth_kernel(int nk, int nth, const float* __restrict__ sig, int npts,
                                                    unsigned long long seed, float bimodal, float* __restrict__ a,
                                                    float* __restrict__ cg, float* __restrict__ kd) {
  __shared__ double sh[4][6 * kSynthMax];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // A wavefront past npts still reaches the barrier below; it computes for point npts-1
  // and writes nothing.
  const bool live = static_cast<int>(blockIdx.x) * 4 + wv < npts;
  const int p = live ? static_cast<int>(blockIdx.x) * 4 + wv : npts - 1;
  double* kk = sh[wv];
  double* cgk = kk + kSynthMax;
  double* ef1 = cgk + kSynthMax;
  double* ef2 = ef1 + kSynthMax;
  double* d1 = ef2 + kSynthMax;
  double* d2 = d1 + kSynthMax;
  const double g = 9.806, pi = 3.141592653589793;
  const double two = uni(seed, p, 0) < static_cast<double>(bimodal) ? 1.0 : 0.0;
  const double hs1 = 0.5 + 7.5 * uni(seed, p, 1);
  const double tp1 = two > 0.0 ? 3.0 + 6.0 * uni(seed, p, 2) : 4.0 + 12.0 * uni(seed, p, 2);
  const double ga1 = 1.0 + 4.0 * uni(seed, p, 3);
  const double th1 = 2.0 * pi * uni(seed, p, 4);
  const double sp1 = 2.0 + 18.0 * uni(seed, p, 5);
  const double hs2 = 0.5 + 3.5 * uni(seed, p, 6);
  const double tp2 = 10.0 + 8.0 * uni(seed, p, 7);
  const double ga2 = 1.0 + 6.0 * uni(seed, p, 8);
  const double th2 = 2.0 * pi * uni(seed, p, 9);
  const double sp2 = 10.0 + 30.0 * uni(seed, p, 10);
  const double depth = exp(log(10.0) + (log(4000.0) - log(10.0)) * uni(seed, p, 11));
  const double dth = 2.0 * pi / nth;
  const double xfr = nk > 1 ? static_cast<double>(sig[1]) / static_cast<double>(sig[0]) : 1.1;

  auto jonswap = [&](double f, double fp, double gam) {
    const double sw = f <= fp ? 0.07 : 0.09;
    const double q = (f - fp) / (sw * fp);
    return g * g * pow(2.0 * pi, -4.0) * pow(f, -5.0) * exp(-1.25 * pow(fp / f, 4.0)) * pow(gam, exp(-0.5 * q * q));
  };
  if (lane < nk) {
    const double s = sig[lane];
    double k = s * s / g;
    for (int it = 0; it < 30; ++it) {
      const double t = tanh(k * depth);
      const double ch = cosh(fmin(k * depth, 300.0));
      k -= (g * k * t - s * s) / (g * t + g * k * depth / (ch * ch));
    }
    const double kd2 = fmin(2.0 * k * depth, 80.0);
    kk[lane] = k;
    cgk[lane] = 0.5 * s / k * (1.0 + kd2 / sinh(kd2));
    const double f = s / (2.0 * pi);
    ef1[lane] = jonswap(f, 1.0 / tp1, ga1);
    ef2[lane] = two > 0.0 ? jonswap(f, 1.0 / tp2, ga2) : 0.0;
  }
  if (lane < nth) {
    const double th = lane * dth;
    const double c1 = cos(0.5 * (th - th1)), c2 = cos(0.5 * (th - th2));
    d1[lane] = pow(c1 * c1, sp1);
    d2[lane] = pow(c2 * c2, sp2);
  }
  __syncthreads();
  double sd1 = 0.0, sd2 = 0.0, m01 = 0.0, m02 = 0.0;
  for (int t = 0; t < nth; ++t) {
    sd1 += d1[t];
    sd2 += d2[t];
  }
  for (int k = 0; k < nk; ++k) {
    const double dsig = static_cast<double>(sig[k]) * 0.5 * (xfr - 1.0 / xfr);
    m01 += ef1[k] * dsig / (2.0 * pi);
    m02 += ef2[k] * dsig / (2.0 * pi);
  }
  const double al1 = (hs1 / 4.0) * (hs1 / 4.0) / fmax(m01, 1e-300);
  const double al2 = two > 0.0 ? (hs2 / 4.0) * (hs2 / 4.0) / fmax(m02, 1e-300) : 0.0;
  const double n1 = 1.0 / (sd1 * dth), n2 = sd2 > 0.0 ? 1.0 / (sd2 * dth) : 0.0;
  if (!live) return;
  const int nspec = nk * nth;
  float* ar = a + static_cast<size_t>(p) * nspec;
  for (int i = lane; i < nspec; i += 64) {
    const int k = i / nth, t = i - k * nth;
    const double e = (al1 * ef1[k] * d1[t] * n1 + al2 * ef2[k] * d2[t] * n2) / (2.0 * pi);
    ar[i] = static_cast<float>(e * cgk[k] / static_cast<double>(sig[k]));
  }
  if (lane < nk) cg[static_cast<size_t>(p) * nk + lane] = static_cast<float>(cgk[lane]);
  if (lane == 0) {
    double num = 0.0, den = 0.0;
    for (int k = 0; k < nk; ++k) {
      const double ek = al1 * ef1[k] + al2 * ef2[k];
      num += ek * kk[k] * depth;
      den += ek;
    }
    kd[p] = static_cast<float>(num / fmax(den, 1e-300));
  }
}

}  // namespace

hipError_t gemm_nt(int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb,
                   const Epi& e, hipStream_t st) {
  if (M <= 0 || N <= 0 || K <= 0 || M % kTile || N % kTile || K % kTileK) return hipErrorInvalidValue;
  const int ntn = N / kTile;
  const long tiles = static_cast<long>(M / kTile) * ntn;
  if (tiles > 0x7fffffffL) return hipErrorInvalidValue;
  static const int cus = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      n = 256;
    return n;
  }();
  const int T = static_cast<int>(tiles);
  const int G = T < 2 * cus ? T : (2 * cus) / 8 * 8;  // two workgroups per CU, a multiple of 8
  const dim3 grid(static_cast<unsigned>(G)), block(256);
  switch (e.mode) {
    case kGeluFwd: gemm_nt_kernel<kGeluFwd><<<grid, block, 0, st>>>(K, ntn, T, A, lda, B, ldb, e); break;
    case kLinearFwd: gemm_nt_kernel<kLinearFwd><<<grid, block, 0, st>>>(K, ntn, T, A, lda, B, ldb, e); break;
    case kPlain: gemm_nt_kernel<kPlain><<<grid, block, 0, st>>>(K, ntn, T, A, lda, B, ldb, e); break;
    case kGeluBwd: gemm_nt_kernel<kGeluBwd><<<grid, block, 0, st>>>(K, ntn, T, A, lda, B, ldb, e); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t prep(const Norm& n, int npts, int mpad, const float* a, const float* cg, const float* kd,
                const float* s_ref, uint16_t* x, uint16_t* xt, int ldt, float* y, float* scale, int* shift,
                int* flag, hipStream_t st) {
  if (mpad % 4 || npts > mpad || n.nspec > kMaxSpec || n.nk > kMaxK) return hipErrorInvalidValue;
  prep_kernel<<<mpad / 4, 256, 4 * n.nspec * sizeof(float), st>>>(n, npts, a, cg, kd, s_ref, x, xt, ldt, y, scale,
                                                                   shift, flag);
  return hipGetLastError();
}

hipError_t post(const Norm& n, int npts, const uint16_t* o, const float* cg, const float* scale,
                const int* shift, int* flag, float* s, int* list, int* nlist, hipStream_t st) {
  if (n.nspec > kMaxSpec || n.nk > kMaxK) return hipErrorInvalidValue;
  post_kernel<<<(npts + 3) / 4, 256, 0, st>>>(n, npts, o, cg, scale, shift, flag, s, list, nlist);
  return hipGetLastError();
}

hipError_t loss_grad(int B, int nout, const float* o, const float* y, float* loss, uint16_t* go, uint16_t* got,
                     hipStream_t st) {
  loss_kernel<<<(B + 3) / 4, 256, 0, st>>>(B, nout, o, y, loss, go, got);
  return hipGetLastError();
}

hipError_t column_sums(int N, int M, const uint16_t* gt, float* db, hipStream_t st) {
  column_sum_kernel<<<N, 64, 0, st>>>(M, gt, db);
  return hipGetLastError();
}

hipError_t adam(int rows, int cols, float* w, const float* g, float* m, float* v, uint16_t* wb, uint16_t* wt,
                float lr, float b1, float b2, float eps, float bc1, float bc2, hipStream_t st) {
  const size_t n = static_cast<size_t>(rows) * cols;
  const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 4096));
  adam_kernel<<<blocks, 256, 0, st>>>(rows, cols, w, g, m, v, wb, wt, lr, b1, b2, eps, bc1, bc2, 1);
  return hipGetLastError();
}

hipError_t to_bf16(int rows, int cols, const float* w, uint16_t* wb, uint16_t* wt, hipStream_t st) {
  const size_t n = static_cast<size_t>(rows) * cols;
  const unsigned blocks = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 4096));
  adam_kernel<<<blocks, 256, 0, st>>>(rows, cols, const_cast<float*>(w), nullptr, nullptr, nullptr, wb, wt,
                                      0.f, 0.f, 0.f, 0.f, 1.f, 1.f, 0);
  return hipGetLastError();
}

hipError_t gather(int nspec, int nk, const int* list, int n, const float* a, const float* cg, const float* kd,
                  float* ap, float* cgp, float* kdp, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  gather_kernel<<<n, 256, 0, st>>>(nspec, nk, list, a, cg, kd, ap, cgp, kdp);
  return hipGetLastError();
}

hipError_t scatter(int nspec, const int* list, int n, const float* sp, float* s, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  scatter_kernel<<<n, 256, 0, st>>>(nspec, list, sp, s);
  return hipGetLastError();
}

hipError_t synth(int nk, int nth, const float* sig, int npts, unsigned long long seed, float bimodal, float* a,
                 float* cg, float* kd, hipStream_t st) {
  if (nk < 2 || nk > kSynthMax || nth < 2 || nth > kSynthMax || npts <= 0) return hipErrorInvalidValue;
  synth_kernel<<<(npts + 3) / 4, 256, 0, st>>>(nk, nth, sig, npts, seed, bimodal, a, cg, kd);
  return hipGetLastError();
}

}  // namespace sur
}  // namespace ww3gpu

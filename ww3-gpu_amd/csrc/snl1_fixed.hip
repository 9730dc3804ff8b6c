// ww3-gpu_amd/csrc/snl1_fixed.hip
// W3SNL1 kernels specialised for the common stencils (LAMBDA = 0.25, XFR = 1.1 with
// NTH = 24 or 36): the quadruplet offsets are compile-time constants, so every LDS
// gather is one ds_read with an immediate offset from the bin's slot.
//
// Same arithmetic, in the same order, as snl1_kernel.hip (reference
// kokkos/src/ww_kokkos/snl1_team.hpp:352-497); only the addressing differs:
//  * the workgroup is R whole frequency rows of NTH threads, so a thread keeps its
//    direction c for the whole kernel and walks rows r0 = tid/NTH + it*R;
//  * the six section-3 arrays are interleaved by row (row stride 6W, array k at +kW), so
//    one base address serves all six with immediate offsets;
//  * CONX and CONS*AF11 are computed once per row into LDS;
//  * the parametric tail is continued by the threads of row NK along their own column.
#include "devlibm.hpp"
#include "snl1_device.hpp"

namespace ww3gpu {
namespace snl1 {
namespace {

__device__ inline float fmaxo(float a, float b) { return a > b ? a : b; }

template <int NTH_, int ITHP, int ITHM, int IFRP, int IFRM, int R_>
struct Stencil {
  static constexpr int NTH = NTH_;
  static constexpr int R = R_;                 // frequency rows per pass
  static constexpr int BS = NTH * R;           // threads
  static constexpr int H = (ITHP > ITHM ? ITHP : ITHM) + 1;  // largest direction shift
  static constexpr int W = NTH + 2 * H;        // padded row
  static constexpr int RS = 6 * W;             // row stride of the interleaved section-3 arrays
  static constexpr int ZR = (1 - IFRM) > (IFRP + 1) ? (1 - IFRM) : (IFRP + 1);  // zero rows
  static constexpr int ifo(int m) {            // IF_m(IFR) - IFR, m = 1..8
    return m == 1 ? IFRP : m == 2 ? IFRP + 1 : m == 3 ? IFRM : m == 4 ? IFRM - 1
         : m == 5 ? -IFRP : m == 6 ? -(IFRP + 1) : m == 7 ? -IFRM : -(IFRM - 1);
  }
  static constexpr int ito(int n) {            // IT_n(ITH) - ITH, n = 1..8
    return n == 1 ? ITHP : n == 2 ? ITHP + 1 : n == 3 ? ITHM : n == 4 ? ITHM + 1
         : n == 5 ? -ITHP : n == 6 ? -(ITHP + 1) : n == 7 ? -ITHM : -(ITHM + 1);
  }
  // (IT, IF) of the section-3 tables IP11..IP14, IM11..IM14, IP21..IP24, IM21..IM24 and of
  // the section-4 tables IC11..IC81, IC12..IC82 (reference snl1_tables.cpp:202-239).
  static constexpr int kx(int j, int which) {
    constexpr int t[16][2] = {{2, 2}, {1, 2}, {2, 1}, {1, 1}, {8, 4}, {7, 4}, {8, 3}, {7, 3},
                              {6, 2}, {5, 2}, {6, 1}, {5, 1}, {4, 4}, {3, 4}, {4, 3}, {3, 3}};
    return t[j][which];
  }
  static constexpr int kc(int j, int which) {
    constexpr int t[16][2] = {{6, 6}, {5, 6}, {6, 5}, {5, 5}, {4, 8}, {3, 8}, {4, 7}, {3, 7},
                              {2, 6}, {1, 6}, {2, 5}, {1, 5}, {8, 8}, {7, 8}, {8, 7}, {7, 7}};
    return t[j][which];
  }
  static constexpr int o3(int j) { return ifo(kx(j, 1)) * W + ito(kx(j, 0)); }
  static constexpr int o4(int j) { return ifo(kc(j, 1)) * RS + ito(kc(j, 0)); }
};

template <class ST, int IT>
__global__ __launch_bounds__(ST::BS) void dia_fixed(const Geom g, const float* __restrict__ af11_row,
                                                    const float* __restrict__ sig,
                                                    const float* __restrict__ a,
                                                    const float* __restrict__ cg,
                                                    const float* __restrict__ kdmean,
                                                    float* __restrict__ s, float* __restrict__ d,
                                                    const int npts, const Layout L) {
  constexpr int NTH = ST::NTH, H = ST::H, W = ST::W, RS = ST::RS, ZR = ST::ZR, R = ST::R,
                BS = ST::BS;
  extern __shared__ float lds[];
  const int p = layout_point(L, npts);
  if (p < 0) return;  // uniform across the workgroup, before any barrier
  const long bs = L.bin_stride, ks = L.cg_k_stride;
  const int nk = g.nk, nfrchg = g.nfrchg, nfrhgh = g.nfrhgh;
  float* const ue = lds;
  float* const sa = ue + (ZR + nfrhgh) * W;
  float* const conx_row = sa + (ZR + nfrchg) * RS;
  float* const caf_row = conx_row + nk;

  const int tid = threadIdx.x;
  const int c = tid % NTH;
  const int rr = tid / NTH;
  const float* const ap = a + p * L.pt_stride;
  const float* const cgp = cg + p * L.cg_pt_stride;

  // 1. Proportionality constant.
  const float x = fmaxo(g.kdcon * kdmean[p], g.kdmn);
  const float x2 = fmaxo(-1.e15f, g.snls3 * x);
  const float cons = g.snlc1 * (1.0f + g.snls1 / x * (1.0f - g.snls2 * x) * libm::expf(x2));

  // Zero rows; CONX and CONS*AF11 per row.
  for (int i = tid; i < ZR * W; i += BS) ue[i] = 0.0f;
  for (int i = tid; i < ZR * RS; i += BS) sa[i] = 0.0f;
  for (int i = tid; i < nk; i += BS) conx_row[i] = g.tpiinv / sig[i] * cgp[i * ks];
  for (int i = tid; i < nfrchg; i += BS) caf_row[i] = cons * af11_row[i];
  __syncthreads();

  // 2. UE = A / CONX on rows 1..NK; the row-NK threads continue the tail up their column.
  for (int r0 = rr; r0 < nk; r0 += R) {
    const float v = ap[(r0 * NTH + c) * bs] / conx_row[r0];
    float* const u = ue + (r0 + ZR) * W + H + c;
    u[0] = v;
    if (c < H) u[NTH] = v;
    if (c >= NTH - H) u[-NTH] = v;
    if (r0 == nk - 1) {
      float t = v;
      for (int r1 = nk; r1 < nfrhgh; ++r1) {
        t = t * g.fachfe;
        float* const ut = ue + (r1 + ZR) * W + H + c;
        ut[0] = t;
        if (c < H) ut[NTH] = t;
        if (c >= NTH - H) ut[-NTH] = t;
      }
    }
  }
  __syncthreads();

  // 3. Interactions over the extended spectrum.
  const float dal1 = g.dal1, dal2 = g.dal2, dal3 = g.dal3;
  const float w0 = g.awg[0], w1 = g.awg[1], w2 = g.awg[2], w3 = g.awg[3];
  const float w4 = g.awg[4], w5 = g.awg[5], w6 = g.awg[6], w7 = g.awg[7];
  float dcs[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int r0 = rr + it * R;
    if (r0 < nfrchg) {
      const float* const u = ue + (r0 + ZR) * W + H + c;
      const float e00 = u[0];
      const float ep1 = w0 * u[ST::o3(0)] + w1 * u[ST::o3(1)] + w2 * u[ST::o3(2)] + w3 * u[ST::o3(3)];
      const float em1 = w4 * u[ST::o3(4)] + w5 * u[ST::o3(5)] + w6 * u[ST::o3(6)] + w7 * u[ST::o3(7)];
      const float ep2 = w0 * u[ST::o3(8)] + w1 * u[ST::o3(9)] + w2 * u[ST::o3(10)] + w3 * u[ST::o3(11)];
      const float em2 = w4 * u[ST::o3(12)] + w5 * u[ST::o3(13)] + w6 * u[ST::o3(14)] + w7 * u[ST::o3(15)];
      const float caf = caf_row[r0];
      const float factor = caf * e00;
      const float sa1a = e00 * (ep1 * dal1 + em1 * dal2);
      const float sa1b = sa1a - ep1 * em1 * dal3;
      const float sa2a = e00 * (ep2 * dal1 + em2 * dal2);
      const float sa2b = sa2a - ep2 * em2 * dal3;
      float v[6];
      v[0] = factor * sa1b;                       // SA1
      v[1] = factor * sa2b;                       // SA2
      const float da1c = caf * (sa1a + sa1b);
      v[2] = factor * (dal1 * e00 - dal3 * em1);  // DA1P
      v[3] = factor * (dal2 * e00 - dal3 * ep1);  // DA1M
      const float da2c = caf * (sa2a + sa2b);
      v[4] = factor * (dal1 * e00 - dal3 * em2);  // DA2P
      v[5] = factor * (dal2 * e00 - dal3 * ep2);  // DA2M
      dcs[it] = da1c + da2c;
      float* const q = sa + (r0 + ZR) * RS + H + c;
#pragma unroll
      for (int k = 0; k < 6; ++k) q[k * W] = v[k];
      if (c < H) {
#pragma unroll
        for (int k = 0; k < 6; ++k) q[k * W + NTH] = v[k];
      }
      if (c >= NTH - H) {
#pragma unroll
        for (int k = 0; k < 6; ++k) q[k * W - NTH] = v[k];
      }
    }
  }
  __syncthreads();

  // 4. Source term and diagonal (same rows as section 3, so dcs[it] is this bin's).
  const float v0 = g.swg[0], v1 = g.swg[1], v2 = g.swg[2], v3 = g.swg[3];
  const float v4 = g.swg[4], v5 = g.swg[5], v6 = g.swg[6], v7 = g.swg[7];
  float* const sp = s + p * L.pt_stride;
  float* const dp = d + p * L.pt_stride;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int r0 = rr + it * R;
    if (r0 < nk) {
      const float* const q = sa + (r0 + ZR) * RS + H + c;
      constexpr int A1 = 0, A2 = W, P1 = 2 * W, M1 = 3 * W, P2 = 4 * W, M2 = 5 * W;
      const float sv = conx_row[r0] *
                       (-2.0f * (q[A1] + q[A2]) +
                        w0 * (q[ST::o4(0) + A1] + q[ST::o4(8) + A2]) +
                        w1 * (q[ST::o4(1) + A1] + q[ST::o4(9) + A2]) +
                        w2 * (q[ST::o4(2) + A1] + q[ST::o4(10) + A2]) +
                        w3 * (q[ST::o4(3) + A1] + q[ST::o4(11) + A2]) +
                        w4 * (q[ST::o4(4) + A1] + q[ST::o4(12) + A2]) +
                        w5 * (q[ST::o4(5) + A1] + q[ST::o4(13) + A2]) +
                        w6 * (q[ST::o4(6) + A1] + q[ST::o4(14) + A2]) +
                        w7 * (q[ST::o4(7) + A1] + q[ST::o4(15) + A2]));
      const float dv = -2.0f * dcs[it] +
                       v0 * (q[ST::o4(0) + P1] + q[ST::o4(8) + P2]) +
                       v1 * (q[ST::o4(1) + P1] + q[ST::o4(9) + P2]) +
                       v2 * (q[ST::o4(2) + P1] + q[ST::o4(10) + P2]) +
                       v3 * (q[ST::o4(3) + P1] + q[ST::o4(11) + P2]) +
                       v4 * (q[ST::o4(4) + M1] + q[ST::o4(12) + M2]) +
                       v5 * (q[ST::o4(5) + M1] + q[ST::o4(13) + M2]) +
                       v6 * (q[ST::o4(6) + M1] + q[ST::o4(14) + M2]) +
                       v7 * (q[ST::o4(7) + M1] + q[ST::o4(15) + M2]);
      sp[(r0 * NTH + c) * bs] = sv;
      dp[(r0 * NTH + c) * bs] = dv;
    }
  }
}

template <class ST>
bool matches(const Geom& g) {
  if (g.nth != ST::NTH || g.h != ST::H || g.zr != ST::ZR || g.w != ST::W) return false;
  for (int m = 1; m <= 8; ++m)
    if (g.drow[m - 1] != ST::ifo(m) || g.sh[m - 1] != ST::ito(m)) return false;
  return true;
}

template <class ST>
size_t fixed_lds(const Geom& g) {
  return sizeof(float) * (static_cast<size_t>(ST::ZR + g.nfrhgh) * ST::W +
                          static_cast<size_t>(ST::ZR + g.nfrchg) * ST::RS + g.nk + g.nfrchg);
}

template <class ST, int IT>
hipError_t go(const Geom& g, const float* af11_row, const float* sig, int npts, const float* a,
              const float* cg, const float* kdmean, float* s, float* d, hipStream_t stream,
              const Layout& L) {
  const size_t lds = fixed_lds<ST>(g);
  if (lds > 64u * 1024u) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&dia_fixed<ST, IT>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((dia_fixed<ST, IT>), dim3(grid_size(L, npts)), dim3(ST::BS), lds, stream, g,
                     af11_row, sig, a, cg, kdmean, s, d, npts, L);
  return hipGetLastError();
}

template <class ST>
bool try_stencil(const Geom& g, const float* af11_row, const float* sig, int npts, const float* a,
                 const float* cg, const float* kdmean, float* s, float* d, hipStream_t stream,
                 hipError_t* err, const Layout& L) {
  if (!matches<ST>(g) || fixed_lds<ST>(g) > 160u * 1024u) return false;
  const int it = (g.nfrchg + ST::R - 1) / ST::R;
  switch (it) {
    case 1: *err = go<ST, 1>(g, af11_row, sig, npts, a, cg, kdmean, s, d, stream, L); return true;
    case 2: *err = go<ST, 2>(g, af11_row, sig, npts, a, cg, kdmean, s, d, stream, L); return true;
    case 3: *err = go<ST, 3>(g, af11_row, sig, npts, a, cg, kdmean, s, d, stream, L); return true;
    case 4: *err = go<ST, 4>(g, af11_row, sig, npts, a, cg, kdmean, s, d, stream, L); return true;
    default: return false;
  }
}

// LAMBDA = 0.25 and XFR = 1.1 give IFRP = 2, IFRM = -3.  The quadruplet angles are
// |DELTH4| = 11.48 and DELTH3 = 33.56 degrees, so NTH = 24 (15 degrees) -> ITHP = 0,
// ITHM = 2 and NTH = 36 (10 degrees) -> ITHP = 1, ITHM = 3.
using St24 = Stencil<24, 0, 2, 2, -3, 18>;
using St36 = Stencil<36, 1, 3, 2, -3, 12>;

}  // namespace

bool launch_specialized(const Geom& g, const float* af11_row, const float* sig, int npts,
                        const float* a, const float* cg, const float* kdmean, float* s, float* d,
                        hipStream_t stream, hipError_t* err, const Layout& L) {
  return try_stencil<St24>(g, af11_row, sig, npts, a, cg, kdmean, s, d, stream, err, L) ||
         try_stencil<St36>(g, af11_row, sig, npts, a, cg, kdmean, s, d, stream, err, L);
}

bool has_specialized(const Geom& g) {
  const auto fits = [&](auto st) {
    using ST = decltype(st);
    return matches<ST>(g) && fixed_lds<ST>(g) <= 160u * 1024u && (g.nfrchg + ST::R - 1) / ST::R <= 4;
  };
  return fits(St24{}) || fits(St36{});
}

}  // namespace snl1
}  // namespace ww3gpu

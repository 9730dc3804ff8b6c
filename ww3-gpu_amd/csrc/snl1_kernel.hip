// ww3-gpu_amd/csrc/snl1_kernel.hip
// W3SNL1 (the DIA) for gfx950: one workgroup per sea point, the point's working arrays
// in LDS.
//
// Reference algorithm: kokkos/src/ww_kokkos/snl1_team.hpp:352-497 (WW3 w3snl1md.F90
// sections 1-4).  Every float32 expression keeps the reference's operation order and is
// compiled with -ffp-contract=off; expf is glibc's algorithm (devlibm.hpp).  No
// reductions: each output is written by one thread, so the bits do not depend on the
// launch shape.
//
// LDS layout per point (snl1.hpp): seven arrays of rows of w = NTH + 2H floats.  Row
// slot q holds Fortran frequency row q + 1 - ZR (rows <= 0 are zero rows); columns
// H..H+NTH-1 hold directions 0..NTH-1, and the H columns either side hold wrapped
// copies, so every quadruplet partner of a bin is at a constant offset (Geom::o3, o4)
// from the bin's slot.  UE has rows up to NFRHGH, the six section-3 arrays up to NFRCHG.
// DA1C+DA2C is only read at the bin's own slot, by the thread that wrote it, so it stays
// in registers.
//
// Traffic per point: A (NSPEC) + CG (NK) + KDMEAN in, S and D (NSPEC each) out; SIG and
// AF11 are per-grid and stay in cache.
#include "devlibm.hpp"
#include "snl1_device.hpp"

namespace ww3gpu {
namespace snl1 {
namespace {

constexpr int kBlock = 256;

// Fortran MAX(a, b) for ordered operands.
__device__ inline float fmaxo(float a, float b) { return a > b ? a : b; }

template <int MAXIT>
__global__ __launch_bounds__(kBlock) void dia_kernel(const Geom g, const float* __restrict__ af11_row,
                                                     const float* __restrict__ sig,
                                                     const float* __restrict__ a,
                                                     const float* __restrict__ cg,
                                                     const float* __restrict__ kdmean,
                                                     float* __restrict__ s, float* __restrict__ d,
                                                     const int npts, const Layout L) {
  extern __shared__ float lds[];
  const int p = layout_point(L, npts);
  if (p < 0) return;  // uniform across the workgroup, before any barrier
  const int tid = threadIdx.x;
  const long bs = L.bin_stride, ks = L.cg_k_stride;
  const int nth = g.nth, w = g.w, h = g.h, zr = g.zr;
  const int sa_len = g.sa_rows * w;
  float* const ue = lds;
  float* const sa1 = ue + g.ue_rows * w;
  float* const sa2 = sa1 + sa_len;
  float* const da1p = sa2 + sa_len;
  float* const da1m = da1p + sa_len;
  float* const da2p = da1m + sa_len;
  float* const da2m = da2p + sa_len;
  const float* const ap = a + p * L.pt_stride;
  const float* const cgp = cg + p * L.cg_pt_stride;

  // Row/column of this thread's first bin and the per-iteration step.
  const int q = kBlock / nth, rem = kBlock - q * nth;
  const int r_first = tid / nth, c_first = tid - r_first * nth;

  // 1. Proportionality constant (every thread, no broadcast needed).
  const float x = fmaxo(g.kdcon * kdmean[p], g.kdmn);
  const float x2 = fmaxo(-1.e15f, g.snls3 * x);
  const float cons = g.snlc1 * (1.0f + g.snls1 / x * (1.0f - g.snls2 * x) * libm::expf(x2));

  // 2. Zero rows, then UE = A / CONX on rows 1..NK with the wrapped halo copies.
  for (int i = tid; i < zr * w; i += kBlock) {
    ue[i] = 0.0f;
    sa1[i] = 0.0f;
    sa2[i] = 0.0f;
    da1p[i] = 0.0f;
    da1m[i] = 0.0f;
    da2p[i] = 0.0f;
    da2m[i] = 0.0f;
  }
  {
    int r0 = r_first, c = c_first;
    for (int isp = tid; isp < g.nspec; isp += kBlock) {
      const float conx = g.tpiinv / sig[r0] * cgp[r0 * ks];
      const float v = ap[isp * bs] / conx;
      const int slot = (r0 + zr) * w + h + c;
      ue[slot] = v;
      if (c < h) ue[slot + nth] = v;
      if (c >= nth - h) ue[slot - nth] = v;
      c += rem;
      r0 += q;
      if (c >= nth) {
        c -= nth;
        ++r0;
      }
    }
  }
  __syncthreads();
  // Parametric tail: row r = row r-1 times FACHFE, sequential in r for each column
  // (halo columns included: they repeat the same products).
  for (int j = tid; j < w; j += kBlock) {
    float v = ue[(g.nk - 1 + zr) * w + j];
    for (int r0 = g.nk; r0 < g.nfrhgh; ++r0) {
      v = v * g.fachfe;
      ue[(r0 + zr) * w + j] = v;
    }
  }
  __syncthreads();

  // 3. Interactions over the extended spectrum.
  const float dal1 = g.dal1, dal2 = g.dal2, dal3 = g.dal3;
  float dcs[MAXIT];
  {
    int r0 = r_first, c = c_first;
#pragma unroll
    for (int it = 0; it < MAXIT; ++it) {
      const int isp = tid + it * kBlock;
      if (isp < g.nspecx) {
        const int b = (r0 + zr) * w + h + c;
        const float e00 = ue[b];
        const float ep1 = g.awg[0] * ue[b + g.o3[0]] + g.awg[1] * ue[b + g.o3[1]] +
                          g.awg[2] * ue[b + g.o3[2]] + g.awg[3] * ue[b + g.o3[3]];
        const float em1 = g.awg[4] * ue[b + g.o3[4]] + g.awg[5] * ue[b + g.o3[5]] +
                          g.awg[6] * ue[b + g.o3[6]] + g.awg[7] * ue[b + g.o3[7]];
        const float ep2 = g.awg[0] * ue[b + g.o3[8]] + g.awg[1] * ue[b + g.o3[9]] +
                          g.awg[2] * ue[b + g.o3[10]] + g.awg[3] * ue[b + g.o3[11]];
        const float em2 = g.awg[4] * ue[b + g.o3[12]] + g.awg[5] * ue[b + g.o3[13]] +
                          g.awg[6] * ue[b + g.o3[14]] + g.awg[7] * ue[b + g.o3[15]];
        const float caf = cons * af11_row[r0];
        const float factor = caf * e00;
        const float sa1a = e00 * (ep1 * dal1 + em1 * dal2);
        const float sa1b = sa1a - ep1 * em1 * dal3;
        const float sa2a = e00 * (ep2 * dal1 + em2 * dal2);
        const float sa2b = sa2a - ep2 * em2 * dal3;
        const float v_sa1 = factor * sa1b;
        const float v_sa2 = factor * sa2b;
        const float v_da1c = caf * (sa1a + sa1b);
        const float v_da1p = factor * (dal1 * e00 - dal3 * em1);
        const float v_da1m = factor * (dal2 * e00 - dal3 * ep1);
        const float v_da2c = caf * (sa2a + sa2b);
        const float v_da2p = factor * (dal1 * e00 - dal3 * em2);
        const float v_da2m = factor * (dal2 * e00 - dal3 * ep2);
        dcs[it] = v_da1c + v_da2c;
        sa1[b] = v_sa1;
        sa2[b] = v_sa2;
        da1p[b] = v_da1p;
        da1m[b] = v_da1m;
        da2p[b] = v_da2p;
        da2m[b] = v_da2m;
        if (c < h) {
          const int bb = b + nth;
          sa1[bb] = v_sa1;
          sa2[bb] = v_sa2;
          da1p[bb] = v_da1p;
          da1m[bb] = v_da1m;
          da2p[bb] = v_da2p;
          da2m[bb] = v_da2m;
        }
        if (c >= nth - h) {
          const int bb = b - nth;
          sa1[bb] = v_sa1;
          sa2[bb] = v_sa2;
          da1p[bb] = v_da1p;
          da1m[bb] = v_da1m;
          da2p[bb] = v_da2p;
          da2m[bb] = v_da2m;
        }
      }
      c += rem;
      r0 += q;
      if (c >= nth) {
        c -= nth;
        ++r0;
      }
    }
  }
  __syncthreads();

  // 4. Source term and diagonal.
  {
    float* const sp = s + p * L.pt_stride;
    float* const dp = d + p * L.pt_stride;
    int r0 = r_first, c = c_first;
#pragma unroll
    for (int it = 0; it < MAXIT; ++it) {
      const int isp = tid + it * kBlock;
      if (isp < g.nspec) {
        const int b = (r0 + zr) * w + h + c;
        const float conx = g.tpiinv / sig[r0] * cgp[r0 * ks];
        sp[isp * bs] = conx * (-2.0f * (sa1[b] + sa2[b]) +
                          g.awg[0] * (sa1[b + g.o4[0]] + sa2[b + g.o4[8]]) +
                          g.awg[1] * (sa1[b + g.o4[1]] + sa2[b + g.o4[9]]) +
                          g.awg[2] * (sa1[b + g.o4[2]] + sa2[b + g.o4[10]]) +
                          g.awg[3] * (sa1[b + g.o4[3]] + sa2[b + g.o4[11]]) +
                          g.awg[4] * (sa1[b + g.o4[4]] + sa2[b + g.o4[12]]) +
                          g.awg[5] * (sa1[b + g.o4[5]] + sa2[b + g.o4[13]]) +
                          g.awg[6] * (sa1[b + g.o4[6]] + sa2[b + g.o4[14]]) +
                          g.awg[7] * (sa1[b + g.o4[7]] + sa2[b + g.o4[15]]));
        dp[isp * bs] = -2.0f * dcs[it] +
                  g.swg[0] * (da1p[b + g.o4[0]] + da2p[b + g.o4[8]]) +
                  g.swg[1] * (da1p[b + g.o4[1]] + da2p[b + g.o4[9]]) +
                  g.swg[2] * (da1p[b + g.o4[2]] + da2p[b + g.o4[10]]) +
                  g.swg[3] * (da1p[b + g.o4[3]] + da2p[b + g.o4[11]]) +
                  g.swg[4] * (da1m[b + g.o4[4]] + da2m[b + g.o4[12]]) +
                  g.swg[5] * (da1m[b + g.o4[5]] + da2m[b + g.o4[13]]) +
                  g.swg[6] * (da1m[b + g.o4[6]] + da2m[b + g.o4[14]]) +
                  g.swg[7] * (da1m[b + g.o4[7]] + da2m[b + g.o4[15]]);
      }
      c += rem;
      r0 += q;
      if (c >= nth) {
        c -= nth;
        ++r0;
      }
    }
  }
}

template <int MAXIT>
hipError_t launch_n(const Geom& g, const float* af11_row, const float* sig, int npts,
                    const float* a, const float* cg, const float* kdmean, float* s, float* d,
                    hipStream_t stream, const Layout& L) {
  const size_t lds = lds_bytes(g);
  if (lds > 64u * 1024u) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&dia_kernel<MAXIT>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(dia_kernel<MAXIT>, dim3(grid_size(L, npts)), dim3(kBlock), lds, stream, g,
                     af11_row, sig, a, cg, kdmean, s, d, npts, L);
  return hipGetLastError();
}

}  // namespace

hipError_t launch(const Geom& g, const float* af11_row, const float* sig, int npts,
                  const float* a, const float* cg, const float* kdmean, float* s, float* d,
                  hipStream_t stream, int variant, const Layout& L) {
  if (npts <= 0) return hipSuccess;
  hipError_t err = hipSuccess;
  if (variant == 0 &&
      launch_specialized(g, af11_row, sig, npts, a, cg, kdmean, s, d, stream, &err, L))
    return err;
  const int it = (g.nspecx + kBlock - 1) / kBlock;
  if (it <= 2) return launch_n<2>(g, af11_row, sig, npts, a, cg, kdmean, s, d, stream, L);
  if (it <= 4) return launch_n<4>(g, af11_row, sig, npts, a, cg, kdmean, s, d, stream, L);
  if (it <= 6) return launch_n<6>(g, af11_row, sig, npts, a, cg, kdmean, s, d, stream, L);
  if (it <= 8) return launch_n<8>(g, af11_row, sig, npts, a, cg, kdmean, s, d, stream, L);
  if (it <= 12) return launch_n<12>(g, af11_row, sig, npts, a, cg, kdmean, s, d, stream, L);
  if (it <= 16) return launch_n<16>(g, af11_row, sig, npts, a, cg, kdmean, s, d, stream, L);
  return hipErrorInvalidValue;
}

}  // namespace snl1
}  // namespace ww3gpu

// ww3-gpu_amd/csrc/prop3_kernels.hip
// gfx950 kernels of the PR3/UQ propagation: W3XYP3 + W3QCK3 over all components of a call,
// W3KTP3 + W3QCK1/W3QCK2 over the sea points of an intra-spectral part.
//
//   W3XYP3  kokkos/src/ww_kokkos/prop_xyp3.hpp:156-584
//   W3QCK3  kokkos/src/ww_kokkos/prop_qck3.hpp:58-214
//   W3KTP3  kokkos/src/ww_kokkos/prop_ktp3.hpp:143-353
//   W3QCK1  kokkos/src/ww_kokkos/prop_qck1.hpp:54-177
//   W3QCK2  kokkos/src/ww_kokkos/prop_qck2.hpp:60-193
//
// Every float operation is the reference's, in its order; MAX/MIN keep the operand order of
// the reference (mx(a,b) = a > b ? a : b, mn(a,b) = a < b ? a : b).
#include <hip/hip_runtime.h>

#include "prop3.hpp"

namespace ww3gpu {
namespace prop {
namespace {

__device__ __forceinline__ float mx(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float mn(float a, float b) { return a < b ? a : b; }
/// INT(): truncation, INT_MIN for NaN or out of range (cvttss2si).
__device__ __forceinline__ int fint(float x) {
  return (x > -2147483904.f && x < 2147483648.f) ? static_cast<int>(x) : static_cast<int>(0x80000000u);
}
__device__ __forceinline__ int isign(int a, int b) { return ((a ^ b) < 0) ? -a : a; }

constexpr int kTile = 64;

// ------------------------------------------------------------------------------------
// W3GATH and the velocity scaling of prop_xyp3.hpp:1080-1084: FIELD = 0, then at each sea
// cell VA(ISP, ISEA) / CG * CLATS.  A 64-cell x 64-component tile goes through LDS so that
// VA's rows (components contiguous) and the planes (cells contiguous) are both read and
// written in whole cache lines.
// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void gather_kernel(XyGrid g, XyState s, const float* va) {
  __shared__ float tile[kTile][kTile + 1];
  const int tx = threadIdx.x, ty = threadIdx.y;  // 64 x 4
  const int k0 = blockIdx.x * kTile, c0 = blockIdx.y * kTile;
  for (int r = ty; r < kTile; r += 4) {
    const int k = k0 + r, c = c0 + tx;
    const int t = k + 1 - g.ny;
    float v = 0.f;
    if (c < s.nlist && t >= 1 && t <= g.nxy) {
      const int isea = g.cellsea[t];
      if (isea > 0) v = va[static_cast<size_t>(isea) * g.nspec + s.comps[c].isp - 1];
    }
    tile[r][tx] = v;
  }
  __syncthreads();
  for (int r = ty; r < kTile; r += 4) {
    const int c = c0 + r, k = k0 + tx;
    if (c >= s.nlist || k >= g.nq) continue;
    float v = tile[tx][r];
    const int t = k + 1 - g.ny;
    if (t >= 1 && t <= g.nxy) {
      const int isea = g.cellsea[t];
      if (isea > 0) {
        const float cgi = g.cg[isea * g.cgld + s.comps[c].ik];
        v = (v / cgi) * g.clats[isea];
      }
    }
    s.q[0][static_cast<size_t>(c) * g.nq + k] = v;
  }
}

/// MAPSTX (prop_xyp3.hpp:949-957): MAPSTA with two ghost columns each side, the far side's
/// on a closed grid, 0 otherwise.
__device__ __forceinline__ int mapstx(const XyGrid& g, int j) {
  if (j >= 1 && j <= g.nxy) return g.mapsta[j - 1];
  if (!g.global) return 0;
  return j <= 0 ? g.mapsta[j + g.nxy - 1] : g.mapsta[j - g.nxy - 1];
}

/// Index of a store's offset from its source in {0, 1, -1, NY, -NY, NY+1, NY-1, -NY+1, -NY-1}
/// (distinct for NY >= 3).
__device__ __forceinline__ int dir_index(int off, int ny) {
  if (off == 0) return 0;
  if (off == 1) return 1;
  if (off == -1) return 2;
  if (off == ny) return 3;
  if (off == -ny) return 4;
  if (off == ny + 1) return 5;
  if (off == ny - 1) return 6;
  if (off == -ny + 1) return 7;
  return 8;
}

// ------------------------------------------------------------------------------------
// The averaging stencil of each sea point for each component (prop_xyp3.hpp:983-1059) with
// the CG factors of :1160-1163, and where each of its nine stores lands.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ int corner_bits(float dx, float dy, float& rda, float& rdb) {
  int b = (dx < 0.f ? 1 : 0) | (dy < 0.f ? 2 : 0);
  if (__builtin_fabsf(dx) > __builtin_fabsf(dy)) {
    b |= 4;
    rda = __builtin_fabsf(dy / dx);
    rdb = __builtin_fabsf(dx);
  } else {
    rda = (__builtin_fabsf(dy) > 1.E-5f) ? __builtin_fabsf(dx / dy) : 1.f;
    rdb = __builtin_fabsf(dy);
  }
  return b;
}

__global__ __launch_bounds__(256) void records_kernel(XyGrid g, XyState s) {
  const int ixy = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  const int c = blockIdx.y;
  if (ixy > g.nxy) return;
  const int ny = g.ny;
  const size_t po = static_cast<size_t>(c) * g.nf;
  const int isea = g.cellsea[ixy];
  if (isea == 0) {  // VLCFLX/VLCFLY are 0 off the sea points (:944-945)
    s.vlcx[po + ixy - 1] = 0.f;
    s.vlcy[po + ixy - 1] = 0.f;
    if (ixy <= ny) {
      s.vlcx[po + g.nxy + ixy - 1] = 0.f;
      s.vlcy[po + g.nxy + ixy - 1] = 0.f;
    }
    return;
  }
  const Comp cp = s.comps[c];
  const int gi = ixy - 1;
  const float clat = g.clats[isea];
  const float dpdx = g.dpdx[gi], dpdy = g.dpdy[gi], dqdx = g.dqdx[gi], dqdy = g.dqdy[gi];
  float tmpx = cp.fthx / clat, tmpy = cp.fthy;
  const float dxcgn = dpdx * tmpx + dpdy * tmpy;
  const float dycgn = dqdx * tmpx + dqdy * tmpy;
  tmpx = cp.fcgx / clat;
  tmpy = cp.fcgy;
  const float dxcgs = dpdx * tmpx + dpdy * tmpy;
  const float dycgs = dqdx * tmpx + dqdy * tmpy;
  float rdi1, rdi2, rdi3, rdi4;
  const int ba = corner_bits(dxcgn + dxcgs, dycgn + dycgs, rdi1, rdi2);
  const int bb = corner_bits(dxcgn - dxcgs, dycgn - dycgs, rdi3, rdi4);
  const float cgi = g.cg[isea * g.cgld + cp.ik];
  const size_t o = static_cast<size_t>(c) * (g.nxy + 1) + ixy;
  s.rec[o] = Rec{rdi1, mn(rdi2 * cgi, 1.f), rdi3, mn(cgi * rdi4, 1.f)};
  const int ixa = (ba & 1) ? -ny : ny, iya = (ba & 2) ? -1 : 1;
  const int ixb = (bb & 1) ? -ny : ny, iyb = (bb & 2) ? -1 : 1;
  const int i1 = ixa + iya, i2 = (ba & 4) ? ixa : iya;
  const int i3 = ixb + iyb, i4 = (bb & 4) ? ixb : iyb;
  int off[9];
  off[0] = 0;
  const bool coast = g.pos[isea] > g.ncent;
  if (!coast) {  // :1166-1180
    off[1] = i1;
    off[2] = i2;
    off[3] = i3;
    off[4] = i4;
    off[5] = -i1;
    off[6] = -i2;
    off[7] = -i3;
    off[8] = -i4;
  } else {  // :1202-1272
    const auto diag = [&](int i) {
      const int ixc = isign(ny, i), iyc = i - ixc;
      return (mapstx(g, ixy + i) >= 1 && !(mapstx(g, ixy + ixc) <= 0 && mapstx(g, ixy + iyc) <= 0)) ? i : 0;
    };
    const auto axial = [&](int i) { return mapstx(g, ixy + i) >= 1 ? i : 0; };
    off[1] = diag(i1);
    off[2] = diag(-i1);
    off[3] = axial(i2);
    off[4] = axial(-i2);
    off[5] = diag(i3);
    off[6] = diag(-i3);
    off[7] = axial(i4);
    off[8] = axial(-i4);
  }
  uint64_t w = coast ? (1ULL << kNearCoastBit) : 0ULL;
  for (int j = 1; j < 9; ++j) {
    const int d = dir_index(off[j], ny);
    const int k = coast ? (j + 1) / 2 : (j - 1) % 4 + 1;
    if (d == 0) {  // sent back to the centre (near-coast only)
      w |= 1ULL << (kCentreShift - 1 + j);
      continue;
    }
    const int sh = 6 * (d - 1);
    const int cnt = static_cast<int>((w >> sh) & 3ULL);
    w |= static_cast<uint64_t>(k - 1) << (sh + 2 + 2 * cnt);
    w = (w & ~(3ULL << sh)) | (static_cast<uint64_t>(cnt + 1) << sh);
  }
  s.tw[o] = w;
  // VLCFLX, VLCFLY at this cell (prop_xyp3.hpp:1080-1137)
  float cxt = (cp.ecosi * cgi) / clat;
  float cyt = cp.esini * cgi;
  if (g.flcur) {
    cxt = cxt + (g.cx[isea] / g.clats[isea]);
    cyt = cyt + g.cy[isea];
  }
  const float vx = cp.dtrad * (cxt * dpdx + cyt * dpdy);
  const float vy = cp.dtrad * (cxt * dqdx + cyt * dqdy);
  s.vlcx[po + ixy - 1] = vx;
  s.vlcy[po + ixy - 1] = vy;
  if (ixy <= ny) {  // the extension column: W3QCK3's closure copy of column 1 (:1105), else 0
    s.vlcx[po + g.nxy + ixy - 1] = g.global ? vx : 0.f;
    s.vlcy[po + g.nxy + ixy - 1] = 0.f;
  }
}

// ------------------------------------------------------------------------------------
// Sub-step: averaging (prop_xyp3.hpp:1147-1291) as an ordered gather, then x GSQRT (:1301).
// ------------------------------------------------------------------------------------

/// The value the reference's two scatter loops leave in VQ(t), starting from the 0. of
/// :1150: the stores of the neighbours listed for t that land on t, each source's stores in
/// its loop's order (prop_xyp3.hpp:1164-1180 central, :1199-1272 near-coast).
__device__ float avg_fold(const XyGrid& g, const XyState& s, int c, const float* aq, int t) {
  const int ny = g.ny;
  uint64_t ord = g.order[t + ny - 1];
  float v = 0.f;
  const size_t rb = static_cast<size_t>(c) * (g.nxy + 1);
  for (;;) {
    const int d = static_cast<int>(ord & 0xFu);
    if (d == 0xF) break;
    ord >>= 4;
    const int off = d == 0 ? 0 : (d == 1 ? 1 : d == 2 ? -1 : d == 3 ? ny : d == 4 ? -ny : d == 5 ? ny + 1 : d == 6 ? ny - 1 : d == 7 ? -ny + 1 : -ny - 1);
    const int src = t - off;
    const float a = aq[src + ny - 1];
    if (g.maptrn[src - 1] != 0) {  // :1157-1158, :1192-1193 (listed at its own cell only)
      v = a;
      continue;
    }
    const uint64_t w = s.tw[rb + src];
    const Rec r = s.rec[rb + src];
    const auto term = [&](int k) {
      if (k == 0) return (((3.f - r.rd2) - r.rd4) * a) / 3.f;
      if (k == 1) return ((r.rd2 * a) * r.rd1) / 6.f;
      if (k == 2) return (((1.f - r.rd1) * a) * r.rd2) / 6.f;
      if (k == 3) return ((r.rd4 * a) * r.rd3) / 6.f;
      return (((1.f - r.rd3) * a) * r.rd4) / 6.f;
    };
    if (d == 0) {  // the centre store, then the near-coast stores sent back, in store order
      v = v + term(0);
      unsigned m = static_cast<unsigned>((w >> kCentreShift) & 0xFFULL);
      while (m != 0u) {
        const int j = __builtin_ctz(m) + 1;
        m &= m - 1u;
        v = v + term((j + 1) / 2);
      }
    } else {
      const unsigned f = static_cast<unsigned>((w >> (6 * (d - 1))) & 0x3FULL);
      if ((f & 3u) != 0u) v = v + term(static_cast<int>((f >> 2) & 3u) + 1);
      if ((f & 3u) == 2u) v = v + term(static_cast<int>((f >> 4) & 3u) + 1);
    }
  }
  return v;
}

__global__ __launch_bounds__(256) void avg_kernel(XyGrid g, XyState s, int itloc) {
  const int c = blockIdx.y;
  if (itloc > s.comps[c].ntloc) return;
  const int k = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= g.nq) return;
  const float* aq = s.q[(itloc - 1) & 1] + static_cast<size_t>(c) * g.nq;
  float* vq = s.q[itloc & 1] + static_cast<size_t>(c) * g.nq;
  const int t = k + 1 - g.ny;
  float v = avg_fold(g, s, c, aq, t);
  if (t >= 1 && t <= g.nxy) {
    if (g.mapsta[t - 1] == 2) v = aq[k];  // :1280-1282
    if (g.global) {                       // :1286-1291
      if (t <= g.ny) v = v + avg_fold(g, s, c, aq, g.nxy + t);
      if (t > g.nxy - g.ny) v = v + avg_fold(g, s, c, aq, t - g.nxy);
    }
    if (g.cellsea[t] > 0) v = v * g.gsqrt[t - 1];  // :1301-1306
  }
  vq[k] = v;
}

// The averaging through LDS.  A workgroup takes kAvgT consecutive plane entries (targets) of
// one component; every source a target can receive from lies within NY+1 cells, so the tile's
// sources are the kAvgT + 2 NY + 2 cells around it.  Each source is read once, coalesced, and
// its five store values (((P * A) * Q) / D, the factor sets of avg_fold) are formed once into
// LDS with its store word; the targets then fold from LDS in the order avg_fold uses, adding
// the same values.  A transparent source (MAPTRN) keeps its A (kTrnBit).  The global closure's
// far-side folds (first and last NY cells) read global memory through avg_fold.
constexpr int kAvgT = 1024;
constexpr int kAvgBS = 256;
constexpr uint64_t kTrnBit = 1ULL << 63;

__device__ __forceinline__ int dir_offset(int d, int ny) {
  return d == 0 ? 0 : (d == 1 ? 1 : d == 2 ? -1 : d == 3 ? ny : d == 4 ? -ny : d == 5 ? ny + 1 : d == 6 ? ny - 1 : d == 7 ? -ny + 1 : -ny - 1);
}

__global__ __launch_bounds__(kAvgBS) void avg_tiled_kernel(XyGrid g, XyState s, int itloc) {
  extern __shared__ uint64_t avg_lds[];
  const int c = blockIdx.y;
  if (itloc > s.comps[c].ntloc) return;
  const int ny = g.ny, nl = kAvgT + 2 * ny + 2;  // nl is even: the store words stay aligned
  uint64_t* tws = avg_lds;                        // [nl]
  float* tm = reinterpret_cast<float*>(avg_lds + nl);  // [5][nl]
  const int k0 = static_cast<int>(blockIdx.x) * kAvgT;
  const int s0 = k0 + 1 - ny - ny - 1;  // the first source cell: first target - (NY + 1)
  const float* aq = s.q[(itloc - 1) & 1] + static_cast<size_t>(c) * g.nq;
  float* vq = s.q[itloc & 1] + static_cast<size_t>(c) * g.nq;
  const size_t rb = static_cast<size_t>(c) * (g.nxy + 1);
#pragma unroll 8
  for (int i = static_cast<int>(threadIdx.x); i < nl; i += kAvgBS) {
    const int src = s0 + i;
    uint64_t w = 0;
    float e0 = 0.f, e1 = 0.f, e2 = 0.f, e3 = 0.f, e4 = 0.f;
    if (src >= 1 && src <= g.nxy) {
      const float a = aq[src + ny - 1];
      if (g.maptrn[src - 1] != 0) {
        w = kTrnBit;
        e0 = a;
      } else {
        w = s.tw[rb + src];
        const Rec r = s.rec[rb + src];
        e0 = (((3.f - r.rd2) - r.rd4) * a) / 3.f;
        e1 = ((r.rd2 * a) * r.rd1) / 6.f;
        e2 = (((1.f - r.rd1) * a) * r.rd2) / 6.f;
        e3 = ((r.rd4 * a) * r.rd3) / 6.f;
        e4 = (((1.f - r.rd3) * a) * r.rd4) / 6.f;
      }
    }
    tws[i] = w;
    tm[i] = e0;
    tm[nl + i] = e1;
    tm[2 * nl + i] = e2;
    tm[3 * nl + i] = e3;
    tm[4 * nl + i] = e4;
  }
  __syncthreads();
  for (int j = static_cast<int>(threadIdx.x); j < kAvgT; j += kAvgBS) {
    const int k = k0 + j;
    if (k >= g.nq) break;
    const int t = k + 1 - ny;
    uint64_t ord = g.order[k];
    float v = 0.f;
    for (;;) {
      const int d = static_cast<int>(ord & 0xFu);
      if (d == 0xF) break;
      ord >>= 4;
      const int i = t - dir_offset(d, ny) - s0;
      const uint64_t w = tws[i];
      if (w & kTrnBit) {  // :1157-1158, :1192-1193 (listed at its own cell only)
        v = tm[i];
        continue;
      }
      if (d == 0) {  // the centre store, then the near-coast stores sent back, in store order
        v = v + tm[i];
        unsigned m = static_cast<unsigned>((w >> kCentreShift) & 0xFFULL);
        while (m != 0u) {
          const int j = __builtin_ctz(m) + 1;
          m &= m - 1u;
          v = v + tm[((j + 1) / 2) * nl + i];
        }
      } else {
        const unsigned f = static_cast<unsigned>((w >> (6 * (d - 1))) & 0x3FULL);
        if ((f & 3u) != 0u) v = v + tm[(static_cast<int>((f >> 2) & 3u) + 1) * nl + i];
        if ((f & 3u) == 2u) v = v + tm[(static_cast<int>((f >> 4) & 3u) + 1) * nl + i];
      }
    }
    if (t >= 1 && t <= g.nxy) {
      if (g.mapsta[t - 1] == 2) v = aq[k];  // :1280-1282
      if (g.global) {                       // :1286-1291
        if (t <= g.ny) v = v + avg_fold(g, s, c, aq, g.nxy + t);
        if (t > g.nxy - g.ny) v = v + avg_fold(g, s, c, aq, t - g.nxy);
      }
      if (g.cellsea[t] > 0) v = v * g.gsqrt[t - 1];  // :1301-1306
    }
    vq[k] = v;
  }
}

/// LDS bytes of an averaging tile (0: the tile does not fit, the plain gather runs instead).
size_t avg_tile_lds(int ny) {
  const size_t nl = static_cast<size_t>(kAvgT) + 2 * static_cast<size_t>(ny) + 2;
  const size_t b = nl * (sizeof(uint64_t) + 5 * sizeof(float));
  return b <= 64 * 1024 ? b : 0;
}

// ------------------------------------------------------------------------------------
// W3QCK3 sweeps (prop_qck3.hpp): the flux of every FLA entry, then the update of the active
// cells.  The closure copies (:1087-1112, :1240-1247) are reads of the wrapped column; the
// CFL numbers VLCFLX/VLCFLY (prop_xyp3.hpp:1080-1137) are formed where they are read.
// ------------------------------------------------------------------------------------
struct Sweep {
  int inc, close, xdir;
  const unsigned char* cls;
  const float* trans;
};

/// The CFL plane of a sweep for component c.
__device__ __forceinline__ const float* cfl_plane(const XyState& s, const Sweep& w, int c, int nf) {
  return (w.xdir ? s.vlcx : s.vlcy) + static_cast<size_t>(c) * nf;
}

__device__ __forceinline__ float cfll(const XyGrid& g, const Sweep& w, const float* cf, int j) {
  return cf[j - 1];
}

__device__ __forceinline__ float qat(const XyGrid& g, const Sweep& w, const float* q, int j) {
  if (w.close) {
    if (j <= 0) j += g.nxy;
    else if (j > g.nxy) j -= g.nxy;
  }
  return q[j + g.ny - 1];
}

__device__ float flux(const XyGrid& g, const float* cf, const Sweep& w, const float* q, int ixy) {
  const int cls = w.cls[ixy];
  if (cls == 0) return 0.f;
  const int inc = w.inc;
  if (cls == 1) {
    const float cfl = (cfll(g, w, cf, ixy + inc) + cfll(g, w, cf, ixy)) * 0.5f;
    const float sg = __builtin_copysignf(1.1f, cfl);
    const int ixyc = ixy - inc * fint(mn(sg, 0.f));
    const float c2 = cfl * cfl;
    const float t1 = (1.f - cfl) * qat(g, w, q, ixy + inc);
    const float t2 = (cfl + 1.f) * qat(g, w, q, ixy);
    const float h1 = (t1 + t2) * 0.5f;
    const float h2 = (1.f - c2) / 6.f;
    const float qixyc = qat(g, w, q, ixyc);
    const float d2 = (qat(g, w, q, ixyc - inc) - qixyc * 2.f) + qat(g, w, q, ixyc + inc);
    const float qb = h1 - h2 * d2;
    const int ixyu = ixyc - inc * fint(sg);
    const int ixyd = 2 * ixyc - ixyu;
    const float qixyu = qat(g, w, q, ixyu);
    const float dq = qat(g, w, q, ixyd) - qixyu;
    const float dqnz = __builtin_copysignf(mx(__builtin_fabsf(dq), 1.E-15f), dq);
    const float qcn = mn(mx((qixyc - qixyu) / dqnz, -0.1f), 1.1f);
    const float qbn0 = mx((qb - qixyu) / dqnz, qcn);
    const float u = mn(qcn / mx(__builtin_fabsf(cfl), 1.E-10f), 1.f);
    const float qbn = mn(qbn0, u);
    const float qbr = qixyu + qbn * dq;
    const float cfac = static_cast<float>(fint(__builtin_fabsf(qcn - 0.5f) * 2.f));
    const float qb2 = qixyc * cfac + (1.f - cfac) * qbr;
    return cfl * qb2;
  }
  const float cfl = cls == 2 ? cfll(g, w, cf, ixy) : cfll(g, w, cf, ixy + inc);
  const int ixyc = ixy - inc * fint(mn(__builtin_copysignf(1.1f, cfl), 0.f));
  return qat(g, w, q, ixyc) * cfl;
}

__global__ __launch_bounds__(256) void flux_kernel(XyGrid g, XyState s, int itloc, Sweep w) {
  const int c = blockIdx.y;
  const Comp cp = s.comps[c];
  if (itloc > cp.ntloc) return;
  const int k = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= g.nf) return;
  const float* q = s.q[itloc & 1] + static_cast<size_t>(c) * g.nq;
  const float* cf = cfl_plane(s, w, c, g.nf);
  const int i = k + 1 - g.ny;  // FLA(i), i = 1-NY .. NY*NX
  float v;
  if (i <= 0)
    v = w.close ? flux(g, cf, w, q, i + g.nxy) : 0.f;
  else
    v = flux(g, cf, w, q, i);
  s.fla[static_cast<size_t>(c) * g.nf + k] = v;
}

__global__ __launch_bounds__(256) void update_kernel(XyGrid g, XyState s, int itloc, Sweep w) {
  const int c = blockIdx.y;
  if (itloc > s.comps[c].ntloc) return;
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  if (i > g.nxy || g.act[i] == 0) return;
  float* q = s.q[itloc & 1] + static_cast<size_t>(c) * g.nq;
  const float* fla = s.fla + static_cast<size_t>(c) * g.nf;
  const float fm = fla[i - w.inc + g.ny - 1];
  const float fp = fla[i + g.ny - 1];
  const int jn = (fm > 0.f) ? -1 : 0;
  const int jp = (fp < 0.f) ? 1 : 0;
  const float t = q[i + g.ny - 1] + fm * w.trans[static_cast<size_t>(jn + 1) * g.nxy + i - 1];
  const float u = t - fp * w.trans[static_cast<size_t>(jp + 1) * g.nxy + i - 1];
  q[i + g.ny - 1] = mx(u, 0.f);
}

// :1368-1415: / GSQRT at the sea cells, then the boundary input (the last listing of a cell).
__global__ __launch_bounds__(256) void end_kernel(XyGrid g, XyState s, int itloc) {
  const int c = blockIdx.y;
  const Comp cp = s.comps[c];
  if (itloc > cp.ntloc) return;
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  if (i > g.nxy) return;
  float* q = s.q[itloc & 1] + static_cast<size_t>(c) * g.nq;
  if (g.cellsea[i] > 0) q[i + g.ny - 1] = q[i + g.ny - 1] / g.gsqrt[i - 1];
  if (g.flbpi && g.bnd[i] > 0) {
    float rd1 = g.dsec_bpi0 - ((g.dtg * static_cast<float>(cp.ntloc - itloc)) / cp.rntloc);
    float rd2 = g.dsec_bpin;
    if (rd2 > 0.001f) {
      rd2 = mn(mx(rd1 / rd2, 0.f), 1.f);
      rd1 = 1.f - rd2;
    } else {
      rd1 = 0.f;
      rd2 = 1.f;
    }
    const int ibi = g.bnd[i];
    const int ib = g.isbpi[ibi - 1];
    const float b0 = g.bbpi0[static_cast<size_t>(ibi) * g.nspec + cp.isp - 1];
    const float bn = g.bbpin[static_cast<size_t>(ibi) * g.nspec + cp.isp - 1];
    q[i + g.ny - 1] = (((b0 * rd1) + (bn * rd2)) / g.cg[ib * g.cgld + cp.ik]) * g.clats[ib];
  }
}

// Section 4 (:1430-1438) and W3SCAT: VA(ISP, ISEA) = the plane's cell where MAPSTA /= 0.
__global__ __launch_bounds__(256) void store_kernel(XyGrid g, XyState s, float* va) {
  __shared__ float tile[kTile][kTile + 1];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = blockIdx.x * kTile + 1, c0 = blockIdx.y * kTile;
  for (int r = ty; r < kTile; r += 4) {
    const int c = c0 + r, i = i0 + tx;
    float v = 0.f;
    if (c < s.nlist && i <= g.nxy) {
      const Comp cp = s.comps[c];
      v = s.q[cp.ntloc & 1][static_cast<size_t>(c) * g.nq + i + g.ny - 1];
      const int isea = g.cellsea[i];
      if (isea > 0 && g.mapsta[i - 1] > 0) {
        const float x = (g.cg[isea * g.cgld + cp.ik] / g.clats[isea]) * v;
        v = mx(x, 0.f);
      }
    }
    tile[r][tx] = v;
  }
  __syncthreads();
  for (int r = ty; r < kTile; r += 4) {
    const int i = i0 + r, c = c0 + tx;
    if (c >= s.nlist || i > g.nxy) continue;
    const int isea = g.cellsea[i];
    if (isea == 0 || g.mapsta[i - 1] == 0) continue;
    va[static_cast<size_t>(isea) * g.nspec + s.comps[c].isp - 1] = tile[tx][r];
  }
}

// ------------------------------------------------------------------------------------
// W3KTP3 (prop_ktp3.hpp:143-353), one 64-lane workgroup per point.
// ------------------------------------------------------------------------------------
struct KtLds {
  int vq, vcflt, cflk, fla, dsdd, frk, frg, fkc, db, dm, total;
  __host__ __device__ KtLds(int nk, int nth) {
    const int nk2 = nk + 2;
    vq = 0;
    vcflt = vq + nk2 * (nth + 3);
    cflk = vcflt + nk2 * (nth + 1);
    fla = cflk + nk2 * (nth + 1);
    dsdd = fla + nk2 * (nth + 1);
    frk = dsdd + nk2;
    frg = frk + nk;
    fkc = frg + nk;
    db = fkc + nth;
    dm = db + nk2;
    total = dm + nk2;
  }
};

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_xor(v, o, 64);
    v = u > v ? u : v;
  }
  return v;
}

/// W3QCK1 with CLOSE on the rotated spectrum: MX = NX = NTH, MY = NY = NK2, INC = NK2,
/// MAPACT = MAPBOU = MAPTH2, NACT = NB0 = NB1 = NB2 = NSPEC.
__device__ void qck1_theta(const KtGrid& g, float* vq, float* cfll, float* fla) {
  const int l = threadIdx.x, nk2 = g.nk2, nth = g.nth, my = nk2, inc = nk2, qoff = my - 1;
  for (int i = l; i < my * (nth + 1); i += 64) fla[i] = 0.f;
  for (int iy = l + 1; iy <= my; iy += 64) {
    vq[iy - my + qoff] = vq[iy - my + my * nth + qoff];
    vq[iy + my * nth + qoff] = vq[iy + qoff];
    vq[iy + my + my * nth + qoff] = vq[iy + my + qoff];
    cfll[iy + my * nth - 1] = cfll[iy - 1];
  }
  __syncthreads();
  for (int ip = l + 1; ip <= g.nspec; ip += 64) {
    const int ixy = g.mapth2[ip - 1];
    const float cfl = (cfll[ixy + inc - 1] + cfll[ixy - 1]) * 0.5f;
    const float sg = __builtin_copysignf(1.1f, cfl);
    const int ixyc = ixy - inc * fint(mn(sg, 0.f));
    const float c2 = cfl * cfl;
    const float t2 = (1.f - cfl) * vq[ixy + inc + qoff];
    const float t4 = (cfl + 1.f) * vq[ixy + qoff];
    const float t6 = (t4 + t2) * 0.5f;
    const float t8 = (1.f - c2) / 6.f;
    const float qixyc = vq[ixyc + qoff];
    const float t11 = vq[ixyc + inc + qoff] + (vq[ixyc - inc + qoff] - qixyc * 2.f);
    const float qb = t6 - t11 * t8;
    const int ixyu = ixyc - inc * fint(sg);
    const int ixyd = 2 * ixyc - ixyu;
    const float qixyu = vq[ixyu + qoff];
    const float dq = vq[ixyd + qoff] - qixyu;
    const float dqnz = __builtin_copysignf(mx(__builtin_fabsf(dq), 1.E-15f), dq);
    const float qcn = mn(mx((qixyc - qixyu) / dqnz, -0.1f), 1.1f);
    const float qbn0 = mx((qb - qixyu) / dqnz, qcn);
    const float t13 = mn(qcn / mx(__builtin_fabsf(cfl), 1.E-10f), 1.f);
    const float qbn = mn(qbn0, t13);
    const float qbr = qbn * dq + qixyu;
    const float cfac = static_cast<float>(fint(__builtin_fabsf(qcn - 0.5f) * 2.f));
    const float qb2 = cfac * qixyc + (1.f - cfac) * qbr;
    fla[ixy + qoff] = qb2 * cfl;
  }
  __syncthreads();
  for (int iy = l + 1; iy <= my; iy += 64) fla[iy - my + qoff] = fla[iy - my + my * nth + qoff];
  __syncthreads();
  for (int ip = l + 1; ip <= g.nspec; ip += 64) {
    const int ixy = g.mapth2[ip - 1];
    const float u = (vq[ixy + qoff] + fla[ixy - inc + qoff]) - fla[ixy + qoff];
    vq[ixy + qoff] = mx(u, 0.f);
  }
  __syncthreads();
}

/// W3QCK2 without closure on the rotated spectrum: INC = 1, DT = FACK, MAPBOU = MAPWN2 with
/// NB0 = NSPEC-NTH, NB1 = NSPEC, NB2 = NSPEC+NTH, MAPACT = MAPTH2.  DX1 and DX2 are the band
/// widths DB and DM, the same in every direction column: db[IK-1], dm[IK-1].
__device__ void qck2_k(const KtGrid& g, float fack, float* vq, const float* velo, const float* db,
                       const float* dm, float* fla) {
  const int l = threadIdx.x, nk2 = g.nk2, nth = g.nth, qoff = nk2 - 1;
  const int nb0 = g.nspec - nth, nb1 = g.nspec, nb2 = g.nspec + nth;
  const auto dx1 = [&](int i) { return db[((i - 1) % nk2 + nk2) % nk2]; };
  const auto dx2 = [&](int i) { return dm[((i - 1) % nk2 + nk2) % nk2]; };
  for (int i = l; i < nk2 * (nth + 1); i += 64) fla[i] = 0.f;
  __syncthreads();
  for (int ip = l + 1; ip <= nb2; ip += 64) {
    const int ixy = g.mapwn2[ip - 1];
    float f;
    if (ip <= nb0) {
      const float vel = (velo[ixy + 1 - 1] + velo[ixy - 1]) * 0.5f;
      const float dx2ixy = dx2(ixy);
      const float cfl = (fack * vel) / dx2ixy;
      const float sg = __builtin_copysignf(1.1f, cfl);
      const int ixyc = ixy - fint(mn(sg, 0.f));
      const float d2 = dx2ixy * dx2ixy;
      const float c2 = cfl * cfl;
      const float t2 = vq[ixy + 1 + qoff] * (1.f - cfl);
      const float t4 = vq[ixy + qoff] * (cfl + 1.f);
      const float t6 = (t4 + t2) * 0.5f;
      const float t8 = ((1.f - c2) * (d2 / dx1(ixyc))) / 6.f;
      const float qixyc = vq[ixyc + qoff];
      const float u1 = (vq[ixyc + 1 + qoff] - qixyc) / dx2(ixyc);
      const float u2 = (qixyc - vq[ixyc - 1 + qoff]) / dx2(ixyc - 1);
      const float qb = t6 - (u1 - u2) * t8;
      const int ixyu = ixyc - fint(sg);
      const int ixyd = 2 * ixyc - ixyu;
      const float qixyu = vq[ixyu + qoff];
      const float dq = vq[ixyd + qoff] - qixyu;
      const float dqnz = __builtin_copysignf(mx(__builtin_fabsf(dq), 1.E-15f), dq);
      const float qcn = mn(mx((qixyc - qixyu) / dqnz, -0.1f), 1.1f);
      const float qbn0 = mx((qb - qixyu) / dqnz, qcn);
      const float t13 = mn(qcn / mx(__builtin_fabsf(cfl), 1.E-10f), 1.f);
      const float qbn = mn(qbn0, t13);
      const float qbr = qbn * dq + qixyu;
      const float cfac = static_cast<float>(fint(__builtin_fabsf(qcn - 0.5f) * 2.f));
      const float qb2 = cfac * qixyc + (1.f - cfac) * qbr;
      f = qb2 * vel;
    } else {
      const float vel = ip <= nb1 ? velo[ixy - 1] : velo[ixy + 1 - 1];
      const int ixyc = ixy - fint(mn(__builtin_copysignf(1.1f, vel), 0.f));
      f = vq[ixyc + qoff] * vel;
    }
    fla[ixy + qoff] = f;
  }
  __syncthreads();
  for (int ip = l + 1; ip <= g.nspec; ip += 64) {
    const int ixy = g.mapth2[ip - 1];
    const float f = fack / dx1(ixy);
    const float u = vq[ixy + qoff] + f * (fla[ixy - 1 + qoff] - fla[ixy + qoff]);
    vq[ixy + qoff] = mx(u, 0.f);
  }
  __syncthreads();
}

__global__ __launch_bounds__(64) void ktp3_kernel(KtGrid g, int itstep, int ntimes, float facth, float fack,
                                                  const float* cg_all, const float* wn_all, const int* jsea,
                                                  const int* isea_l, const float* pscal, float* va,
                                                  float* cflthmax, float* cflkmax) {
  extern __shared__ float lds[];
  const KtLds o(g.nk, g.nth);
  float* vq = lds + o.vq;
  float* vcflt = lds + o.vcflt;
  float* cflk = lds + o.cflk;
  float* fla = lds + o.fla;
  float* dsdd = lds + o.dsdd;
  float* frk = lds + o.frk;
  float* frg = lds + o.frg;
  float* fkc = lds + o.fkc;
  float* db = lds + o.db;
  float* dm = lds + o.dm;
  const int j = blockIdx.x, l = threadIdx.x;
  const int nk = g.nk, nth = g.nth, nspec = g.nspec, nk2 = g.nk2, vqoff = nk2 - 1;
  const int js = jsea[j], is = isea_l[j];
  const float* ps = pscal + 10 * static_cast<size_t>(j);
  const float cthg0 = ps[0], dw = ps[1], dddx = ps[2], dddy = ps[3], cx = ps[4], cy = ps[5];
  const float dcxdx = ps[6], dcxdy = ps[7], dcydx = ps[8], dcydy = ps[9];
  const float* cg = cg_all + static_cast<size_t>(is) * nk2;
  const float* wn = wn_all + static_cast<size_t>(is) * nk2;
  float* vcol = va + static_cast<size_t>(js) * nspec;
  float acc_th = 0.f, acc_k = 0.f;
  for (int tt = 0; tt < ntimes; ++tt) {
    // 1-2: depth, zero fills, DSDD, the rotated spectrum (:1688-1725)
    const float depth = mx(g.dmin, dw);
    for (int i = l; i < nk2 * (nth + 3); i += 64) vq[i] = 0.f;
    for (int i = l; i < nk2 * (nth + 1); i += 64) {
      vcflt[i] = 0.f;
      cflk[i] = 0.f;
    }
    for (int ik = l; ik <= nk + 1; ik += 64) {
      if (wn[ik] * depth < 5.f) {
        const float t = wn[ik] * cg[ik] - g.sig[ik] * 0.5f;
        dsdd[ik] = mx(t, 0.f) / depth;
      } else {
        dsdd[ik] = 0.f;
      }
    }
    __syncthreads();
    for (int isp = l + 1; isp <= nspec; isp += 64) vq[g.mapth2[isp - 1] + vqoff] = vcol[isp - 1];
    acc_th = 0.f;
    acc_k = 0.f;
    // 3: refraction velocities and CFLTHMAX (:1729-1805)
    if (g.flcth) {
      const float fdg = cthg0 * facth;
      for (int ik = l + 1; ik <= nk; ik += 64) {
        frk[ik - 1] = (facth * dsdd[ik]) / wn[ik];
        frg[ik - 1] = fdg * cg[ik];
      }
      if (g.flcur) {
        const float dcyx = dcydx * facth;
        const float dcxxyy = (dcxdx - dcydy) * facth;
        const float dcxy = dcxdy * facth;
        for (int isp = l + 1; isp <= nspec; isp += 64)
          vcflt[g.mapth2[isp - 1] - 1] =
              ((g.es2[isp - 1] * dcyx) + (g.esc[isp - 1] * dcxxyy)) - (g.ec2[isp - 1] * dcxy);
      }
      __syncthreads();
      int lastnan = -1;
      for (int isp = l + 1; isp <= nspec; isp += 64) {
        const int m2 = g.mapth2[isp - 1], mw = g.mapwn[isp - 1];
        const float v = (vcflt[m2 - 1] + frg[mw - 1] * g.ecos[isp - 1]) +
                        frk[mw - 1] * ((g.esin[isp - 1] * dddx) - (g.ecos[isp - 1] * dddy));
        const float av = __builtin_fabsf(v);
        fla[isp - 1] = av;  // FLA is free until the passes
        if (av != av) lastnan = isp;
        vcflt[m2 - 1] = __builtin_copysignf(mn(av, g.ctmax), v);
      }
      // The in-order MAX from 0 restarts after a NaN: its result is the largest entry after
      // the last NaN (at least 0), or that NaN when it is the last entry.
      const int mlast = wave_max_i(lastnan);
      float m = 0.f;
      for (int isp = l + 1; isp <= nspec; isp += 64)
        if (isp > mlast) m = fmaxf(m, fla[isp - 1]);
      m = wave_max(m);
      __syncthreads();
      acc_th = (mlast == nspec) ? fla[nspec - 1] : m;
    }
    // 4: wavenumber shift velocities and CFLKMAX (:1811-1856)
    if (g.flck) {
      const float dcxx = -dcxdx, dcxyyx = -(dcxdy + dcydx), dcyy = -dcydy;
      const float fkd = (cx * dddx) + (cy * dddy);
      for (int ith = l + 1; ith <= nth; ith += 64)
        fkc[ith - 1] = ((g.ec2[ith - 1] * dcxx) + (g.esc[ith - 1] * dcxyyx)) + (g.es2[ith - 1] * dcyy);
      for (int ik = l; ik <= nk + 1; ik += 64) {
        db[ik] = g.dsip[ik] / cg[ik];
        dm[ik] = ik <= nk ? wn[ik + 1] - wn[ik] : 0.f;
      }
      __syncthreads();
      bool nan = false;
      float m = 0.f;
      for (int e = l; e < nk2 * nth; e += 64) {
        const int ik = e % nk2, ith = e / nk2 + 1;
        const float fkd0 = (fkd / cg[ik]) * dsdd[ik];
        const float velfac = fack / db[ik];
        const float v = ((wn[ik] * fkc[ith - 1]) + fkd0) * velfac;
        cflk[ik + (ith - 1) * nk2] = __builtin_copysignf(mn(g.ctmax, __builtin_fabsf(v)), v) / velfac;
        const float av = __builtin_fabsf(v);
        if (av != av) nan = true;
        else m = fmaxf(m, av);
      }
      m = wave_max(m);
      const bool anynan = __any(nan);
      __syncthreads();
      if (!anynan) {
        acc_k = m;
      } else {
        // the grouped order of the reference's vectorised MAX (prop_ktp3.hpp:277-310)
        float a = 0.f;
        if (l == 0) {
          for (int ik = 0; ik <= nk + 1; ++ik) {
            const float fkd0 = (fkd / cg[ik]) * dsdd[ik];
            const float velfac = fack / db[ik];
            const auto av = [&](int ith) {
              return __builtin_fabsf(((wn[ik] * fkc[ith - 1]) + fkd0) * velfac);
            };
            int ith = 1;
            float m0 = a, m1 = a;
            if (nth >= 4) {
              float l0 = a, l1 = a, l2 = a, l3 = a;
              const int n4 = (nth / 4) * 4;
              for (; ith <= n4; ith += 4) {
                l0 = mx(l0, av(ith));
                l1 = mx(l1, av(ith + 1));
                l2 = mx(l2, av(ith + 2));
                l3 = mx(l3, av(ith + 3));
              }
              m0 = mx(l2, l0);
              m1 = mx(l3, l1);
              a = mx(m1, m0);
            }
            if (nth - ith + 1 >= 2) {
              m0 = mx(m0, av(ith));
              m1 = mx(m1, av(ith + 1));
              ith += 2;
              a = mx(m1, m0);
            }
            if (ith <= nth) a = mx(a, av(ith));
          }
        }
        acc_k = __shfl(a, 0, 64);
      }
    }
    __syncthreads();
    // 5: the two passes, k first on an even ITSTEP (:1861-1933)
    for (int pass = 0; pass < 2; ++pass) {
      const bool kpass = (pass == 0) == (itstep % 2 == 0);
      if (kpass && g.flck) {
        for (int ith = l + 1; ith <= nth; ith += 64)
          vq[nk + 2 + (ith - 1) * nk2 + vqoff] = g.fachfa * vq[nk + 1 + (ith - 1) * nk2 + vqoff];
        __syncthreads();
        qck2_k(g, fack, vq, cflk, db, dm, fla);
      }
      if (!kpass && g.flcth) qck1_theta(g, vq, vcflt, fla);
    }
    // 6: back to VA (:1937-1939)
    for (int isp = l + 1; isp <= nspec; isp += 64) vcol[isp - 1] = vq[g.mapth2[isp - 1] + vqoff];
    __syncthreads();
  }
  if (l == 0) {
    cflthmax[js - 1] = acc_th;
    cflkmax[js - 1] = acc_k;
  }
}

}  // namespace
}  // namespace prop
}  // namespace ww3gpu

namespace ww3gpu {
namespace prop {

hipError_t xyp3_gather(const XyGrid& g, const XyState& s, const float* va, hipStream_t st) {
  const dim3 grid((g.nq + kTile - 1) / kTile, (s.nlist + kTile - 1) / kTile);
  hipLaunchKernelGGL(gather_kernel, grid, dim3(kTile, 4), 0, st, g, s, va);
  return hipGetLastError();
}

hipError_t xyp3_records(const XyGrid& g, const XyState& s, hipStream_t st) {
  hipLaunchKernelGGL(records_kernel, dim3((g.nxy + 255) / 256, s.nlist), dim3(256), 0, st, g, s);
  return hipGetLastError();
}

hipError_t xyp3_substep(const XyGrid& g, const XyState& s, int itloc, bool yfirst, bool flcx, bool flcy,
                        hipStream_t st) {
  const dim3 b(256);
  const size_t lds = avg_tile_lds(g.ny);
  if (lds > 0)
    hipLaunchKernelGGL(avg_tiled_kernel, dim3((g.nq + kAvgT - 1) / kAvgT, s.nlist), dim3(kAvgBS),
                       static_cast<unsigned>(lds), st, g, s, itloc);
  else
    hipLaunchKernelGGL(avg_kernel, dim3((g.nq + 255) / 256, s.nlist), b, 0, st, g, s, itloc);
  const Sweep ys{1, 0, 0, g.clsy, g.atrny};
  const Sweep xs{g.ny, g.global, 1, g.clsx, g.atrnx};
  for (int pass = 0; pass < 2; ++pass) {
    const bool ydir = (pass == 0) == yfirst;
    if (ydir ? !flcy : !flcx) continue;
    const Sweep& w = ydir ? ys : xs;
    hipLaunchKernelGGL(flux_kernel, dim3((g.nf + 255) / 256, s.nlist), b, 0, st, g, s, itloc, w);
    hipLaunchKernelGGL(update_kernel, dim3((g.nxy + 255) / 256, s.nlist), b, 0, st, g, s, itloc, w);
  }
  hipLaunchKernelGGL(end_kernel, dim3((g.nxy + 255) / 256, s.nlist), b, 0, st, g, s, itloc);
  return hipGetLastError();
}

hipError_t xyp3_store(const XyGrid& g, const XyState& s, float* va, hipStream_t st) {
  const dim3 grid((g.nxy + kTile - 1) / kTile, (s.nlist + kTile - 1) / kTile);
  hipLaunchKernelGGL(store_kernel, grid, dim3(kTile, 4), 0, st, g, s, va);
  return hipGetLastError();
}

size_t ktp3_lds_bytes(int nk, int nth) { return static_cast<size_t>(KtLds(nk, nth).total) * sizeof(float); }

hipError_t ktp3_batch(const KtGrid& g, int itstep, int ntimes, float facth, float fack, const float* cg,
                      const float* wn, int npts, const int* jsea, const int* isea, const float* pscal,
                      float* va, float* cflthmax, float* cflkmax, hipStream_t st) {
  if (npts <= 0) return hipSuccess;
  hipLaunchKernelGGL(ktp3_kernel, dim3(npts), dim3(64), ktp3_lds_bytes(g.nk, g.nth), st, g, itstep, ntimes,
                     facth, fack, cg, wn, jsea, isea, pscal, va, cflthmax, cflkmax);
  return hipGetLastError();
}

}  // namespace prop
}  // namespace ww3gpu

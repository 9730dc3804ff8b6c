// ww3-gpu_amd/csrc/snl1.hpp
// The DIA (W3SNL1) on gfx950: geometry of the quadruplet stencil, host set-up and the
// launch interface.
//
// INSNL1 (reference: kokkos/src/ww_kokkos/snl1_tables.cpp:82-329) produces 32 index
// tables.  Every entry is IT_n(ITH) + (IF_m(IFR)-1)*NTH: one of eight direction shifts
// (wrapped once) and one of eight frequency-row maps, the rows clamped at 0 (the zero
// row).  The GPU kernel does not gather through those tables.  It keeps each point's
// working arrays in LDS as rows of NTH values padded by a halo of H wrapped copies on
// both sides, and below row 1 as many zero rows as the deepest clamped row reference
// needs.  A table entry is then "row r + drow[m], column c + sh[n]": a constant offset
// from the bin's own slot.  The set-up builds the Fortran tables too, and proves, entry
// by entry, that this form reads the same values (same bin, or a zero row for a clamped
// one) before the kernel is allowed to run.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace ww3gpu {
namespace snl1 {

/// W3GDATMD's DIA inputs under the WW3 names (reference snl1_config.hpp:54-70).
struct Config {
  int nk, nth;
  float xfr, dth, lam, snlc1, kdcon, kdmn, snls1, snls2, snls3, fachfe;
};

/// Everything the kernel needs that is uniform across points.  Passed by value.
struct Geom {
  int nk, nth, nspec;
  int nfrchg, nfrhgh, nspecx;
  int h;        ///< halo columns on each side of a row
  int w;        ///< row stride in LDS: nth + 2h
  int zr;       ///< zero rows below Fortran row 1
  int ue_rows;  ///< zr + nfrhgh
  int sa_rows;  ///< zr + nfrchg
  int drow[8];  ///< IF_m(IFR) - IFR for an unclamped row, m = 1..8
  int sh[8];    ///< IT_n(ITH) - ITH before the wrap, n = 1..8
  /// LDS offsets (in floats) from a bin's own slot to each table's source:
  /// o3 = IP11..IP14, IM11..IM14, IP21..IP24, IM21..IM24 (section 3, into UE);
  /// o4 = IC11..IC81, IC12..IC82 (section 4, into the SA/DA arrays).
  int o3[16];
  int o4[16];
  float dal1, dal2, dal3;
  float awg[8], swg[8];
  float snlc1, kdcon, kdmn, snls1, snls2, snls3, fachfe;
  float tpiinv;
};

/// Where a launch finds its points.  The C ABI's arrays are point-major (a point's NSPEC
/// bins contiguous); the resident W3SRCE state is bin-major (point fastest).
struct Layout {
  long bin_stride;    ///< a, s, d: distance between consecutive bins of one point
  long pt_stride;     ///< a, s, d: distance between consecutive points
  long cg_k_stride;   ///< cg: distance between frequencies
  long cg_pt_stride;  ///< cg: distance between points
  const int* skip;    ///< when set, points with skip[p] != 0 are left untouched
  int xcd_chunk;      ///< > 0: workgroup i takes slot (i % 8) * xcd_chunk + i / 8
  const int* list;    ///< when set, slot i is point list[i] (npts counts the slots)
};

/// The C ABI's layout: a, s, d (NSPEC, npts) and cg (NK, npts), column-major.
inline Layout point_major(int nk, int nspec) {
  return Layout{1, nspec, 1, nk, nullptr, 0, nullptr};
}

/// Host-side result of INSNL1 for one grid.
struct Tables {
  Geom g{};
  bool structured = false;     ///< the halo/offset form reproduces every table entry
  std::vector<float> af11_row; ///< AF11 per frequency row, NFRCHG values
  std::vector<float> sig;      ///< SIG(1:NK)
  // Fortran-convention tables (1-based addresses), kept for export and verification.
  std::vector<int> ipim;       ///< 16 x NSPECX: IP11..IP14, IM11..IM14, IP21..IP24, IM21..IM24
  std::vector<int> ic;         ///< 16 x NSPEC : IC11..IC81, IC12..IC82
  float scal[19];              ///< DAL1..3, AWG1..8, SWG1..8
  int ifr_off[4];              ///< IFRP, IFRP1, IFRM, IFRM1
  int ith_off[4];              ///< ITHP, ITHP1, ITHM, ITHM1
};

/// INSNL1 on the host.  Returns false (and leaves `t` unusable) on a bad config.
bool build_tables(const Config& c, const float* sig, Tables& t);

/// LDS bytes one point's working set takes in the kernel.
size_t lds_bytes(const Geom& g);

/// Launch W3SNL1 over `npts` points on `stream`.  a, s, d: (NSPEC, npts) column-major
/// device arrays; cg: (NK, npts); kdmean: (npts).  af11_row, sig: device copies of the
/// grid's tables.  variant 0 picks a stencil-specialised kernel when one matches the
/// geometry (snl1_fixed.hip), else the general one; variant 1 forces the general one.
/// Returns hipSuccess or the launch error.
hipError_t launch(const Geom& g, const float* af11_row, const float* sig, int npts,
                  const float* a, const float* cg, const float* kdmean, float* s, float* d,
                  hipStream_t stream, int variant, const Layout& L);

/// The specialised kernels: returns false when none matches `g` (nothing launched), else
/// launches and stores the launch status in *err.
bool launch_specialized(const Geom& g, const float* af11_row, const float* sig, int npts,
                        const float* a, const float* cg, const float* kdmean, float* s, float* d,
                        hipStream_t stream, hipError_t* err, const Layout& L);

/// Workgroups a launch over `npts` points takes under `L`.
inline unsigned grid_size(const Layout& L, int npts) {
  return static_cast<unsigned>(L.xcd_chunk > 0 ? 8 * L.xcd_chunk : npts);
}

/// Whether launch_specialized() has a kernel for `g`.
bool has_specialized(const Geom& g);

/// Grid igrid's DIA as another module (W3SRCE) runs it: geometry, device tables, kernel
/// choice.  Valid until the grid is re-initialised or the runtime finalised.
struct GridView {
  const Geom* g = nullptr;
  const float* af11_row = nullptr;
  const float* sig = nullptr;
  int variant = 0;
};
/// False when grid igrid has no ready DIA context.
bool grid_view(int igrid, GridView& v);

}  // namespace snl1
}  // namespace ww3gpu

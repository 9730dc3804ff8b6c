// ww3-gpu_amd/csrc/srce.hpp
// ST4 source terms and the W3SRCE integration on gfx950: parameters, tables and the
// per-point state of one call.
//
// Reference: W3SRCE kokkos/src/ww_kokkos/srce_kernel.hpp:212-610, ST4
// kokkos/src/ww_kokkos/st4_*.hpp, packets st4_params.hpp:79-137 and srce_params.hpp:39-87.
//
// Work mapping.  A point is integrated by a group of 16 lanes (one DPP row; four points per
// wavefront).  The lanes share a point's per-bin work (growth rates, saturation, the
// cumulative breaking term, the update) and run its fixed-order sums together: every lane
// of the row adds the row's values in bin order, each operand broadcast from the lane that
// holds it (v_add_f32 with a row_newbcast DPP source), so each sum keeps the reference's
// order and every lane ends with the same bits.
//
// Layout.  Everything per point is point-major, as the C ABI passes it: a point's NSPEC bins
// are contiguous, so a row's 16 lanes read 64 consecutive bytes.  The caller's arrays are
// used in place; the library adds the per-bin arrays that live across kernels (initial
// spectrum, input and DIA terms, W3SDB1's source, LLWS) and a scalar block per point.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace ww3gpu {
namespace srce {

/// Default active-point counts at or below which a sub-step runs two 16-lane rows per point
/// instead of one (kHalfBelow), and a wavefront per point (kWideBelow).  At two wavefronts
/// per SIMD (the step kernel's register use) the GPU holds 4096 two-row points or 2048
/// wavefront points at once.  set_width_thresholds replaces them (the autotuner's choice).
constexpr int kHalfBelow = 4096;
constexpr int kWideBelow = 2048;

/// Row length of the mirrored cumulative table: 2*NTH-1 rounded up to a multiple of 4.
__host__ __device__ constexpr int kMirStride(int nth) { return (2 * nth - 1 + 3) / 4 * 4; }

/// ST4 scalars (W3GDATMD SRCP and W3SRC4MD module variables) and device table pointers.
struct St4Dev {
  int nk, nth, nspec, stab3, ig1, nsat;
  float dth, xfr, fte, ftf, aaircmin, aairgb, igpars5, wwnmeanptail, sstxftftail;
  float aalpha, bbeta, zz0max, zz0rat, zzalp, ssinthp, ttauwshelter;
  float sswellf[7], ssdsc[21], ssdsbr, sintailpar[5], ssdsp, wwnmeanp, sstxftwn, ssdsbrf1;
  float capchnk[10], zzwnd, ssdsdth, ssdsbm[5], delab;
  const float* sig;     ///< SIG(0:NK+1)
  const float* sig2;    ///< SIG2(1:NSPEC)
  const float* dsip;    ///< DSIP(0:NK+1)
  const float* dden;    ///< DDEN(1:NK)
  const float* dden2;   ///< DDEN2(1:NSPEC)
  const float* esin;    ///< ESIN(1:NSPEC+NTH)
  const float* ecos;    ///< ECOS(1:NSPEC+NTH)
  const float* satw;    ///< SATWEIGHTS(1:NSAT,1:NTH)
  const int* sati;      ///< SATINDICES(1:NSAT,1:NTH)
  /// CUMULW(IS2,IS) as a function of (IK, IK2, |ITH2-ITH|): cumulk[((IK-1)*NK + IK2-1)*NTH + D]
  /// (INSIN4's table has this structure; the commit checks every entry against it).
  const float* cumulk;
  /// The same rows mirrored: cumulm[((IK-1)*NK + IK2-1)*kMirStride(NTH) + M] = CUMULW for
  /// |ITH2-ITH| = |M - (NTH-1)|, M = 0 .. 2*NTH-2, so that ITH2-ITH+NTH-1 indexes it directly.
  const float* cumulm;
  const float* fwtable; ///< FWTABLE(0:300), or null when SWELLF(2) == 0
};

/// W3SRCE scalars and grid arrays (ww_srce_init).
struct SrceDev {
  int nk, nth, nspec, ic_numerics, fdonly, ln1;
  float dmin, dtmax, dtmin, facti1, facti2, fachfa, facp, xflt, xrel, fhmax, ffxfm, ffxpm, ffxfa;
  float icescales[4], slnc1, fspm, fshf, sbtc1, sdbc1, sdbc2;
  const float* sig;   ///< SIG(1:NK)
  const float* dden;  ///< DDEN(1:NK)
  const float* ecos;  ///< ECOS(1:NSPEC)
  const float* esin;  ///< ESIN(1:NSPEC)
  const int* mapwn;   ///< MAPWN(1:NSPEC), values 1-based
};

/// Flags of one call (ww_srce_batch's by-value arguments).
struct StepFlags {
  int it, inflags24, flogrd57, flogrd58;
  float dtg;
};

/// Per-point scalar fields; the first 22 are the C ABI's psc rows.
enum Scalar : int {
  S_USTAR = 0, S_USTDIR, S_TAUWX, S_TAUWY, S_DTDYN, S_FCUT, S_TAUOX, S_TAUOY, S_TAUWIX,
  S_TAUWIY, S_TAUWNX, S_TAUWNY, S_PHIAW, S_CHARN, S_TWS, S_PHIOC, S_PHIBBL, S_PHICE, S_TAUOCX,
  S_TAUOCY, S_WNMEAN, S_FPI,
  N_PSC,
  S_EMEAN = N_PSC, S_FMEAN, S_FMEAN1, S_AMAX, S_CD, S_Z0, S_FMEANWS, S_DLWMEAN, S_FHIGH,
  S_DTTOT, S_TAUWAX, S_TAUWAY, S_DEPTH, S_DRAT, S_DTDB,
  N_SCAL
};
/// pin rows (the C ABI's WwSrceBatchIn).
enum PinRow : int { P_D_INP = 0, P_U10ABS, P_U10DIR, P_AS, P_ICE, P_DAIR, P_COEF, N_PIN };
/// Per-point integer fields.
enum IScalar : int { I_NSTEPS = 0, I_DONE, I_SHAVE, I_NSCUT, N_ISCAL };

/// The arrays of one call over npts points.
struct State {
  int npts;
  // the caller's arrays (C ABI shapes, point-major)
  float* spec;            ///< [npts][NSPEC] INOUT
  const float* wn1;       ///< [npts][NK]
  const float* cg1;       ///< [npts][NK]
  const float* pin;       ///< [npts][N_PIN]
  float* psc;             ///< [npts][N_PSC] INOUT
  float* whitecap;        ///< [npts][4] INOUT
  float* taubbl;          ///< [npts][2] INOUT
  float* tauice;          ///< [npts][2] INOUT
  int* shave;             ///< [npts] OUT
  // the library's arrays
  float* specinit;        ///< [npts][NSPEC]
  float* vsin;            ///< [npts][NSPEC] W3SIN4's S (DSTAB(1) under STAB3 while it runs)
  float* vdin;            ///< [npts][NSPEC] W3SIN4's D
  float* vsnl;            ///< [npts][NSPEC] DIA S
  float* vdnl;            ///< [npts][NSPEC] DIA D
  float* vsds;            ///< [npts][NSPEC] W3SDS4's S of the current sub-step
  float* vdds;            ///< [npts][NSPEC] W3SDS4's D of the current sub-step
  float* vsdb;            ///< [npts][NSPEC] W3SDB1's S, kept across sub-steps (its early return)
  float* rnw;             ///< [npts][NSPEC] W3SDS4's cumulative term RENEW of the sub-step
  unsigned char* llws;    ///< [npts][NSPEC]
  float* mft;             ///< [npts][NK] the whitecap's MFT integral (fixed for the call)
  float* sc;              ///< [npts][N_SCAL]
  int* isc;               ///< [npts][N_ISCAL]
  float* kdmean;          ///< [npts] the DIA's KDMEAN of the next sub-step
  int* list[2];           ///< points still integrating, alternate sub-steps
  int* count;             ///< [kMaxIter] length of the list each sub-step leaves
};

/// Bytes of library memory a call over npts points needs (State minus the caller's arrays).
size_t state_bytes(int nk, int nth, int npts);
/// Carve the library's arrays of a State out of `base` (state_bytes() long, 256-aligned).
State carve_state(void* base, int nk, int nth, int npts);

/// LDS bytes one workgroup (four points) of the group kernels takes.
size_t group_lds_bytes(int nk, int nth);

/// The kernels (srce_group.hip).
/// init: W3SRCE's section 1 for every point, then the first sub-step's input terms and
/// KDMEAN; lists every point in list[0] and clears the sub-step counters.
hipError_t k_init(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f,
                  hipStream_t st);
/// One sub-step for the n points of list[iter % 2] (after the DIA of that sub-step): the
/// dissipation terms, the time-step limit, the update, the fluxes, W3SPR4, the tail, the
/// next sub-step's input terms.  Points not done are appended to list[(iter + 1) % 2] and
/// counted in count[iter].
hipError_t k_step(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f,
                  int iter, int n, hipStream_t st);
/// The group width of k_step: 0 chooses it per sub-step from the active-point count
/// (kHalfBelow, kWideBelow or set_width_thresholds); 16, 32 or 64 lanes per point use that
/// width for every sub-step.
void set_group_lanes(int lanes);
/// The active-point thresholds of the automatic width choice (wide_below <= half_below).
void set_width_thresholds(int half_below, int wide_below);
/// W3SRCE's section 3 for every point and the outputs.
hipError_t k_finish(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f,
                    hipStream_t st);

/// One ww_spr4 / ww_sin4 / ww_sds4 call on npts points (device copies, point-major).
struct St4Call {
  int npts, which;  ///< which: 0 W3SPR4, 1 W3SIN4, 2 W3SDS4
  int flogrd57, flogrd58;
  const float* a;   ///< [npts][NSPEC]
  const float* k;   ///< W3SPR4, W3SDS4: [npts][NK]; W3SIN4: [npts][NSPEC]
  const float* cg;  ///< [npts][NK]
  /// per-point scalars read, in the entry's argument order: W3SPR4 U, UDIR, TAUWX, TAUWY;
  /// W3SIN4 U, USTAR, DRAT, AS, USDIR, Z0; W3SDS4 USTAR, USDIR, DEPTH, DAIR
  const float* in[6];
  /// per-point scalars written: W3SPR4 USTAR, USDIR, EMEAN, FMEAN, FMEAN1, WNMEAN, AMAX, CD,
  /// Z0, CHARN, FMEANWS, DLWMEAN (USTAR and CHARN are also read); W3SIN4 TAUWX, TAUWY,
  /// TAUWNX, TAUWNY
  float* out[12];
  float* s;         ///< W3SIN4 S, W3SDS4 SRHS [npts][NSPEC]
  float* d;         ///< W3SIN4 D, W3SDS4 DDIAG
  float* brlambda;  ///< W3SDS4 BRLAMBDA
  float* whitecap;  ///< W3SDS4 [npts][4] INOUT
  int* llws;        ///< W3SPR4 in, W3SIN4 out, [npts][NSPEC] 0/1
  // the library's scratch
  float* pin;             ///< [npts][N_PIN]
  unsigned char* llws8;   ///< [npts][NSPEC]
  float* rnw;             ///< [npts][NSPEC]
  float* mft;             ///< [npts][NK]
  float* kband;           ///< [npts][NK]
};
/// The routine of io.which on every point (16-lane rows, four points per workgroup).
hipError_t k_routine(const St4Call& io, const St4Dev& q, hipStream_t st);

/// Counter slots of State::count: the most sub-steps one call may take.  W3SRCE's
/// sub-steps are at least min(DTMIN, rest) long, so a call takes about DTG/DTMIN of them;
/// a call that would need more is stopped with an error instead of looping on.
constexpr int kMaxIter = 4096;

}  // namespace srce
}  // namespace ww3gpu

// ww3-gpu_amd/csrc/srce.hpp
// ST4 source terms and the W3SRCE integration on gfx950: parameters, tables and the
// resident per-point state.
//
// Reference: W3SRCE kokkos/src/ww_kokkos/srce_kernel.hpp:212-610, ST4
// kokkos/src/ww_kokkos/st4_*.hpp, packets st4_params.hpp:79-137 and srce_params.hpp:39-87.
//
// Layout.  The C ABI hands over point-major arrays (a point's NSPEC bins contiguous).  The
// kernels run one point per lane, so the resident state is bin-major with the point index
// fastest ("[bin][P]", P = npts padded to 64): every load of one bin by a wavefront of 64
// points is one contiguous 256-byte line.  Transposes happen only at the C ABI boundary.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace ww3gpu {
namespace srce {

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
  const float* cumulw;  ///< CUMULW(1:NSPEC,1:NSPEC)
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

/// Per-point scalar fields of the resident state; the first 22 are the C ABI's psc rows.
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

/// Device pointers of the resident state of one batch of npts points, leading dimension P.
struct State {
  int npts, P;
  float* spec;      ///< [NSPEC][P]
  float* specinit;  ///< [NSPEC][P]
  float* vsin;      ///< [NSPEC][P]
  float* vdin;
  float* vsnl;
  float* vdnl;
  float* vsds;
  float* vdds;
  float* vsdb;      ///< W3SDB1's S, kept across sub-steps (its early return keeps it)
  float* brlambda;
  int* llws;        ///< [NSPEC][P]
  float* wn1;       ///< [NK][P]
  float* cg1;       ///< [NK][P]
  float* bth0;      ///< [NK][P]  W3SDS4 saturation maxima
  float* wnf;       ///< [NK][P]  W3SLN1 frequency factor
  float* cbeta;     ///< [NK][P]  W3SBT1 coefficient
  float* dirf;      ///< [NTH][P] W3SLN1 direction factor
  float* sc;        ///< [N_SCAL][P]
  int* isc;         ///< [N_ISCAL][P]
  float* pin;       ///< [N_PIN][P]
  float* whitecap;  ///< [4][P]
  float* taubbl;    ///< [2][P]
  float* tauice;    ///< [2][P]
  float* kdmean;    ///< [P] the DIA's KDMEAN of the current sub-step
  int* active;      ///< [kMaxIter] points still integrating after each sub-step
};

/// Bytes of device memory a State of npts points takes (with padding).
size_t state_bytes(int nk, int nth, int npts);
/// Carve a State out of `base` (state_bytes() long, 256-byte aligned).
State carve_state(void* base, int nk, int nth, int npts);

/// The kernels (srce_kernels.hip).  Every launch covers the batch's npts points, one per
/// lane; points that finished their integration return at once.
hipError_t k_pack(const State& s, int nk, int nth, const float* spec, const float* wn1,
                  const float* cg1, const float* pin, const float* psc, const float* whitecap,
                  const float* taubbl, const float* tauice, hipStream_t st);
hipError_t k_unpack(const State& s, int nk, int nth, float* spec, float* psc, float* whitecap,
                    float* taubbl, float* tauice, int* shave, hipStream_t st);
hipError_t k_init(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f,
                  hipStream_t st);
hipError_t k_step_a(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f,
                    int iter, hipStream_t st);
hipError_t k_step_c(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f,
                    int iter, hipStream_t st);
hipError_t k_finish(const State& s, const St4Dev& q, const SrceDev& p, const StepFlags& f,
                    hipStream_t st);

/// Counter slots of State::active: the most sub-steps one call may take.  W3SRCE's
/// sub-steps are at least min(DTMIN, rest) long, so a call takes about DTG/DTMIN of them;
/// a call that would need more is stopped with an error instead of looping on.
constexpr int kMaxIter = 4096;

}  // namespace srce
}  // namespace ww3gpu

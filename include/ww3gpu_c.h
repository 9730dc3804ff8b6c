/*
 * include/ww3gpu_c.h -- the C ABI of libww3gpu.so, the MI355X implementation of the
 * WAVEWATCH III spectral-step hot path.
 *
 * The first group of entry points has exactly the names, signatures and semantics of the
 * reference's Fortran-facing ABI (kokkos/src/fortran_iface/ww_kokkos_c.hpp), so the WW3
 * fork's ISO_C_BINDING interface blocks bind to this library unchanged.  Each declaration
 * cites the reference declaration it replaces.  The second group (prefix ww3gpu_) has no
 * counterpart in the reference: device-resident and inspection entry points.
 *
 * Rules, as in the reference (ww_kokkos_c.hpp:6-33):
 *  - Host arrays are Fortran-ordered (column-major), caller-owned, borrowed for the call.
 *  - float is WW3's default REAL.
 *  - Nothing throws or aborts.  Set-up entries return a status; compute entries are void
 *    and record it for the *_last_error() of their module.  A call that fails validation
 *    writes no output.
 *  - Not thread-safe: one caller at a time.
 */
#ifndef WW3GPU_C_H
#define WW3GPU_C_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: ww_kokkos_c.hpp:42-50. */
enum WwKokkosStatus {
  WW_KOKKOS_OK = 0,
  WW_KOKKOS_ERR_NOT_INITIALISED = 1,
  WW_KOKKOS_ERR_BAD_SHAPE = 2,
  WW_KOKKOS_ERR_KERNEL = 3,
  WW_KOKKOS_ERR_RUNTIME = 4,
  WW_KOKKOS_ERR_UNSUPPORTED = 5,
  WW_KOKKOS_ERR_TABLE = 6
};

/* ------------------------------------------------------------------------------------ */
/* Runtime                                                                               */
/* ------------------------------------------------------------------------------------ */

/* Replaces ww_kokkos_c.hpp:73 (ww_kokkos_init).  Selects the HIP device named by
 * WW_KOKKOS_DEVICE_ID (default 0), creates the library's stream, reads the run-time
 * flags.  comm_f (an MPI_Comm_c2f handle or -1) is ignored, as in the reference.
 * Idempotent.  Returns WW_KOKKOS_OK or WW_KOKKOS_ERR_RUNTIME. */
int ww_kokkos_init(int comm_f);

/* Replaces ww_kokkos_c.hpp:79 (ww_kokkos_finalize).  Frees every grid context and the
 * library's device buffers.  Safe to call twice or before init. */
void ww_kokkos_finalize(void);

/* Replaces ww_kokkos_c.hpp:875 (ww_kokkos_flag; resolution rules
 * kokkos/src/fortran_iface/part_flags.cpp:99-112), read once per name: 1 when the variable
 * `name` is "1" or unset and named by the WW_KOKKOS_PRESET preset, 2 when unset and
 * WW_KOKKOS_ALL is "1" (except the opt-in parts XYP1, KTP1, XYP2, KTP2, XYPUG, XYP3, KTP3, OUTG,
 * IOPE, UPDV, CPRT, ST4_TABLES), else 0. */
int ww_kokkos_flag(const char* name);

/* Replaces ww_kokkos_c.hpp:879: WW_KOKKOS_PRESET 0 unset or empty, 1 a known preset
 * ("resident": SRCE, XYP3, KTP3, OUTG, IOPE, UPDV, CPRT), -1 any other name. */
int ww_kokkos_preset_status(void);
/* Replaces ww_kokkos_c.hpp:883: 1 when WW_KOKKOS_STRICT is "1" (read once). */
int ww_kokkos_strict(void);
/* Replaces ww_kokkos_c.hpp:888: the counter of a part's calls left on the Fortran (name at
 * most 24 characters), created on first use, printed at exit; -1 when the name is too long or
 * the table (64) is full.  Call serially. */
int ww_kokkos_fallback_slot(const char* part);
/* Replaces ww_kokkos_c.hpp:891: one more count for the slot. */
void ww_kokkos_fallback_add(int slot);

/* ------------------------------------------------------------------------------------ */
/* W3SNL1, the DIA                                                                       */
/* ------------------------------------------------------------------------------------ */

/* Replaces ww_kokkos_c.hpp:98-100 (ww_snl1_init).  INSNL1 for grid igrid; sig is
 * SIG(1:nk), copied.  Returns OK, NOT_INITIALISED, BAD_SHAPE, UNSUPPORTED (a stencil
 * geometry the kernel does not cover, or a working set beyond LDS) or RUNTIME. */
int ww_snl1_init(int igrid, int nk, int nth, float xfr, float dth, float lam, float snlc1,
                 float kdcon, float kdmn, float snls1, float snls2, float snls3, float fachfe,
                 const float* sig);

/* Replaces ww_kokkos_c.hpp:116-117 (ww_snl1).  W3SNL1 for npts points: a, s, d are
 * (nspec, npts), cg (nk, npts), kdmean (npts), host memory.  s and d are overwritten.
 * Copies in, runs, copies out, synchronises.  npts == 0 is a no-op. */
void ww_snl1(int igrid, int npts, const float* a, const float* cg, const float* kdmean,
             float* s, float* d);

/* Replaces ww_kokkos_c.hpp:122 (ww_snl1_enabled): WW_KOKKOS_SNL1 as read at init. */
int ww_snl1_enabled(void);

/* Replaces ww_kokkos_c.hpp:133 (ww_snl1_last_error). */
int ww_snl1_last_error(void);

/* ------------------------------------------------------------------------------------ */
/* ST4: W3SPR4 with CALC_USTAR, W3SIN4, W3SDS4 and their tables                           */
/* ------------------------------------------------------------------------------------ */

/* ST4 RPAR indices, the reference's wire format (kokkos/src/ww_kokkos/st4_params.hpp:79-123).
 * Array parameters occupy consecutive slots: SSWELLF(1:7), SSDSC(1:21), SINTAILPAR(1:5),
 * CAPCHNK(1:10), SSDSBM(0:4). */
enum WwSt4Rpar {
  WW_ST4_R_DTH = 0, WW_ST4_R_XFR, WW_ST4_R_FTE, WW_ST4_R_FTF, WW_ST4_R_AAIRCMIN,
  WW_ST4_R_AAIRGB, WW_ST4_R_IGPARS5, WW_ST4_R_WWNMEANPTAIL, WW_ST4_R_SSTXFTFTAIL,
  WW_ST4_R_AALPHA, WW_ST4_R_BBETA, WW_ST4_R_ZZ0MAX, WW_ST4_R_ZZ0RAT, WW_ST4_R_ZZALP,
  WW_ST4_R_SSINTHP, WW_ST4_R_TTAUWSHELTER, WW_ST4_R_SSWELLF,
  WW_ST4_R_SSDSC = WW_ST4_R_SSWELLF + 7,
  WW_ST4_R_SSDSBR = WW_ST4_R_SSDSC + 21, WW_ST4_R_SINTAILPAR,
  WW_ST4_R_SSDSP = WW_ST4_R_SINTAILPAR + 5, WW_ST4_R_WWNMEANP, WW_ST4_R_SSTXFTWN,
  WW_ST4_R_SSDSBRF1, WW_ST4_R_SSDSBINT, WW_ST4_R_SSDSBCK, WW_ST4_R_SSDSHCK, WW_ST4_R_SSDSABK,
  WW_ST4_R_SSDSPBK, WW_ST4_R_CAPCHNK,
  WW_ST4_R_ZZWND = WW_ST4_R_CAPCHNK + 10, WW_ST4_R_SSDSCOS, WW_ST4_R_SSDSDTH, WW_ST4_R_SSDSBT,
  WW_ST4_R_SSDSBM,
  WW_ST4_R_DELUST = WW_ST4_R_SSDSBM + 5, WW_ST4_R_DELALP, WW_ST4_R_DELTAUW, WW_ST4_R_DELU,
  WW_ST4_R_DELTAIL, WW_ST4_R_DELAB,
  WW_ST4_NRPAR
};

/* ST4 IPAR indices (st4_params.hpp:126-137). */
enum WwSt4Ipar {
  WW_ST4_I_NK = 0, WW_ST4_I_NTH, WW_ST4_I_NSPEC, WW_ST4_I_MK, WW_ST4_I_MTH, WW_ST4_I_MSPEC,
  WW_ST4_I_STAB3, WW_ST4_I_IG1, WW_ST4_I_FLX5,
  WW_ST4_NIPAR
};

/* Array identifiers of ww_st4_set_real/int (kokkos/src/ww_kokkos/st4_arrays.hpp:42-62), with
 * the LBOUND every dimension must have: SIG(0:MK+1), SIG2(1:), DSIP(0:), DDEN(1:),
 * DDEN2(1:), ESIN(1:), ECOS(1:), TH(1:), SATWEIGHTS(1:,1:), CUMULW(1:,1:), DCKI(1:,1:),
 * QBI(1:,1:), SATINDICES(1:,1:) int, IKTAB(1:,1:) int, TAUT(0:,0:), TAUHFT(0:,0:),
 * TAUHFT2(0:,0:,0:), FWTABLE(0:).  The last four are module arrays: one copy per process,
 * the last registration wins. */
enum WwSt4ArrayId {
  WW_ST4_A_SIG = 0, WW_ST4_A_SIG2, WW_ST4_A_DSIP, WW_ST4_A_DDEN, WW_ST4_A_DDEN2, WW_ST4_A_ESIN,
  WW_ST4_A_ECOS, WW_ST4_A_TH, WW_ST4_A_SATWEIGHTS, WW_ST4_A_CUMULW, WW_ST4_A_DCKI,
  WW_ST4_A_QBI, WW_ST4_A_SATINDICES, WW_ST4_A_IKTAB, WW_ST4_A_TAUT, WW_ST4_A_TAUHFT,
  WW_ST4_A_TAUHFT2, WW_ST4_A_FWTABLE,
  WW_ST4_NARRAYS
};

/* Replaces ww_kokkos_c.hpp:158 (ww_st4_init).  Start (or restart) the configuration of grid
 * igrid; nrpar/nipar must equal WW_ST4_NRPAR/WW_ST4_NIPAR.  Returns OK, NOT_INITIALISED,
 * BAD_SHAPE, or UNSUPPORTED when ipar[WW_ST4_I_FLX5] is set. */
int ww_st4_init(int igrid, const float* rpar, int nrpar, const int* ipar, int nipar);

/* Replaces ww_kokkos_c.hpp:163-166 (ww_st4_set_real, ww_st4_set_int).  Register one array of
 * grid igrid under WwSt4ArrayId id, copied; lbound and extent have rank entries each, as
 * LBOUND() and SHAPE() give them.  Returns OK, NOT_INITIALISED or TABLE. */
int ww_st4_set_real(int igrid, int id, const float* data, int rank, const int* lbound,
                    const int* extent);
int ww_st4_set_int(int igrid, int id, const int* data, int rank, const int* lbound,
                   const int* extent);

/* Replaces ww_kokkos_c.hpp:172 (ww_st4_commit).  Check grid igrid's arrays against its
 * scalars (every array the reference checks, same extents) and copy them to the device.
 * With WW_KOKKOS_ST4_TABLES=1 the saturation and cumulative tables are rebuilt from the grid
 * arrays first (ww3gpu_insin4_tables) and one stderr line says how many values differ from
 * the registered ones.  Returns OK, NOT_INITIALISED, TABLE, RUNTIME, or UNSUPPORTED for a
 * configuration the kernels do not cover: SINTAILPAR(1) > 0.5 (tabulated u* and tail
 * stress) or SDSBCHOICE = NINT(SSDSC(1)) other than 1. */
int ww_st4_commit(int igrid);

/* Replaces ww_kokkos_c.hpp:200 (ww_st4_enabled): which = 1 WW_KOKKOS_SDS4, 2 WW_KOKKOS_SIN4,
 * 3 WW_KOKKOS_SPR4, 4 WW_KOKKOS_ST4_TABLES, each 1 when set to "1" in the environment. */
int ww_st4_enabled(int which);

/* Replaces ww_kokkos_c.hpp:203 (ww_st4_last_error). */
int ww_st4_last_error(void);

/* ------------------------------------------------------------------------------------ */
/* W3SRCE: the source-term integration of one time step                                 */
/* ------------------------------------------------------------------------------------ */

/* W3SRCE RPAR indices (kokkos/src/ww_kokkos/srce_params.hpp:39-67). */
enum WwSrceRpar {
  WW_SRCE_R_DMIN = 0, WW_SRCE_R_DTMAX, WW_SRCE_R_DTMIN, WW_SRCE_R_FACTI1, WW_SRCE_R_FACTI2,
  WW_SRCE_R_FACHFA, WW_SRCE_R_FACP, WW_SRCE_R_XFLT, WW_SRCE_R_XREL, WW_SRCE_R_FHMAX,
  WW_SRCE_R_FFXFM, WW_SRCE_R_FFXPM, WW_SRCE_R_FFXFA, WW_SRCE_R_ICESCALES,
  WW_SRCE_R_SLNC1 = WW_SRCE_R_ICESCALES + 4, WW_SRCE_R_FSPM, WW_SRCE_R_FSHF, WW_SRCE_R_SBTC1,
  WW_SRCE_R_SDBC1, WW_SRCE_R_SDBC2, WW_SRCE_R_FACSD, WW_SRCE_R_FTTR, WW_SRCE_R_FLDALPHA,
  WW_SRCE_R_TAIL_LEV, WW_SRCE_R_TAIL_TRAN1, WW_SRCE_R_TAIL_TRAN2,
  WW_SRCE_NRPAR
};

/* W3SRCE IPAR indices (srce_params.hpp:70-84). */
enum WwSrceIpar {
  WW_SRCE_I_NK = 0, WW_SRCE_I_NTH, WW_SRCE_I_NSPEC, WW_SRCE_I_IQTPE, WW_SRCE_I_IICEDISP,
  WW_SRCE_I_IC_NUMERICS, WW_SRCE_I_FDONLY, WW_SRCE_I_LAB_SWITCH, WW_SRCE_I_LN1,
  WW_SRCE_I_SEED, WW_SRCE_I_FLD2, WW_SRCE_I_TAIL_ID,
  WW_SRCE_NIPAR
};

/* Rows of ww_srce_batch's per-point input packet pin(WW_SRCE_NPIN, npts)
 * (ww_kokkos_c.hpp:236-245). */
enum WwSrceBatchIn {
  WW_SRCE_IN_D_INP = 0, WW_SRCE_IN_U10ABS, WW_SRCE_IN_U10DIR, WW_SRCE_IN_AS, WW_SRCE_IN_ICE,
  WW_SRCE_IN_DAIR, WW_SRCE_IN_COEF,
  WW_SRCE_NPIN
};

/* Rows of ww_srce_batch's INOUT packet psc(WW_SRCE_NPSC, npts): W3SRCE's pointer scalars in
 * argument order, SHAVE excepted (ww_kokkos_c.hpp:249-256). */
enum WwSrceBatchScal {
  WW_SRCE_SC_USTAR = 0, WW_SRCE_SC_USTDIR, WW_SRCE_SC_TAUWX, WW_SRCE_SC_TAUWY,
  WW_SRCE_SC_DTDYN, WW_SRCE_SC_FCUT, WW_SRCE_SC_TAUOX, WW_SRCE_SC_TAUOY, WW_SRCE_SC_TAUWIX,
  WW_SRCE_SC_TAUWIY, WW_SRCE_SC_TAUWNX, WW_SRCE_SC_TAUWNY, WW_SRCE_SC_PHIAW,
  WW_SRCE_SC_CHARN, WW_SRCE_SC_TWS, WW_SRCE_SC_PHIOC, WW_SRCE_SC_PHIBBL, WW_SRCE_SC_PHICE,
  WW_SRCE_SC_TAUOCX, WW_SRCE_SC_TAUOCY, WW_SRCE_SC_WNMEAN, WW_SRCE_SC_FPI,
  WW_SRCE_NPSC
};

/* Replaces ww_kokkos_c.hpp:216-218 (ww_srce_init).  Configure grid igrid: rpar/ipar at the
 * indices above; sig (SIG(1:nk)), dden (nk), ecos, esin, mapwn (nspec), copied.  Needs
 * ww_snl1_init and ww_st4_commit done for the grid, with the same NK and NTH.  Returns OK,
 * NOT_INITIALISED, BAD_SHAPE, or UNSUPPORTED when ipar[WW_SRCE_I_LAB_SWITCH] is not 1, or
 * when SEED or FLD2 is set (source terms the kernels do not cover). */
int ww_srce_init(int igrid, const float* rpar, int nrpar, const int* ipar, int nipar,
                 const float* sig, const float* dden, const float* ecos, const float* esin,
                 const int* mapwn);

/* Replaces ww_kokkos_c.hpp:225-231 (ww_srce): W3SRCE for one point, the arguments in the
 * order of the Fortran dummy list; every pointer argument is INOUT except wn1, cg1 (nk).
 * fpi is left as it is (written only under FLD2).  Status in ww_srce_last_error(). */
void ww_srce(int igrid, int it, int inflags24, int flogrd57, int flogrd58, float* spec,
             const float* wn1, const float* cg1, float d_inp, float u10abs, float u10dir,
             float as, float ice, float dtg, float dair, float coef, float* ustar,
             float* ustdir, float* tauwx, float* tauwy, float* dtdyn, float* fcut,
             float* tauox, float* tauoy, float* tauwix, float* tauwiy, float* tauwnx,
             float* tauwny, float* phiaw, float* charn, float* tws, float* phioc,
             float* whitecap, float* phibbl, float* taubbl, float* tauice, float* phice,
             float* tauocx, float* tauocy, float* wnmean, float* fpi, int* shave);

/* Replaces ww_kokkos_c.hpp:267-270 (ww_srce_batch): W3SRCE for npts points, host arrays
 * packed point by point: spec (nspec, npts), psc (WW_SRCE_NPSC, npts), whitecap (4, npts),
 * taubbl (2, npts), tauice (2, npts) INOUT; wn1, cg1 (nk, npts), pin (WW_SRCE_NPIN, npts)
 * IN; shave (npts) OUT as 0/1.  Copies in, runs every point's sub-step loop to DTG, copies
 * out, synchronises.  npts == 0 is a no-op.  IT > 0 with SINTAILPAR(4) <= 0.5 (TAUWBUG off)
 * is UNSUPPORTED: that W3SRCE reads LLWS before setting it, and the reference hands the
 * case to the Fortran (srce_kernel.hpp:288-290).  Status in ww_srce_last_error(). */
void ww_srce_batch(int igrid, int it, int inflags24, int flogrd57, int flogrd58, float dtg,
                   int npts, float* spec, const float* wn1, const float* cg1,
                   const float* pin, float* psc, float* whitecap, float* taubbl,
                   float* tauice, int* shave);

/* Replaces ww_kokkos_c.hpp:273 (ww_srce_enabled): WW_KOKKOS_SRCE as read at init. */
int ww_srce_enabled(void);

/* Replaces ww_kokkos_c.hpp:276 (ww_srce_last_error). */
int ww_srce_last_error(void);

/* ------------------------------------------------------------------------------------ */
/* Extensions (no counterpart in the reference ABI)                                     */
/* ------------------------------------------------------------------------------------ */

/* ww_srce_batch on device-resident arrays (same shapes, device pointers), enqueued on
 * `stream` (NULL = the default stream).  The host waits on one small counter per sub-step
 * (the number of points still integrating), so the call returns after the last sub-step's
 * kernels are enqueued; the outputs are complete when `stream` is.  Returns the status
 * (also recorded for ww_srce_last_error). */
int ww3gpu_srce_batch_dev(int igrid, int it, int inflags24, int flogrd57, int flogrd58,
                          float dtg, int npts, float* spec, const float* wn1, const float* cg1,
                          const float* pin, float* psc, float* whitecap, float* taubbl,
                          float* tauice, int* shave, void* stream);

/* Sub-steps the last ww_srce_batch / ww3gpu_srce_batch_dev call of grid igrid took (the
 * largest over its points), or -1. */
int ww3gpu_srce_substeps(int igrid);

/* Point-sub-steps of the last call of grid igrid (the number of points that took part in
 * each sub-step, summed over its sub-steps), or -1. */
long ww3gpu_srce_point_substeps(int igrid);

/* Test control, with no counterpart in the reference: the lanes per point of the W3SRCE
 * sub-step kernel for every later call of every grid, 16, 32 or 64, or 0 (the default: the
 * width follows each sub-step's active-point count).  Results do not depend on it.  Returns
 * OK or BAD_SHAPE. */
int ww3gpu_srce_group_lanes(int lanes);

/* Tuning control, with no counterpart in the reference: the active-point counts at or below
 * which the automatic width choice takes two rows per point (half_below) and a wavefront per
 * point (wide_below), for every later call of every grid; 0, 0 restores the defaults (4096,
 * 2048).  ww3gpu/autotune.py measures candidates and sets the fastest.  Results do not depend
 * on it.  Returns BAD_SHAPE for negative values or wide_below > half_below. */
int ww3gpu_srce_width_thresholds(int half_below, int wide_below);

/* Per-kernel timing of grid igrid's W3SRCE calls.  enable = 1 starts it (and clears the
 * timers), 0 stops it.  While on, every launch of a call is bracketed by HIP events on the
 * call's stream.  Returns OK or NOT_INITIALISED. */
int ww3gpu_srce_timing(int igrid, int enable);

/* The timers since ww3gpu_srce_timing(igrid, 1), waiting for the recorded launches: ms[k]
 * is the summed device time and launches[k] the number of launches of group k, for k < n:
 * 0 pack (point-major to bin-major), 1 init, 2 input terms (W3SLN1, W3SIN4), 3 DIA,
 * 4 dissipation and update (W3SDS4, W3SDB1, W3SBT1, limiter, update, W3SPR4, tail, W3SIN4),
 * 5 finish, 6 unpack.  Returns OK, NOT_INITIALISED or BAD_SHAPE. */
int ww3gpu_srce_timers(int igrid, double* ms, long* launches, int n);

/* INSIN4's saturation and cumulative tables on the host, from the ST4 packets and the grid
 * arrays, as ww_st4_commit rebuilds them under WW_KOKKOS_ST4_TABLES=1
 * (kokkos/src/ww_kokkos/st4_insin4.cpp:110-138, 239-262): satw, sati (2*(MTH/2-1)+1, MTH),
 * cumulw (MSPEC, MSPEC), Fortran order; sig0 is SIG(0:NK+1), dsip0 DSIP(0:NK+1), ecos
 * ECOS(1:), th TH(1:NTH).  Returns OK or BAD_SHAPE. */
int ww3gpu_insin4_tables(const float* rpar, int nrpar, const int* ipar, int nipar,
                         const float* sig0, const float* dsip0, const float* ecos,
                         const float* th, float* satw, int* sati, float* cumulw);

/* W3SNL1 on device-resident arrays (same shapes as ww_snl1, device pointers), enqueued
 * on `stream` (a hipStream_t; NULL = the default stream) without synchronising.  The
 * device-side counterpart of the reference's ww::snl1::snl1() on device Views
 * (kokkos/src/ww_kokkos/snl1_dia.hpp; timed by kokkos/tests/bench_snl1.cpp:93-126).
 * Returns the status (also recorded for ww_snl1_last_error). */
int ww3gpu_snl1_dev(int igrid, int npts, const float* a, const float* cg, const float* kdmean,
                    float* s, float* d, void* stream);

/* Export grid igrid's INSNL1 output in the Fortran convention (1-based addresses), in
 * the order of the reference fixture (kokkos/tests/fixtures/README.md:40-68):
 * dims9 = NFR, NFRHGH, NFRCHG, NSPECX, NSPECY, IFRP, IFRP1, IFRM, IFRM1;
 * scal19 = DAL1..3, AWG1..8, SWG1..8; ipim (16 x NSPECX); ic (16 x NSPEC);
 * af11 (NSPECX).  Any output pointer may be NULL.  Returns OK or NOT_INITIALISED. */
int ww3gpu_snl1_tables(int igrid, int* dims9, float* scal19, int* ipim, int* ic, float* af11);

/* Replace grid igrid's AF11 with the caller's NSPECX values (constant along each
 * frequency row, as INSNL1 makes them), e.g. to replay answers computed with another
 * x**11 evaluation order.  Returns OK, NOT_INITIALISED, BAD_SHAPE or TABLE. */
int ww3gpu_snl1_set_af11(int igrid, const float* af11, int n);

/* INSNL1 alone, on the host, with no device and no init: the tables the library would
 * build for this configuration, exported as ww3gpu_snl1_tables() does, plus
 * structured = 1 when the kernel's row/shift form reproduces every table entry.
 * Returns OK or BAD_SHAPE. */
int ww3gpu_insnl1(int nk, int nth, float xfr, float dth, float lam, const float* sig,
                  int* dims9, float* scal19, int* ipim, int* ic, float* af11, int* structured);

/* Choose grid igrid's DIA kernel: 0 (default) uses a kernel specialised for the grid's
 * quadruplet stencil when one exists, 1 always uses the general kernel.  Both compute the
 * same bits.  Returns OK, NOT_INITIALISED or BAD_SHAPE. */
int ww3gpu_snl1_set_variant(int igrid, int variant);

/* The DIA kernel grid igrid's next call will run: 0 specialised, 1 general, -1 no grid. */
int ww3gpu_snl1_kernel_variant(int igrid);

/* Bytes of LDS one point of grid igrid takes in the DIA kernel, or -1. */
long ww3gpu_snl1_lds_bytes(int igrid);

/* Evaluate device math function `fn` (0 = expf) on n device floats x (x2 is the second
 * argument of two-argument functions, else unused), writing y, on `stream` (NULL = the
 * default stream), without synchronising.  The device functions reproduce the host
 * glibc's results; this entry lets a caller check that on any inputs.
 * Returns OK, NOT_INITIALISED, BAD_SHAPE or KERNEL. */
int ww3gpu_devlibm_eval(int fn, const float* x, const float* x2, float* y, long n, void* stream);

/* The per-routine ST4 entries (ww_kokkos_c.hpp:177-196): one routine on npts points of a
 * committed grid, every per-point scalar an array of npts, point-major arrays; every OUT
 * argument is copied in as well as out.  Void; the status is read by ww_st4_last_error():
 * OK, NOT_INITIALISED (grid not committed), BAD_SHAPE (npts < 0 or a null array), RUNTIME,
 * KERNEL.  The same device routines as W3SRCE's.
 * W3SDS4: a, srhs, ddiag, brlambda (nspec, npts); k, cg (nk, npts); whitecap (4, npts);
 * ustar, usdir, depth, dair, dlwmean (npts; DLWMEAN is read only under SDSBCHOICE = 2, which
 * the commit refuses).  Replaces ww_kokkos_c.hpp:177-180. */
void ww_sds4(int igrid, int npts, const float* a, const float* k, const float* cg,
             const float* ustar, const float* usdir, const float* depth, const float* dair,
             const float* dlwmean, int flogrd57, int flogrd58, float* srhs, float* ddiag,
             float* brlambda, float* whitecap);
/* W3SIN4 in the argument order of W3SRCE's call: a, k, s, d, llws (nspec, npts); cg (nk,
 * npts); the scalars (npts).  Replaces ww_kokkos_c.hpp:184-187. */
void ww_sin4(int igrid, int npts, const float* a, const float* cg, const float* k,
             const float* u, const float* ustar, const float* drat, const float* as,
             const float* usdir, const float* z0, const float* cd, float* tauwx, float* tauwy,
             float* tauwnx, float* tauwny, float* s, float* d, int* llws);
/* W3SPR4 with CALC_USTAR: a, llws (nspec, npts); cg, wn (nk, npts); ustar, usdir INOUT.
 * Replaces ww_kokkos_c.hpp:191-196. */
void ww_spr4(int igrid, int npts, const float* a, const float* cg, const float* wn,
             const float* u, const float* udir, const float* tauwx, const float* tauwy,
             const int* llws, float* ustar, float* usdir, float* emean, float* fmean,
             float* fmean1, float* wnmean, float* amax, float* cd, float* z0, float* charn,
             float* fmeanws, float* dlwmean);

/* ------------------------------------------------------------------------------------ */
/* The resident spectrum (kokkos/src/fortran_iface/resident_c.hpp:38-103).  One context per  */
/* grid: VA(NSPEC,0:NSEALM) bound on the grid's first call (va = VA(1,0), nspec = SIZE(VA,1), */
/* nsealm = UBOUND(VA,2)), a device copy and the state HOST / DEVICE / BOTH saying which copy */
/* holds the newer bits.  Every call returns the status (BAD_SHAPE when va, nspec or nsealm   */
/* differ from the first call, or a column is outside 0..nsealm).                            */
/* ------------------------------------------------------------------------------------ */
/* Before a host read of VA: DEVICE -> one device-to-host copy, state BOTH.  Replaces :45. */
int ww_va_host_read(int igrid, float* va, int nspec, int nsealm);
/* Before a host write of VA: ww_va_host_read, then state HOST.  Replaces :48. */
int ww_va_host_write(int igrid, float* va, int nspec, int nsealm);
/* With state DEVICE, copy back only the columns cols[0..ncol) (0-based JSEA, duplicates
 * allowed); the state is not changed.  Replaces :57. */
int ww_va_host_read_cols(int igrid, float* va, int nspec, int nsealm, int ncol, const int* cols);
/* After host writes of exactly the columns cols[0..ncol): with state DEVICE or BOTH those host
 * columns are copied into the device copy; the state is not changed.  Replaces :68. */
int ww_va_host_wrote_cols(int igrid, float* va, int nspec, int nsealm, int ncol, const int* cols);
/* Before a host statement whose only writes of VA zero whole columns: no copy, no change of
 * state.  Replaces :74. */
int ww_va_host_update(int igrid, float* va, int nspec, int nsealm);
/* After it: the device copy's columns cols[0..ncol) = 0 when it holds the newer or equal bits
 * (DEVICE or BOTH); state unchanged.  Replaces :82-83. */
int ww_va_device_zero_cols(int igrid, float* va, int nspec, int nsealm, int ncol,
                           const int* cols);
/* Inside a kernel entry, before its launch reads the device copy: HOST -> one host-to-device
 * copy, state BOTH; write = 1 then marks DEVICE.  NOT_INITIALISED for an unbound grid.
 * Replaces :87. */
int ww_va_device(int igrid, int write);
/* The traffic counters, to stderr (also at exit under WW_KOKKOS_VA_COUNT=1).  Replaces :90. */
void ww_va_report(void);
/* ww_srce_batch on the resident spectrum: the spectra are the columns jsea[0..npts) (W3WAVE's
 * JSEA, 1..NSEALM, each once) of VA, read and written in the device copy (state DEVICE
 * afterwards); every other argument as ww_srce_batch's.  Status in ww_srce_last_error().
 * Replaces :98-101. */
void ww_srce_batch_va(int igrid, int it, int inflags24, int flogrd57, int flogrd58, float dtg,
                      float* va, int nspec, int nsealm, int npts, const int* jsea,
                      const float* wn1, const float* cg1, const float* pin, float* psc,
                      float* whitecap, float* taubbl, float* tauice, int* shave);
/* Extension: the state of grid igrid's resident VA, 0 HOST, 1 DEVICE, 2 BOTH, -1 unbound. */
int ww3gpu_va_state(int igrid);

/* ------------------------------------------------------------------------------------ */
/* PR3/UQ propagation (kokkos/src/fortran_iface/prop3_c.hpp:17-94).                       */
/* ------------------------------------------------------------------------------------ */

/* Packet lengths of ww_xyp3_batch (RPAR DTCFL, DTH, XFR, WDCG, WDTH; IPAR NX, NY, NSEA, NK,
 * NTH, NSPEC, ICLOSE, FLCX, FLCY, FLAGLL, FLCUR; IMAPS NMX0, NMX1, NMX2, NMY0, NMY1, NMY2,
 * NACT, NCENT; TSCAL DSEC21 of the day, of TBPI0 to TIME, of TBPI0 to TBPIN) and of
 * ww_ktp3_batch (RPAR DMIN, CTMAX, FACHFA; IPAR NK, NTH, NSPEC, NK2, FLCTH, FLCK, FLCUR;
 * PSCAL CTHG0S, DEPTH, DDDX, DDDY, CX, CY, DCXDX, DCXDY, DCYDX, DCYDY per point).
 * Replaces prop3_c.hpp:21 and :75. */
enum { WW_XYP3_NRPAR = 5, WW_XYP3_NIPAR = 11, WW_XYP3_NIMAPS = 8, WW_XYP3_NTSCAL = 3 };
enum { WW_KTP3_NRPAR = 3, WW_KTP3_NIPAR = 7, WW_KTP3_NPSCAL = 10 };

/* W3GATH, W3XYP3, W3SCAT for every component ISPEC with on[ISPEC-1] /= 0 in one call: the
 * ISPEC loop of W3WAVE section 3.6.3 on a regular or curvilinear grid.  Arguments in the
 * Fortran's layouts (prop3_shim.cpp:936-968 lists them); VA(NSPEC, 0:ncol-1) INOUT.
 * Returns OK, NOT_INITIALISED, BAD_SHAPE, UNSUPPORTED (tripole grid), TABLE (maps that
 * W3MAP3 cannot produce: a cell listed twice, a row outside the map's range, MAPCXY not a
 * permutation), RUNTIME or KERNEL; a refused call writes nothing.
 * Replaces prop3_c.hpp:25-32. */
int ww_xyp3_batch(const float* rpar, int nrpar, const int* ipar, int nipar, const int* imaps,
                  int nimaps, const float* tscal, int ntscal, float dtg, const int* mapsf,
                  const float* clats, const float* sig, const float* ecos, const float* esin,
                  const float* dpdx, const float* dpdy, const float* dqdx, const float* dqdy,
                  const float* gsqrt, const float* atrnx, const float* atrny, const int* mapx2,
                  const int* mapy2, const int* mapaxy, const int* mapcxy, const int* maptrn,
                  const int* mapsta, const float* cg, const float* cx, const float* cy,
                  int flbpi, int nbi, const int* isbpi, const float* bbpi0, const float* bbpin,
                  const int* on, int ncol, float* va);

/* ww_xyp3_batch on the resident VA of grid igrid (ww_va_*; ncol = nsealm + 1): the kernels
 * read and write the device copy, synced first; the state is DEVICE afterwards.  Replaces
 * prop3_c.hpp:37-44.  The MPI forms (prop3_c.hpp:50-68) return UNSUPPORTED: one GPU, no
 * exchange. */
int ww_xyp3_batch_va(int igrid, const float* rpar, int nrpar, const int* ipar, int nipar,
                     const int* imaps, int nimaps, const float* tscal, int ntscal, float dtg,
                     const int* mapsf, const float* clats, const float* sig, const float* ecos,
                     const float* esin, const float* dpdx, const float* dpdy, const float* dqdx,
                     const float* dqdy, const float* gsqrt, const float* atrnx,
                     const float* atrny, const int* mapx2, const int* mapy2, const int* mapaxy,
                     const int* mapcxy, const int* maptrn, const int* mapsta, const float* cg,
                     const float* cx, const float* cy, int flbpi, int nbi, const int* isbpi,
                     const float* bbpi0, const float* bbpin, const int* on, float* va, int nspec,
                     int nsealm);
int ww_xyp3_batch_mpi(int igrid, const float* rpar, int nrpar, const int* ipar, int nipar,
                      const int* imaps, int nimaps, const float* tscal, int ntscal, float dtg,
                      const int* mapsf, const float* clats, const float* sig, const float* ecos,
                      const float* esin, const float* dpdx, const float* dpdy, const float* dqdx,
                      const float* dqdy, const float* gsqrt, const float* atrnx,
                      const float* atrny, const int* mapx2, const int* mapy2, const int* mapaxy,
                      const int* mapcxy, const int* maptrn, const int* mapsta, const float* cg,
                      const float* cx, const float* cy, int flbpi, int nbi, const int* isbpi,
                      const float* bbpi0, const float* bbpin, const int* on, float* va,
                      int nspec, int nsealm, int naproc, int iaproc, const int* poff,
                      const float* rbuf, float* sbuf, int nbuf);
int ww_xyp3_mpi_pack(int igrid, float* va, int nspec, int nsealm, const int* isps, int nisp,
                     float* buf);
int ww_xyp3_mpi_unpack(int igrid, float* va, int nspec, int nsealm, const int* isps, int nisp,
                       const float* buf);

/* W3KTP3 NTIMES times on each of NPTS sea points (W3WAVE sections 3.6.2/3.6.4), VA(NSPEC,
 * 0:ncol-1), CFLTHMAX/CFLKMAX(ncol-1) INOUT.  Status codes as ww_xyp3_batch.
 * Replaces prop3_c.hpp:80-85. */
int ww_ktp3_batch(const float* rpar, int nrpar, const int* ipar, int nipar, const float* sig,
                  const float* dsip, const float* ecos, const float* esin, const float* ec2,
                  const float* esc, const float* es2, const int* mapwn, const int* mapth2,
                  const int* mapwn2, int itstep, int ntimes, float facth, float fack, int nsea,
                  const float* cg, const float* wn, int npts, const int* jsea, const int* isea,
                  const float* pscal, int ncol, float* va, float* cflthmax, float* cflkmax);
/* ww_ktp3_batch on the resident VA of grid igrid, as ww_xyp3_batch_va; CFLTHMAX and CFLKMAX
 * stay host arrays (NSEALM each).  Replaces prop3_c.hpp:89-94. */
int ww_ktp3_batch_va(int igrid, const float* rpar, int nrpar, const int* ipar, int nipar,
                     const float* sig, const float* dsip, const float* ecos, const float* esin,
                     const float* ec2, const float* esc, const float* es2, const int* mapwn,
                     const int* mapth2, const int* mapwn2, int itstep, int ntimes, float facth,
                     float fack, int nsea, const float* cg, const float* wn, int npts,
                     const int* jsea, const int* isea, const float* pscal, float* va, int nspec,
                     int nsealm, float* cflthmax, float* cflkmax);

/* Extensions (no counterpart in the reference): ww_xyp3_batch / ww_ktp3_batch with VA (and
 * the CFL maxima) in device memory, enqueued on `stream` (NULL = the library's stream); the
 * call returns when the work is done.  The other arrays are host arrays, as in the ABI. */
int ww3gpu_xyp3_batch_dev(const float* rpar, int nrpar, const int* ipar, int nipar,
                          const int* imaps, int nimaps, const float* tscal, int ntscal,
                          float dtg, const int* mapsf, const float* clats, const float* sig,
                          const float* ecos, const float* esin, const float* dpdx,
                          const float* dpdy, const float* dqdx, const float* dqdy,
                          const float* gsqrt, const float* atrnx, const float* atrny,
                          const int* mapx2, const int* mapy2, const int* mapaxy,
                          const int* mapcxy, const int* maptrn, const int* mapsta,
                          const float* cg, const float* cx, const float* cy, int flbpi, int nbi,
                          const int* isbpi, const float* bbpi0, const float* bbpin,
                          const int* on, int ncol, float* d_va, void* stream);
int ww3gpu_ktp3_batch_dev(const float* rpar, int nrpar, const int* ipar, int nipar,
                          const float* sig, const float* dsip, const float* ecos,
                          const float* esin, const float* ec2, const float* esc,
                          const float* es2, const int* mapwn, const int* mapth2,
                          const int* mapwn2, int itstep, int ntimes, float facth, float fack,
                          int nsea, const float* cg, const float* wn, int npts, const int* jsea,
                          const int* isea, const float* pscal, int ncol, float* d_va,
                          float* d_cflthmax, float* d_cflkmax, void* stream);
/* The status of the last propagation call. */
int ww3gpu_prop3_last_error(void);
/* out[0] GPU ms of the last W3XYP3 call, out[1] its cell-component sub-steps (NX*NY per
 * component per CFL sub-step), out[2] the most sub-steps of a component, out[3] GPU ms of
 * the last W3KTP3 call.  n >= 4. */
int ww3gpu_prop3_stats(double* out, int n);

/* ---- Learned W3SNL1 surrogate -------------------------------------------------------------
 * No counterpart in the reference, whose DIA is exact only.  A multilayer perceptron
 * (NSPEC + 1 inputs padded to a multiple of 128, three GELU layers of 512, NSPEC outputs
 * padded likewise; bf16 weights and activations with fp32 accumulation on the matrix cores)
 * maps a point's normalised spectrum and relative depth to S_nl.  It is trained here on the
 * exact DIA of the same grid; at inference every point whose prediction fails an energy-
 * conservation check (or whose spectrum cannot be normalised) gets the exact DIA instead.
 * Each entry needs ww_snl1_init(igrid) first.  Arrays are the DIA's: a, s (npts, NSPEC),
 * cg (npts, NK), kdmean (npts), point-major. */

/* Random initial network for grid igrid (replaces an existing one). */
int ww3gpu_snlsur_init(int igrid, unsigned long long seed);
/* Parameter count; dims4 (optional) gets the padded input width, hidden width, padded output
 * width and NSPEC.  Negative when igrid has no network. */
long long ww3gpu_snlsur_nparams(int igrid, int* dims4);
/* set != 0: load fp32 parameters from host; 0: copy them out.  Order: W1 [512][kin], b1, W2,
 * b2, W3, b3, W4 [nout][512], b4, each weight row-major [out][in]. */
int ww3gpu_snlsur_params(int igrid, int set, float* host, long long count);
/* Training data on the device: seeded JONSWAP sea states (with probability `bimodal` a swell
 * partition added), a (npts, NSPEC), cg (npts, NK), kdmean (npts). */
int ww3gpu_snlsur_synth_dev(int igrid, int npts, unsigned long long seed, float bimodal, float* a,
                            float* cg, float* kdmean, void* stream);
/* One Adam step on a batch of B points (B a multiple of 128) with DIA targets s; mean_loss
 * (optional, host) gets the batch's mean relative squared error in the normalised space. */
int ww3gpu_snlsur_train_dev(int igrid, int B, const float* a, const float* cg, const float* kdmean,
                            const float* s, float lr, float* mean_loss, void* stream);
/* Generate nsamples spectra, their exact DIA, and train for `epochs` passes in batches of
 * `batch` (cosine learning-rate decay from lr to lr/100).  epoch_loss (optional, host,
 * `epochs` values): mean loss over the last batch of each pass. */
int ww3gpu_snlsur_fit(int igrid, int nsamples, int batch, int epochs, float lr,
                      unsigned long long seed, float bimodal, float* epoch_loss, void* stream);
/* S_nl of npts points on device arrays; flags (optional, device, npts ints) gets 1 where the
 * exact DIA was used; *nfallback (optional, host) the number of such points.  Synchronises
 * the stream once per 2^20 points (the fallback count). */
int ww3gpu_snlsur_dev(int igrid, int npts, const float* a, const float* cg, const float* kdmean,
                      float* s, int* flags, int* nfallback, void* stream);
/* Test entries: loss per point and the fp32 parameter gradients for B host feature rows
 * x (B, kin) and normalised targets y (B, nout); the raw network output o (M, nout) for
 * M rows x (M, kin); C = A B^T for bf16 A (M, K), B (N, K) on device (M, N multiples of 128,
 * K of 64, row strides lda, ldb >= K and multiples of 8, ldc >= N; BAD_SHAPE otherwise). */
int ww3gpu_snlsur_grads(int igrid, int B, const float* x, const float* y, float* loss, float* grad);
int ww3gpu_snlsur_forward(int igrid, int M, const float* x, float* o);
int ww3gpu_sur_gemm_dev(int M, int N, int K, const void* a, int lda, const void* b, int ldb, float* c,
                        int ldc, void* stream);
/* The status of the last surrogate call. */
int ww3gpu_snlsur_last_error(void);

/* Library version string. */
const char* ww3gpu_version(void);

#ifdef __cplusplus
}
#endif

#endif /* WW3GPU_C_H */

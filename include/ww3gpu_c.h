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
 * kokkos/src/fortran_iface/part_flags.cpp:1-60): 1 if the environment variable
 * `name` was "1" (or WW_KOKKOS_ALL=1 and name is a part flag) when first asked, else 0. */
int ww_kokkos_flag(const char* name);

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
/* Extensions (no counterpart in the reference ABI)                                     */
/* ------------------------------------------------------------------------------------ */

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

/* Library version string. */
const char* ww3gpu_version(void);

#ifdef __cplusplus
}
#endif

#endif /* WW3GPU_C_H */

// ww3-gpu_amd/csrc/snlsur.hpp
// A learned surrogate for W3SNL1: a multilayer perceptron that maps a point's normalised
// spectrum (plus its relative depth) to the DIA's S_nl, in bf16 with fp32 accumulation on
// the gfx950 matrix cores, trained on DIA outputs this library computes.  It has no
// counterpart in the reference (whose DIA is exact only); the exact DIA stays the judge of
// its error and the fallback for points whose prediction fails a conservation check.
//
// What the network learns.  W3SNL1 computes UE = A / CONX with CONX_k = CG_k / (2 pi SIG_k),
// forms S_int = CONS(kd) * F(UE) from fixed quadruplet tables and returns S = CONX * S_int.
// F is cubic and homogeneous (F(c UE) = c^3 F(UE)), uses deep-water geometry and commutes
// with a rotation by whole direction bins.  The features are therefore u = UE / max(UE),
// rotated so the peak bin sits in direction 0 (linear values: bf16 keeps 2^-9 relative
// precision at every magnitude); the target is S_int / (CONS m^3 q) rotated the same way,
// with m = max(UE) and q = c0 sum_k,theta f_k^11 u^3 (a cubic proxy of the DIA's size that
// follows the energetic partition of a mixed sea).  Depth enters W3SNL1 only through CONS
// and CONX, both applied exactly around the network; the relative depth is kept as one more
// input feature.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ww3gpu {
namespace sur {

constexpr int kTile = 128;     ///< GEMM workgroup tile (rows and columns); every dimension is a multiple
constexpr int kTileK = 64;     ///< GEMM reduction step; every reduction length is a multiple
constexpr int kHidden = 512;   ///< width of the three hidden layers
constexpr float kScaleC0 = 1.0f / 64.0f;  ///< c0 in the target scale (targets of order one per bin)
constexpr float kEnergyTol = 0.12f;     ///< |sum S sig dsig| <= tol * sum |S| sig dsig, else exact DIA

inline int round_up(int n, int m) { return (n + m - 1) / m * m; }

/// Epilogue of the GEMM C[m][n] = sum_k A[m][k] B[n][k] (+ bias[n]).
enum EpiMode : int {
  kGeluFwd = 0,   ///< z = C + bias -> zf (fp32, optional), gelu(z) -> hb (bf16), its transpose -> ht (optional)
  kLinearFwd = 1, ///< C + bias -> of (fp32), or -> hb (bf16) when of is null
  kPlain = 2,     ///< C -> of (fp32)
  kGeluBwd = 3,   ///< C * gelu'(zf) -> hb (bf16) and its transpose -> ht
};

struct Epi {
  int mode;
  const float* bias;
  float* of; int ldo;        ///< fp32 output [M][ldo]
  uint16_t* hb; int ldh;     ///< bf16 output [M][ldh]
  uint16_t* ht; int ldt;     ///< bf16 transposed output [N][ldt]
  float* zf; int ldz;        ///< pre-activation: written by kGeluFwd, read by kGeluBwd
};

/// C = A B^T for bf16 A [M][lda], B [N][ldb] (row-major); M, N multiples of kTile, K of kTileK.
hipError_t gemm_nt(int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb,
                   const Epi& e, hipStream_t st);

/// The DIA-side constants the normalisation needs (a copy of the grid's W3SNL1 set-up).
struct Norm {
  int nk, nth, nspec, kin, nout;
  float tpiinv, snlc1, kdcon, kdmn, snls1, snls2, snls3;
  const float* sig;   ///< device SIG(1:NK)
  const float* wsig;  ///< device sig_k * dsig_k (energy weights of the conservation check)
};

/// Features (bf16 [Mpad][kin], and its transpose [kin][ldt] when xt is set) and per-point
/// (NSPEC <= 2048, NK <= 64)
/// scale/shift/flag from A (npts, NSPEC), CG (npts, NK), KDMEAN (npts).  With s_ref set
/// also the fp32 targets y [Mpad][nout].  Rows npts..mpad-1 are zero.
hipError_t prep(const Norm& n, int npts, int mpad, const float* a, const float* cg,
                const float* kd, const float* s_ref, uint16_t* x, uint16_t* xt, int ldt, float* y,
                float* scale, int* shift, int* flag, hipStream_t st);

/// S from the network output o (bf16 [Mpad][nout]); points that were flagged by prep or fail the
/// energy check are appended to list (count in *nlist, which must start at 0) and marked in
/// flag.
hipError_t post(const Norm& n, int npts, const uint16_t* o, const float* cg, const float* scale,
                const int* shift, int* flag, float* s, int* list, int* nlist, hipStream_t st);

/// Relative squared error per point: loss[p] = |o-y|^2 / |y|^2, and the gradient of
/// mean_p loss[p] as bf16 go [B][ldg] and its transpose got [nout][B].
hipError_t loss_grad(int B, int nout, const float* o, const float* y, float* loss, uint16_t* go,
                     uint16_t* got, hipStream_t st);

/// db[n] = sum_m gt[n][m] for bf16 gt [N][M].
hipError_t column_sums(int N, int M, const uint16_t* gt, float* db, hipStream_t st);

/// One Adam step on fp32 parameters w (rows x cols), with bf16 copies wb [rows][cols] and
/// wt [cols][rows] refreshed when given.
hipError_t adam(int rows, int cols, float* w, const float* g, float* m, float* v, uint16_t* wb,
                uint16_t* wt, float lr, float b1, float b2, float eps, float bc1, float bc2,
                hipStream_t st);

/// Refresh bf16 copies of fp32 w (rows x cols) without an update.
hipError_t to_bf16(int rows, int cols, const float* w, uint16_t* wb, uint16_t* wt, hipStream_t st);

/// Copy rows of the listed points: a (nspec), cg (nk), kd -> packed buffers.
hipError_t gather(int nspec, int nk, const int* list, int n, const float* a, const float* cg,
                  const float* kd, float* ap, float* cgp, float* kdp, hipStream_t st);
/// Scatter packed rows (nspec) back to the listed points.
hipError_t scatter(int nspec, const int* list, int n, const float* sp, float* s, hipStream_t st);

/// Seeded training spectra on the device: JONSWAP sea states drawn as the plain benchmark's
/// (Hs, Tp, gamma, direction, spreading, log-uniform depth), and with probability
/// `bimodal` a swell partition added.  a (npts, NSPEC), cg (npts, NK), kd (npts).
hipError_t synth(int nk, int nth, const float* sig, int npts, unsigned long long seed,
                 float bimodal, float* a, float* cg, float* kd, hipStream_t st);

}  // namespace sur
}  // namespace ww3gpu

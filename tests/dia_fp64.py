"""W3SNL1 evaluated in float64 with numpy, for the error of the float32 results against a
float64 evaluation of the same routine (test and benchmark infrastructure).

Follows the reference's operation sequence (kokkos/src/ww_kokkos/snl1_team.hpp:383-496):
CONS from KDMEAN, UE = A / CONX with the zero row below frequency 1 and the parametric tail
UE(row + 1) = FACHFE UE(row) above NK, the quadruplet sums over the extended spectrum, then
S = CONX (-2 (SA1 + SA2) + sum_j AWG_j (SA1(IC1j) + SA2(IC2j))).  The interaction tables,
DAL1-3 and AWG are INSNL1's (the float32 oracle's, taken exactly); SIG, CG, A and KDMEAN are
the float32 inputs; every operation after that, AF11 = (SIG/2pi)^11 included, is float64.
"""
from __future__ import annotations

import numpy as np


def w3snl1_fp64(cfg: dict, tables, sig, a, cg, kdmean) -> np.ndarray:
    nk, nth = int(cfg["nk"]), int(cfg["nth"])
    nspec = nk * nth
    dims = tables.dims
    nfr, nfrhgh, nfrchg, nspecx = (int(dims[i]) for i in range(4))
    scal = tables.scal.astype(np.float64)
    dal1, dal2, dal3 = scal[0], scal[1], scal[2]
    awg = scal[3:11]
    a = np.asarray(a, np.float64)
    npts = a.shape[0]
    sig64 = np.asarray(sig, np.float64)
    conx = (1.0 / (2.0 * np.pi)) / sig64[None, :] * np.asarray(cg, np.float64)  # (npts, nk)
    x = np.maximum(float(cfg["kdcon"]) * np.asarray(kdmean, np.float64), float(cfg["kdmn"]))
    cons = float(cfg["snlc1"]) * (1.0 + float(cfg["snls1"]) / x * (1.0 - float(cfg["snls2"]) * x) *
                                  np.exp(np.maximum(-1e15, float(cfg["snls3"]) * x)))
    # AF11 per extended row: (SIG/2pi)^11, the rows above NK continued by FR = FR * XFR.
    fr = np.empty(nfrchg)
    fr[:nfr] = sig64[:nfr] / (2.0 * np.pi)
    for r in range(nfr, nfrchg):
        fr[r] = fr[r - 1] * float(cfg["xfr"])
    af11 = np.repeat(fr ** 11, nth)[:nspecx]
    # UE with Fortran addresses 1-NTH .. NFRHGH*NTH at column J - 1 + NTH.
    ue = np.zeros((npts, nth + nfrhgh * nth))
    ue[:, nth:nth + nspec] = (a.reshape(npts, nk, nth) / conx[:, :, None]).reshape(npts, nspec)
    for r in range(nk, nfrhgh):
        ue[:, nth + r * nth:nth + (r + 1) * nth] = ue[:, nth + (r - 1) * nth:nth + r * nth] * float(cfg["fachfe"])

    def g(arr, idx):
        return arr[:, idx - 1 + nth]

    T = tables.ipim
    e00 = ue[:, nth:nth + nspecx]
    ep1 = sum(awg[j] * g(ue, T[j]) for j in range(4))
    em1 = sum(awg[4 + j] * g(ue, T[4 + j]) for j in range(4))
    ep2 = sum(awg[j] * g(ue, T[8 + j]) for j in range(4))
    em2 = sum(awg[4 + j] * g(ue, T[12 + j]) for j in range(4))
    factor = cons[:, None] * af11[None, :] * e00
    sa1 = factor * (e00 * (ep1 * dal1 + em1 * dal2) - ep1 * em1 * dal3)
    sa2 = factor * (e00 * (ep2 * dal1 + em2 * dal2) - ep2 * em2 * dal3)
    SA1 = np.zeros((npts, nth + nspecx))
    SA2 = np.zeros((npts, nth + nspecx))
    SA1[:, nth:] = sa1
    SA2[:, nth:] = sa2
    C = tables.ic
    acc = -2.0 * (SA1[:, nth:nth + nspec] + SA2[:, nth:nth + nspec])
    for j in range(8):
        acc = acc + awg[j] * (g(SA1, C[j]) + g(SA2, C[8 + j]))
    con = np.repeat(conx, nth, axis=1)
    return con * acc


def error_vs_fp64(s32: np.ndarray, s64: np.ndarray) -> dict:
    """Relative L2 error over all points, per point (median, max), and the largest error of
    a value relative to its point's largest |S|."""
    s32 = np.asarray(s32, np.float64)
    d = s32 - s64
    per = np.sqrt((d ** 2).sum(1)) / np.maximum(np.sqrt((s64 ** 2).sum(1)), 1e-300)
    scale = np.maximum(np.abs(s64).max(1, keepdims=True), 1e-300)
    return {"rel_l2_global": float(np.sqrt((d ** 2).sum() / (s64 ** 2).sum())),
            "rel_l2_point_median": float(np.median(per)), "rel_l2_point_max": float(per.max()),
            "max_err_over_point_max": float((np.abs(d) / scale).max())}

"""Seeded synthetic sea states for tests and benchmarks.

Spectral grid: sigma_k = 2 pi f1 XFR^(k-1) (f1 = 0.0418 Hz, XFR = 1.1), theta_j =
(j-1) 2 pi / NTH.  Each point gets a JONSWAP frequency spectrum with a cos^(2s) spreading
(Hs ~ U(0.5, 8) m, Tp ~ U(4, 16) s, gamma ~ U(1, 5), mean direction ~ U(0, 2 pi),
s ~ U(2, 20)), scaled so that 4 sqrt(m0) = Hs, then converted to the action density the
source terms take, A = E(sigma, theta) Cg / sigma, with k and Cg from linear dispersion
(30 Newton steps) -- the recipe of the reference's synthetic DIA input
(bench/dia/bench_dia.F90:81-107).  KDMEAN is the energy-weighted mean of k d.

Everything is computed in float64 with numpy's PCG64 and rounded to float32 once, so the
same seed gives the same bits on every machine.
"""
from __future__ import annotations

import numpy as np

GRAV = 9.806
PI32 = np.float32(3.141592653589793)
TPI32 = np.float32(2.0) * PI32


def dia_config(nk: int = 32, nth: int = 24, xfr: float = 1.1, f1: float = 0.0418) -> dict:
    """W3GDATMD's DIA inputs with WW3's NL1 defaults (LAMBDA 0.25, NLPROP 2.5e7,
    KDCONV 0.75, KDMIN 0.5, SNLCS1..3 = 5.5, 0.833, -1.25, FACHF 5)."""
    f32 = np.float32
    xfr32 = f32(xfr)
    cfg = {
        "nk": nk, "nth": nth, "xfr": xfr32, "dth": f32(TPI32 / f32(nth)), "lam": f32(0.25),
        "snlc1": f32(2.5e7 / GRAV ** 4), "kdcon": f32(0.75), "kdmn": f32(0.5),
        "snls1": f32(5.5), "snls2": f32(0.833), "snls3": f32(-1.25),
        "fachfe": f32(float(xfr32) ** -5.0),
    }
    sig = (2.0 * np.pi * f1 * float(xfr32) ** np.arange(nk)).astype(np.float32)
    return cfg, sig


def wavenumber(sig: np.ndarray, depth: np.ndarray) -> np.ndarray:
    """k solving sigma^2 = g k tanh(k d) (float64), shape (npts, nk)."""
    s = sig[None, :].astype(np.float64)
    d = np.asarray(depth, np.float64)[:, None]
    k = s * s / GRAV
    for _ in range(30):
        t = np.tanh(k * d)
        k = k - (GRAV * k * t - s * s) / (GRAV * t + GRAV * k * d / np.cosh(np.minimum(k * d, 300.0)) ** 2)
    return k


def group_velocity(sig: np.ndarray, k: np.ndarray, depth: np.ndarray) -> np.ndarray:
    d = np.asarray(depth, np.float64)[:, None]
    kd2 = np.minimum(2.0 * k * d, 80.0)
    return 0.5 * sig[None, :] / k * (1.0 + kd2 / np.sinh(kd2))


def jonswap_points(npts: int, nk: int, nth: int, sig: np.ndarray, depth, seed: int):
    """Returns (a (npts, NSPEC) float32, cg (npts, NK) float32, kdmean (npts,) float32,
    wn (npts, NK) float32)."""
    rng = np.random.default_rng(seed)
    hs = rng.uniform(0.5, 8.0, npts)
    tp = rng.uniform(4.0, 16.0, npts)
    gam = rng.uniform(1.0, 5.0, npts)
    thm = rng.uniform(0.0, 2.0 * np.pi, npts)
    spread = rng.uniform(2.0, 20.0, npts)
    depth = np.broadcast_to(np.asarray(depth, np.float64), (npts,))

    s64 = sig.astype(np.float64)
    f = s64 / (2.0 * np.pi)
    th = np.arange(nth) * (2.0 * np.pi / nth)
    xfr = s64[1] / s64[0] if nk > 1 else 1.1
    dsig = s64 * 0.5 * (xfr - 1.0 / xfr)
    dth = 2.0 * np.pi / nth

    fp = 1.0 / tp
    ff = f[None, :]
    sw = np.where(ff <= fp[:, None], 0.07, 0.09)
    shape = (GRAV ** 2 * (2.0 * np.pi) ** -4 * ff ** -5 * np.exp(-1.25 * (fp[:, None] / ff) ** 4) *
             gam[:, None] ** np.exp(-(ff - fp[:, None]) ** 2 / (2.0 * sw ** 2 * fp[:, None] ** 2)))
    dirs = (np.cos(0.5 * (th[None, :] - thm[:, None])) ** 2) ** spread[:, None]
    dirs = dirs / (dirs.sum(axis=1, keepdims=True) * dth)
    # E(sigma, theta) = E(f) / (2 pi) D(theta), then scale alpha so 4 sqrt(m0) = Hs.
    e = shape[:, :, None] / (2.0 * np.pi) * dirs[:, None, :]
    m0 = (e * dsig[None, :, None]).sum(axis=(1, 2)) * dth
    e *= ((hs / 4.0) ** 2 / np.maximum(m0, 1e-300))[:, None, None]

    k = wavenumber(sig, depth)
    cg = group_velocity(s64, k, depth)
    a = e * (cg / s64[None, :])[:, :, None]
    ek = e.sum(axis=2)
    kdmean = (ek * k * depth[:, None]).sum(axis=1) / np.maximum(ek.sum(axis=1), 1e-300)
    return (a.reshape(npts, nk * nth).astype(np.float32), cg.astype(np.float32),
            kdmean.astype(np.float32), k.astype(np.float32))


def dia_workload(npts: int, nk: int, nth: int, seed: int, depth=4000.0, log_depth=None):
    """Inputs of a DIA call: (cfg, sig, a, cg, kdmean).  depth is fixed unless
    log_depth=(lo, hi) asks for log-uniform depths."""
    cfg, sig = dia_config(nk, nth)
    if log_depth is not None:
        rng = np.random.default_rng(seed + 7919)
        depth = np.exp(rng.uniform(np.log(log_depth[0]), np.log(log_depth[1]), npts))
    a, cg, kd, _ = jonswap_points(npts, nk, nth, sig, depth, seed)
    return cfg, sig, a, cg, kd

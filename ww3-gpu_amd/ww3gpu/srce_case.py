"""A seeded, self-contained W3SRCE case: grid arrays, ST4 / W3SRCE / DIA parameter packets
and per-point forcing and spectra, in the reference ABI's layouts.

Grid arrays follow WW3's set-up (W3GRID/W3DIMX): SIG(0:NK+1) by the recurrence
SIGMA = SIGMA*XFR from FR1*TPI/XFR**2, DSIP = SIGMA*(XFR-1/XFR)/2, DSII with the
half-bands at both ends, DDEN = DTH*DSII*SIG, TH = (ITH-1)*DTH, ECOS/ESIN with |.|<1e-5
set to 0 and repeated for the wrap row, SIG2/DDEN2/MAPWN per bin.

Parameter set (documented, synthetic): ST4 close to the T471 tables of the WW3 manual
(BETAMAX 1.43, ZALP 0.006, TAUWSHELTER 0.3, SWELLF 0.66 / 0 / 0.022 / 1.5e5 / 1.2 / 0 /
3.6e5, Z0RAT 0.04, SDSC2 -2.2e-5, SDSCUM -0.40344, SDSC6 0.3, SDSBR 9e-4, SDSDTH 80,
SDSCOS 2, SDSP 2), analytic tail (SINTAILPAR(1) = 0), TAUWBUG on; W3SRCE with LN1, BT1
(GAMMA -0.067), DB1 (ALPHA 1, GAMMA 0.73), MLIM, DTMAX = DTG = 300 s, DTMIN 10 s.
SWELLF(2) = 0 keeps the friction factor constant (no FWTABLE).  Point forcing: U10
Weibull(k=2, mean 9 m/s) capped at 35 m/s, random direction, depth log-uniform 10-4000 m,
ice 0, air density 1.225, JONSWAP spectra as in synth.jonswap_points.
"""
from __future__ import annotations

import math

import numpy as np

from . import synth

f32 = np.float32
GRAV = f32(9.806)
TPI = f32(2.0) * f32(3.141592653589793)

# ST4 RPAR/IPAR and W3SRCE RPAR/IPAR index layouts (the reference's wire format:
# kokkos/src/ww_kokkos/st4_params.hpp:79-137, srce_params.hpp:39-87).
ST4_R = {k: i for i, k in enumerate(
    ["DTH", "XFR", "FTE", "FTF", "AAIRCMIN", "AAIRGB", "IGPARS5", "WWNMEANPTAIL", "SSTXFTFTAIL",
     "AALPHA", "BBETA", "ZZ0MAX", "ZZ0RAT", "ZZALP", "SSINTHP", "TTAUWSHELTER"])}
ST4_R["SSWELLF"] = 16
ST4_R["SSDSC"] = ST4_R["SSWELLF"] + 7
ST4_R["SSDSBR"] = ST4_R["SSDSC"] + 21
ST4_R["SINTAILPAR"] = ST4_R["SSDSBR"] + 1
_n = ST4_R["SINTAILPAR"] + 5
for k in ["SSDSP", "WWNMEANP", "SSTXFTWN", "SSDSBRF1", "SSDSBINT", "SSDSBCK", "SSDSHCK", "SSDSABK",
          "SSDSPBK"]:
    ST4_R[k] = _n
    _n += 1
ST4_R["CAPCHNK"] = _n
_n += 10
for k in ["ZZWND", "SSDSCOS", "SSDSDTH", "SSDSBT"]:
    ST4_R[k] = _n
    _n += 1
ST4_R["SSDSBM"] = _n
_n += 5
for k in ["DELUST", "DELALP", "DELTAUW", "DELU", "DELTAIL", "DELAB"]:
    ST4_R[k] = _n
    _n += 1
ST4_NRPAR = _n
ST4_I = {k: i for i, k in enumerate(["NK", "NTH", "NSPEC", "MK", "MTH", "MSPEC", "STAB3", "IG1", "FLX5"])}
ST4_NIPAR = len(ST4_I)

SRCE_R = {k: i for i, k in enumerate(
    ["DMIN", "DTMAX", "DTMIN", "FACTI1", "FACTI2", "FACHFA", "FACP", "XFLT", "XREL", "FHMAX", "FFXFM",
     "FFXPM", "FFXFA"])}
SRCE_R["ICESCALES"] = 13
_n = 17
for k in ["SLNC1", "FSPM", "FSHF", "SBTC1", "SDBC1", "SDBC2", "FACSD", "FTTR", "FLDALPHA", "TAIL_LEV",
          "TAIL_TRAN1", "TAIL_TRAN2"]:
    SRCE_R[k] = _n
    _n += 1
SRCE_NRPAR = _n
SRCE_I = {k: i for i, k in enumerate(
    ["NK", "NTH", "NSPEC", "IQTPE", "IICEDISP", "IC_NUMERICS", "FDONLY", "LAB_SWITCH", "LN1", "SEED",
     "FLD2", "TAIL_ID"])}
SRCE_NIPAR = len(SRCE_I)

PIN = ["D_INP", "U10ABS", "U10DIR", "AS", "ICE", "DAIR", "COEF"]
PSC = ["USTAR", "USTDIR", "TAUWX", "TAUWY", "DTDYN", "FCUT", "TAUOX", "TAUOY", "TAUWIX", "TAUWIY",
       "TAUWNX", "TAUWNY", "PHIAW", "CHARN", "TWS", "PHIOC", "PHIBBL", "PHICE", "TAUOCX", "TAUOCY",
       "WNMEAN", "FPI"]


def grid_arrays(nk: int, nth: int, xfr: float = 1.1, fr1: float = 0.0418) -> dict:
    xfr = f32(xfr)
    dth = f32(TPI / f32(nth))
    sig = np.zeros(nk + 2, np.float32)
    dsip = np.zeros(nk + 2, np.float32)
    sigma = f32(f32(fr1) * TPI / (xfr * xfr))
    sxfr = f32(f32(0.5) * (xfr - f32(1.0) / xfr))
    for ik in range(nk + 2):
        sigma = f32(sigma * xfr)
        sig[ik] = sigma
        dsip[ik] = f32(sigma * sxfr)
    dsii = np.zeros(nk, np.float32)
    dsii[0] = f32(f32(0.5) * sig[1] * (xfr - f32(1.0)))
    for ik in range(1, nk - 1):
        dsii[ik] = dsip[ik + 1]
    dsii[nk - 1] = f32(f32(0.5) * sig[nk] * (xfr - f32(1.0)) / xfr)
    dden = (dth * dsii * sig[1:nk + 1]).astype(np.float32)
    th = (np.arange(nth, dtype=np.float32) * dth).astype(np.float32)
    ec = np.cos(th.astype(np.float64)).astype(np.float32)
    es = np.sin(th.astype(np.float64)).astype(np.float32)
    ec[np.abs(ec) < 1e-5] = 0.0
    es[np.abs(es) < 1e-5] = 0.0
    nspec = nk * nth
    ecos = np.concatenate([np.tile(ec, nk), ec]).astype(np.float32)
    esin = np.concatenate([np.tile(es, nk), es]).astype(np.float32)
    sig2 = np.repeat(sig[1:nk + 1], nth).astype(np.float32)
    dden2 = np.repeat(dden, nth).astype(np.float32)
    mapwn = np.repeat(np.arange(1, nk + 1, dtype=np.int32), nth)
    return {"nk": nk, "nth": nth, "nspec": nspec, "xfr": xfr, "dth": dth, "sig0": sig, "dsip0": dsip,
            "dden": dden, "th": th, "ecos": ecos, "esin": esin, "sig2": sig2, "dden2": dden2,
            "mapwn": mapwn}


def st4_packets(g: dict) -> tuple:
    nk, nth = g["nk"], g["nth"]
    r = np.zeros(ST4_NRPAR, np.float32)
    i = np.zeros(ST4_NIPAR, np.int32)
    sign = g["sig0"][nk]
    dth = g["dth"]
    fachf = 5.0
    wnp, wnpt = 0.5, -0.5
    r[ST4_R["DTH"]] = dth
    r[ST4_R["XFR"]] = g["xfr"]
    r[ST4_R["FTE"]] = f32(0.25) * sign * dth * sign
    r[ST4_R["FTF"]] = f32(0.20) * dth * sign
    r[ST4_R["AAIRCMIN"]] = 2.0
    r[ST4_R["AAIRGB"]] = 0.8
    r[ST4_R["WWNMEANPTAIL"]] = wnpt
    r[ST4_R["SSTXFTFTAIL"]] = f32(1.0 / (fachf - 1.0 - wnpt * 2.0)) * dth * sign
    r[ST4_R["AALPHA"]] = 0.0095
    r[ST4_R["BBETA"]] = 1.43
    r[ST4_R["ZZ0MAX"]] = 0.0
    r[ST4_R["ZZ0RAT"]] = 0.04
    r[ST4_R["ZZALP"]] = 0.006
    r[ST4_R["SSINTHP"]] = 2.0
    r[ST4_R["TTAUWSHELTER"]] = 0.3
    for j, v in enumerate([0.66, 0.0, 0.022, 1.5e5, 1.2, 0.0, 3.6e5]):
        r[ST4_R["SSWELLF"] + j] = v
    ssdsc = [1.0, -2.2e-5, -0.40344, 1.0, 0.0, 0.3, 0.3, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0,
             0.0, 0.0, 1.0, 0.0]
    for j, v in enumerate(ssdsc):
        r[ST4_R["SSDSC"] + j] = v
    r[ST4_R["SSDSBR"]] = 9.0e-4
    for j, v in enumerate([0.0, 1.0, 0.0, 1.0, 1.0]):
        r[ST4_R["SINTAILPAR"] + j] = v
    r[ST4_R["SSDSP"]] = 2.0
    r[ST4_R["WWNMEANP"]] = wnp
    r[ST4_R["SSTXFTWN"]] = f32(1.0 / (fachf - 1.0 - wnp * 2.0) * (float(sign) ** (2 * wnp)) /
                               (9.806 ** wnp)) * sign * dth
    r[ST4_R["SSDSBRF1"]] = 0.5
    r[ST4_R["SSDSBINT"]] = 0.3
    r[ST4_R["SSDSBCK"]] = 0.0
    r[ST4_R["SSDSHCK"]] = 1.5
    r[ST4_R["SSDSABK"]] = 1.5
    r[ST4_R["SSDSPBK"]] = 4.0
    r[ST4_R["ZZWND"]] = 10.0
    r[ST4_R["SSDSCOS"]] = 2.0
    r[ST4_R["SSDSDTH"]] = 80.0
    r[ST4_R["SSDSBT"]] = 0.0
    r[ST4_R["SSDSBM"]] = 1.0
    r[ST4_R["DELAB"]] = 0.0
    i[ST4_I["NK"]] = nk
    i[ST4_I["NTH"]] = nth
    i[ST4_I["NSPEC"]] = nk * nth
    i[ST4_I["MK"]] = nk
    i[ST4_I["MTH"]] = nth
    i[ST4_I["MSPEC"]] = nk * nth
    return r, i


def srce_packets(g: dict, dtg: float = 300.0) -> tuple:
    nk, nth = g["nk"], g["nth"]
    r = np.zeros(SRCE_NRPAR, np.float32)
    i = np.zeros(SRCE_NIPAR, np.int32)
    xfr = float(g["xfr"])
    facti1 = f32(1.0) / f32(math.log(xfr))
    r[SRCE_R["DMIN"]] = 2.5
    r[SRCE_R["DTMAX"]] = dtg
    r[SRCE_R["DTMIN"]] = 10.0
    r[SRCE_R["FACTI1"]] = facti1
    r[SRCE_R["FACTI2"]] = f32(1.0) - f32(math.log(float(g["sig0"][1]))) * facti1
    r[SRCE_R["FACHFA"]] = f32(xfr ** -5.0)
    xp = 0.15
    r[SRCE_R["FACP"]] = f32(xp / math.pi * 0.62e-3 * (2 * math.pi) ** 4 / 9.806 ** 2)
    r[SRCE_R["XFLT"]] = 0.05
    r[SRCE_R["XREL"]] = 0.1
    r[SRCE_R["FHMAX"]] = 0.5
    r[SRCE_R["FFXFM"]] = f32(2.5) * TPI
    r[SRCE_R["FFXPM"]] = f32(4.0) * GRAV
    r[SRCE_R["FFXFA"]] = 0.0
    for j in range(4):
        r[SRCE_R["ICESCALES"] + j] = 1.0
    r[SRCE_R["SLNC1"]] = f32(80.0 * (1.225e-3) ** 2 / 9.806 ** 2)
    r[SRCE_R["FSPM"]] = 1.0
    r[SRCE_R["FSHF"]] = 0.5
    r[SRCE_R["SBTC1"]] = f32(2.0 * -0.067 / 9.806)
    r[SRCE_R["SDBC1"]] = 1.0
    r[SRCE_R["SDBC2"]] = 0.73
    i[SRCE_I["NK"]] = nk
    i[SRCE_I["NTH"]] = nth
    i[SRCE_I["NSPEC"]] = nk * nth
    i[SRCE_I["IQTPE"]] = 1
    i[SRCE_I["LAB_SWITCH"]] = 1
    i[SRCE_I["LN1"]] = 1
    return r, i


def make_case(npts: int, nk: int = 32, nth: int = 36, seed: int = 0, dtg: float = 300.0,
              depth_range=(10.0, 4000.0)) -> dict:
    g = grid_arrays(nk, nth)
    st4_r, st4_i = st4_packets(g)
    srce_r, srce_i = srce_packets(g, dtg)
    cfg, _ = synth.dia_config(nk, nth)
    sig = g["sig0"][1:nk + 1].copy()
    rng = np.random.default_rng(seed + 104729)
    depth = np.exp(rng.uniform(np.log(depth_range[0]), np.log(depth_range[1]), npts))
    lam = 9.0 / math.gamma(1.5)
    u10 = np.minimum(rng.weibull(2.0, npts) * lam, 35.0)
    udir = rng.uniform(0.0, 2.0 * np.pi, npts)
    a, cg, _, wn = synth.jonswap_points(npts, nk, nth, sig, depth, seed)
    pin = np.zeros((npts, len(PIN)), np.float32)
    pin[:, 0] = depth
    pin[:, 1] = u10
    pin[:, 2] = udir
    pin[:, 5] = 1.225
    pin[:, 6] = 1.0
    psc = np.zeros((npts, len(PSC)), np.float32)
    psc[:, 0] = (np.sqrt(1.2e-3) * u10).astype(np.float32)
    psc[:, 1] = udir
    return {"grid": g, "st4_rpar": st4_r, "st4_ipar": st4_i, "srce_rpar": srce_r, "srce_ipar": srce_i,
            "dia_cfg": cfg, "sig": sig, "npts": npts, "nk": nk, "nth": nth, "spec": a, "wn1": wn,
            "cg1": cg, "pin": pin, "psc": psc, "whitecap": np.zeros((npts, 4), np.float32),
            "taubbl": np.zeros((npts, 2), np.float32), "tauice": np.zeros((npts, 2), np.float32),
            "shave": np.zeros(npts, np.int32),
            "flags": {"it": 1, "inflags24": 0, "flogrd57": 1, "flogrd58": 1, "dtg": float(dtg)}}

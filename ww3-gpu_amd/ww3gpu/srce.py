"""ST4 + W3SRCE through the C ABI: the reference's ww_st4_init / ww_st4_set_real/int /
ww_st4_commit and ww_srce_init / ww_srce / ww_srce_batch calls
(kokkos/src/fortran_iface/ww_kokkos_c.hpp:135-276), plus the device-resident batch entry.

A grid is configured as the WW3 fork's W3INIT configures it: the DIA first (ww_snl1_init),
then ST4 (packets, every array the commit checks, commit), then W3SRCE.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import srce_case as sc
from .capi import check, lib, ptr
from .snl1 import Snl1, init_runtime

# ArrayIds of ww_st4_set_real/int (include/ww3gpu_c.h, WwSt4ArrayId).
A_SIG, A_SIG2, A_DSIP, A_DDEN, A_DDEN2, A_ESIN, A_ECOS, A_TH, A_SATWEIGHTS, A_CUMULW, A_DCKI, \
    A_QBI, A_SATINDICES, A_IKTAB, A_TAUT, A_TAUHFT, A_TAUHFT2, A_FWTABLE = range(18)
NKHS, NKD, NDTAB, SIZEFWTABLE = 2000, 1300, 2000, 300


def insin4_tables(case: dict) -> dict:
    """The library's INSIN4 saturation and cumulative tables for the case's grid (host):
    satw, sati as (MTH, nsat) C arrays = (nsat, MTH) Fortran; cumulw (MSPEC, MSPEC)."""
    g = case["grid"]
    r, i = case["st4_rpar"], case["st4_ipar"]
    mth, mspec = int(i[sc.ST4_I["MTH"]]), int(i[sc.ST4_I["MSPEC"]])
    nsat = 2 * (mth // 2 - 1) + 1
    satw = np.zeros((mth, nsat), np.float32)
    sati = np.zeros((mth, nsat), np.int32)
    cumulw = np.zeros((mspec, mspec), np.float32)
    check("ww3gpu_insin4_tables",
          lib().ww3gpu_insin4_tables(ptr(r), r.size, ptr(i), i.size, ptr(g["sig0"]), ptr(g["dsip0"]),
                                     ptr(g["ecos"]), ptr(g["th"]), ptr(satw), ptr(sati), ptr(cumulw)))
    return {"satw": satw, "sati": sati, "cumulw": cumulw, "nsat": nsat}


def _set(igrid: int, aid: int, data: np.ndarray, lbound, extent, is_int=False) -> int:
    lb = np.asarray(lbound, np.int32)
    ex = np.asarray(extent, np.int32)
    fn = lib().ww_st4_set_int if is_int else lib().ww_st4_set_real
    return fn(igrid, aid, ptr(data), len(extent), ptr(lb), ptr(ex))


def register_st4(igrid: int, case: dict, tables: dict, skip=()) -> None:
    """ww_st4_init and one registration per array, as W3INIT makes them (no commit)."""
    g = case["grid"]
    nk, nth, nspec = g["nk"], g["nth"], g["nspec"]
    r = np.ascontiguousarray(case["st4_rpar"], np.float32)
    i = np.ascontiguousarray(case["st4_ipar"], np.int32)
    check("ww_st4_init", lib().ww_st4_init(igrid, ptr(r), r.size, ptr(i), i.size))
    mk, mth, mspec = int(i[sc.ST4_I["MK"]]), int(i[sc.ST4_I["MTH"]]), int(i[sc.ST4_I["MSPEC"]])
    nsat = tables["nsat"]
    arrays = [
        (A_SIG, g["sig0"], [0], [nk + 2], False),
        (A_SIG2, g["sig2"], [1], [nspec], False),
        (A_DSIP, g["dsip0"], [0], [nk + 2], False),
        (A_DDEN, g["dden"], [1], [nk], False),
        (A_DDEN2, g["dden2"], [1], [nspec], False),
        (A_ESIN, g["esin"], [1], [nspec + nth], False),
        (A_ECOS, g["ecos"], [1], [nspec + nth], False),
        (A_TH, g["th"], [1], [nth], False),
        (A_SATWEIGHTS, tables["satw"], [1, 1], [nsat, mth], False),
        (A_SATINDICES, tables["sati"], [1, 1], [nsat, mth], True),
        (A_CUMULW, tables["cumulw"], [1, 1], [mspec, mspec], False),
        (A_DCKI, np.zeros((NKD, NKHS), np.float32), [1, 1], [NKHS, NKD], False),
        (A_QBI, np.zeros((NKD, NKHS), np.float32), [1, 1], [NKHS, NKD], False),
        (A_IKTAB, np.zeros((NDTAB, mk), np.int32), [1, 1], [mk, NDTAB], True),
        (A_FWTABLE, case.get("fwtable", np.zeros(SIZEFWTABLE + 1, np.float32)), [0], [SIZEFWTABLE + 1],
         False),
    ]
    for aid, data, lb, ex, is_int in arrays:
        if aid in skip:
            continue
        dt = np.int32 if is_int else np.float32
        check(f"ww_st4_set array {aid}", _set(igrid, aid, np.ascontiguousarray(data, dt), lb, ex, is_int))


class Srce:
    """One grid configured for W3SRCE (DIA + ST4 + W3SRCE contexts)."""

    def __init__(self, igrid: int, case: dict, tables: dict = None):
        init_runtime()
        self.igrid = igrid
        self.case = case
        g = case["grid"]
        self.nk, self.nth, self.nspec = g["nk"], g["nth"], g["nspec"]
        self.dia = Snl1(igrid, case["dia_cfg"], case["sig"])
        self.tables = tables if tables is not None else insin4_tables(case)
        register_st4(igrid, case, self.tables)
        check("ww_st4_commit", lib().ww_st4_commit(igrid))
        r = np.ascontiguousarray(case["srce_rpar"], np.float32)
        i = np.ascontiguousarray(case["srce_ipar"], np.int32)
        self._keep = (r, i)
        sig = np.ascontiguousarray(g["sig0"][1:self.nk + 1], np.float32)
        ecos = np.ascontiguousarray(g["ecos"][:self.nspec], np.float32)
        esin = np.ascontiguousarray(g["esin"][:self.nspec], np.float32)
        check("ww_srce_init", lib().ww_srce_init(igrid, ptr(r), r.size, ptr(i), i.size, ptr(sig),
                                                 ptr(np.ascontiguousarray(g["dden"], np.float32)),
                                                 ptr(ecos), ptr(esin),
                                                 ptr(np.ascontiguousarray(g["mapwn"], np.int32))))

    def batch(self, case: dict = None, npts: int = None, flags: dict = None) -> dict:
        """ww_srce_batch on host arrays: returns the INOUT arrays after the call."""
        c = case if case is not None else self.case
        n = c["npts"] if npts is None else npts
        fl = dict(c["flags"]) if flags is None else flags
        out = {k: np.ascontiguousarray(c[k][:n]).copy() for k in
               ("spec", "psc", "whitecap", "taubbl", "tauice", "shave")}
        wn1 = np.ascontiguousarray(c["wn1"][:n], np.float32)
        cg1 = np.ascontiguousarray(c["cg1"][:n], np.float32)
        pin = np.ascontiguousarray(c["pin"][:n], np.float32)
        lib().ww_srce_batch(self.igrid, fl["it"], fl["inflags24"], fl["flogrd57"], fl["flogrd58"],
                            float(fl["dtg"]), n, ptr(out["spec"]), ptr(wn1), ptr(cg1), ptr(pin),
                            ptr(out["psc"]), ptr(out["whitecap"]), ptr(out["taubbl"]), ptr(out["tauice"]),
                            ptr(out["shave"]))
        check("ww_srce_batch", lib().ww_srce_last_error())
        return out

    def point(self, c: dict, ip: int, flags: dict = None) -> dict:
        """ww_srce (one point, scalar arguments) for point ip of case c."""
        fl = dict(c["flags"]) if flags is None else flags
        spec = np.ascontiguousarray(c["spec"][ip]).copy()
        wn1 = np.ascontiguousarray(c["wn1"][ip], np.float32)
        cg1 = np.ascontiguousarray(c["cg1"][ip], np.float32)
        pin = c["pin"][ip]
        psc = [ctypes.c_float(float(v)) for v in c["psc"][ip]]
        wc = np.ascontiguousarray(c["whitecap"][ip]).copy()
        tb = np.ascontiguousarray(c["taubbl"][ip]).copy()
        ti = np.ascontiguousarray(c["tauice"][ip]).copy()
        sh = ctypes.c_int(int(c["shave"][ip]))
        P = sc.PSC.index
        refs = [ctypes.byref(x) for x in psc]
        args = (refs[P("USTAR")], refs[P("USTDIR")], refs[P("TAUWX")], refs[P("TAUWY")], refs[P("DTDYN")],
                refs[P("FCUT")], refs[P("TAUOX")], refs[P("TAUOY")], refs[P("TAUWIX")], refs[P("TAUWIY")],
                refs[P("TAUWNX")], refs[P("TAUWNY")], refs[P("PHIAW")], refs[P("CHARN")], refs[P("TWS")],
                refs[P("PHIOC")], ptr(wc), refs[P("PHIBBL")], ptr(tb), ptr(ti), refs[P("PHICE")],
                refs[P("TAUOCX")], refs[P("TAUOCY")], refs[P("WNMEAN")], refs[P("FPI")], ctypes.byref(sh))
        f = [float(pin[k]) for k in range(3)] + [float(pin[3]), float(pin[4]), float(fl["dtg"]),
                                                 float(pin[5]), float(pin[6])]
        lib().ww_srce(self.igrid, fl["it"], fl["inflags24"], fl["flogrd57"], fl["flogrd58"], ptr(spec),
                      ptr(wn1), ptr(cg1), *f, *args)
        check("ww_srce", lib().ww_srce_last_error())
        return {"spec": spec, "psc": np.array([x.value for x in psc], np.float32), "whitecap": wc,
                "taubbl": tb, "tauice": ti, "shave": sh.value}

    def batch_dev(self, t: dict, flags: dict, stream=None) -> int:
        """ww3gpu_srce_batch_dev on device tensors: t holds spec, wn1, cg1, pin, psc, whitecap,
        taubbl, tauice, shave (point-major, the C ABI's shapes).  Enqueued on `stream`."""
        import torch
        st = stream if stream is not None else torch.cuda.current_stream()
        n = t["spec"].shape[0]
        rc = lib().ww3gpu_srce_batch_dev(self.igrid, flags["it"], flags["inflags24"], flags["flogrd57"],
                                         flags["flogrd58"], float(flags["dtg"]), n, ptr(t["spec"]),
                                         ptr(t["wn1"]), ptr(t["cg1"]), ptr(t["pin"]), ptr(t["psc"]),
                                         ptr(t["whitecap"]), ptr(t["taubbl"]), ptr(t["tauice"]),
                                         ptr(t["shave"]), st.cuda_stream)
        check("ww3gpu_srce_batch_dev", rc)
        return lib().ww3gpu_srce_substeps(self.igrid)

"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the CPU-baseline
leg of bench.py, always as the checker.  The GPU library never loads it.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int)
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.or_snl1_dims.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, _i32p]
        L.or_snl1_dims.restype = None
        L.or_insnl1.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                ctypes.c_float, _f32p, ctypes.c_int, _f32p, _i32p, _i32p, _f32p]
        L.or_insnl1.restype = ctypes.c_int
        L.or_w3snl1_batch.argtypes = ([ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 9 +
                                      [_f32p, _i32p, _i32p, _f32p, _f32p, ctypes.c_int,
                                       _f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_int])
        L.or_w3snl1_batch.restype = ctypes.c_int
        L.or_expf_many.argtypes = [_f32p, _f32p, ctypes.c_long]
        L.or_expf_many.restype = None
        _lib = L
    return _lib


def fp(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_f32p)


def ip(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_i32p)


@dataclass
class Snl1Tables:
    """INSNL1's output in the Fortran convention (1-based addresses)."""
    dims: np.ndarray   # NFR, NFRHGH, NFRCHG, NSPECX, NSPECY, IFRP, IFRP1, IFRM, IFRM1
    scal: np.ndarray   # DAL1..3, AWG1..8, SWG1..8
    ipim: np.ndarray   # (16, NSPECX) IP11..IP14, IM11..IM14, IP21..IP24, IM21..IM24
    ic: np.ndarray     # (16, NSPEC)  IC11..IC81, IC12..IC82
    af11: np.ndarray   # (NSPECX,)


def snl1_dims(nk, nth, xfr, lam) -> np.ndarray:
    d = np.zeros(9, np.int32)
    lib().or_snl1_dims(nk, nth, float(xfr), float(lam), ip(d))
    return d


def insnl1(nk, nth, xfr, dth, lam, sig, chain=0) -> Snl1Tables:
    """chain 0: AF11 as x^5*x^6 (-O3 WW3); 1: x^3*x^8 (__powisf2, the fixture's build)."""
    sig = np.ascontiguousarray(sig, np.float32)
    dims = snl1_dims(nk, nth, xfr, lam)
    nspecx = int(dims[3])
    scal = np.zeros(19, np.float32)
    ipim = np.zeros((16, nspecx), np.int32)
    ic = np.zeros((16, nk * nth), np.int32)
    af11 = np.zeros(nspecx, np.float32)
    rc = lib().or_insnl1(nk, nth, float(xfr), float(dth), float(lam), fp(sig), chain, fp(scal), ip(ipim),
                         ip(ic), fp(af11))
    if rc != 0:
        raise ValueError("or_insnl1 failed")
    return Snl1Tables(dims, scal, ipim, ic, af11)


def w3snl1(cfg: dict, tables: Snl1Tables, sig, a, cg, kdmean, nthreads=1):
    """W3SNL1 over points. a: (npts, NSPEC) float32 (a point's spectrum contiguous, i.e.
    Fortran (NSPEC, npts)); cg: (npts, NK); kdmean: (npts,). Returns (s, d)."""
    a = np.ascontiguousarray(a, np.float32)
    cg = np.ascontiguousarray(cg, np.float32)
    kdmean = np.ascontiguousarray(kdmean, np.float32)
    sig = np.ascontiguousarray(sig, np.float32)
    npts = a.shape[0]
    s = np.empty_like(a)
    d = np.empty_like(a)
    c = cfg
    f = [float(c[k]) for k in ("xfr", "lam", "snlc1", "kdcon", "kdmn", "snls1", "snls2", "snls3",
                               "fachfe")]
    rc = lib().or_w3snl1_batch(int(c["nk"]), int(c["nth"]), *f,
                               fp(tables.scal), ip(tables.ipim), ip(tables.ic), fp(tables.af11),
                               fp(sig), npts, fp(a), fp(cg), fp(kdmean), fp(s), fp(d), nthreads)
    if rc != 0:
        raise RuntimeError("or_w3snl1_batch failed")
    return s, d


def expf(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    lib().or_expf_many(fp(x), fp(y), x.size)
    return y


# ----------------------------------------------------------------------------------------
# ST4 / W3SRCE
# ----------------------------------------------------------------------------------------
def _vp(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _declare_srce(L):
    if hasattr(L, "_srce_declared"):
        return
    L.or_insin4_tables.argtypes = [ctypes.c_void_p] * 9
    L.or_insin4_tables.restype = None
    L.or_srce_batch.argtypes = ([ctypes.c_void_p] * 12 + [ctypes.c_int] + [ctypes.c_void_p] * 10 +
                                [ctypes.c_int] * 4 + [ctypes.c_float, ctypes.c_int] +
                                [ctypes.c_void_p] * 9 + [ctypes.c_int])
    L.or_srce_batch.restype = ctypes.c_int
    L._srce_declared = True


def insin4_tables(case: dict):
    """SATWEIGHTS (nsat, NTH), SATINDICES (nsat, NTH), CUMULW (NSPEC, NSPEC), Fortran order,
    returned as C-contiguous arrays of the transposed shapes (NTH, nsat) and (NSPEC, NSPEC)."""
    L = lib()
    _declare_srce(L)
    g = case["grid"]
    nth, nspec = g["nth"], g["nspec"]
    nsat = 2 * (nth // 2 - 1) + 1
    satw = np.zeros((nth, nsat), np.float32)
    sati = np.zeros((nth, nsat), np.int32)
    cumulw = np.zeros((nspec, nspec), np.float32)
    L.or_insin4_tables(_vp(case["st4_rpar"]), _vp(case["st4_ipar"]), _vp(g["sig0"]), _vp(g["dsip0"]),
                       _vp(g["ecos"]), _vp(g["th"]), _vp(satw), _vp(sati), _vp(cumulw))
    return {"satw": satw, "sati": sati, "cumulw": cumulw, "nsat": nsat}


def dia_tables_for(case: dict):
    c = case["dia_cfg"]
    return insnl1(case["nk"], case["nth"], c["xfr"], c["dth"], c["lam"], case["sig"], chain=0)


def srce_batch(case: dict, tables: dict, dia: Snl1Tables = None, nthreads: int = 1,
               npts: int = None):
    """W3SRCE over the case's points (or the first npts).  Returns a dict of the INOUT arrays
    after the call; the case is not modified."""
    L = lib()
    _declare_srce(L)
    g = case["grid"]
    n = case["npts"] if npts is None else npts
    if dia is None:
        dia = dia_tables_for(case)
    c = case["dia_cfg"]
    cfg9 = np.array([c[k] for k in ("xfr", "lam", "snlc1", "kdcon", "kdmn", "snls1", "snls2", "snls3",
                                    "fachfe")], np.float32)
    out = {k: np.ascontiguousarray(case[k][:n]).copy() for k in
           ("spec", "psc", "whitecap", "taubbl", "tauice", "shave")}
    wn1 = np.ascontiguousarray(case["wn1"][:n])
    cg1 = np.ascontiguousarray(case["cg1"][:n])
    pin = np.ascontiguousarray(case["pin"][:n])
    fl = case["flags"]
    rc = L.or_srce_batch(_vp(case["st4_rpar"]), _vp(case["st4_ipar"]), _vp(g["sig0"]), _vp(g["sig2"]),
                         _vp(g["dsip0"]), _vp(g["dden"]), _vp(g["dden2"]), _vp(g["esin"]), _vp(g["ecos"]),
                         _vp(g["th"]), _vp(tables["satw"]), _vp(tables["sati"]), tables["nsat"],
                         _vp(tables["cumulw"]), None, _vp(case["srce_rpar"]), _vp(case["srce_ipar"]),
                         _vp(g["mapwn"]), _vp(cfg9), _vp(dia.scal), _vp(dia.ipim), _vp(dia.ic),
                         _vp(dia.af11), fl["it"], fl["inflags24"], fl["flogrd57"], fl["flogrd58"],
                         fl["dtg"], n, _vp(out["spec"]), _vp(wn1), _vp(cg1), _vp(pin), _vp(out["psc"]),
                         _vp(out["whitecap"]), _vp(out["taubbl"]), _vp(out["tauice"]), _vp(out["shave"]),
                         nthreads)
    if rc != 0:
        raise RuntimeError("or_srce_batch: unsupported configuration or allocation failure")
    return out


# ----------------------------------------------------------------------------------------
# PR3/UQ propagation: W3XYP3 + W3QCK3, W3KTP3 + W3QCK1/W3QCK2
# ----------------------------------------------------------------------------------------
_XYP3_INTS = {"dtg": ctypes.c_float, "flbpi": ctypes.c_int, "nbi": ctypes.c_int, "ncol": ctypes.c_int}
_KTP3_INTS = {"itstep": ctypes.c_int, "ntimes": ctypes.c_int, "facth": ctypes.c_float,
              "fack": ctypes.c_float, "nsea": ctypes.c_int, "npts": ctypes.c_int, "ncol": ctypes.c_int}


def _declare_prop(L, xyp3_args, ktp3_args):
    if hasattr(L, "_prop_declared"):
        return
    L.or_xyp3_batch.argtypes = [_XYP3_INTS.get(k, ctypes.c_void_p) for k in xyp3_args] + [ctypes.c_int]
    L.or_xyp3_batch.restype = ctypes.c_int
    L.or_ktp3_batch.argtypes = [_KTP3_INTS.get(k, ctypes.c_void_p) for k in ktp3_args] + [ctypes.c_int]
    L.or_ktp3_batch.restype = ctypes.c_int
    L.or_xyp3_ntloc.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_float] + [ctypes.c_void_p] * 6
    L.or_xyp3_ntloc.restype = None
    L.or_prop_const.argtypes = [ctypes.c_int]
    L.or_prop_const.restype = ctypes.c_float
    L._prop_declared = True


def _argv(case, names, scalars):
    out = []
    for k in names:
        v = case[k]
        out.append(scalars[k](v) if k in scalars else _vp(v))
    return out


def xyp3_batch(case: dict, nthreads: int = 1):
    """ww_xyp3_batch on the CPU; returns VA after the call (the case is not modified)."""
    from ww3gpu import prop_case as pc
    L = lib()
    _declare_prop(L, pc.XYP3_ARGS, pc.KTP3_ARGS)
    c = dict(case)
    c["va"] = case["va"].copy()
    L.or_xyp3_batch(*_argv(c, pc.XYP3_ARGS, _XYP3_INTS), nthreads)
    return c["va"]


def xyp3_ntloc(case: dict) -> np.ndarray:
    from ww3gpu import prop_case as pc
    L = lib()
    _declare_prop(L, pc.XYP3_ARGS, pc.KTP3_ARGS)
    out = np.zeros(case["nspec"], np.int32)
    L.or_xyp3_ntloc(_vp(case["rpar"]), _vp(case["ipar"]), case["dtg"], _vp(case["sig"]),
                    _vp(case["ecos"]), _vp(case["esin"]), _vp(case["cx"]), _vp(case["cy"]), _vp(out))
    return out


def prop_const(which: int) -> float:
    from ww3gpu import prop_case as pc
    L = lib()
    _declare_prop(L, pc.XYP3_ARGS, pc.KTP3_ARGS)
    return L.or_prop_const(which)


def ktp3_batch(case: dict, nthreads: int = 1):
    """ww_ktp3_batch on the CPU; returns (VA, CFLTHMAX, CFLKMAX) after the call."""
    from ww3gpu import prop_case as pc
    L = lib()
    _declare_prop(L, pc.XYP3_ARGS, pc.KTP3_ARGS)
    c = dict(case)
    for k in ("va", "cflthmax", "cflkmax"):
        c[k] = case[k].copy()
    L.or_ktp3_batch(*_argv(c, pc.KTP3_ARGS, _KTP3_INTS), nthreads)
    return c["va"], c["cflthmax"], c["cflkmax"]


def qck3(mx, my, nx, ny, cfll, trans, q, close, inc, mapact, nact, mapbou, nb0, nb1, nb2):
    """W3QCK3 on one field; cfll and q are updated in place."""
    L = lib()
    L.or_qck3.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 3 + [ctypes.c_int] * 2 + \
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p] + [ctypes.c_int] * 3 + [ctypes.c_void_p]
    L.or_qck3.restype = None
    fla = np.zeros(my * (mx + 1), np.float32)
    L.or_qck3(mx, my, nx, ny, _vp(cfll), _vp(trans), _vp(q), int(close), inc, _vp(mapact), nact,
              _vp(mapbou), nb0, nb1, nb2, _vp(fla))


def qck1(mx, my, nx, ny, cfll, q, close, inc, mapact, nact, mapbou, nb0, nb1, nb2):
    """W3QCK1 on one field; cfll and q are updated in place."""
    L = lib()
    L.or_qck1.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 2 + [ctypes.c_int] * 2 + \
        [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p] + [ctypes.c_int] * 3 + [ctypes.c_void_p]
    L.or_qck1.restype = None
    fla = np.zeros(my * (mx + 1), np.float32)
    L.or_qck1(mx, my, nx, ny, _vp(cfll), _vp(q), int(close), inc, _vp(mapact), nact,
              _vp(mapbou), nb0, nb1, nb2, _vp(fla))


def _st4_table_args(case: dict, tables: dict):
    g = case["grid"]
    return [_vp(case["st4_rpar"]), _vp(case["st4_ipar"]), _vp(g["sig0"]), _vp(g["sig2"]), _vp(g["dsip0"]),
            _vp(g["dden"]), _vp(g["dden2"]), _vp(g["esin"]), _vp(g["ecos"]), _vp(g["th"]),
            _vp(tables["satw"]), _vp(tables["sati"]), tables["nsat"], _vp(tables["cumulw"]), None]


def st4_routine(which: str, case: dict, tables: dict, npts: int, arrays: list, ints=()):
    """or_spr4_batch / or_sin4_batch / or_sds4_batch: `arrays` in the C entry's order after
    npts (numpy, updated in place), `ints` the trailing int arguments (W3SDS4's flags)."""
    L = lib()
    fn = getattr(L, f"or_{which}_batch")
    n_tab = 15
    argt = [ctypes.c_void_p] * 12 + [ctypes.c_int] + [ctypes.c_void_p] * 2 + [ctypes.c_int]
    if which == "sds4":
        argt += [ctypes.c_void_p] * 8 + [ctypes.c_int] * 2 + [ctypes.c_void_p] * 4
        args = arrays[:8] + list(ints) + arrays[8:]
    else:
        argt += [ctypes.c_void_p] * len(arrays)
        args = arrays
    fn.argtypes = argt
    fn.restype = ctypes.c_int
    vals = [a if isinstance(a, int) else _vp(a) for a in args]
    rc = fn(*_st4_table_args(case, tables), npts, *vals)
    assert n_tab == 15
    if rc != 0:
        raise RuntimeError(f"or_{which}_batch refused the configuration")

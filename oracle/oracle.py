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

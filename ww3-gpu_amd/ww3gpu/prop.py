"""Host-side calls of the propagation entries (ww_xyp3_batch, ww_ktp3_batch) on the cases of
prop_case: the reference's W3WAVE loops (prop3_shim.cpp:936-1040) as one call each.  The GPU
library does the work; this module only passes the arrays."""
from __future__ import annotations

import numpy as np

from . import capi
from .prop_case import KTP3_ARGS, XYP3_ARGS

_XYP3_SCALARS = {"dtg", "flbpi", "nbi", "ncol"}
_KTP3_SCALARS = {"itstep", "ntimes", "facth", "fack", "nsea", "npts", "ncol"}
_LENS = {"rpar": None, "ipar": None, "imaps": None, "tscal": None}


def _args(case, names, scalars, packets):
    out = []
    for k in names:
        v = case[k]
        if k in scalars:
            out.append(v)
        else:
            out.append(capi.ptr(v) if v is not None else None)
            if k in packets:
                out.append(int(np.asarray(v).size))
    return out


def xyp3_batch(case: dict, va=None) -> np.ndarray:
    """ww_xyp3_batch on the case; returns VA after the call (the case's VA is not modified)."""
    c = dict(case)
    c["va"] = case["va"].copy() if va is None else va
    rc = capi.lib().ww_xyp3_batch(*_args(c, XYP3_ARGS, _XYP3_SCALARS, ("rpar", "ipar", "imaps", "tscal")))
    capi.check("ww_xyp3_batch", rc)
    return c["va"]


def xyp3_batch_dev(case: dict, d_va, stream=None) -> None:
    """ww3gpu_xyp3_batch_dev: VA is a device tensor, updated in place."""
    c = dict(case)
    c["va"] = d_va
    args = _args(c, XYP3_ARGS, _XYP3_SCALARS, ("rpar", "ipar", "imaps", "tscal"))
    rc = capi.lib().ww3gpu_xyp3_batch_dev(*args, None if stream is None else stream.cuda_stream)
    capi.check("ww3gpu_xyp3_batch_dev", rc)


def ktp3_batch(case: dict):
    """ww_ktp3_batch on the case; returns (VA, CFLTHMAX, CFLKMAX) after the call."""
    c = dict(case)
    for k in ("va", "cflthmax", "cflkmax"):
        c[k] = case[k].copy()
    rc = capi.lib().ww_ktp3_batch(*_args(c, KTP3_ARGS, _KTP3_SCALARS, ("rpar", "ipar")))
    capi.check("ww_ktp3_batch", rc)
    return c["va"], c["cflthmax"], c["cflkmax"]


def ktp3_batch_dev(case: dict, d_va, d_cth, d_ck, stream=None) -> None:
    c = dict(case)
    c["va"], c["cflthmax"], c["cflkmax"] = d_va, d_cth, d_ck
    args = _args(c, KTP3_ARGS, _KTP3_SCALARS, ("rpar", "ipar"))
    rc = capi.lib().ww3gpu_ktp3_batch_dev(*args, None if stream is None else stream.cuda_stream)
    capi.check("ww3gpu_ktp3_batch_dev", rc)


def stats() -> dict:
    out = np.zeros(4, np.float64)
    capi.check("ww3gpu_prop3_stats", capi.lib().ww3gpu_prop3_stats(capi.ptr(out), 4))
    return {"xyp3_ms": out[0], "xyp3_cell_substeps": int(out[1]), "xyp3_substeps": int(out[2]),
            "ktp3_ms": out[3]}

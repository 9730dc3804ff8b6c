"""W3SNL1 (DIA) through the C ABI: the reference's ww_snl1_init / ww_snl1 calls
(kokkos/src/fortran_iface/ww_kokkos_c.hpp:98-133), plus the device-resident entry."""
from __future__ import annotations

import numpy as np

from .capi import Status, WW3GpuError, check, lib, ptr


def init_runtime() -> None:
    check("ww_kokkos_init", lib().ww_kokkos_init(-1))


def insnl1_host(nk, nth, xfr, dth, lam, sig):
    """The library's INSNL1 on the host (no device): dict of Fortran-convention tables."""
    sig = np.ascontiguousarray(sig, np.float32)
    dims = np.zeros(9, np.int32)
    scal = np.zeros(19, np.float32)
    # extents first, from a dry run with only dims requested
    xfr, dth, lam = float(xfr), float(dth), float(lam)
    check("ww3gpu_insnl1", lib().ww3gpu_insnl1(nk, nth, xfr, dth, lam, ptr(sig), ptr(dims),
                                               None, None, None, None, None))
    nspecx = int(dims[3])
    ipim = np.zeros((16, nspecx), np.int32)
    ic = np.zeros((16, nk * nth), np.int32)
    af11 = np.zeros(nspecx, np.float32)
    structured = np.zeros(1, np.int32)
    check("ww3gpu_insnl1", lib().ww3gpu_insnl1(nk, nth, xfr, dth, lam, ptr(sig), ptr(dims), ptr(scal),
                                               ptr(ipim), ptr(ic), ptr(af11), ptr(structured)))
    return {"dims": dims, "scal": scal, "ipim": ipim, "ic": ic, "af11": af11,
            "structured": bool(structured[0])}


class Snl1:
    """One configured grid of the DIA."""

    def __init__(self, igrid: int, cfg: dict, sig: np.ndarray):
        init_runtime()
        self.igrid = igrid
        self.cfg = cfg
        self.nk, self.nth = cfg["nk"], cfg["nth"]
        self.nspec = self.nk * self.nth
        self.sig = np.ascontiguousarray(sig, np.float32)
        c = cfg
        f = [float(c[k]) for k in ("xfr", "dth", "lam", "snlc1", "kdcon", "kdmn", "snls1", "snls2",
                                   "snls3", "fachfe")]
        rc = lib().ww_snl1_init(igrid, int(c["nk"]), int(c["nth"]), *f, ptr(self.sig))
        check("ww_snl1_init", rc)

    def host(self, a: np.ndarray, cg: np.ndarray, kdmean: np.ndarray):
        """ww_snl1 on host arrays; a (npts, NSPEC), cg (npts, NK), kdmean (npts,)."""
        a = np.ascontiguousarray(a, np.float32)
        cg = np.ascontiguousarray(cg, np.float32)
        kdmean = np.ascontiguousarray(kdmean, np.float32)
        npts = a.shape[0]
        s = np.full_like(a, np.nan)
        d = np.full_like(a, np.nan)
        lib().ww_snl1(self.igrid, npts, ptr(a), ptr(cg), ptr(kdmean), ptr(s), ptr(d))
        check("ww_snl1", lib().ww_snl1_last_error())
        return s, d

    def device(self, a, cg, kdmean, s, d, stream=None) -> None:
        """ww3gpu_snl1_dev on device tensors (torch), enqueued on `stream` (a torch stream,
        default: torch's current stream)."""
        import torch
        npts = a.shape[0]
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = lib().ww3gpu_snl1_dev(self.igrid, npts, ptr(a), ptr(cg), ptr(kdmean), ptr(s), ptr(d),
                                   st.cuda_stream)
        check("ww3gpu_snl1_dev", rc)

    def tables(self):
        dims = np.zeros(9, np.int32)
        check("ww3gpu_snl1_tables", lib().ww3gpu_snl1_tables(self.igrid, ptr(dims), None, None, None, None))
        nspecx = int(dims[3])
        scal = np.zeros(19, np.float32)
        ipim = np.zeros((16, nspecx), np.int32)
        ic = np.zeros((16, self.nspec), np.int32)
        af11 = np.zeros(nspecx, np.float32)
        check("ww3gpu_snl1_tables", lib().ww3gpu_snl1_tables(self.igrid, ptr(dims), ptr(scal), ptr(ipim),
                                                             ptr(ic), ptr(af11)))
        return {"dims": dims, "scal": scal, "ipim": ipim, "ic": ic, "af11": af11}

    def set_af11(self, af11: np.ndarray) -> None:
        af11 = np.ascontiguousarray(af11, np.float32)
        check("ww3gpu_snl1_set_af11", lib().ww3gpu_snl1_set_af11(self.igrid, ptr(af11), af11.size))

    def set_variant(self, variant: int) -> None:
        """0: stencil-specialised kernel when available (default); 1: general kernel."""
        check("ww3gpu_snl1_set_variant", lib().ww3gpu_snl1_set_variant(self.igrid, variant))

    def lds_bytes(self) -> int:
        return int(lib().ww3gpu_snl1_lds_bytes(self.igrid))


__all__ = ["Snl1", "init_runtime", "insnl1_host", "Status", "WW3GpuError"]

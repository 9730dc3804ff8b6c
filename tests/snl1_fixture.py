"""Reader for tests/golden/snl1_nk25_nth24.bin, the reference's frozen W3SNL1/INSNL1
answers (layout: kokkos/tests/fixtures/README.md:40-68 of the reference; reader there:
kokkos/src/ww_kokkos/fixture_io.cpp:63-129).  Little-endian int32/float32 stream."""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SNL1_FIXTURE = os.path.join(GOLDEN, "snl1_nk25_nth24.bin")
MAGIC = 0x534E4C31  # 'SNL1'


@dataclass
class Snl1Fixture:
    nk: int
    nth: int
    npts: int
    cfg: dict
    sig: np.ndarray
    dims5: np.ndarray          # NFR, NFRHGH, NFRCHG, NSPECX, NSPECY
    scal: np.ndarray           # DAL1..3, AWG1..8, SWG1..8
    ipim: np.ndarray           # (16, NSPECX), 1-based
    ic: np.ndarray             # (16, NSPEC), 1-based
    af11: np.ndarray
    kdmean: np.ndarray = field(default=None)   # (npts,)
    cg: np.ndarray = field(default=None)       # (npts, NK)
    a: np.ndarray = field(default=None)        # (npts, NSPEC)
    s: np.ndarray = field(default=None)
    d: np.ndarray = field(default=None)


def load(path: str = SNL1_FIXTURE) -> Snl1Fixture:
    raw = open(path, "rb").read()
    off = 0

    def ints(n):
        nonlocal off
        v = np.frombuffer(raw, "<i4", n, off).astype(np.int32)
        off += 4 * n
        return v

    def flts(n):
        nonlocal off
        v = np.frombuffer(raw, "<f4", n, off).astype(np.float32)
        off += 4 * n
        return v

    magic, nk, nth, npts = (int(v) for v in ints(4))
    if magic != MAGIC:
        raise ValueError("not an SNL1 fixture")
    names = ["xfr", "dth", "lam", "snlc1", "kdcon", "kdmn", "snls1", "snls2", "snls3", "fachfe"]
    vals = flts(10)
    cfg = {"nk": nk, "nth": nth}
    cfg.update({k: np.float32(v) for k, v in zip(names, vals)})
    sig = flts(nk)
    dims5 = ints(5)
    nspecx = int(dims5[3])
    nspec = nk * nth
    scal = np.concatenate([flts(3), flts(8), flts(8)])
    ipim = ints(16 * nspecx).reshape(16, nspecx)
    ic = ints(16 * nspec).reshape(16, nspec)
    af11 = flts(nspecx)
    kd, cg, a, s, d = [], [], [], [], []
    for _ in range(npts):
        kd.append(flts(1)[0])
        cg.append(flts(nk))
        a.append(flts(nspec))
        s.append(flts(nspec))
        d.append(flts(nspec))
    if off != len(raw):
        raise ValueError(f"fixture length mismatch: read {off} of {len(raw)} bytes")
    f = Snl1Fixture(nk, nth, npts, cfg, sig, dims5, scal, ipim, ic, af11)
    f.kdmean = np.array(kd, np.float32)
    f.cg = np.stack(cg)
    f.a = np.stack(a)
    f.s = np.stack(s)
    f.d = np.stack(d)
    return f

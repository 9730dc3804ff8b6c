"""The per-routine ST4 entries ww_spr4, ww_sin4, ww_sds4 on the GPU, bit for bit against the
oracle's W3SPR4, W3SIN4, W3SDS4 on the same points: NTH 24 and 36, STAB3, the isotropic
cumulative term, the whitecap outputs on and off; an uncommitted grid and a null array are
refused with nothing written."""
import numpy as np
import pytest

import oracle as O
from ww3gpu import capi, srce_case as sc
from ww3gpu.srce import insin4_tables, register_st4

pytestmark = pytest.mark.gpu

_grid = [700]


def _new_grid():
    _grid[0] += 1
    return _grid[0]


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _same(got, ref, label):
    g, r = _bits(got), _bits(ref)
    if not np.array_equal(g, r):
        bad = np.argwhere(g != r)
        raise AssertionError(f"{label}: {len(bad)} differ, first {tuple(bad[0])}: "
                             f"gpu {np.asarray(got)[tuple(bad[0])]!r} oracle {np.asarray(ref)[tuple(bad[0])]!r}")


@pytest.fixture(scope="module")
def rt(gpu_runtime):
    return gpu_runtime


def _case(nk, nth, npts, seed, st4=None, st4i=None):
    c = sc.make_case(npts, nk=nk, nth=nth, seed=seed)
    for k, v in (st4 or {}).items():
        c["st4_rpar"][sc.ST4_R[k[0]] + k[1] if isinstance(k, tuple) else sc.ST4_R[k]] = v
    for k, v in (st4i or {}).items():
        c["st4_ipar"][sc.ST4_I[k]] = v
    tab = insin4_tables(c)
    g = _new_grid()
    register_st4(g, c, tab)
    assert capi.lib().ww_st4_commit(g) == capi.Status.OK
    return c, tab, g


def _inputs(c, seed):
    rng = np.random.default_rng(seed)
    n, nk, nth = c["npts"], c["nk"], c["nth"]
    nspec = nk * nth
    f = lambda lo, hi: rng.uniform(lo, hi, n).astype(np.float32)
    pin = c["pin"]
    d = {
        "a": np.ascontiguousarray(c["spec"]), "k": np.ascontiguousarray(c["wn1"]),
        "cg": np.ascontiguousarray(c["cg1"]),
        "kbin": np.ascontiguousarray(np.repeat(c["wn1"], nth, axis=1)).astype(np.float32),
        "u": np.ascontiguousarray(pin[:, sc.PIN.index("U10ABS")]),
        "udir": np.ascontiguousarray(pin[:, sc.PIN.index("U10DIR")]),
        "ustar": f(0.05, 1.2), "usdir": f(0, 6.2), "drat": np.full(n, 1.225e-3, np.float32),
        "as": f(-2, 2), "z0": f(1e-5, 3e-3), "cd": f(5e-4, 3e-3), "depth": f(5, 3000),
        "dair": np.full(n, 1.225, np.float32), "dlwmean": f(0, 6), "tauwx": f(-0.05, 0.05),
        "tauwy": f(-0.05, 0.05), "llws": (rng.random((n, nspec)) < 0.5).astype(np.int32),
    }
    return d


def _run_spr4(c, tab, g, d):
    n = c["npts"]
    outs = {k: np.zeros(n, np.float32) for k in ("emean", "fmean", "fmean1", "wnmean", "amax", "cd", "z0",
                                                  "charn", "fmeanws", "dlwmean")}
    res = []
    for side in ("gpu", "cpu"):
        ustar, usdir = d["ustar"].copy(), d["usdir"].copy()
        o = {k: v.copy() for k, v in outs.items()}
        arrs = [d["a"], d["cg"], d["k"], d["u"], d["udir"], d["tauwx"], d["tauwy"], d["llws"], ustar, usdir,
                o["emean"], o["fmean"], o["fmean1"], o["wnmean"], o["amax"], o["cd"], o["z0"], o["charn"],
                o["fmeanws"], o["dlwmean"]]
        if side == "gpu":
            capi.lib().ww_spr4(g, n, *[capi.ptr(x) for x in arrs])
            assert capi.lib().ww_st4_last_error() == capi.Status.OK
        else:
            O.st4_routine("spr4", c, tab, n, arrs)
        res.append([ustar, usdir] + [o[k] for k in sorted(o)])
    return res


def _run_sin4(c, tab, g, d):
    n, nspec = c["npts"], c["nk"] * c["nth"]
    res = []
    for side in ("gpu", "cpu"):
        tx, ty, tnx, tny = (np.zeros(n, np.float32) for _ in range(4))
        s, dd = np.zeros((n, nspec), np.float32), np.zeros((n, nspec), np.float32)
        llws = np.zeros((n, nspec), np.int32)
        arrs = [d["a"], d["cg"], d["kbin"], d["u"], d["ustar"], d["drat"], d["as"], d["usdir"], d["z0"], d["cd"],
                tx, ty, tnx, tny, s, dd, llws]
        if side == "gpu":
            capi.lib().ww_sin4(g, n, *[capi.ptr(x) for x in arrs])
            assert capi.lib().ww_st4_last_error() == capi.Status.OK
        else:
            O.st4_routine("sin4", c, tab, n, arrs)
        res.append([tx, ty, tnx, tny, s, dd, llws])
    return res


def _run_sds4(c, tab, g, d, fl57=1, fl58=1):
    n, nspec = c["npts"], c["nk"] * c["nth"]
    res = []
    for side in ("gpu", "cpu"):
        srhs, ddiag, brl = (np.zeros((n, nspec), np.float32) for _ in range(3))
        wc = np.full((n, 4), 7.0, np.float32)
        arrs = [d["a"], d["k"], d["cg"], d["ustar"], d["usdir"], d["depth"], d["dair"], d["dlwmean"],
                srhs, ddiag, brl, wc]
        if side == "gpu":
            ptrs = [capi.ptr(x) for x in arrs]
            capi.lib().ww_sds4(g, n, *ptrs[:8], fl57, fl58, *ptrs[8:])
            assert capi.lib().ww_st4_last_error() == capi.Status.OK
        else:
            O.st4_routine("sds4", c, tab, n, arrs, ints=(fl57, fl58))
        res.append([srhs, ddiag, brl, wc])
    return res


CASES = {
    "nk32_nth36": dict(nk=32, nth=36, npts=37),
    "nk25_nth24": dict(nk=25, nth=24, npts=21),
    "nk20_nth12": dict(nk=20, nth=12, npts=9),
    "stab3": dict(nk=25, nth=24, npts=16, st4i={"STAB3": 1}),
    "cumulative_isotropic": dict(nk=25, nth=24, npts=12, st4={("SSDSC", 2): 0.5}),
    "tail_level": dict(nk=32, nth=36, npts=8, st4={("SINTAILPAR", 2): 0.04}),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_routines(rt, name):
    kw = dict(CASES[name])
    c, tab, g = _case(kw.pop("nk"), kw.pop("nth"), kw.pop("npts"), seed=41, **kw)
    d = _inputs(c, 42)
    for lab, (got, ref) in (("spr4", _run_spr4(c, tab, g, d)), ("sin4", _run_sin4(c, tab, g, d)),
                            ("sds4", _run_sds4(c, tab, g, d))):
        for i, (x, y) in enumerate(zip(got, ref)):
            _same(x, y, f"{name} {lab} output {i}")
    got, ref = _run_sds4(c, tab, g, d, fl57=0, fl58=0)
    for i, (x, y) in enumerate(zip(got, ref)):
        _same(x, y, f"{name} sds4 (no whitecap) output {i}")


def test_routine_refusals(rt):
    L = capi.lib()
    c, tab, g = _case(25, 24, 4, seed=43)
    d = _inputs(c, 44)
    n = c["npts"]
    srhs = np.full((n, 600), 3.0, np.float32)
    L.ww_sds4(99999, n, *[capi.ptr(d[k]) for k in ("a", "k", "cg", "ustar", "usdir", "depth", "dair", "dlwmean")],
              1, 1, capi.ptr(srhs), capi.ptr(srhs), capi.ptr(srhs), capi.ptr(srhs))
    assert L.ww_st4_last_error() == capi.Status.NOT_INITIALISED
    assert np.all(srhs == 3.0)
    L.ww_sds4(g, n, None, *[capi.ptr(d[k]) for k in ("k", "cg", "ustar", "usdir", "depth", "dair", "dlwmean")],
              1, 1, capi.ptr(srhs), capi.ptr(srhs), capi.ptr(srhs), capi.ptr(srhs))
    assert L.ww_st4_last_error() == capi.Status.BAD_SHAPE
    assert np.all(srhs == 3.0)

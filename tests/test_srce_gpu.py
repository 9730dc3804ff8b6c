"""W3SRCE with ST4 on the GPU through the C ABI, bit for bit against the CPU oracle.

Every output of ww_srce_batch (spectrum, the 22 pointer scalars, WHITECAP, TAUBBL, TAUICE,
SHAVE) must equal the oracle's bits.  Cases: the 9 520-point NK=32/NTH=36 step at DTG =
300 s, small cases for each configuration branch (STAB3, IG1, isotropic saturation,
isotropic cumulative term, depth-limited breaking, CAPCHNK, no LN1, FDONLY, no whitecap
output, ice, IT = 0), calm and shallow points, the single-point entry, the device-resident
entry, every group width of the sub-step kernel, and 64 uneven domains in one batch against
each domain alone."""
import numpy as np
import pytest

import oracle as O
from ww3gpu import capi, srce_case as sc
from ww3gpu.srce import Srce, insin4_tables

pytestmark = pytest.mark.gpu

KEYS = ("spec", "psc", "whitecap", "taubbl", "tauice", "shave")
_next_grid = [100]


def _grid():
    _next_grid[0] += 1
    return _next_grid[0]


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _assert_same(got, ref, label):
    for k in KEYS:
        g, r = _bits(got[k]), _bits(ref[k])
        if not np.array_equal(g, r):
            bad = np.argwhere(g != r)
            i = tuple(bad[0])
            raise AssertionError(f"{label}: {k} differs at {len(bad)} places, first {i}: "
                                 f"gpu {np.asarray(got[k])[i]!r} oracle {np.asarray(ref[k])[i]!r}")


def _run(c, nthreads=16):
    print(f"[srce] npts={c['npts']} nk={c['nk']} nth={c['nth']} flags={c['flags']}", flush=True)
    tab = insin4_tables(c)
    s = Srce(_grid(), c, tab)
    got = s.batch()
    ref = O.srce_batch(c, tab, nthreads=nthreads)
    return got, ref, s


@pytest.fixture(scope="module")
def rt(gpu_runtime):
    return gpu_runtime


def test_full_step_9520_points(rt):
    c = sc.make_case(9520, nk=32, nth=36, seed=2024)
    got, ref, s = _run(c)
    _assert_same(got, ref, "9520 points")
    n = capi.lib().ww3gpu_srce_substeps(s.igrid)
    print(f"[srce] sub-steps {n}, SHAVE set on {int(got['shave'].sum())} points", flush=True)
    assert n >= 1
    # the step changed the spectra
    assert not np.array_equal(_bits(got["spec"]), _bits(c["spec"]))


def _variant(nk=25, nth=24, npts=64, seed=7, st4=None, st4i=None, srce=None, srcei=None, flags=None,
             pin=None, **kw):
    c = sc.make_case(npts, nk=nk, nth=nth, seed=seed, **kw)
    for k, v in (st4 or {}).items():
        c["st4_rpar"][sc.ST4_R[k[0]] + k[1] if isinstance(k, tuple) else sc.ST4_R[k]] = v
    for k, v in (st4i or {}).items():
        c["st4_ipar"][sc.ST4_I[k]] = v
    for k, v in (srce or {}).items():
        c["srce_rpar"][sc.SRCE_R[k]] = v
    for k, v in (srcei or {}).items():
        c["srce_ipar"][sc.SRCE_I[k]] = v
    c["flags"].update(flags or {})
    for k, v in (pin or {}).items():
        c["pin"][:, sc.PIN.index(k)] = v
    return c


VARIANTS = {
    "nk32_nth36": dict(nk=32, nth=36),
    "nk25_nth24": dict(),
    "nk20_nth12": dict(nk=20, nth=12),
    "nk36_nth36": dict(nk=36, nth=36),
    "stab3": dict(st4i={"STAB3": 1}, pin={"AS": -1.5}),
    "ig1": dict(st4i={"IG1": 1}, st4={"IGPARS5": 3.0}),
    "sat_isotropic": dict(st4={"SSDSDTH": 180.0}),
    "cumulative_isotropic": dict(st4={("SSDSC", 2): 0.5}),
    "breaking_depth_factor": dict(st4={("SSDSBM", 0): 0.0, ("SSDSBM", 1): 0.9, ("SSDSBM", 2): 0.1,
                                ("SSDSBM", 3): -0.05, ("SSDSBM", 4): 0.01}, depth_range=(3.0, 40.0)),
    "capchnk": dict(st4={("CAPCHNK", 0): 1.0, ("CAPCHNK", 1): 0.0095, ("CAPCHNK", 2): 0.0157,
                         ("CAPCHNK", 3): 20.0, ("CAPCHNK", 4): 5.0}),
    "no_ln1": dict(srcei={"LN1": 0}),
    "fdonly": dict(srcei={"FDONLY": 1}, depth_range=(2.5, 15.0)),
    "no_whitecap": dict(flags={"flogrd57": 0, "flogrd58": 0}),
    "whitecap57": dict(flags={"flogrd57": 1, "flogrd58": 0}),
    "ice": dict(flags={"inflags24": 1}, pin={"ICE": 0.4}),
    "it0": dict(flags={"it": 0}),
    "shallow": dict(depth_range=(2.5, 12.0), seed=11),
    "dtg_900": dict(flags={"dtg": 900.0}, srce={"DTMAX": 900.0}),
    "swell_turb": dict(st4={("SSWELLF", 2): -0.01, ("SSWELLF", 6): 0.0}),
    # SINTAILPAR(3) > 0: the tail's Phillips-level terms (seven per-step arrays instead of three)
    "tail_level": dict(nk=32, nth=36, st4={("SINTAILPAR", 2): 0.04}),
}


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_variant(rt, name):
    c = _variant(**VARIANTS[name])
    got, ref, _ = _run(c)
    _assert_same(got, ref, name)


@pytest.mark.parametrize("lanes", [16, 32, 64])
@pytest.mark.parametrize("nk,nth", [(32, 36), (25, 24), (20, 12)])
def test_group_widths(rt, lanes, nk, nth):
    """Every sub-step at one group width (16, 32 or 64 lanes per point; the compiled NTH = 36,
    24 and the generic path): the same bits as the oracle."""
    L = capi.lib()
    c = _variant(nk=nk, nth=nth, npts=96, seed=31 + lanes)
    assert L.ww3gpu_srce_group_lanes(lanes) == capi.Status.OK
    try:
        got, ref, _ = _run(c)
    finally:
        L.ww3gpu_srce_group_lanes(0)
    _assert_same(got, ref, f"{lanes} lanes, nk {nk} nth {nth}")


def test_uneven_domains_in_one_batch(rt):
    """The batched-grids configuration at test size: 64 regional domains, each with its own
    seeded winds, depths and spectra and a log-uniform size (8 to 300 points), integrated as one
    batch and each alone.  Every domain's outputs are the same bits either way, and the batch
    equals the oracle."""
    rng = np.random.default_rng(77)
    sizes = np.exp(rng.uniform(np.log(8.0), np.log(300.0), 64)).astype(int)
    cases = [sc.make_case(int(n), nk=32, nth=36, seed=1000 + i) for i, n in enumerate(sizes)]
    big = dict(cases[0])
    big["npts"] = int(sizes.sum())
    for k in ("spec", "wn1", "cg1", "pin", "psc", "whitecap", "taubbl", "tauice", "shave"):
        big[k] = np.concatenate([c[k] for c in cases])
    print(f"[srce] {len(cases)} domains, {big['npts']} points in one batch", flush=True)
    tab = insin4_tables(cases[0])
    s = Srce(_grid(), big, tab)
    got = s.batch()
    _assert_same(got, O.srce_batch(big, tab, nthreads=16), "batch of 64 domains")
    off = 0
    for i, c in enumerate(cases):
        one = s.batch(case=c)
        n = c["npts"]
        _assert_same({k: got[k][off:off + n] for k in KEYS}, one, f"domain {i} ({n} points)")
        off += n


def test_calm_and_extreme_points(rt):
    """U10 = 0 (u* = 0: FHIGH = +Inf on the first sub-step), 35 m/s, a zero spectrum, very
    shallow water."""
    c = _variant(nk=32, nth=36, npts=16, seed=3)
    u = c["pin"][:, sc.PIN.index("U10ABS")]
    u[0:4] = 0.0
    u[4:8] = 35.0
    c["spec"][8:10] = 0.0
    c["pin"][10:13, sc.PIN.index("D_INP")] = 1.0
    got, ref, _ = _run(c)
    _assert_same(got, ref, "calm/extreme")


def test_single_point_entry_equals_batch(rt):
    c = _variant(npts=5, seed=9)
    tab = insin4_tables(c)
    s = Srce(_grid(), c, tab)
    b = s.batch()
    for ip in range(c["npts"]):
        p = s.point(c, ip)
        for k in ("spec", "psc", "whitecap", "taubbl", "tauice"):
            assert np.array_equal(_bits(p[k]), _bits(b[k][ip])), (ip, k)
        assert p["shave"] == b["shave"][ip]


def test_device_entry_equals_host_entry(rt):
    torch = rt
    c = _variant(nk=32, nth=36, npts=300, seed=13)
    tab = insin4_tables(c)
    s = Srce(_grid(), c, tab)
    host = s.batch()
    dev = {k: torch.from_numpy(np.ascontiguousarray(c[k])).cuda() for k in
           ("spec", "wn1", "cg1", "pin", "psc", "whitecap", "taubbl", "tauice", "shave")}
    st = torch.cuda.Stream()
    torch.cuda.synchronize()
    n = s.batch_dev(dev, c["flags"], st)
    st.synchronize()
    assert n >= 1
    for k in KEYS:
        assert np.array_equal(_bits(dev[k].cpu().numpy()), _bits(host[k])), k
    # a second call on the outputs continues the integration, as a second time step would
    s.batch_dev(dev, c["flags"], st)
    st.synchronize()
    c2 = dict(c)
    for k in ("spec", "psc", "whitecap", "taubbl", "tauice", "shave"):
        c2[k] = host[k]
    ref2 = O.srce_batch(c2, tab, nthreads=16)
    for k in KEYS:
        assert np.array_equal(_bits(dev[k].cpu().numpy()), _bits(ref2[k])), k


def test_batch_sizes(rt):
    """npts = 1, 63, 64, 65 (the lane padding) and npts = 0 (a no-op)."""
    for n in (1, 63, 64, 65):
        c = _variant(npts=n, seed=17 + n)
        got, ref, _ = _run(c)
        _assert_same(got, ref, f"npts={n}")
    c = _variant(npts=3)
    s = Srce(_grid(), c, insin4_tables(c))
    before = c["spec"].copy()
    out = s.batch(npts=0)
    assert out["spec"].shape[0] == 0 and np.array_equal(c["spec"], before)


def test_errors_write_nothing(rt):
    L = capi.lib()
    c = _variant(npts=4)
    s = Srce(_grid(), c, insin4_tables(c))
    # IT > 0 with TAUWBUG off: refused, nothing written
    c["st4_rpar"][sc.ST4_R["SINTAILPAR"] + 3] = 0.0
    s2 = Srce(_grid(), c, insin4_tables(c))
    spec = c["spec"].copy()
    p = capi.ptr
    L.ww_srce_batch(s2.igrid, 1, 0, 1, 1, 300.0, 4, p(spec), p(c["wn1"]), p(c["cg1"]), p(c["pin"]),
                    p(c["psc"].copy()), p(c["whitecap"].copy()), p(c["taubbl"].copy()),
                    p(c["tauice"].copy()), p(c["shave"].copy()))
    assert L.ww_srce_last_error() == capi.Status.UNSUPPORTED
    assert np.array_equal(spec, c["spec"])
    # null array
    L.ww_srce_batch(s.igrid, 1, 0, 1, 1, 300.0, 4, None, p(c["wn1"]), p(c["cg1"]), p(c["pin"]),
                    p(c["psc"]), p(c["whitecap"]), p(c["taubbl"]), p(c["tauice"]), p(c["shave"]))
    assert L.ww_srce_last_error() == capi.Status.BAD_SHAPE
    # an unknown grid
    L.ww_srce_batch(99999, 1, 0, 1, 1, 300.0, 4, p(spec), p(c["wn1"]), p(c["cg1"]), p(c["pin"]),
                    p(c["psc"]), p(c["whitecap"]), p(c["taubbl"]), p(c["tauice"]), p(c["shave"]))
    assert L.ww_srce_last_error() == capi.Status.NOT_INITIALISED


def test_st4_commit_checks(rt):
    """The commit refuses a missing array (TABLE), a wrong LBOUND (TABLE at registration),
    and configurations the kernels do not translate (UNSUPPORTED)."""
    from ww3gpu import srce as S
    L = capi.lib()
    c = _variant(npts=2)
    tab = insin4_tables(c)
    g = _grid()
    S.register_st4(g, c, tab, skip=(S.A_CUMULW,))
    assert L.ww_st4_commit(g) == capi.Status.TABLE
    a = np.zeros(10, np.float32)
    lb = np.array([1], np.int32)
    ex = np.array([10], np.int32)
    assert L.ww_st4_set_real(g, S.A_SIG, capi.ptr(a), 1, capi.ptr(lb), capi.ptr(ex)) == capi.Status.TABLE
    # tabulated u* (SINTAILPAR(1) > 0.5): the commit first wants the module's stress tables
    # (TABLE), then refuses the configuration (UNSUPPORTED)
    c["st4_rpar"][sc.ST4_R["SINTAILPAR"]] = 1.0
    g2 = _grid()
    S.register_st4(g2, c, tab)
    assert L.ww_st4_commit(g2) == capi.Status.TABLE
    for aid, ext in ((S.A_TAUT, [201, 201]), (S.A_TAUHFT, [101, 201]), (S.A_TAUHFT2, [101, 201, 51])):
        z = np.zeros(int(np.prod(ext)), np.float32)
        lb0 = np.zeros(len(ext), np.int32)
        e = np.array(ext, np.int32)
        assert L.ww_st4_set_real(g2, aid, capi.ptr(z), len(ext), capi.ptr(lb0), capi.ptr(e)) == capi.Status.OK
    assert L.ww_st4_commit(g2) == capi.Status.UNSUPPORTED
    # SDSBCHOICE other than 1
    c["st4_rpar"][sc.ST4_R["SINTAILPAR"]] = 0.0
    c["st4_rpar"][sc.ST4_R["SSDSC"]] = 2.0
    g3 = _grid()
    S.register_st4(g3, c, tab)
    assert L.ww_st4_commit(g3) == capi.Status.UNSUPPORTED
    # W3SRCE before the ST4 commit of its grid
    c3 = _variant(npts=2)
    r, i = c3["srce_rpar"], c3["srce_ipar"]
    gg = c3["grid"]
    p = capi.ptr
    rc = L.ww_srce_init(_grid(), p(r), r.size, p(i), i.size, p(gg["sig0"][1:]), p(gg["dden"]), p(gg["ecos"]),
                        p(gg["esin"]), p(gg["mapwn"]))
    assert rc == capi.Status.NOT_INITIALISED

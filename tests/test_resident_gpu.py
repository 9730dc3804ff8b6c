"""The resident spectrum (ww_va_*) and the entries that work on it (ww_srce_batch_va,
ww_xyp3_batch_va, ww_ktp3_batch_va) on the GPU.

A VA(NSPEC, 0:NSEALM) is bound per grid; kernels update the device copy and the host copy is
stale until a sync.  Checked: the HOST/DEVICE/BOTH transitions, W3SRCE on scattered columns
equal to the host-array entry (and so to the oracle) with the other columns untouched, the
partial read and write of columns, zeroed columns, the propagation entries on the resident
copy equal to their host-array forms, the refusal of a second binding, and three coupled time
steps (W3XYP3 then W3SRCE on the resident copy) against the oracle's same sequence."""
import numpy as np
import pytest

import oracle as O
from ww3gpu import capi, prop, prop_case as pc, srce_case as sc
from ww3gpu.srce import Srce, insin4_tables

pytestmark = pytest.mark.gpu
HOST, DEVICE, BOTH = 0, 1, 2


@pytest.fixture(scope="module")
def rt(gpu_runtime):
    return gpu_runtime


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_srce_batch_va_and_state(rt):
    L = capi.lib()
    c = sc.make_case(40, nk=32, nth=36, seed=61)
    tab = insin4_tables(c)
    s = Srce(801, c, tab)
    nspec, nsealm, n = 32 * 36, 60, 40
    rng = np.random.default_rng(62)
    jsea = rng.permutation(np.arange(1, nsealm + 1))[:n].astype(np.int32)
    va = np.full((nsealm + 1, nspec), 0.5, np.float32)
    va[jsea] = c["spec"]
    keep = va.copy()
    p = capi.ptr
    outs = {k: c[k].copy() for k in ("psc", "whitecap", "taubbl", "tauice", "shave")}
    fl = c["flags"]
    L.ww_srce_batch_va(s.igrid, fl["it"], fl["inflags24"], fl["flogrd57"], fl["flogrd58"], fl["dtg"], p(va),
                       nspec, nsealm, n, p(jsea), p(c["wn1"]), p(c["cg1"]), p(c["pin"]), p(outs["psc"]),
                       p(outs["whitecap"]), p(outs["taubbl"]), p(outs["tauice"]), p(outs["shave"]))
    assert L.ww_srce_last_error() == capi.Status.OK
    assert L.ww3gpu_va_state(s.igrid) == DEVICE
    assert np.array_equal(va, keep)  # the host copy is stale until a sync
    assert L.ww_va_host_read(s.igrid, p(va), nspec, nsealm) == capi.Status.OK
    assert L.ww3gpu_va_state(s.igrid) == BOTH
    ref = O.srce_batch(c, tab, nthreads=16)
    assert np.array_equal(_bits(va[jsea]), _bits(ref["spec"]))
    others = np.setdiff1d(np.arange(nsealm + 1), jsea)
    assert np.array_equal(va[others], keep[others])
    for k in ("psc", "whitecap", "taubbl", "tauice"):
        assert np.array_equal(_bits(outs[k]), _bits(ref[k])), k
    assert np.array_equal(outs["shave"], ref["shave"])
    # a host write, then a kernel: the host bits go to the device first
    assert L.ww_va_host_write(s.igrid, p(va), nspec, nsealm) == capi.Status.OK
    assert L.ww3gpu_va_state(s.igrid) == HOST
    va[others[:3]] = 9.0
    assert L.ww_va_device(s.igrid, 1) == capi.Status.OK
    assert L.ww3gpu_va_state(s.igrid) == DEVICE
    # zeroed columns, then a partial read: only the listed columns come back
    z = np.array([others[0], jsea[0]], np.int32)
    assert L.ww_va_device_zero_cols(s.igrid, p(va), nspec, nsealm, 2, p(z)) == capi.Status.OK
    before = va.copy()
    cols = np.array([jsea[0], others[1]], np.int32)
    assert L.ww_va_host_read_cols(s.igrid, p(va), nspec, nsealm, 2, p(cols)) == capi.Status.OK
    assert L.ww3gpu_va_state(s.igrid) == DEVICE
    assert np.all(va[jsea[0]] == 0.0) and np.all(va[others[1]] == 9.0)
    assert np.array_equal(va[others[0]], before[others[0]])  # zeroed on the device, not read back
    # host writes of listed columns go to the device copy
    va[others[2]] = 4.0
    w = np.array([others[2]], np.int32)
    assert L.ww_va_host_wrote_cols(s.igrid, p(va), nspec, nsealm, 1, p(w)) == capi.Status.OK
    assert L.ww_va_host_read(s.igrid, p(va), nspec, nsealm) == capi.Status.OK
    assert np.all(va[others[0]] == 0.0) and np.all(va[others[2]] == 4.0)
    # a second binding with another shape is refused
    assert L.ww_va_host_read(s.igrid, p(va), nspec, nsealm - 1) == capi.Status.BAD_SHAPE
    bad = np.array([nsealm + 1], np.int32)
    assert L.ww_va_host_read_cols(s.igrid, p(va), nspec, nsealm, 1, p(bad)) == capi.Status.BAD_SHAPE


def test_propagation_va(rt):
    L = capi.lib()
    c = pc.make_xyp3(nx=40, ny=30, nk=6, nth=8, seed=63, boundary=True)
    nspec, nsea = c["nspec"], c["nsea"]
    va = c["va"].copy()
    args = prop._args(dict(c, va=va), pc.XYP3_ARGS, prop._XYP3_SCALARS, ("rpar", "ipar", "imaps", "tscal"))
    # drop ncol and va (the last two) and pass the resident form's tail
    rc = L.ww_xyp3_batch_va(902, *args[:-2], capi.ptr(va), nspec, nsea)
    assert rc == capi.Status.OK
    assert L.ww_va_host_read(902, capi.ptr(va), nspec, nsea) == capi.Status.OK
    ref = O.xyp3_batch(c, nthreads=16)
    assert np.array_equal(_bits(va), _bits(ref))
    k = pc.make_ktp3(nsea=50, seed=64)
    kva = k["va"].copy()
    cth, ck = np.zeros(50, np.float32), np.zeros(50, np.float32)
    kargs = prop._args(dict(k, va=kva), pc.KTP3_ARGS, prop._KTP3_SCALARS, ("rpar", "ipar"))
    rc = L.ww_ktp3_batch_va(903, *kargs[:-4], capi.ptr(kva), k["nspec"], 50, capi.ptr(cth), capi.ptr(ck))
    assert rc == capi.Status.OK
    assert L.ww_va_host_read(903, capi.ptr(kva), k["nspec"], 50) == capi.Status.OK
    rva, rth, rk = O.ktp3_batch(k, nthreads=16)
    assert np.array_equal(_bits(kva), _bits(rva))
    assert np.array_equal(_bits(cth), _bits(rth)) and np.array_equal(_bits(ck), _bits(rk))


def test_coupled_resident_steps(rt):
    """Three time steps of the resident loop on one grid: W3XYP3 on the resident VA, then W3SRCE
    on every sea column of it (ww_srce_batch_va), with the spectrum never leaving the device
    between the calls, bit for bit against the oracle running the same sequence (the
    propagation plus source-term step at test size)."""
    L = capi.lib()
    p = capi.ptr
    x = pc.make_xyp3(nx=14, ny=12, nk=20, nth=12, seed=71, boundary=True)
    nspec, nsea = x["nspec"], x["nsea"]
    c = sc.make_case(nsea, nk=20, nth=12, seed=72)
    tab = insin4_tables(c)
    igrid = 911
    Srce(igrid, c, tab)
    va = x["va"].copy()
    ref_va = x["va"].copy()
    jsea = np.arange(1, nsea + 1, dtype=np.int32)
    inout = ("psc", "whitecap", "taubbl", "tauice", "shave")
    outs = {k: c[k].copy() for k in inout}
    ref = {k: c[k].copy() for k in inout}
    fl = c["flags"]
    xargs = prop._args(dict(x, va=va), pc.XYP3_ARGS, prop._XYP3_SCALARS, ("rpar", "ipar", "imaps", "tscal"))
    for step in range(3):
        print(f"[resident] step {step}: {nsea} sea points, {nspec} components", flush=True)
        assert L.ww_xyp3_batch_va(igrid, *xargs[:-2], p(va), nspec, nsea) == capi.Status.OK
        L.ww_srce_batch_va(igrid, fl["it"], fl["inflags24"], fl["flogrd57"], fl["flogrd58"], fl["dtg"], p(va),
                           nspec, nsea, nsea, p(jsea), p(c["wn1"]), p(c["cg1"]), p(c["pin"]), p(outs["psc"]),
                           p(outs["whitecap"]), p(outs["taubbl"]), p(outs["tauice"]), p(outs["shave"]))
        assert L.ww_srce_last_error() == capi.Status.OK
        assert L.ww3gpu_va_state(igrid) == DEVICE
        ref_va = O.xyp3_batch(dict(x, va=ref_va), nthreads=16)
        rv = ref_va.reshape(nsea + 1, nspec)
        r = O.srce_batch(dict(c, spec=np.ascontiguousarray(rv[1:]), **ref), tab, nthreads=16)
        rv[1:] = r["spec"]
        ref_va = rv.reshape(-1).copy()
        ref = {k: r[k] for k in inout}
    assert L.ww_va_host_read(igrid, p(va), nspec, nsea) == capi.Status.OK
    assert np.array_equal(_bits(va), _bits(ref_va))
    for k in ("psc", "whitecap", "taubbl", "tauice"):
        assert np.array_equal(_bits(outs[k]), _bits(ref[k])), k
    assert np.array_equal(outs["shave"], ref["shave"])

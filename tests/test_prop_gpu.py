"""The PR3/UQ propagation on the GPU through the C ABI, bit for bit against the CPU oracle.

ww_xyp3_batch (W3GATH + W3XYP3 + W3SCAT over all components) on regional and global grids,
lon-lat and Cartesian, with currents, boundary input (a point listed twice: the last write
wins), obstructions (MAPTRN), a subset of components, and a second call on the first's
output; ww_ktp3_batch (W3KTP3 NTIMES times per point) for both pass orders, each velocity
alone and both, NTIMES 0/1/3, a subset of points, and NaN inputs (the CFL maxima's restart
after a NaN).  Refused calls write nothing."""
import numpy as np
import pytest

import oracle as O
from ww3gpu import capi, prop
from ww3gpu import prop_case as pc

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _same(got, ref, label, nan_equal=False):
    g, r = np.asarray(got, np.float32), np.asarray(ref, np.float32)
    if nan_equal:
        both = np.isnan(g) & np.isnan(r)
        g = np.where(both, np.float32(0), g)
        r = np.where(both, np.float32(0), r)
    gb, rb = _bits(g), _bits(r)
    if not np.array_equal(gb, rb):
        bad = np.nonzero(gb != rb)[0]
        i = bad[0]
        raise AssertionError(f"{label}: {bad.size} of {gb.size} differ, first at {i}: gpu {g[i]!r} "
                             f"oracle {r[i]!r}")


@pytest.fixture(scope="module")
def rt(gpu_runtime):
    return gpu_runtime


XYP3 = {
    "regional_ll": dict(),
    "global_ll": dict(global_=True, seed=2),
    "cartesian": dict(flagll=False, seed=3),
    "currents": dict(flcur=True, seed=4),
    "boundary": dict(boundary=True, seed=5),
    "no_obstructions": dict(obstructions=0.0, seed=6),
    "many_obstructions": dict(obstructions=0.4, seed=7, global_=True),
    "no_land": dict(land=0.0, seed=8, global_=True),
    "nk8_nth12_60x45": dict(nx=60, ny=45, nk=8, nth=12, seed=9, boundary=True, flcur=True),
    "odd_day": dict(dsec_day=4500.0, seed=10),
}


@pytest.mark.parametrize("name", sorted(XYP3))
def test_xyp3(rt, name):
    c = pc.make_xyp3(**XYP3[name])
    print(f"[xyp3] {name}: nsea={c['nsea']} nspec={c['nspec']} imaps={c['imaps'].tolist()} "
          f"ntloc={sorted(set(O.xyp3_ntloc(c).tolist()))}", flush=True)
    got = prop.xyp3_batch(c)
    ref = O.xyp3_batch(c, nthreads=16)
    _same(got, ref, name)
    assert not np.array_equal(got, c["va"])


def test_xyp3_subset_and_repeat(rt):
    c = pc.make_xyp3(nx=30, ny=24, nk=5, nth=8, seed=11, boundary=True)
    on = np.zeros(c["nspec"], np.int32)
    on[::3] = 1
    c["on"] = on
    got = prop.xyp3_batch(c)
    ref = O.xyp3_batch(c, nthreads=16)
    _same(got, ref, "subset")
    va = c["va"].reshape(-1, c["nspec"])
    assert np.array_equal(got.reshape(-1, c["nspec"])[:, on == 0], va[:, on == 0])
    # a second call on the output (inputs cached on the device), then with other maps
    c2 = dict(c)
    c2["va"] = got
    _same(prop.xyp3_batch(c2), O.xyp3_batch(c2, nthreads=16), "second call")
    c3 = pc.make_xyp3(nx=30, ny=24, nk=5, nth=8, seed=12, land=0.35)
    _same(prop.xyp3_batch(c3), O.xyp3_batch(c3, nthreads=16), "new maps")


def test_xyp3_device_entry(rt):
    torch = rt
    c = pc.make_xyp3(nx=40, ny=30, nk=6, nth=8, seed=13, global_=True, flcur=True)
    d = torch.from_numpy(c["va"].copy()).cuda()
    st = torch.cuda.Stream()
    prop.xyp3_batch_dev(c, d, st)
    _same(d.cpu().numpy(), O.xyp3_batch(c, nthreads=16), "device entry")
    s = prop.stats()
    assert s["xyp3_cell_substeps"] > 0 and s["xyp3_ms"] > 0


def test_xyp3_refusals_write_nothing(rt):
    L = capi.lib()
    c = pc.make_xyp3(nx=16, ny=12, nk=3, nth=4, seed=14)
    va = c["va"].copy()
    bad = dict(c)
    bad["ipar"] = c["ipar"].copy()
    bad["ipar"][6] = 8  # tripole
    with pytest.raises(capi.WW3GpuError) as e:
        prop.xyp3_batch(bad, va=va)
    assert e.value.code == capi.Status.UNSUPPORTED
    bad = dict(c)
    bad["mapcxy"] = c["mapcxy"].copy()
    bad["mapcxy"][0] = bad["mapcxy"][1]  # not a permutation
    with pytest.raises(capi.WW3GpuError) as e:
        prop.xyp3_batch(bad, va=va)
    assert e.value.code == capi.Status.TABLE
    assert np.array_equal(va, c["va"])
    assert L.ww_xyp3_mpi_pack(1, None, 1, 1, None, 0, None) == capi.Status.UNSUPPORTED


KTP3 = {
    "both_odd": dict(itstep=1),
    "both_even": dict(itstep=2),
    "theta_only": dict(flck=False),
    "k_only": dict(flcth=False),
    "no_currents": dict(flcur=False),
    "ntimes1": dict(ntimes=1),
    "ntimes3_nk32_nth24": dict(ntimes=3, nk=32, nth=24, nsea=300),
    "nk32_nth36": dict(nk=32, nth=36, nsea=150),
    "subset": dict(npts=37, nsea=120),
    "tight_ctmax": dict(ctmax=0.05),
}


@pytest.mark.parametrize("name", sorted(KTP3))
def test_ktp3(rt, name):
    c = pc.make_ktp3(**KTP3[name])
    print(f"[ktp3] {name}: npts={c['npts']} nk={c['ipar'][0]} nth={c['ipar'][1]}", flush=True)
    got = prop.ktp3_batch(c)
    ref = O.ktp3_batch(c, nthreads=16)
    for g, r, lab in zip(got, ref, ("va", "cflthmax", "cflkmax")):
        _same(g, r, f"{name} {lab}")
    assert not np.array_equal(got[0], c["va"])


def test_ktp3_nan_inputs(rt):
    c = pc.make_ktp3(nsea=40, seed=15)
    ps = c["pscal"].reshape(-1, 10)
    ps[3, 2] = np.nan       # DDDX: every refraction velocity of the point NaN
    ps[5, 4] = np.nan       # CX: every wavenumber velocity of the point NaN
    c["va"].reshape(-1, c["nspec"])[c["jsea"][7], 5] = np.nan
    got = prop.ktp3_batch(c)
    ref = O.ktp3_batch(c, nthreads=16)
    for g, r, lab in zip(got, ref, ("va", "cflthmax", "cflkmax")):
        _same(g, r, f"nan {lab}", nan_equal=True)


def test_ktp3_zero_work_and_refusal(rt):
    c = pc.make_ktp3(nsea=20, seed=16, ntimes=0)
    got = prop.ktp3_batch(c)
    assert np.array_equal(got[0], c["va"])
    bad = dict(c)
    bad["ntimes"] = 1
    bad["jsea"] = c["jsea"].copy()
    bad["jsea"][0] = c["ncol"]  # outside VA
    with pytest.raises(capi.WW3GpuError) as e:
        prop.ktp3_batch(bad)
    assert e.value.code == capi.Status.BAD_SHAPE

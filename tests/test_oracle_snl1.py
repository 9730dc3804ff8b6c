"""The CPU oracle for INSNL1/W3SNL1, pinned before it is trusted.

Pins, in order of strength:
  1. the reference's golden vectors (tests/golden/snl1_nk25_nth24.bin: Fortran INSNL1
     tables and W3SNL1 answers for three points), bit for bit;
  2. exact symmetries of the DIA that hold in float32 arithmetic, at other grid sizes:
     rotation by one direction bin, mirror theta -> -theta (the two quadruplet images swap,
     and float addition is commutative), and scaling A -> 2A (S -> 8S, D -> 4D, exact
     for a power of two away from over/underflow);
  3. closed forms: zero spectrum -> zero output; weights of each interpolation stencil
     sum to one.
A wrong table index, a swapped image or a dropped factor fails at least one of them.
"""
import numpy as np
import pytest

import oracle as O
from ww3gpu import synth


def bits(x):
    return np.ascontiguousarray(x).view(np.uint32)


def test_tables_equal_fixture(fixture_snl1):
    f = fixture_snl1
    c = f.cfg
    t = O.insnl1(f.nk, f.nth, c["xfr"], c["dth"], c["lam"], f.sig, chain=1)
    assert list(t.dims[:5]) == list(f.dims5)
    assert np.array_equal(bits(t.scal), bits(f.scal))
    assert np.array_equal(t.ipim, f.ipim)
    assert np.array_equal(t.ic, f.ic)
    assert np.array_equal(bits(t.af11), bits(f.af11))


def test_af11_chain_o3_differs_only_in_last_bit(fixture_snl1):
    """-O3 WW3 computes x**11 as x^5*x^6; the fixture's -O0 build as x^3*x^8 (reference
    docs/log/2026-10-14-snl1-bitwise-af11.md).  Both are chains of correctly rounded
    products, so they agree to a few ulps and differ on part of the rows."""
    f = fixture_snl1
    c = f.cfg
    t0 = O.insnl1(f.nk, f.nth, c["xfr"], c["dth"], c["lam"], f.sig, chain=0)
    d = np.abs(bits(t0.af11).astype(np.int64) - bits(f.af11).astype(np.int64))
    assert d.max() <= 4 and d.sum() > 0


def test_w3snl1_equals_fixture(fixture_snl1):
    f = fixture_snl1
    c = f.cfg
    t = O.insnl1(f.nk, f.nth, c["xfr"], c["dth"], c["lam"], f.sig, chain=1)
    s, d = O.w3snl1(c, t, f.sig, f.a, f.cg, f.kdmean)
    assert np.array_equal(bits(s), bits(f.s))
    assert np.array_equal(bits(d), bits(f.d))


def test_w3snl1_fixture_tables_give_fixture_answers(fixture_snl1):
    f = fixture_snl1
    c = f.cfg
    dims = O.snl1_dims(f.nk, f.nth, c["xfr"], c["lam"])
    ft = O.Snl1Tables(dims, f.scal, np.ascontiguousarray(f.ipim), np.ascontiguousarray(f.ic), f.af11)
    s, d = O.w3snl1(c, ft, f.sig, f.a, f.cg, f.kdmean)
    assert np.array_equal(bits(s), bits(f.s)) and np.array_equal(bits(d), bits(f.d))


def test_stencil_weights_sum_to_one():
    cfg, sig = synth.dia_config(32, 36)
    t = O.insnl1(32, 36, cfg["xfr"], cfg["dth"], cfg["lam"], sig)
    awg = t.scal[3:11].astype(np.float64)
    assert abs(awg[:4].sum() - 1.0) < 1e-6 and abs(awg[4:].sum() - 1.0) < 1e-6
    assert np.array_equal(bits(t.scal[11:]), bits((t.scal[3:11] * t.scal[3:11]).astype(np.float32)))


@pytest.mark.parametrize("nk,nth", [(32, 24), (32, 36), (6, 8)])
def test_zero_spectrum_gives_zero(nk, nth):
    cfg, sig = synth.dia_config(nk, nth)
    t = O.insnl1(nk, nth, cfg["xfr"], cfg["dth"], cfg["lam"], sig)
    a = np.zeros((2, nk * nth), np.float32)
    cg = np.ones((2, nk), np.float32)
    s, d = O.w3snl1(cfg, t, sig, a, cg, np.array([3.0, 300.0], np.float32))
    assert not s.any() and not d.any()


def _case(nk, nth, npts=6, seed=11):
    cfg, sig, a, cg, kd = synth.dia_workload(npts, nk, nth, seed, log_depth=(5.0, 4000.0))
    t = O.insnl1(nk, nth, cfg["xfr"], cfg["dth"], cfg["lam"], sig)
    return cfg, sig, t, a, cg, kd


@pytest.mark.parametrize("nk,nth", [(32, 24), (32, 36), (25, 24), (10, 12)])
def test_rotation_equivariance(nk, nth):
    cfg, sig, t, a, cg, kd = _case(nk, nth)
    s, d = O.w3snl1(cfg, t, sig, a, cg, kd)
    rot = lambda x: np.roll(x.reshape(-1, nk, nth), 1, axis=2).reshape(x.shape)
    s2, d2 = O.w3snl1(cfg, t, sig, rot(a), cg, kd)
    assert np.array_equal(bits(s2), bits(rot(s))) and np.array_equal(bits(d2), bits(rot(d)))


@pytest.mark.parametrize("nk,nth", [(32, 24), (32, 36), (25, 24), (10, 12)])
def test_mirror_equivariance(nk, nth):
    cfg, sig, t, a, cg, kd = _case(nk, nth)
    s, d = O.w3snl1(cfg, t, sig, a, cg, kd)
    idx = (-np.arange(nth)) % nth
    mir = lambda x: x.reshape(-1, nk, nth)[:, :, idx].reshape(x.shape)
    s2, d2 = O.w3snl1(cfg, t, sig, mir(a), cg, kd)
    assert np.array_equal(bits(s2), bits(mir(s))) and np.array_equal(bits(d2), bits(mir(d)))


@pytest.mark.parametrize("nk,nth", [(32, 24), (32, 36)])
def test_cubic_scaling(nk, nth):
    cfg, sig, t, a, cg, kd = _case(nk, nth)
    s, d = O.w3snl1(cfg, t, sig, a, cg, kd)
    s2, d2 = O.w3snl1(cfg, t, sig, a * np.float32(2.0), cg, kd)
    # exact for a power of two unless a product is subnormal: compare far above that range
    ms, md = np.abs(s) > 1e-20, np.abs(d) > 1e-20
    assert ms.mean() > 0.5 and md.mean() > 0.5
    assert np.array_equal(bits(s2[ms]), bits((s * np.float32(8.0))[ms]))
    assert np.array_equal(bits(d2[md]), bits((d * np.float32(4.0))[md]))


def test_thread_count_does_not_change_bits():
    cfg, sig, t, a, cg, kd = _case(32, 24, npts=64)
    s1, d1 = O.w3snl1(cfg, t, sig, a, cg, kd, nthreads=1)
    s4, d4 = O.w3snl1(cfg, t, sig, a, cg, kd, nthreads=4)
    assert np.array_equal(bits(s1), bits(s4)) and np.array_equal(bits(d1), bits(d4))


def test_fp64_evaluation_tracks_the_float32_oracle():
    """The float64 restatement of W3SNL1 (tests/dia_fp64.py) against the float32 oracle on
    the 1000-point deep-water case and on log-uniform depths: they differ only by float32
    rounding (relative L2 below 1e-5), so the float64 evaluation follows the same routine
    and the float32 path's error against it is rounding."""
    import dia_fp64
    from ww3gpu import synth
    for nth, kw in ((24, {}), (36, {"log_depth": (10.0, 4000.0)})):
        cfg, sig, a, cg, kd = synth.dia_workload(300, 32, nth, seed=17, **kw)
        t = O.insnl1(32, nth, cfg["xfr"], cfg["dth"], cfg["lam"], sig)
        s32, _ = O.w3snl1(cfg, t, sig, a, cg, kd)
        s64 = dia_fp64.w3snl1_fp64(cfg, t, sig, a, cg, kd)
        err = dia_fp64.error_vs_fp64(s32, s64)
        assert err["rel_l2_global"] < 1e-5, err
        assert err["rel_l2_point_max"] < 1e-4, err

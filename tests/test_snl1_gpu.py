"""W3SNL1 (DIA) on the GPU through the C ABI, against the reference's golden vectors and
the CPU oracle, bit for bit.

Sizes: the fixture (3 points, NK=25 NTH=24), the 1000-point NK=32 NTH=24 deep-water DIA
case, and the 9520-point NK=32 NTH=36 shape of the full source-term step (log-uniform
depths 10-4000 m).  At full size the exact symmetries of the DIA (rotation, mirror, A->2A)
are checked on the device as well.
"""
import numpy as np
import pytest

import oracle as O
from ww3gpu import capi, synth
from ww3gpu.snl1 import Snl1

pytestmark = pytest.mark.gpu


def bits(x):
    return np.ascontiguousarray(x).view(np.uint32)


def assert_bits(got, want, what):
    g, w = bits(got), bits(want)
    if not np.array_equal(g, w):
        bad = np.flatnonzero(g.ravel() != w.ravel())
        i = bad[0]
        ulp = np.abs(g.ravel()[bad].astype(np.int64) - w.ravel()[bad].astype(np.int64)).max()
        raise AssertionError(f"{what}: {bad.size} values differ (max {ulp} ulp), first at {i}: "
                             f"{got.ravel()[i]!r} vs {want.ravel()[i]!r}")


def test_fixture_bit_exact(gpu_runtime, fixture_snl1):
    f = fixture_snl1
    op = Snl1(101, f.cfg, f.sig)
    t = op.tables()
    assert np.array_equal(t["ipim"], f.ipim) and np.array_equal(t["ic"], f.ic)
    op.set_af11(f.af11)  # the fixture's AF11 (x^3*x^8 chain of its -O0 build)
    for variant in (0, 1):
        op.set_variant(variant)
        s, d = op.host(f.a, f.cg, f.kdmean)
        assert_bits(s, f.s, f"variant {variant} S vs Fortran")
        assert_bits(d, f.d, f"variant {variant} D vs Fortran")


CASES = [
    ("dia_1000pts_nk32_nth24_deep", 1000, 32, 24, 1, None),
    ("step_9520pts_nk32_nth36", 9520, 32, 36, 2, (10.0, 4000.0)),
    ("nk25_nth24_shallow", 300, 25, 24, 3, (1.0, 60.0)),
    ("nk10_nth12", 77, 10, 12, 4, (5.0, 500.0)),
    ("nk30_nth72", 40, 30, 72, 5, (20.0, 3000.0)),
]


@pytest.mark.parametrize("name,npts,nk,nth,seed,logd", CASES, ids=[c[0] for c in CASES])
def test_matches_oracle(gpu_runtime, name, npts, nk, nth, seed, logd):
    print(f"case {name}", flush=True)
    if logd is None:
        cfg, sig, a, cg, kd = synth.dia_workload(npts, nk, nth, seed, depth=4000.0)
    else:
        cfg, sig, a, cg, kd = synth.dia_workload(npts, nk, nth, seed, log_depth=logd)
    t = O.insnl1(nk, nth, cfg["xfr"], cfg["dth"], cfg["lam"], sig)
    s_ref, d_ref = O.w3snl1(cfg, t, sig, a, cg, kd, nthreads=8)
    op = Snl1(200 + seed, cfg, sig)
    L = capi.lib()
    # the NTH = 24 and 36 grids (LAMBDA 0.25, XFR 1.1) have a specialised kernel
    assert L.ww3gpu_snl1_kernel_variant(op.igrid) == (0 if nth in (24, 36) else 1)
    for variant in (0, 1):  # stencil-specialised kernel (where one exists), general kernel
        op.set_variant(variant)
        s, d = op.host(a, cg, kd)
        assert_bits(s, s_ref, f"{name} variant {variant} S")
        assert_bits(d, d_ref, f"{name} variant {variant} D")


def test_device_entry_equals_host_entry(gpu_runtime):
    torch = gpu_runtime
    cfg, sig, a, cg, kd = synth.dia_workload(2048, 32, 36, 9, log_depth=(10.0, 4000.0))
    op = Snl1(301, cfg, sig)
    s_h, d_h = op.host(a, cg, kd)
    A, CG, KD = (torch.from_numpy(x).cuda() for x in (a, cg, kd))
    S = torch.full_like(A, float("nan"))
    D = torch.full_like(A, float("nan"))
    op.device(A, CG, KD, S, D)
    torch.cuda.synchronize()
    assert_bits(S.cpu().numpy(), s_h, "device S")
    assert_bits(D.cpu().numpy(), d_h, "device D")


def test_edge_cases(gpu_runtime):
    nk, nth = 32, 24
    cfg, sig = synth.dia_config(nk, nth)
    t = O.insnl1(nk, nth, cfg["xfr"], cfg["dth"], cfg["lam"], sig)
    op = Snl1(302, cfg, sig)
    rng = np.random.default_rng(5)
    npts = 257  # ragged: not a multiple of anything the launch uses
    a = (rng.random((npts, nk * nth), np.float32) * np.float32(1e-3))
    a[0] = 0.0                                    # zero spectrum
    a[1] = np.float32(1e-38)                      # denormal-range products
    a[2, ::7] = 0.0                               # sparse
    a[3] = rng.random(nk * nth, np.float32) * np.float32(1e3)  # large
    cg = (rng.random((npts, nk), np.float32) + np.float32(0.1))
    kd = np.exp(rng.uniform(np.log(1e-3), np.log(1e4), npts)).astype(np.float32)
    kd[4] = 0.0                                   # KDMIN clamp
    kd[5] = np.float32(1e30)                      # expf underflow
    s_ref, d_ref = O.w3snl1(cfg, t, sig, a, cg, kd, nthreads=8)
    s, d = op.host(a, cg, kd)
    assert_bits(s, s_ref, "edge S")
    assert_bits(d, d_ref, "edge D")
    # one point
    s1, d1 = op.host(a[7:8], cg[7:8], kd[7:8])
    assert_bits(s1, s_ref[7:8], "1-point S")
    # npts == 0 is a legal no-op
    L = capi.lib()
    L.ww_snl1(302, 0, None, None, None, None, None)
    assert L.ww_snl1_last_error() == capi.Status.OK


def test_errors_write_nothing(gpu_runtime):
    cfg, sig = synth.dia_config(32, 24)
    op = Snl1(303, cfg, sig)
    L = capi.lib()
    a = np.ones((2, 768), np.float32)
    cg = np.ones((2, 32), np.float32)
    kd = np.ones(2, np.float32)
    s = np.full_like(a, 7.0)
    L.ww_snl1(303, 2, a.ctypes.data, cg.ctypes.data, None, s.ctypes.data, s.ctypes.data)
    assert L.ww_snl1_last_error() == capi.Status.BAD_SHAPE
    L.ww_snl1(303, -1, a.ctypes.data, cg.ctypes.data, kd.ctypes.data, s.ctypes.data, s.ctypes.data)
    assert L.ww_snl1_last_error() == capi.Status.BAD_SHAPE
    L.ww_snl1(99999, 2, a.ctypes.data, cg.ctypes.data, kd.ctypes.data, s.ctypes.data, s.ctypes.data)
    assert L.ww_snl1_last_error() == capi.Status.NOT_INITIALISED
    assert np.all(s == 7.0)
    # init errors
    assert L.ww_snl1_init(304, 0, 24, 1.1, 0.26, 0.25, 1., 1., 1., 1., 1., 1., 1., sig.ctypes.data) \
        == capi.Status.BAD_SHAPE
    assert L.ww_snl1_init(304, 32, 24, 1.1, 0.26, 0.25, 1., 1., 1., 1., 1., 1., 1., None) \
        == capi.Status.BAD_SHAPE


def test_symmetries_at_full_size(gpu_runtime):
    """Size-independent properties on the 9520-point NK=32 NTH=36 shape, device only."""
    torch = gpu_runtime
    nk, nth = 32, 36
    cfg, sig, a, cg, kd = synth.dia_workload(9520, nk, nth, 21, log_depth=(10.0, 4000.0))
    op = Snl1(305, cfg, sig)
    s, d = op.host(a, cg, kd)
    rot = lambda x: np.roll(x.reshape(-1, nk, nth), 3, axis=2).reshape(x.shape)
    idx = (-np.arange(nth)) % nth
    mir = lambda x: x.reshape(-1, nk, nth)[:, :, idx].reshape(x.shape)
    s2, d2 = op.host(rot(a), cg, kd)
    assert_bits(s2, rot(s), "rotation S")
    assert_bits(d2, rot(d), "rotation D")
    s3, d3 = op.host(mir(a), cg, kd)
    assert_bits(s3, mir(s), "mirror S")
    assert_bits(d3, mir(d), "mirror D")
    # A -> 2A scales every product by a power of two, exactly, unless a product falls into
    # the subnormal range; compare where the outputs are far above it.
    s4, d4 = op.host(a * np.float32(2.0), cg, kd)
    ms, md = np.abs(s) > 1e-20, np.abs(d) > 1e-20
    assert ms.mean() > 0.5 and md.mean() > 0.5
    assert_bits(s4[ms], (s * np.float32(8.0))[ms], "scaling S")
    assert_bits(d4[md], (d * np.float32(4.0))[md], "scaling D")


def test_dia_error_against_fp64_evaluation(gpu_runtime):
    """The device DIA's S_nl (bit-identical to the float32 oracle) against a float64
    evaluation of the same routine (tests/dia_fp64.py) on the source-term step's shape
    (NK=32, NTH=36, log-uniform depths): float32 rounding only, relative L2 below 2e-6."""
    import dia_fp64
    torch = gpu_runtime
    cfg, sig, a, cg, kd = synth.dia_workload(2000, 32, 36, seed=2024, log_depth=(10.0, 4000.0))
    op = Snl1(23, cfg, sig)
    A, CG, KD = (torch.from_numpy(x).cuda() for x in (a, cg, kd))
    S = torch.empty_like(A)
    D = torch.empty_like(A)
    op.device(A, CG, KD, S, D)
    torch.cuda.synchronize()
    t = O.insnl1(32, 36, cfg["xfr"], cfg["dth"], cfg["lam"], sig)
    err = dia_fp64.error_vs_fp64(S.cpu().numpy(), dia_fp64.w3snl1_fp64(cfg, t, sig, a, cg, kd))
    print("DIA error vs float64:", err)
    assert err["rel_l2_global"] < 2e-6 and err["rel_l2_point_max"] < 1e-5, err

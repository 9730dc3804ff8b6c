"""The W3SNL1 surrogate on the CPU: the reference network against torch's own bfloat16
conversion and autograd, the exact properties of the DIA that the surrogate's
normalisation relies on (checked on the oracle), and the entry points' error paths."""
import numpy as np
import pytest

import oracle as O
import snlsur_ref as R
from ww3gpu import capi, synth


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 20000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, np.inf, -np.inf], np.float32)])
    want = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(R.bf16(x).view(np.uint32), want.view(np.uint32))


def test_reference_gradients_match_finite_differences():
    """Gradients of the reference's loss against central differences in float64 on a small
    network without bf16 rounding (weights and inputs already on the bf16 grid)."""
    import torch
    rng = np.random.default_rng(1)
    sizes = [(8, 6), (8, 8), (8, 8), (5, 8)]
    layers = [(R.bf16(rng.standard_normal(s).astype(np.float32) * 0.5), rng.standard_normal(s[0]).astype(np.float32) * 0.1)
              for s in sizes]
    x = R.bf16(rng.standard_normal((4, 6)).astype(np.float32))
    y = rng.standard_normal((4, 5)).astype(np.float32)

    def loss64(ls):
        h = torch.tensor(x, dtype=torch.float64)
        for i, (w, b) in enumerate(ls):
            z = h @ torch.tensor(w, dtype=torch.float64).T + torch.tensor(b, dtype=torch.float64)
            h = torch.nn.functional.gelu(z, approximate="tanh") if i < 3 else z
        yt = torch.tensor(y, dtype=torch.float64)
        return float((((h - yt) ** 2).sum(1) / (yt * yt).sum(1)).mean())

    _, _, grads = R.forward_backward(layers, x, y)
    eps = 1e-4
    for li in range(4):
        for wi in (0, 1):
            arr = layers[li][wi]
            for idx in [(0,) * arr.ndim, tuple(s - 1 for s in arr.shape)]:
                plus = [(w.astype(np.float64).copy(), b.astype(np.float64).copy()) for w, b in layers]
                minus = [(w.astype(np.float64).copy(), b.astype(np.float64).copy()) for w, b in layers]
                plus[li][wi][idx] += eps
                minus[li][wi][idx] -= eps
                fd = (loss64(plus) - loss64(minus)) / (2 * eps)
                got = grads[li][wi][idx]
                assert abs(got - fd) <= 2e-3 * max(1.0, abs(fd)), (li, wi, idx, got, fd)


@pytest.fixture(scope="module")
def dia36():
    cfg, sig = synth.dia_config(32, 36)
    t = O.insnl1(32, 36, cfg["xfr"], cfg["dth"], cfg["lam"], sig)
    return cfg, sig, t


def test_dia_depth_enters_only_through_cons_and_conx(dia36):
    """S / (CONX CONS) depends on UE = A / CONX alone: the same UE at two depths gives the
    same S / (CONX CONS) (to rounding of the A <-> UE conversions)."""
    cfg, sig, t = dia36
    _, _, a, cg, kd = synth.dia_workload(6, 32, 36, seed=5, log_depth=(10, 4000))
    tpiinv = np.float32(1.0) / (np.float32(2.0) * np.float32(np.pi))
    conx = (tpiinv / sig[None, :] * cg).astype(np.float32)
    ue = a.reshape(6, 32, 36) / conx[:, :, None]
    cg2 = (cg * np.float32(1.37)).astype(np.float32)  # another depth's group velocities
    conx2 = (tpiinv / sig[None, :] * cg2).astype(np.float32)
    a2 = (ue * conx2[:, :, None]).reshape(6, -1).astype(np.float32)
    kd2 = (kd * np.float32(0.31)).astype(np.float32)

    def cons(k):
        x = np.maximum(cfg["kdcon"] * k, cfg["kdmn"]).astype(np.float64)
        return cfg["snlc1"] * (1 + cfg["snls1"] / x * (1 - cfg["snls2"] * x) * np.exp(np.maximum(-1e15, cfg["snls3"] * x)))

    s1, _ = O.w3snl1(cfg, t, sig, a, cg, kd)
    s2, _ = O.w3snl1(cfg, t, sig, a2, cg2, kd2)
    f1 = s1.reshape(6, 32, 36) / conx[:, :, None] / cons(kd)[:, None, None]
    f2 = s2.reshape(6, 32, 36) / conx2[:, :, None] / cons(kd2)[:, None, None]
    scale = np.abs(f1).max(axis=(1, 2), keepdims=True)
    assert np.abs(f1 - f2).max() <= 1e-5 * scale.max()
    assert np.all(np.abs(f1 - f2) <= 1e-5 * scale)


def test_dia_is_cubic_and_rotation_equivariant(dia36):
    """F(2 UE) = 8 F(UE) and a rotation of A by whole direction bins rotates S, bit for bit:
    the normalisation's max-scaling and peak rotation lose nothing."""
    cfg, sig, t = dia36
    _, _, a, cg, kd = synth.dia_workload(4, 32, 36, seed=6)
    s, _ = O.w3snl1(cfg, t, sig, a, cg, kd)
    s2, _ = O.w3snl1(cfg, t, sig, (a * np.float32(2)).astype(np.float32), cg, kd)
    # Exact unless a product falls into the subnormal range: compare far above it.
    big = np.abs(s) > 1e-20
    assert big.mean() > 0.5
    assert np.array_equal((s * np.float32(8))[big].view(np.uint32), s2[big].view(np.uint32))
    ar = np.roll(a.reshape(4, 32, 36), 7, axis=2).reshape(4, -1).copy()
    sr, _ = O.w3snl1(cfg, t, sig, ar, cg, kd)
    assert np.array_equal(np.roll(s.reshape(4, 32, 36), 7, axis=2).reshape(4, -1).view(np.uint32),
                          sr.view(np.uint32))


def test_surrogate_entries_before_init_report_not_initialised():
    L = capi.lib()
    assert L.ww3gpu_snlsur_init(977, 0) == capi.Status.NOT_INITIALISED
    assert L.ww3gpu_snlsur_nparams(977, None) < 0
    assert L.ww3gpu_snlsur_dev(977, 1, None, None, None, None, None, None, None) == capi.Status.NOT_INITIALISED
    assert L.ww3gpu_snlsur_last_error() == capi.Status.NOT_INITIALISED


def test_gemm_entry_rejects_bad_shapes():
    """Shapes the GEMM tiles do not cover, short or unaligned row strides and null pointers
    are BAD_SHAPE before anything is launched."""
    L = capi.lib()
    dummy = 0x1000  # never dereferenced: every call below fails validation first
    bad = [(100, 128, 64, 64, 64, 128), (128, 100, 64, 64, 64, 100), (128, 128, 60, 60, 60, 128),
           (128, 128, 64, 32, 64, 128), (128, 128, 64, 64, 68, 128), (128, 128, 64, 64, 64, 64)]
    for m, n, k, lda, ldb, ldc in bad:
        assert L.ww3gpu_sur_gemm_dev(m, n, k, dummy, lda, dummy, ldb, dummy, ldc, None) == capi.Status.BAD_SHAPE
    assert L.ww3gpu_sur_gemm_dev(128, 128, 64, None, 64, dummy, 64, dummy, 128, None) == capi.Status.BAD_SHAPE

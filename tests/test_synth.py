"""The seeded workload generator: deterministic, finite, physically sensible."""
import numpy as np
import pytest

from ww3gpu import synth


@pytest.mark.parametrize("nk,nth,depth", [(32, 24, 4000.0), (32, 36, None), (25, 24, 10.0)])
def test_jonswap_finite_and_hs(nk, nth, depth):
    cfg, sig = synth.dia_config(nk, nth)
    d = depth if depth is not None else np.exp(np.linspace(np.log(10.0), np.log(4000.0), 200))
    n = 200
    a, cg, kd, k = synth.jonswap_points(n, nk, nth, sig, d, seed=3)
    for x in (a, cg, kd, k):
        assert np.isfinite(x).all()
    assert (a >= 0).all() and (cg > 0).all() and (kd > 0).all()
    # energy back from action: E = A sigma / Cg; 4 sqrt(m0) within the drawn Hs range
    xfr = float(sig[1] / sig[0])
    dsig = sig.astype(np.float64) * 0.5 * (xfr - 1 / xfr)
    e = a.reshape(n, nk, nth).astype(np.float64) * (sig / cg)[:, :, None]
    hs = 4 * np.sqrt((e * dsig[None, :, None]).sum(axis=(1, 2)) * (2 * np.pi / nth))
    assert hs.min() > 0.45 and hs.max() < 8.1


def test_deterministic():
    a1 = synth.dia_workload(50, 32, 24, seed=7, log_depth=(10, 4000))[2]
    a2 = synth.dia_workload(50, 32, 24, seed=7, log_depth=(10, 4000))[2]
    assert np.array_equal(a1.view(np.uint32), a2.view(np.uint32))

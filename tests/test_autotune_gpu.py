"""The W3SRCE width autotuner: thresholds change the launch shapes, never the results, and
the tuner keeps the fastest measured candidate."""
import numpy as np
import pytest

from ww3gpu import autotune, srce_case as sc
from ww3gpu.srce import Srce

pytestmark = pytest.mark.gpu

KEYS = ("spec", "wn1", "cg1", "pin", "psc", "whitecap", "taubbl", "tauice", "shave")


def _device_case(torch, npts, seed):
    c = sc.make_case(npts, nk=32, nth=36, seed=seed)
    src = {k: torch.from_numpy(np.ascontiguousarray(c[k])).cuda() for k in KEYS}
    return c, src


def test_thresholds_do_not_change_results(gpu_runtime):
    torch = gpu_runtime
    c, src = _device_case(torch, 1500, 31)
    op = Srce(61, c)
    outs = []
    try:
        for th in ((0, 0), (1000, 500), (200, 0), (100000, 100000)):
            autotune.set_thresholds(*th)
            work = {k: v.clone() for k, v in src.items()}
            op.batch_dev(work, c["flags"])
            torch.cuda.synchronize()
            outs.append({k: work[k].cpu().numpy() for k in ("spec", "psc", "whitecap", "taubbl", "tauice", "shave")})
    finally:
        autotune.set_thresholds(0, 0)
    for o in outs[1:]:
        for k in o:
            assert np.array_equal(o[k].view(np.uint8), outs[0][k].view(np.uint8)), k


def test_tuner_keeps_the_fastest_candidate(gpu_runtime):
    torch = gpu_runtime
    c, src = _device_case(torch, 3000, 32)
    op = Srce(62, c)
    work = {k: v.clone() for k, v in src.items()}
    cands = [(4096, 2048), (1024, 512), (8192, 4096)]
    try:
        r = autotune.tune_srce_widths(op, work, src, c["flags"], candidates=cands, steps=3, warmup=1)
    finally:
        autotune.set_thresholds(0, 0)
    assert set(r["ms"]) == set(cands)
    assert r["best"] == min(r["ms"], key=r["ms"].get)
    assert all(v > 0 for v in r["ms"].values())

"""Measured tuning of the W3SRCE sub-step kernel's occupancy choice.

Each sub-step launch picks the lanes per point from its active-point count: one 16-lane row
above `half_below` points, two rows down to `wide_below`, a wavefront at or below it
(ww3gpu_srce_width_thresholds).  The right crossover depends on how many points of each
width fill the GPU in one round, on the sub-step's latency at each width and on the
distribution of active counts over a step.  This tuner measures instead of modelling: it
times whole W3SRCE steps of the caller's own workload (device arrays, each timed step
restarting from the same state, HIP events on one stream) for every candidate pair and
keeps the fastest.  Results do not depend on the choice (every width is bit-identical), so
tuning never changes an answer.
"""
from __future__ import annotations

from .capi import check, lib

DEFAULT = (4096, 2048)
CANDIDATES = [(h, w) for h in (3072, 4096, 5120, 6144) for w in (1024, 1536, 2048, 2560, 3072) if w <= h]


def set_thresholds(half_below: int, wide_below: int) -> None:
    check("ww3gpu_srce_width_thresholds", lib().ww3gpu_srce_width_thresholds(half_below, wide_below))


def tune_srce_widths(op, work: dict, src: dict, flags: dict, candidates=None, steps: int = 8,
                     warmup: int = 2, apply: bool = True) -> dict:
    """Time `steps` W3SRCE steps per candidate (half_below, wide_below) on device tensors.

    op: a configured ww3gpu.srce.Srce; work/src: device tensors as Srce.batch_dev takes them
    (src holds the input state, copied into work before every step).  Returns {"best":
    (h, w), "ms": {(h, w): ms per step}}; with apply the best pair stays set."""
    import torch

    cands = list(candidates) if candidates is not None else list(CANDIDATES)
    if DEFAULT not in cands:
        cands.insert(0, DEFAULT)
    inout = ("spec", "psc", "whitecap", "taubbl", "tauice", "shave")
    stream = torch.cuda.Stream()

    def step():
        with torch.cuda.stream(stream):
            for k in inout:
                work[k].copy_(src[k])
        op.batch_dev(work, flags, stream)

    ms = {}
    try:
        for c in cands:
            set_thresholds(*c)
            for _ in range(warmup):
                step()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record(stream)
            for _ in range(steps):
                step()
            e1.record(stream)
            torch.cuda.synchronize()
            ms[c] = e0.elapsed_time(e1) / steps
    finally:
        best = min(ms, key=ms.get) if ms else DEFAULT
        set_thresholds(*(best if apply else DEFAULT))
    return {"best": best, "ms": ms}


__all__ = ["tune_srce_widths", "set_thresholds", "CANDIDATES", "DEFAULT"]

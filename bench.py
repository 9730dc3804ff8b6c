"""Benchmark of the spectral-step hot path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload dia]

A step is one pass of the hot path over one batch of synthetic sea points, with every
input already resident in HBM.  Workload "dia": W3SNL1 over 9520 points, NK=32, NTH=36,
log-uniform depths 10-4000 m (the point set and spectral grid of the full source-term
step).  Prints one JSON line.  The CPU baseline is the oracle (a scalar C restatement,
OpenMP over points) timed on a bounded sample on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "ww3-gpu_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "grid-point-steps/s for SNL1+ST4 step; surrogate vs DIA speedup at ≤5% S_nl err"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def dia_flops_per_point(nk, nth, nspecx):
    # Operations counted from the kernel's expressions: 67 per extended bin (section 3),
    # 54 + 3 per bin (section 4 and the UE fill).  Divisions count as one.
    return 67 * nspecx + 57 * nk * nth


def dia_bytes_per_point(nk, nth):
    # A in, S and D out (NSPEC floats each), CG (NK) and KDMEAN in.
    return 4 * (3 * nk * nth + nk + 1)


def cpu_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def bench_dia(args):
    import numpy as np
    import torch

    from ww3gpu import synth
    from ww3gpu.snl1 import Snl1

    nk, nth, npts = 32, 36, 9520
    cfg, sig, a, cg, kd = synth.dia_workload(npts, nk, nth, seed=2024, log_depth=(10.0, 4000.0))
    op = Snl1(1, cfg, sig)
    nspecx = int(op.tables()["dims"][3])
    from ww3gpu import capi
    kernel_name = "dia_fixed" if capi.lib().ww3gpu_snl1_kernel_variant(op.igrid) == 0 else "dia_kernel"
    A, CG, KD = (torch.from_numpy(x).cuda() for x in (a, cg, kd))
    S = torch.empty_like(A)
    D = torch.empty_like(A)
    # One stream for the launches and the events that time them.
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        op.device(A, CG, KD, S, D, stream)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        op.device(A, CG, KD, S, D, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    ms_step = gpu_ms / args.steps
    value = npts * args.steps / (gpu_ms * 1e-3)

    bpp = dia_bytes_per_point(nk, nth)
    fpp = dia_flops_per_point(nk, nth, nspecx)
    achieved_gbs = bpp * npts / (ms_step * 1e-3) / 1e9
    achieved_tf = fpp * npts / (ms_step * 1e-3) / 1e12

    # CPU baseline: the oracle on a bounded sample of the same workload.
    import oracle as O
    t = O.insnl1(nk, nth, cfg["xfr"], cfg["dth"], cfg["lam"], sig)
    cores = cpu_cores()
    sample = min(npts, max(cores * 64, 1024))
    O.w3snl1(cfg, t, sig, a[:cores], cg[:cores], kd[:cores], nthreads=cores)  # warm
    reps, c0 = 0, time.perf_counter()
    while True:
        O.w3snl1(cfg, t, sig, a[:sample], cg[:sample], kd[:sample], nthreads=cores)
        reps += 1
        if time.perf_counter() - c0 > args.cpu_seconds:
            break
    cpu_s = time.perf_counter() - c0
    cpu_value = sample * reps / cpu_s

    return {
        "metric": METRIC,
        "value": value,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded JONSWAP spectra, log-uniform depths)",
        "config": {
            "workload": "W3SNL1 DIA only (ST4 and W3SRCE integration not in this line): 9520 points, "
                        "NK=32, NTH=36, depth 10-4000 m; a step is one DIA evaluation of every point",
            "npts": npts, "nk": nk, "nth": nth, "parallelism": "single GPU",
            "wall_s": wall,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": None,
            "bytes_per_point": bpp, "kernel": kernel_name,
        },
        "roofline_valu": {
            "achieved": achieved_tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved_tf / FP32_PEAK_TFLOPS, "flops_per_point": fpp,
            "note": "non-fused FP32 operations counted from the kernel source; the kernel may "
                    "not fuse or pack them (bit parity), so its ceiling is 1/4 of this peak",
        },
        "cpu_baseline": {
            "value": cpu_value, "unit": "grid-point-steps/s", "cores": cores, "kind": "port",
            "sample": f"{sample} of the {npts} points x {reps} repetitions, {cpu_s:.1f} s, "
                      f"oracle W3SNL1 (C, -O3, OpenMP over points)",
        },
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="dia", choices=["dia"])
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(source terms shard over points with no exchange; see the design notes)",
              file=sys.stderr)
        sys.exit(2)
    res = bench_dia(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

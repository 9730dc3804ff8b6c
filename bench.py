"""Benchmark of the spectral-step hot path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload srce|dia|prop|surrogate]

A step is one pass of the hot path over one batch of synthetic sea points, with every
input already resident in HBM.

Workload "srce" (default): the full source-term time step, W3SRCE with ST4 input and
dissipation, the DIA, W3SLN1, W3SBT1, W3SDB1 and the semi-implicit sub-step integration to
DTG = 300 s, over 9520 points, NK=32, NTH=36, U10 Weibull(k=2, mean 9 m/s) capped at
35 m/s, depths log-uniform 10-4000 m, JONSWAP spectra.  Every timed step starts from the
same input state (a device-to-device copy of the INOUT arrays, inside the timed region).

Workload "dia": W3SNL1 alone over the same 9520 points.

Workload "surrogate": the learned W3SNL1 surrogate (trained first, on the device, on 8M
exact-DIA pairs, 12 passes) against the exact DIA on 1M held-out points: throughput, speed-up and the
S_nl error.

Prints one JSON line.  The CPU baseline is the oracle (a scalar C restatement, OpenMP over
points) timed on a bounded sample on this host's cores.
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


# Bytes each kernel group must move per point it processes in one sub-step, for the roofline
# of the dominant group (NSPEC-sized float arrays unless noted).  DIA: read A, write S_nl and
# D_nl (and CG, KDMEAN).  Sub-step: read A, A_init, S_in, D_in, S_nl, D_nl, S_db and LLWS
# (bytes); write A, S_in, D_in, LLWS, and S_ds, D_ds (written, then read back once).
def srce_group_bytes(nk, nspec):
    f = 4
    return {
        "DIA (W3SNL1)": f * (3 * nspec + nk + 1),
        "sub-step (W3SLN1, W3SDS4, W3SDB1, W3SBT1, integration, W3SPR4, W3SIN4)":
            f * (7 * nspec + 3 * nspec + 4 * nspec + 2 * nk) + 2 * nspec,
    }


SRCE_GROUPS = ["init (W3SRCE section 1, first W3SIN4)", "DIA (W3SNL1)",
               "sub-step (W3SLN1, W3SDS4, W3SDB1, W3SBT1, integration, W3SPR4, W3SIN4)",
               "finish (W3SRCE section 3)"]
SRCE_KERNELS = ["init_kernel", "dia_fixed", "step_kernel", "finish_kernel"]
SRCE_STEP_BYTES = 9744  # per grid-point-step: spec in+out, WN1/CG1, pin, psc in+out, whitecap..


def bench_srce(args):
    import ctypes

    import numpy as np
    import torch

    from ww3gpu import capi, srce_case as sc
    from ww3gpu.srce import Srce

    nk, nth, npts = 32, 36, args.npts
    c = sc.make_case(npts, nk=nk, nth=nth, seed=2024)
    op = Srce(1, c)
    L = capi.lib()
    fl = c["flags"]
    keys = ("spec", "wn1", "cg1", "pin", "psc", "whitecap", "taubbl", "tauice", "shave")
    src = {k: torch.from_numpy(np.ascontiguousarray(c[k])).cuda() for k in keys}
    work = {k: v.clone() for k, v in src.items()}
    inout = ("spec", "psc", "whitecap", "taubbl", "tauice", "shave")
    tuned = None
    if args.autotune:
        from ww3gpu import autotune
        r = autotune.tune_srce_widths(op, work, src, fl)
        tuned = {"best": list(r["best"]), "ms_per_step": {f"{h},{w}": v for (h, w), v in r["ms"].items()}}
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()

    def step():
        with torch.cuda.stream(stream):
            for k in inout:
                work[k].copy_(src[k])
        return op.batch_dev(work, fl, stream)

    for _ in range(args.warmup):
        step()
    stream.synchronize()
    substeps = L.ww3gpu_srce_substeps(op.igrid)
    check_timers = L.ww3gpu_srce_timing(op.igrid, 1)
    capi.check("ww3gpu_srce_timing", check_timers)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    ms_step = gpu_ms / args.steps
    value = npts * args.steps / (gpu_ms * 1e-3)
    nt = len(SRCE_GROUPS)
    ms = (ctypes.c_double * nt)()
    nl = (ctypes.c_long * nt)()
    capi.check("ww3gpu_srce_timers", L.ww3gpu_srce_timers(op.igrid, ctypes.addressof(ms),
                                                            ctypes.addressof(nl), nt))
    L.ww3gpu_srce_timing(op.igrid, 0)
    groups = {SRCE_GROUPS[i]: {"kernel": SRCE_KERNELS[i], "ms_per_step": ms[i] / args.steps,
                               "launches_per_step": nl[i] / args.steps,
                               "avg_launch_us": 1e3 * ms[i] / max(nl[i], 1)} for i in range(nt)}
    pts_substeps = L.ww3gpu_srce_point_substeps(op.igrid)
    dom = max(groups, key=lambda g: groups[g]["ms_per_step"])
    gb = srce_group_bytes(nk, nk * nth)
    if dom in gb:
        # bytes per launch: the group's bytes per point-sub-step times the point-sub-steps of
        # one step, over the group's launches per step
        per_launch = gb[dom] * pts_substeps / groups[dom]["launches_per_step"]
        achieved = per_launch / (groups[dom]["avg_launch_us"] * 1e-6) / 1e9
    else:
        per_launch, achieved = None, None

    # CPU baseline: the oracle's W3SRCE on a bounded sample of the same points.
    import oracle as O
    cores = cpu_cores()
    tab = op.tables
    sample = min(npts, cores * 32)
    O.srce_batch(c, tab, nthreads=cores, npts=cores)  # warm
    reps, c0 = 0, time.perf_counter()
    while True:
        O.srce_batch(c, tab, nthreads=cores, npts=sample)
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
        "data": "synthetic (seeded JONSWAP spectra, Weibull winds, log-uniform depths)",
        "config": {
            "workload": "full ST4 source-term step (W3SRCE: W3SIN4 + W3SDS4 + W3SNL1 + W3SLN1 + "
                        f"W3SBT1 + W3SDB1, semi-implicit sub-steps to DTG): {npts} points, NK=32, "
                        "NTH=36, DTG=300 s, IT=1",
            "npts": npts, "nk": nk, "nth": nth, "dtg_s": fl["dtg"], "parallelism": "single GPU",
            "substeps_max": substeps, "point_substeps_per_step": pts_substeps,
            "wall_s": wall, "kernel_groups": groups,
            "width_thresholds_autotuned": tuned,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": measured_traffic("srce_step_kernel_hbm_traffic") if npts == 9520 else None,
            "kernel": groups[dom]["kernel"], "group": dom, "bytes_per_launch": per_launch,
            "step_bytes_per_point": SRCE_STEP_BYTES,
            "step_achieved_gbs": SRCE_STEP_BYTES * npts / (ms_step * 1e-3) / 1e9,
        },
        "cpu_baseline": {
            "value": cpu_value, "unit": "grid-point-steps/s", "cores": cores, "kind": "port",
            "sample": f"{sample} of the {npts} points x {reps} repetitions, {cpu_s:.1f} s, "
                      f"oracle W3SRCE (C, -O3, OpenMP over points)",
        },
    }


# W3XYP3: bytes a cell-component sub-step must move at least (read and write VQ, the two CFL
# numbers of its faces); W3KTP3: per point-call, VA in and out, CG and WN, the 10 scalars.
PROP_CELL_BYTES = 16


def bench_prop(args):
    import numpy as np
    import torch

    import oracle as O
    from ww3gpu import prop, prop_case as pc
    from ww3gpu.snl1 import init_runtime

    init_runtime()
    nx, ny, nk, nth = 480, 360, 32, 24
    c = pc.make_xyp3(nx=nx, ny=ny, nk=nk, nth=nth, seed=2024, land=0.30, dtg=900.0, cfl=0.8,
                     obstructions=0.02)
    nsea, nspec = c["nsea"], c["nspec"]
    ntloc = O.xyp3_ntloc(c)
    d_va = torch.from_numpy(c["va"].copy()).cuda()
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        prop.xyp3_batch_dev(c, d_va, stream)
    gpu_ms, wall = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        prop.xyp3_batch_dev(c, d_va, stream)
        wall.append(time.perf_counter() - t0)
        gpu_ms.append(prop.stats()["xyp3_ms"])
    st = prop.stats()
    cells = st["xyp3_cell_substeps"]
    ms = float(np.median(gpu_ms))
    value = cells / (ms * 1e-3)
    achieved = PROP_CELL_BYTES * cells / (ms * 1e-3) / 1e9
    # W3KTP3 over the same sea points, one call per part
    k = pc.make_ktp3(nsea=nsea, nk=nk, nth=nth, seed=2025, ntimes=1)
    kva = torch.from_numpy(k["va"].copy()).cuda()
    kth = torch.zeros(nsea, dtype=torch.float32, device="cuda")
    kk = torch.zeros(nsea, dtype=torch.float32, device="cuda")
    for _ in range(args.warmup):
        prop.ktp3_batch_dev(k, kva, kth, kk, stream)
    kms = []
    for _ in range(args.steps):
        prop.ktp3_batch_dev(k, kva, kth, kk, stream)
        kms.append(prop.stats()["ktp3_ms"])
    kt_ms = float(np.median(kms))
    kt_bytes = 4 * (2 * nspec + 2 * (nk + 2) + 10)
    # CPU baseline: the oracle's W3XYP3 on a bounded sample of the components
    cores = cpu_cores()
    sub = dict(c)
    on = np.zeros(nspec, np.int32)
    pick = np.arange(0, nspec, max(1, nspec // cores))[:cores]
    on[pick] = 1
    sub["on"] = on
    reps, c0 = 0, time.perf_counter()
    while True:
        O.xyp3_batch(sub, nthreads=cores)
        reps += 1
        if time.perf_counter() - c0 > args.cpu_seconds:
            break
    cpu_s = time.perf_counter() - c0
    cpu_cells = int(sum(ntloc[i] for i in pick)) * nx * ny * reps
    return {
        "metric": "cell-component sub-steps/s for the W3XYP3 propagation step",
        "value": value,
        "unit": "cell-component-sub-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (smoothed random land, log-uniform depths, gamma-distributed spectra)",
        "config": {
            "workload": f"W3XYP3 + W3QCK3 over all {nspec} components (W3GATH/W3SCAT included), "
                        f"{nx}x{ny} lon-lat grid at 0.25 deg, NK={nk}, NTH={nth}, DTG=900 s, "
                        f"CFL 0.8, VA resident in HBM",
            "nsea": nsea, "nx": nx, "ny": ny, "nk": nk, "nth": nth,
            "cfl_substeps_max": st["xyp3_substeps"], "cell_component_substeps": cells,
            "call_wall_ms_median": 1e3 * float(np.median(wall)),
            "ktp3": {"ms_per_call": kt_ms, "point_calls_per_s": nsea / (kt_ms * 1e-3),
                     "achieved_gbs": kt_bytes * nsea / (kt_ms * 1e-3) / 1e9,
                     "workload": f"W3KTP3 + W3QCK1/W3QCK2 on {nsea} points, refraction, "
                                 f"wavenumber shift and currents on, NTIMES=1"},
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": "W3XYP3 call (gather, records, 5 passes per sub-step, store)",
            "bytes_per_cell_component_substep": PROP_CELL_BYTES,
        },
        "cpu_baseline": {
            "value": cpu_cells / cpu_s, "unit": "cell-component-sub-steps/s", "cores": cores,
            "kind": "port",
            "sample": f"{len(pick)} of the {nspec} components x {reps} calls, {cpu_s:.1f} s, "
                      f"oracle W3XYP3 (C, -O3, OpenMP over components)",
        },
    }


def surrogate_flops_per_point(nspec, hidden=512):
    # Multiply-adds of the four layers at their algorithmic widths (NSPEC + 1 inputs, NSPEC
    # outputs), two operations each; the padding to multiples of 128 is not counted.
    return 2 * ((nspec + 1) * hidden + 2 * hidden * hidden + hidden * nspec)


def bench_surrogate(args):
    """The learned W3SNL1 surrogate against the exact DIA on the same points.

    Trains the network on the device (spectra from the library's seeded generator, half of
    them bimodal swell + sea; targets from the exact DIA), then times both on --sur-npts
    held-out points and reports the surrogate's S_nl error against the DIA."""
    import numpy as np
    import torch

    from ww3gpu import synth
    from ww3gpu.snl1 import Snl1
    from ww3gpu.snlsur import SnlSurrogate

    nk, nth = 32, 36
    cfg, sig = synth.dia_config(nk, nth)
    dia = Snl1(5, cfg, sig)
    sur = SnlSurrogate(dia, seed=1)
    t0 = time.perf_counter()
    losses = sur.fit(args.sur_train, 512, args.sur_epochs, 5e-4, seed=5, bimodal=0.5)
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t0
    print(f"trained on {args.sur_train} DIA pairs x {args.sur_epochs} passes in {fit_s:.1f} s; "
          f"last-batch loss per pass {np.array2string(losses, precision=4)}", file=sys.stderr, flush=True)

    def errors(a, cg, kd):
        s = torch.empty_like(a)
        d = torch.empty_like(a)
        dia.device(a, cg, kd, s, d)
        p, f, nfb = sur.predict(a, cg, kd)
        torch.cuda.synchronize()
        rel = (((p - s) ** 2).sum(1).sqrt() / (s ** 2).sum(1).sqrt()).float()
        q = torch.quantile(rel[:200000], torch.tensor([0.5, 0.9, 0.99], device=rel.device)).tolist()
        glob = float((((p - s) ** 2).sum() / (s ** 2).sum()).sqrt())
        return {"rel_l2_global": glob, "rel_l2_point_median": q[0], "rel_l2_point_p90": q[1],
                "rel_l2_point_p99": q[2], "fallback_fraction": nfb / a.shape[0]}

    npts = args.sur_npts
    a, cg, kd = sur.synth(npts, seed=4242, bimodal=0.5)
    err_mixed = errors(a, cg, kd)
    _, _, an, cgn, kdn = synth.dia_workload(9520, nk, nth, seed=2024, log_depth=(10.0, 4000.0))
    err_c1 = errors(*(torch.from_numpy(x).cuda() for x in (an, cgn, kdn)))

    stream = torch.cuda.current_stream()
    s = torch.empty_like(a)
    d = torch.empty_like(a)
    for _ in range(args.warmup):
        dia.device(a, cg, kd, s, d)
        sur.predict(a, cg, kd)
    torch.cuda.synchronize()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record(stream)
    for _ in range(args.steps):
        dia.device(a, cg, kd, s, d)
    e1.record(stream)
    nfb = 0
    for _ in range(args.steps):
        _, _, nfb = sur.predict(a, cg, kd)
    e2.record(stream)
    torch.cuda.synchronize()
    dia_ms = e0.elapsed_time(e1) / args.steps
    sur_ms = e1.elapsed_time(e2) / args.steps
    fpp = surrogate_flops_per_point(nk * nth)
    achieved = fpp * npts / (sur_ms * 1e-3) / 1e12

    # CPU baseline: the oracle DIA (what the surrogate replaces) on a bounded sample.
    import oracle as O
    t = O.insnl1(nk, nth, cfg["xfr"], cfg["dth"], cfg["lam"], sig)
    cores = cpu_cores()
    sample = max(cores * 64, 1024)
    reps, c0 = 0, time.perf_counter()
    while True:
        O.w3snl1(cfg, t, sig, an[:sample], cgn[:sample], kdn[:sample], nthreads=cores)
        reps += 1
        if time.perf_counter() - c0 > args.cpu_seconds:
            break
    cpu_s = time.perf_counter() - c0
    return {
        "metric": METRIC,
        "value": npts / (sur_ms * 1e-3),
        "unit": "grid-point S_nl evaluations/s (surrogate, fallback included)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": sur_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic (seeded JONSWAP and bimodal swell + sea spectra, log-uniform depths)",
        "config": {
            "workload": f"W3SNL1 surrogate MLP (1153 -> 3 x 512 GELU -> 1152, bf16 matrix cores, fp32 "
                        f"accumulation) with the energy check and exact-DIA fallback, {npts} points, "
                        f"NK=32, NTH=36; a step is one S_nl evaluation of every point",
            "npts": npts, "train_pairs": args.sur_train, "train_passes": args.sur_epochs,
            "train_s": fit_s, "dia_ms_per_step": dia_ms, "speedup_vs_dia": dia_ms / sur_ms,
            "fallback_points": nfb, "error_vs_dia_mixed_sea": err_mixed,
            "error_vs_dia_jonswap_9520": err_c1, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "mfma", "achieved": achieved, "peak": 2500.0, "unit": "TFLOP/s",
            "frac": achieved / 2500.0, "traffic": None, "flops_per_point": fpp,
            "kernel": "whole surrogate call (four GEMM launches, normalisation, check, fallback)",
        },
        "cpu_baseline": {
            "value": sample * reps / cpu_s, "unit": "grid-point S_nl evaluations/s (exact DIA)",
            "cores": cores, "kind": "port",
            "sample": f"{sample} JONSWAP points x {reps} repetitions, {cpu_s:.1f} s, oracle W3SNL1 "
                      f"(C, -O3, OpenMP over points)",
        },
    }


def bench_copy(args):
    """A plain device copy (torch copy_ of 1 GiB): the calibration run for the HBM counters
    FETCH_SIZE and WRITE_SIZE, whose readings per launch are set against the 1 GiB each way
    this copy is known to move."""
    import torch

    n = 1 << 28  # float32: 1 GiB
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    return {
        "metric": "device copy bandwidth (HBM counter calibration)", "value": 2 * n * 4 / dt / 1e9,
        "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "torch copy_ of 1 GiB (1 GiB read and 1 GiB written per launch)",
                   "bytes_read_per_launch": n * 4, "bytes_written_per_launch": n * 4},
    }


def measured_traffic(name):
    """HBM bytes per launch of a kernel as measured offline by the two counter passes and the
    copy calibration (profiles/<name>.json, written from the rocprofv3 summaries there), or
    None."""
    path = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="srce", choices=["srce", "dia", "prop", "copy", "surrogate"])
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--npts", type=int, default=9520,
                    help="points of the W3SRCE workload (9520: the metric's configuration)")
    ap.add_argument("--autotune", action="store_true",
                    help="srce workload: measure the sub-step width thresholds first and use the fastest")
    ap.add_argument("--sur-npts", type=int, default=1 << 20, help="surrogate workload: points timed")
    ap.add_argument("--sur-train", type=int, default=8 << 20, help="surrogate workload: training pairs")
    ap.add_argument("--sur-epochs", type=int, default=12, help="surrogate workload: training passes")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(source terms shard over points with no exchange; see the design notes)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "dia":
        args.steps = 500 if args.steps is None else args.steps
        args.warmup = 50 if args.warmup is None else args.warmup
        res = bench_dia(args)
    elif args.workload == "copy":
        args.steps = 5 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        res = bench_copy(args)
    elif args.workload == "surrogate":
        args.steps = 10 if args.steps is None else args.steps
        args.warmup = 2 if args.warmup is None else args.warmup
        res = bench_surrogate(args)
    elif args.workload == "prop":
        args.steps = 10 if args.steps is None else args.steps
        args.warmup = 2 if args.warmup is None else args.warmup
        res = bench_prop(args)
    else:
        args.steps = 20 if args.steps is None else args.steps
        args.warmup = 3 if args.warmup is None else args.warmup
        res = bench_srce(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark of the tridsolver_amd hot path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload adi|plain] [--dtype f64|f32]

Prints ONE JSON line (rank 0).  Workloads (synthetic data, seeded):
  adi    ADI heat diffusion (apps/adi), 256^3, lambda = 1, a = c = -lambda,
         b = 1 + 2 lambda, Dirichlet-zero boundary, interior u0 ~ U(0,1).
         One step = one ADI time step: 7-point right-hand side, then x, y and
         z sweeps (z adds into u).  value = 3 * 256^3 unknowns per step / time.
  plain  batched scalar solves of 256^3 along solvedim 0, 1 and 2 (one
         step = the three solves), a,c ~ U(-1,-0.1), b = |a|+|c|+U(0.5,1.5),
         d ~ U(-1,1), seed 0.  value = 3 * 256^3 unknowns per step / time.
Inputs are resident in HBM before the timed region.  The timed region is
bracketed by torch.cuda.synchronize(); per-kernel durations come from HIP
events recorded on the library's stream around every launch.

"roofline" reports the dominant kernel's achieved algorithmic bandwidth
(bytes every launch must move at minimum, see ALGO_BYTES) against MI355X's
8 TB/s HBM peak.  "cpu_baseline" times the CPU oracle (a C/OpenMP port of
the reference CPU solver) on the host cores on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["adi", "plain"], default="adi")
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--autotune", type=int, default=1)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (profiler calibration)")
    return ap.parse_args()


def traffic_from_profiles(kernel_key: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (profiles/traffic.json)."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            t = json.load(f)
        v = t.get(kernel_key)
        return None if v is None else float(v["bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def plain_workload(args, torch, np, api, ta):
    n = args.n
    dims = [n, n, n]
    dt = np.float64 if args.dtype == "f64" else np.float32
    wsize = 8 if args.dtype == "f64" else 4
    rng = np.random.default_rng(0)
    shape = (n, n, n)
    a = rng.uniform(-1.0, -0.1, shape)
    c = rng.uniform(-1.0, -0.1, shape)
    b = np.abs(a) + np.abs(c) + rng.uniform(0.5, 1.5, shape)
    d = rng.uniform(-1.0, 1.0, shape)
    host = [np.ascontiguousarray(x.astype(dt)).ravel() for x in (a, b, c, d)]
    del a, b, c, d
    dev = [torch.from_numpy(x).to("cuda") for x in host]
    fn = api.tridDmtsvStridedBatch if dt == np.float64 else api.tridSmtsvStridedBatch
    unknowns = n ** 3
    # algorithmic bytes of one solve: read a, b, c, d and write d
    algo_bytes = 5 * wsize * unknowns
    configs = {}

    def step(events=None):
        for sd in (0, 1, 2):
            if events is not None:
                events[sd][0].record()
            st = fn(dev[0], dev[1], dev[2], dev[3], None, 3, sd, dims, dims, None, 0)
            if st != 0:
                raise RuntimeError(f"solve failed: {st} {ta.last_error()}")
            if events is not None:
                events[sd][1].record()
            configs[sd] = ta.last_config()

    return dict(step=step, kernels=("solvedim0", "solvedim1", "solvedim2"),
                unknowns_per_step=3 * unknowns, algo_bytes=[algo_bytes] * 3,
                host=host, dims=dims, configs=configs,
                workload=f"plain batched scalar solve {n}^3 along solvedim 0,1,2")


def adi_workload(args, torch, np, api, ta):
    from tridsolver_amd.adi import AdiPlan
    n = args.n
    dt = np.float64 if args.dtype == "f64" else np.float32
    tdt = torch.float64 if args.dtype == "f64" else torch.float32
    wsize = 8 if args.dtype == "f64" else 4
    u0 = np.zeros((n, n, n), dt)
    u0[1:-1, 1:-1, 1:-1] = np.random.default_rng(0).uniform(0, 1, (n - 2, n - 2, n - 2))
    u = torch.from_numpy(u0).to("cuda")
    du = torch.empty_like(u)
    plan = AdiPlan(n, n, n, tdt, 1.0, 1)
    pts = n ** 3
    fused = "fused" in plan.describe()
    # algorithmic bytes per launch: rhs (+x when fused) reads u and writes du;
    # the x and y sweeps read and write du; the z sweep reads du and u and
    # writes u
    if fused:
        stages = (4, 2, 3)
        names = ("rhs_x_fused", "sweep_y", "sweep_z_inc")
        algo = [2 * wsize * pts, 2 * wsize * pts, 3 * wsize * pts]
    else:
        stages = (0, 1, 2, 3)
        names = ("rhs", "sweep_x", "sweep_y", "sweep_z_inc")
        algo = [2 * wsize * pts, 2 * wsize * pts, 2 * wsize * pts, 3 * wsize * pts]

    def step(events=None):
        for i, g in enumerate(stages):
            if events is not None:
                events[i][0].record()
            plan.stage(u, du, g)
            if events is not None:
                events[i][1].record()

    def graph_ms(k):
        plan.capture(u, du, k)
        plan.replay()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        plan.replay(sync=0)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    return dict(step=step, kernels=names,
                unknowns_per_step=3 * pts, algo_bytes=algo, host=[u0], dims=[n, n, n],
                configs={i: plan.describe() for i in range(len(names))}, graph_ms=graph_ms,
                workload=f"ADI heat diffusion {n}^3, lambda=1, a=c=-lambda, b=1+2lambda, "
                         f"Dirichlet-zero, u0~U(0,1) seed 0 (apps/adi time step)")


def cpu_baseline_adi(args, np, host, dims, seconds):
    import oracle
    n = dims[0]
    u = host[0].ravel().copy()
    t_total = 0.0
    steps = 0
    while t_total < seconds and steps < 20:
        t0 = time.perf_counter()
        assert oracle.adi(u, n, n, n, 1, 1.0, 1) == 0
        t_total += time.perf_counter() - t0
        steps += 1
    return dict(value=3 * n ** 3 * steps / t_total / 1e9, unit="Gunknowns/s",
                cores=oracle.num_threads(), kind="port",
                sample=f"{steps} ADI time steps of the same {n}^3 workload, oracle C/OpenMP port "
                       f"of apps/adi/src/adi_orig.c (materialised coefficients, as the reference)")


def cpu_baseline_plain(args, np, host, dims, seconds):
    import oracle
    n = dims[0]
    # bounded sample: a slab of the same workload (all x-lines of nz_s planes
    # for solvedim 0 and 1, all z-lines of ny_s rows for solvedim 2)
    t_total = 0.0
    unk_total = 0
    reps = 0
    while t_total < seconds and reps < 50:
        for sd in (0, 1, 2):
            a, b, c, d = (x.copy() for x in host)
            t0 = time.perf_counter()
            assert oracle.mtsv(a, b, c, d, None, 3, sd, dims) == 0
            t_total += time.perf_counter() - t0
            unk_total += n ** 3
        reps += 1
    return dict(value=unk_total / t_total / 1e9, unit="Gunknowns/s", cores=oracle.num_threads(),
                kind="port",
                sample=f"{reps} x (solvedim 0,1,2 on the full {n}^3 grid), oracle C/OpenMP port")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    import tridsolver_amd as ta
    from tridsolver_amd import api

    if args.calibrate_copy:
        # reference kernel for rocprofv3 --pmc: reads and writes exactly 1 GiB
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": 1 << 30, "launches": 5}))
        return
    ta.set_stream(None)  # null stream == torch's default stream
    ta.set_autotune(args.autotune)
    W = (adi_workload if args.workload == "adi" else plain_workload)(args, torch, np, api, ta)
    for _ in range(max(args.warmup, 1)):
        W["step"]()
    torch.cuda.synchronize()

    nk = len(W["kernels"])
    ev = [[[torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)]
           for _ in range(nk)] for _ in range(args.steps)]
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t_start.record()
    for s in range(args.steps):
        W["step"](ev[s])
    t_end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = t_start.elapsed_time(t_end)
    per_kernel_ms = [sum(ev[s][k][0].elapsed_time(ev[s][k][1]) for s in range(args.steps))
                     for k in range(nk)]
    # whole-step time on the GPU clock (launch gaps included)
    step_ms = gpu_ms / args.steps
    value = W["unknowns_per_step"] / (step_ms * 1e-3) / 1e9
    avg = [t / args.steps for t in per_kernel_ms]
    dom = max(range(nk), key=lambda k: avg[k])
    achieved = W["algo_bytes"][dom] / (avg[dom] * 1e-3) / 1e9
    kernels = {}
    for k in range(nk):
        kernels[W["kernels"][k]] = {
            "avg_ms": round(avg[k], 5),
            "GB_s": round(W["algo_bytes"][k] / (avg[k] * 1e-3) / 1e9, 1),
            "frac": round(W["algo_bytes"][k] / (avg[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "config": W["configs"].get(k, ""),
        }
    out = {
        "metric": "Gunknowns/s + %HBM peak at 256^3 FP32/FP64",
        "value": round(value, 3),
        "unit": "Gunknowns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": W["workload"], "grid": f"{args.n}^3", "parallelism": "single GPU"},
        "roofline": {
            "bound": "hbm",
            "kernel": W["kernels"][dom],
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(f"{W['workload']}|{W['kernels'][dom]}|{args.dtype}"),
        },
        "kernels": kernels,
        "graph_ms_per_step": (round(W["graph_ms"](args.steps), 5) if "graph_ms" in W else None),
        "wall_s_timed_region": round(wall, 3),
        "gpu_ms_timed_region": round(gpu_ms, 3),
    }
    if not args.no_cpu:
        fnb = cpu_baseline_adi if args.workload == "adi" else cpu_baseline_plain
        out["cpu_baseline"] = fnb(args, np, W["host"], W["dims"], args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

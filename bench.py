#!/usr/bin/env python3
"""Benchmark of the tridsolver_amd hot path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload adi|plain] [--dtype f64|f32]
                    [--no-extras] [--no-cpu]

Prints ONE JSON line.  Headline workloads (synthetic data, seeded):
  adi    ADI heat diffusion (apps/adi), 256^3, lambda = 1, a = c = -lambda,
         b = 1 + 2 lambda, Dirichlet-zero boundary, interior u0 ~ U(0,1).
         One step = one ADI time step: 7-point right-hand side, then x, y and
         z sweeps (z adds into u).  value = 3 * 256^3 unknowns per step / time.
  plain  batched scalar solves of 256^3 along solvedim 0, 1 and 2 (one
         step = the three solves), a,c ~ U(-1,-0.1), b = |a|+|c|+U(0.5,1.5),
         d ~ U(-1,1), seed 0.  value = 3 * 256^3 unknowns per step / time.
Inputs are resident in HBM before the timed region.  The timed region is
bracketed by torch.cuda.synchronize(); per-kernel durations come from HIP
events that every launch in it records around its own execution on the
library's stream (tridSetLaunchEvents -> hipExtLaunchKernel), the kernel's
duration as the profiler's kernel trace reports it.

"roofline" reports the dominant kernel's achieved algorithmic bandwidth
(bytes every launch must move at minimum) against MI355X's 8 TB/s HBM peak;
"traffic" is the HBM bytes per launch that rocprofv3 counters measured for
that kernel in the same configuration (profiles/traffic.json; null when
this run's kernel configuration differs from the measured one).
"cpu_baseline" times the CPU oracle (a C/OpenMP port of the reference CPU
solver) on the host cores, OMP_PROC_BIND=close OMP_PLACES=cores, on the same
workload, and states every result's error against the FP64 oracle
(err_vs_fp64_oracle: normwise relative error); the oracle runs only in this
leg, after the GPU timing, never inside a timed region.  "extras" (same
line) holds the other measured workloads, each with its own CPU baseline and
error against the FP64 oracle:
  plain     256^3 solves per solvedim, FP64 and FP32;
  ragged    200,000 systems, lengths floor(exp(U(ln 8, ln 4097))) (seed 0),
            one persistent launch vs one strided-batch launch per distinct
            length on length-grouped copies (gather outside the timing),
            eager and captured in a HIP graph;
  vadi_grad 128^3 FP64, 16 steps, checkpoint every 4: loss + gradient vs the
            forward pass;
  block     65,536 systems x 128 block rows, M in {2,3,4,5,8}, FP64, and all
            five in one mixed-M launch;
  split     systems longer than one work-group (split-line solves): 8
            systems of 2^22 rows (solvedim 0) and 64 systems of 2^19 rows
            (solvedim 1), FP64 and FP32;
  adi_sizes the ADI step at 256^3 FP32 and 512^3 FP64 / FP32.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

# the CPU baseline's OpenMP threads: one per core, packed (set before any
# OpenMP runtime starts)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Gunknowns/s + %HBM peak at 256\u00b3 FP32/FP64; ragged & fwd+bwd vs plain"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["adi", "plain"], default="adi")
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--autotune", type=int, default=1)
    ap.add_argument("--no-extras", action="store_true", help="headline workload only")
    ap.add_argument("--no-baselines", action="store_true",
                    help="skip the ragged workload's per-length launches and graph (for "
                         "profiler runs that should hold the ragged kernels only)")
    ap.add_argument("--extras", default="plain,ragged,vadi,block,split,adi_sizes",
                    help="comma-separated subset of plain,ragged,vadi,block,split,adi_sizes")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (profiler calibration)")
    return ap.parse_args()


def kernel_of_config(cfg: str, dtype: str, inc: int, coef: str, nx: int = 0):
    """The kernel (template instance) that runs a configuration string of the
    library (tridGetLastConfig, an ADI plan's stage), as rocprofv3 names it.
    nx: the grid's x extent, for the fused right-hand side + x kernel (its
    last template argument, YS, is the number of 64-lane rows of 16-byte
    pieces in an x-line when a wave holds two or more lines, else 0)."""
    import re
    T = "double" if dtype == "f64" else "float"
    # the last template argument is the cache policy (0 non-temporal, 1 ordinary,
    # 2 non-temporal coefficients with ordinary d and u)
    m = re.search(r"cc<R(\d+),cps\d+((?:,\w+)*)>", cfg)
    if m:
        fl = m.group(2).split(",")
        ns = 2 if "pair" in fl else 1
        return f"k_cc_sweep<{T}, {inc}, {m.group(1)}, unsigned int, {ns}, {int('cached' in fl)}>"
    m = re.search(r"part<\w+,o(\d),R(\d+),cps(\d+),spb\d+((?:,\w+)*)>", cfg)
    if m:
        fl = m.group(4).split(",")
        staged = "staged" in fl
        vec = 4 if "gstaged" in fl else (3 if staged and int(m.group(3)) > 64 else
                                         (2 if staged else (1 if "vec" in fl else 0)))
        cp = 1 if "cached" in fl else (2 if "mixed" in fl else 0)
        return (f"k_part<{T}, {inc}, {m.group(1)}, {m.group(2)}, {vec}, unsigned int, "
                f"{coef}<{T}>, {cp}>")
    m = re.search(r"rhs\+x\(fused,R(\d+)((?:,\w+)*)\)", cfg)
    if m:
        fl = m.group(2).split(",")
        R, E = int(m.group(1)), (2 if T == "double" else 4)
        ys = nx // (64 * E) if nx and nx % (64 * E) == 0 else 0
        ys = ys if ys in (1, 2, 4) and ys < R // E else 0
        return f"k_adi_rhs_x<{T}, {R}, {int('cached' in fl)}, {int('cc' in fl)}, {ys}>"
    return None


def traffic_from_profiles(key: str):
    """HBM bytes per launch of a kernel as measured by rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE (calibrated against a plain device copy), from profiles/traffic.json;
    keys are "<workload><n>|<dtype>|<kernel>" (profiles/traffic_counters.md).
    Returns (bytes or None, note)."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            t = json.load(f)
        v = t.get(key)
        if v is None:
            return None, f"no counter measurement of {key.split('|')[-1]} on this workload"
        return float(v["bytes_per_launch"]), f"rocprofv3 counters of {v['kernel']} ({v.get('note', '')})"
    except (OSError, ValueError, KeyError):
        return None, "profiles/traffic.json unreadable"


def stage_config(desc: str, stage: str) -> str:
    """The configuration of one ADI stage from the plan's description
    ("rhs+x(fused,R8)<double> x:... y:... z:...")."""
    toks = desc.split()
    pre = {"sweep_x": "x:", "sweep_y": "y:", "sweep_z_inc": "z:"}.get(stage)
    if pre is None:
        return toks[0] if toks else ""
    for t in toks:
        if t.startswith(pre):
            return t
    return ""


def kernel_events(ta, pair):
    """The library's next kernel launch records this (start, stop) pair of
    torch events around its own execution (tridSetLaunchEvents: HIP
    timestamps the dispatch packet itself)."""
    st = ta.lib().tridSetLaunchEvents(pair[0].cuda_event, pair[1].cuda_event)
    if st:
        raise RuntimeError(ta.last_error())


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
    kernel_cfg = {}

    def step(events=None):
        for sd in (0, 1, 2):
            if events is not None:
                kernel_events(ta, events[sd])
            st = fn(dev[0], dev[1], dev[2], dev[3], None, 3, sd, dims, dims, None, 0)
            if st != 0:
                raise RuntimeError(f"solve failed: {st} {ta.last_error()}")
            kernel_cfg[sd] = ta.last_config()

    def result():
        """one step (the three solves in turn) from the original d, on the host"""
        dev[3].copy_(torch.from_numpy(host[3]).to("cuda"))
        step()
        torch.cuda.synchronize()
        return dev[3].cpu().numpy()

    return dict(step=step, kernels=("solvedim0", "solvedim1", "solvedim2"),
                unknowns_per_step=3 * unknowns, algo_bytes=[algo_bytes] * 3,
                host=host, dims=dims, kernel_cfg=kernel_cfg, result=result,
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
    u0_dev = u.clone()
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
                kernel_events(ta, events[i])
            plan.stage(u, du, g)

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

    def result(k):
        """k eager steps (the timed stages) from u0, on the host"""
        u.copy_(u0_dev)
        for _ in range(k):
            step()
        torch.cuda.synchronize()
        return u.cpu().numpy()

    return dict(step=step, kernels=names,
                unknowns_per_step=3 * pts, algo_bytes=algo, host=[u0], dims=[n, n, n],
                kernel_cfg={}, describe=plan.describe, graph_ms=graph_ms, result=result,
                workload=f"ADI heat diffusion {n}^3, lambda=1, a=c=-lambda, b=1+2lambda, "
                         f"Dirichlet-zero, u0~U(0,1) seed 0 (apps/adi time step)")


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def rel_err(np, expected, got) -> float:
    """normwise relative error ||got - expected|| / ||expected|| in FP64"""
    e = np.asarray(expected, dtype=np.float64).ravel()
    g = np.asarray(got, dtype=np.float64).ravel()
    den = float(np.linalg.norm(e))
    return float(np.linalg.norm(g - e) / (den if den > 0 else 1.0))


def cpu_meta(oracle) -> dict:
    return {"cores": oracle.num_threads(), "cpu_model": cpu_model(),
            "omp": f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')} "
                   f"OMP_PLACES={os.environ.get('OMP_PLACES')}"}


def cpu_baseline_adi(np, u0, gpu_uT, n, steps):
    """The oracle's ADI (C/OpenMP port of apps/adi/src/adi_orig.c, with
    materialised coefficients as the reference) from the same u0 over the
    same steps, one call per step on a reused work buffer; value from the
    best step.  Also the GPU result's error against it."""
    import oracle
    u = np.ascontiguousarray(u0, dtype=np.float64).ravel().copy()
    ws = np.empty(4 * u.size, np.float64)
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        assert oracle.adi(u, n, n, n, 1, 1.0, 1, ws=ws) == 0
        times.append(time.perf_counter() - t0)
    best = min(times)
    return dict(value=3 * n ** 3 / best / 1e9, unit="Gunknowns/s", **cpu_meta(oracle),
                kind="port",
                GB_s_9words=round(9 * 8 * n ** 3 / best / 1e9, 2),
                mean_step_ms=round(1e3 * sum(times) / len(times), 2),
                best_step_ms=round(1e3 * best, 2),
                err_vs_fp64_oracle=rel_err(np, u, gpu_uT),
                sample=f"{steps} ADI time steps of the same {n}^3 workload from the same u0 "
                       f"(FP64; value from the fastest step, all steps checked), oracle C/OpenMP "
                       f"port of apps/adi/src/adi_orig.c (materialised coefficients, as the "
                       f"reference)")


def cpu_baseline_plain_step(np, host, dims, gpu_d, reps=5):
    """The plain headline step on the oracle: the three solves in turn on the
    same batch (FP64), best of `reps`, and the GPU step's error against it."""
    import oracle
    n = dims[0]
    h = [np.ascontiguousarray(x, dtype=np.float64) for x in host]
    times = []
    for _ in range(reps):
        d = h[3].copy()
        t0 = time.perf_counter()
        for sd in (0, 1, 2):
            assert oracle.mtsv(h[0], h[1], h[2], d, None, 3, sd, dims) == 0
        times.append(time.perf_counter() - t0)
    best = min(times)
    return dict(value=3 * n ** 3 / best / 1e9, unit="Gunknowns/s", **cpu_meta(oracle),
                kind="port", GB_s=round(15 * 8 * n ** 3 / best / 1e9, 2),
                err_vs_fp64_oracle=rel_err(np, d, gpu_d),
                sample=f"solvedim 0, 1, 2 in turn on the full {n}^3 batch (FP64), oracle "
                       f"C/OpenMP port of scalar/src/cpu/trid_cpu.cpp, best of {reps}")


def cpu_baseline_plain(np, host, dims, gpu_out, reps=5):
    """Oracle strided solves (C/OpenMP port of scalar/src/cpu/trid_cpu.cpp)
    of the same 256^3 batch along each solvedim, best of `reps`, and the GPU
    results' errors against the FP64 oracle."""
    import oracle
    n = dims[0]
    out = dict(unit="Gunknowns/s", **cpu_meta(oracle), kind="port",
               sample=f"the full {n}^3 batch per solvedim, best of {reps}")
    h64 = [np.ascontiguousarray(x, dtype=np.float64) for x in host]
    for sd in (0, 1, 2):
        times = []
        for _ in range(reps):
            d = h64[3].copy()
            t0 = time.perf_counter()
            assert oracle.mtsv(h64[0], h64[1], h64[2], d, None, 3, sd, dims) == 0
            times.append(time.perf_counter() - t0)
        best = min(times)
        out[f"solvedim{sd}"] = {"ms": round(1e3 * best, 2),
                                "Gunknowns_s": round(n ** 3 / best / 1e9, 3),
                                "GB_s": round(5 * 8 * n ** 3 / best / 1e9, 2)}
        for name, got in gpu_out.get(sd, {}).items():
            out[f"solvedim{sd}"][f"err_{name}_vs_fp64_oracle"] = rel_err(np, d, got)
    return out


# ---------------------------------------------------------------- workloads


def ragged_lengths(nsys=200_000, seed=0):
    """lengths log-uniform in [8, 4096]: floor(exp(U(ln 8, ln 4097)))"""
    import numpy as np
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(np.log(8), np.log(4097), nsys))).astype(np.int64)


def ragged_arrays(lengths, seed=0):
    """int64 offsets and FP64 a, b, c, d (a, c ~ U(-1,-0.1), b = |a|+|c|+U(0.5,1.5),
    d ~ U(-1,1); drawn in the order a, c, b, d from numpy PCG64 `seed`)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    n = int(off[-1])
    a = rng.uniform(-1.0, -0.1, n)
    c = rng.uniform(-1.0, -0.1, n)
    b = np.abs(a) + np.abs(c) + rng.uniform(0.5, 1.5, n)
    d = rng.uniform(-1.0, 1.0, n)
    return off, [a, b, c, d]


def fourier_kappa(n, seed, modes=8, amp=0.3):
    """kappa = exp(amp * g), g = sum of `modes` Fourier modes
    A_m cos(2 pi k_m . x + phi_m), k_m integer in [-3, 3]^3, A_m ~ N(0, 1/modes),
    phi_m ~ U(0, 2 pi), x = grid index / n; numpy PCG64 seed `seed`.
    Returned with shape (nz, ny, nx), x fastest."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ks = rng.integers(-3, 4, (modes, 3))
    A = rng.normal(0, 1 / np.sqrt(modes), modes)
    ph = rng.uniform(0, 2 * np.pi, modes)
    z, y, x = np.meshgrid(*(np.arange(n) / n,) * 3, indexing="ij")
    g = np.zeros((n, n, n))
    for m in range(modes):
        g += A[m] * np.cos(2 * np.pi * (ks[m, 0] * x + ks[m, 1] * y + ks[m, 2] * z) + ph[m])
    return np.exp(amp * g)


def _timed(torch, fn, reps):
    """mean GPU milliseconds of fn() over `reps` calls, each bracketed by
    events on the null stream (the library's stream here)"""
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    torch.cuda.synchronize()
    for e0, e1 in ev:
        e0.record()
        fn()
        e1.record()
    torch.cuda.synchronize()
    return sum(e0.elapsed_time(e1) for e0, e1 in ev) / reps


def _kernel_timed(torch, ta, fn, reps):
    """mean GPU milliseconds of the one kernel fn() launches, over `reps`
    calls, from events that each launch records around its own execution"""
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for e0, e1 in ev:
        e0.record()
        e1.record()
    torch.cuda.synchronize()
    for pair in ev:
        kernel_events(ta, pair)
        fn()
    torch.cuda.synchronize()
    return sum(e0.elapsed_time(e1) for e0, e1 in ev) / reps


def extra_plain(torch, np, ta, cpu, n=256, reps=10):
    lib = ta.lib()
    from tridsolver_amd.api import _iv
    rng = np.random.default_rng(0)
    shape = (n, n, n)
    a = rng.uniform(-1.0, -0.1, shape)
    c = rng.uniform(-1.0, -0.1, shape)
    b = np.abs(a) + np.abs(c) + rng.uniform(0.5, 1.5, shape)
    d = rng.uniform(-1.0, 1.0, shape)
    dims = _iv([n, n, n])
    out = {}
    gpu_out = {0: {}, 1: {}, 2: {}}
    for name, dt, ws, fn in (("f64", np.float64, 8, lib.tridDmtsvStridedBatch),
                             ("f32", np.float32, 4, lib.tridSmtsvStridedBatch)):
        dev = [torch.from_numpy(x.astype(dt)).to("cuda") for x in (a, b, c, d)]
        d0 = dev[3].clone()
        p = [t.data_ptr() for t in dev]
        for sd in (0, 1, 2):
            def solve():
                st = fn(p[0], p[1], p[2], p[3], None, 3, sd, dims, dims, None, 0)
                if st:
                    raise RuntimeError(ta.last_error())
            solve()
            solve()
            ms = _kernel_timed(torch, ta, solve, reps)
            gbs = 5 * ws * n ** 3 / (ms * 1e-3) / 1e9
            out[f"{name}_solvedim{sd}"] = {
                "ms": round(ms, 5), "Gunknowns_s": round(n ** 3 / (ms * 1e-3) / 1e9, 2),
                "GB_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                "config": ta.last_config()}
            # one solve from the original right-hand side, for the error check
            dev[3].copy_(d0)
            solve()
            gpu_out[sd][name] = dev[3].cpu().numpy()
            dev[3].copy_(d0)
        del dev, d0
    res = {"workload": f"{n}^3, one solve per solvedim, inputs as the plain workload",
           "bytes_model": "5 words per unknown (read a, b, c, d; write d)", **out}
    if cpu:
        cb = cpu_baseline_plain(np, [a, b, c, d], [n, n, n], gpu_out)
        for sd in (0, 1, 2):
            for name in ("f64", "f32"):
                res[f"{name}_solvedim{sd}"]["err_vs_fp64_oracle"] = \
                    cb[f"solvedim{sd}"].pop(f"err_{name}_vs_fp64_oracle")
        res["cpu_baseline"] = cb
    return res


def extra_split(torch, np, ta, cpu, reps=10):
    """Split-line solves: 2^25 unknowns in few, very long systems."""
    lib = ta.lib()
    from tridsolver_amd.api import _iv
    out = {"workload": "2^25 unknowns: [2^22 rows x 8] solvedim 0, [64 x 2^19 rows] solvedim 1",
           "bytes_model": "5 words per unknown (one-pass minimum); the split-line solve "
                          "moves 9 (a, b, c, d read twice, d written once)"}
    rng = np.random.default_rng(0)
    n = 1 << 25
    a = rng.uniform(-1.0, -0.1, n)
    c = rng.uniform(-1.0, -0.1, n)
    b = np.abs(a) + np.abs(c) + rng.uniform(0.5, 1.5, n)
    d = rng.uniform(-1.0, 1.0, n)
    cases = (("sd0_4Mx8", [1 << 22, 8], 0), ("sd1_64x512K", [64, 1 << 19], 1))
    gpu_out = {}
    for name, dt, ws, fn in (("f64", np.float64, 8, lib.tridDmtsvStridedBatch),
                             ("f32", np.float32, 4, lib.tridSmtsvStridedBatch)):
        dev = [torch.from_numpy(x.astype(dt)).to("cuda") for x in (a, b, c, d)]
        d0 = dev[3].clone()
        p = [t.data_ptr() for t in dev]
        for tag, dims, sd in cases:
            dv = _iv(dims)

            def solve():
                st = fn(p[0], p[1], p[2], p[3], None, 2, sd, dv, dv, None, 0)
                if st:
                    raise RuntimeError(ta.last_error())
            solve()
            solve()
            ms = _timed(torch, solve, reps)
            gbs = 5 * ws * n / (ms * 1e-3) / 1e9
            out[f"{name}_{tag}"] = {
                "ms": round(ms, 5), "Gunknowns_s": round(n / (ms * 1e-3) / 1e9, 2),
                "GB_s_5words": round(gbs, 1), "frac_5words": round(gbs / HBM_PEAK_GBS, 4),
                "frac_9words": round(gbs * 9 / 5 / HBM_PEAK_GBS, 4),
                "config": ta.last_config()}
            dev[3].copy_(d0)
            solve()
            gpu_out[(name, tag)] = dev[3].cpu().numpy()
            dev[3].copy_(d0)
        del dev, d0
    if cpu:
        import oracle
        cb = dict(unit="Gunknowns/s", **cpu_meta(oracle), kind="port",
                  sample="the same 2^25 unknowns per layout (FP64), best of 3")
        for tag, dims, sd in cases:
            times = []
            for _ in range(3):
                dd = d.copy()
                t0 = time.perf_counter()
                assert oracle.mtsv(a, b, c, dd, None, 2, sd, dims) == 0
                times.append(time.perf_counter() - t0)
            best = min(times)
            cb[tag] = {"ms": round(1e3 * best, 2), "Gunknowns_s": round(n / best / 1e9, 3),
                       "GB_s": round(5 * 8 * n / best / 1e9, 2)}
            for name in ("f64", "f32"):
                out[f"{name}_{tag}"]["err_vs_fp64_oracle"] = rel_err(np, dd, gpu_out[(name, tag)])
        out["cpu_baseline"] = cb
    return out


RAGGED_BASELINES = True  # --no-baselines sets False


def extra_ragged(torch, np, ta, cpu, reps=10):
    import ctypes as C
    from tridsolver_amd.ragged import RaggedPlan
    lib = ta.lib()
    lengths = ragged_lengths()
    off, host = ragged_arrays(lengths)
    n = int(off[-1])
    nsys = len(lengths)
    # length-grouped copy: systems sorted by length (stable), contiguous
    order = np.argsort(lengths, kind="stable")
    slen = lengths[order]
    soff = np.concatenate([[0], np.cumsum(slen)])
    idx = (np.arange(n, dtype=np.int64) - np.repeat(soff[:-1], slen)
           + np.repeat(off[:-1][order], slen))
    uniq, first = np.unique(slen, return_index=True)
    counts = np.diff(np.append(first, nsys))
    out = {"workload": f"{nsys} systems, lengths floor(exp(U(ln 8, ln 4097))) seed 0, "
                       f"{n} unknowns, {len(uniq)} distinct lengths",
           "bytes_model": "5 words per unknown + 8 B per system (offsets)"}
    t_off = torch.from_numpy(off).to("cuda")
    t_idx = torch.from_numpy(idx).to("cuda")
    gpu_out = {}
    for name, tdt, ws, fn in (("f64", torch.float64, 8, lib.tridDmtsvStridedBatch),
                              ("f32", torch.float32, 4, lib.tridSmtsvStridedBatch)):
        dev = [torch.from_numpy(x).to("cuda").to(tdt) for x in host]
        d0 = dev[3].clone()
        t0 = time.perf_counter()
        plan = RaggedPlan(t_off, tdt)
        t_plan = time.perf_counter() - t0
        srt = [x[t_idx] for x in dev]
        sd0 = srt[3].clone()
        # agreement of the two paths on the same systems
        plan.solve(dev[0], dev[1], dev[2], dev[3])
        gpu_out[name] = dev[3].cpu().numpy()
        p = [x.data_ptr() for x in srt]
        dims = [(C.c_int * 2)(int(L), int(k)) for L, k in zip(uniq, counts)]
        base = [int(s) * ws for s in soff[first]]

        def per_length():
            for j in range(len(uniq)):
                o = base[j]
                st = fn(p[0] + o, p[1] + o, p[2] + o, p[3] + o, None, 2, 0, dims[j], dims[j],
                        None, 0)
                if st:
                    raise RuntimeError(ta.last_error())
        diff = t_eager = t_graph = None
        if RAGGED_BASELINES:
            per_length()
            torch.cuda.synchronize()
            diff = float((dev[3][t_idx] - srt[3]).abs().max() / srt[3].abs().max())
        dev[3].copy_(d0)
        srt[3].copy_(sd0)
        t_rag = _timed(torch, lambda: plan.solve(dev[0], dev[1], dev[2], dev[3], sync=0), reps)
        if RAGGED_BASELINES:
            t_eager = _timed(torch, per_length, 3)
            # the per-length launches captured once into a graph and replayed
            g = torch.cuda.CUDAGraph()
            s = torch.cuda.Stream()
            try:
                with torch.cuda.graph(g, stream=s):
                    ta.set_stream(s)
                    per_length()
                ta.set_stream(None)
                g.replay()
                torch.cuda.synchronize()
                t_graph = _timed(torch, g.replay, reps)
            finally:
                ta.set_stream(None)
            del g
        gbs = (5 * ws * n + 8 * (nsys + 1)) / (t_rag * 1e-3) / 1e9
        out[name] = {
            "ragged_ms": round(t_rag, 4), "Gunknowns_s": round(n / (t_rag * 1e-3) / 1e9, 2),
            "GB_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
            "per_length_eager_ms": None if t_eager is None else round(t_eager, 3),
            "per_length_graph_ms": None if t_graph is None else round(t_graph, 3),
            "speedup_vs_eager": None if t_eager is None else round(t_eager / t_rag, 2),
            "speedup_vs_graph": None if t_graph is None else round(t_graph / t_rag, 2),
            "max_rel_diff_ragged_vs_per_length": diff,
            "plan_create_s": round(t_plan, 3),
            "plan": plan.describe()}
        plan.close()
        del dev, srt
        torch.cuda.empty_cache()
    if cpu:
        import oracle
        times = []
        for _ in range(5):
            dd = host[3].copy()
            t0 = time.perf_counter()
            assert oracle.ragged(host[0], host[1], host[2], dd, None, off) == 0
            times.append(time.perf_counter() - t0)
        best = min(times)
        out["cpu_baseline"] = dict(
            unit="Gunknowns/s", value=round(n / best / 1e9, 3),
            GB_s=round((5 * 8 * n + 8 * (nsys + 1)) / best / 1e9, 2), ms=round(1e3 * best, 1),
            **cpu_meta(oracle), kind="port",
            sample="the same 200,000 systems (FP64), oracle restatement of trid_scalarD per "
                   "system (scalar/src/cpu/trid_cpu.cpp:597-602), best of 5")
        for name in ("f64", "f32"):
            out[name]["err_vs_fp64_oracle"] = rel_err(np, dd, gpu_out[name])
    return out


def extra_vadi(torch, np, ta, cpu, n=128, steps=16, ckpt=4, reps=5):
    from tridsolver_amd.autograd import vadi_loss_grad
    lib = ta.lib()
    kap_h = fourier_kappa(n, seed=0)
    kap = torch.from_numpy(kap_h).to("cuda")
    rng = np.random.default_rng(0)
    u0_h = rng.uniform(0, 1, (n, n, n))
    tgt_h = rng.uniform(0, 1, (n, n, n))
    u0 = torch.from_numpy(u0_h).to("cuda")
    tgt = torch.from_numpy(tgt_h).to("cuda")
    u = u0.clone()

    def fwd():
        u.copy_(u0)
        st = lib.tridVadiForward(kap.data_ptr(), u.data_ptr(), n, n, n, steps, 1.0, 0)
        if st:
            raise RuntimeError(ta.last_error())
    fwd()
    t_copy = _timed(torch, lambda: u.copy_(u0), reps)
    t_fwd = _timed(torch, fwd, reps) - t_copy
    res = {}

    def grad():
        res["r"] = vadi_loss_grad(kap, u0, tgt, steps, 1.0, ckpt)
    grad()
    t_grad = _timed(torch, grad, reps)
    loss, gk, gu = res["r"]
    out = {"workload": f"variable-diffusivity ADI {n}^3 FP64, {steps} steps, lambda 1, kappa = "
                       f"exp(0.3 g), g = 8 Fourier modes seed 0, checkpoint every {ckpt} steps",
           "forward_ms": round(t_fwd, 4), "loss_and_grad_ms": round(t_grad, 4),
           "grad_over_forward": round(t_grad / t_fwd, 3),
           "forward_Gunknowns_s": round(3 * n ** 3 * steps / (t_fwd * 1e-3) / 1e9, 2),
           "loss": loss}
    if cpu:
        import oracle
        gk_h, gu_h = gk.cpu().numpy().ravel(), gu.cpu().numpy().ravel()
        t0 = time.perf_counter()
        oracle.vadi_forward(kap_h.ravel(), u0_h.ravel(), n, steps, 1.0)
        t_cf = time.perf_counter() - t0
        t0 = time.perf_counter()
        eloss, egk, egu = oracle.vadi_grad(kap_h.ravel(), u0_h.ravel(), tgt_h.ravel(), n, steps,
                                           1.0)
        t_cg = time.perf_counter() - t0
        out["err_loss_vs_fp64_oracle"] = abs(loss - eloss) / abs(eloss)
        out["err_grad_kappa_vs_fp64_oracle"] = rel_err(np, egk, gk_h)
        out["err_grad_u0_vs_fp64_oracle"] = rel_err(np, egu, gu_h)
        out["cpu_baseline"] = dict(
            unit="ms", forward_ms=round(1e3 * t_cf, 1), loss_and_grad_ms=round(1e3 * t_cg, 1),
            **cpu_meta(oracle), kind="port",
            sample="the same forward pass and loss + gradient (FP64, the oracle's reverse mode "
                   "without checkpoints), one run each")
    return out


def block_inputs(torch, N, P, M, gen):
    """Off-diagonal blocks U(-1,1)/M; B = U(-1,1)/M + (row-sum |A|+|C| + 1) I;
    d ~ U(-1,1); A_0 and C_{N-1} zero; reference GPU layout (N, P, M, M)."""
    def U(*shape):
        return (torch.rand(*shape, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1)
    A = U(N, P, M, M) / M
    Cm = U(N, P, M, M) / M
    A[0].zero_()
    Cm[N - 1].zero_()
    B = U(N, P, M, M) / M
    B += torch.diag_embed(A.abs().sum(-1) + Cm.abs().sum(-1) + 1.0)
    d = U(N, P, M)
    return A, B, Cm, d


def extra_block(torch, np, ta, cpu, N=128, P=65536, reps=5, sample=2048):
    from tridsolver_amd.block import btsv, btsv_adjoint, btsv_mixed
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    out = {"workload": f"{P} systems x {N} block rows, FP64, blocks as the block tests",
           "bytes_model": "(3M^2 + 2M) words per block row (read A, B, C, d; write u); "
                          "ref_model adds the C'/d' scratch round trip "
                          "((5N-3)M^2 + (5N-1)M words per system); adjoint: the transposed "
                          "solve's ref_model plus reading u, g and writing grad_A, grad_B, "
                          "grad_C, grad_d ((3M^2 + 3M) words per block row)"}
    groups = []
    t_sep = 0.0
    samples = {}
    for M in (2, 3, 4, 5, 8):
        A, B, Cm, d = block_inputs(torch, N, P, M, gen)
        u = torch.empty_like(d)
        Cs = torch.empty_like(A)
        ds = torch.empty_like(d)
        btsv(A, B, Cm, d, u, Cs, ds)
        ms = _timed(torch, lambda: btsv(A, B, Cm, d, u, Cs, ds, sync=0), reps)
        t_sep += ms
        algo = (3 * M * M + 2 * M) * N * P * 8
        ref = ((5 * N - 3) * M * M + (5 * N - 1) * M) * P * 8
        out[f"M{M}"] = {"ms": round(ms, 4),
                        "Gunknowns_s": round(N * P * M / (ms * 1e-3) / 1e9, 2),
                        "GB_s": round(algo / (ms * 1e-3) / 1e9, 1),
                        "frac": round(algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "ref_model_GB_s": round(ref / (ms * 1e-3) / 1e9, 1),
                        "ref_model_frac": round(ref / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        if M == 3:
            # the block adjoint (transposed solve plus the outer-product
            # gradients) on the same batch
            g = torch.rand(N, P, M, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
            btsv_adjoint(A, B, Cm, u, g)
            t_adj = _timed(torch, lambda: btsv_adjoint(A, B, Cm, u, g, sync=0), reps)
            adj = ref + (3 * M * M + 3 * M) * N * P * 8
            out["adjoint_M3"] = {"ms": round(t_adj, 4),
                                 "ref_model_GB_s": round(adj / (t_adj * 1e-3) / 1e9, 1),
                                 "ref_model_frac": round(adj / (t_adj * 1e-3) / 1e9
                                                         / HBM_PEAK_GBS, 4),
                                 "over_forward": round(t_adj / ms, 3)}
            del g
            torch.cuda.empty_cache()
        if cpu:
            samples[M] = [x[:, :sample].contiguous().cpu().numpy() for x in (A, B, Cm, d, u)]
        groups.append(dict(A=A, B=B, C=Cm, d=d, Cstar=Cs, dstar=ds, u=u))
    ms = _timed(torch, lambda: btsv_mixed(groups, sync=0), reps)
    out["mixed_all_M_one_launch_ms"] = round(ms, 4)
    out["separate_launches_sum_ms"] = round(t_sep, 4)
    if cpu:
        import oracle
        cb = dict(unit="Gunknowns/s", **cpu_meta(oracle), kind="port",
                  sample=f"the first {sample} of the {P} systems of each batch (FP64), "
                         f"oracle block Thomas (block/src/blktrid_cpu.cpp:388-504 with the "
                         f"index fix), best of 5")
        for gi, M in enumerate((2, 3, 4, 5, 8)):
            A_h, B_h, C_h, d_h, u_h = samples[M]
            times = []
            for _ in range(5):
                t0 = time.perf_counter()
                ref_u = oracle.btsv(A_h, B_h, C_h, d_h, N, sample, M, layout=0)
                times.append(time.perf_counter() - t0)
            best = min(times)
            algo = (3 * M * M + 2 * M) * N * sample * 8
            cb[f"M{M}"] = {"ms": round(1e3 * best, 2),
                           "Gunknowns_s": round(N * sample * M / best / 1e9, 3),
                           "GB_s": round(algo / best / 1e9, 2)}
            out[f"M{M}"]["err_vs_fp64_oracle"] = rel_err(np, ref_u, u_h)
            mixed_u = groups[gi]["u"][:, :sample].contiguous().cpu().numpy()
            out[f"M{M}"]["err_mixed_launch_vs_fp64_oracle"] = rel_err(np, ref_u, mixed_u)
        out["cpu_baseline"] = cb
    del groups
    torch.cuda.empty_cache()
    return out


def extra_adi_sizes(torch, np, ta, cpu, steps=10, check_steps=4):
    """The ADI time step at the other grid sizes and precisions of the ADI
    configuration: 256^3 FP32, 512^3 FP64 and FP32 (autotuned plans, eager
    steps and graph replay).  With the CPU leg, `check_steps` steps from u0
    are compared with the FP64 oracle's."""
    from tridsolver_amd.adi import AdiPlan
    out = {"workload": "ADI heat diffusion as the headline (lambda=1, a=c=-lambda, "
                       "b=1+2lambda, Dirichlet-zero, u0~U(0,1) seed 0), other grids",
           "bytes_model": "9 words per point per step (rhs: u in, du out; x, y: du in/out; "
                          "z: du, u in, u out)"}
    for n, dname in ((256, "f32"), (512, "f64"), (512, "f32")):
        tdt = torch.float64 if dname == "f64" else torch.float32
        ws = 8 if dname == "f64" else 4
        u = torch.zeros((n, n, n), dtype=tdt, device="cuda")
        g = torch.Generator(device="cuda").manual_seed(0)
        u[1:-1, 1:-1, 1:-1] = torch.rand((n - 2, n - 2, n - 2), generator=g, device="cuda",
                                         dtype=tdt)
        du = torch.empty_like(u)
        u0 = u.clone()
        plan = AdiPlan(n, n, n, tdt, 1.0, 1)
        plan.step(u, du, 3)  # tunes at first use
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        plan.step(u, du, steps, sync=0)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        plan.capture(u, du, steps)
        plan.replay()
        torch.cuda.synchronize()
        e0.record()
        plan.replay(sync=0)
        e1.record()
        torch.cuda.synchronize()
        gms = e0.elapsed_time(e1) / steps
        pts = n ** 3
        out[f"{n}^3_{dname}"] = {
            "ms_per_step": round(ms, 4), "graph_ms_per_step": round(gms, 4),
            "Gunknowns_s": round(3 * pts / (ms * 1e-3) / 1e9, 1),
            "graph_Gunknowns_s": round(3 * pts / (gms * 1e-3) / 1e9, 1),
            "frac_9words": round(9 * ws * pts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "config": plan.describe()}
        if cpu:
            import oracle
            u.copy_(u0)
            plan.step(u, du, check_steps)
            got = u.cpu().numpy()
            ref = u0.cpu().numpy().astype(np.float64).ravel()
            del u0
            ws_h = np.empty(4 * ref.size, np.float64)
            t0 = time.perf_counter()
            assert oracle.adi(ref, n, n, n, check_steps, 1.0, 1, ws=ws_h) == 0
            t_c = (time.perf_counter() - t0) / check_steps
            del ws_h
            out[f"{n}^3_{dname}"]["err_vs_fp64_oracle"] = rel_err(np, ref, got)
            out[f"{n}^3_{dname}"]["cpu_baseline"] = dict(
                unit="Gunknowns/s", value=round(3 * pts / t_c / 1e9, 3),
                ms_per_step=round(1e3 * t_c, 1), **cpu_meta(oracle), kind="port",
                sample=f"{check_steps} FP64 steps of the same grid from the same u0, mean")
        plan.close()
        del u, du
        torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1",
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
    global RAGGED_BASELINES
    RAGGED_BASELINES = not args.no_baselines
    ta.set_stream(None)  # null stream == torch's default stream
    ta.set_autotune(args.autotune)
    W = (adi_workload if args.workload == "adi" else plain_workload)(args, torch, np, api, ta)
    for _ in range(max(args.warmup, 1)):
        W["step"]()
    torch.cuda.synchronize()
    if "describe" in W:  # the plan tunes at its first use
        W["kernel_cfg"] = {k: stage_config(W["describe"](), W["kernels"][k])
                           for k in range(len(W["kernels"]))}
        W["plan"] = W["describe"]()

    nk = len(W["kernels"])
    ev = [[[torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)]
           for _ in range(nk)] for _ in range(args.steps)]
    # create the events (torch creates them at their first record) before the
    # launches record into them
    for s_ev in ev:
        for pair in s_ev:
            pair[0].record()
            pair[1].record()
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
            "config": W["kernel_cfg"].get(k, ""),
        }
    out = {
        "metric": METRIC,
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
            "traffic": None,
        },
        "kernels": kernels,
        "kernel_timing": "per launch, HIP events recorded by the launch itself "
                         "(hipExtLaunchKernel start/stop: kernel execution)",
        "graph_ms_per_step": (round(W["graph_ms"](args.steps), 5) if "graph_ms" in W else None),
        "wall_s_timed_region": round(wall, 3),
        "gpu_ms_timed_region": round(gpu_ms, 3),
    }
    dom_cfg = W["kernel_cfg"].get(dom, "")
    out["roofline"]["config"] = dom_cfg
    if "plan" in W:
        out["config"]["plan"] = W["plan"]
    kname = kernel_of_config(dom_cfg, args.dtype, 1 if W["kernels"][dom] == "sweep_z_inc" else 0,
                             "CoefConst" if args.workload == "adi" else "CoefArrays", args.n)
    out["roofline"]["kernel_name"] = kname
    out["roofline"]["traffic"], out["roofline"]["traffic_source"] = traffic_from_profiles(
        f"{args.workload}{args.n}|{args.dtype}|{kname}")
    if not args.no_cpu:
        # the GPU result of the same work the oracle runs, then the oracle
        # (outside every timed region)
        if args.workload == "adi":
            gpu_uT = W["result"](args.steps)
            out["cpu_baseline"] = cpu_baseline_adi(np, W["host"][0], gpu_uT, args.n, args.steps)
        else:
            gpu_d = W["result"]()
            out["cpu_baseline"] = cpu_baseline_plain_step(np, W["host"], W["dims"], gpu_d)
    if not args.no_extras:
        del W
        torch.cuda.empty_cache()
        fns = {"plain": extra_plain, "ragged": extra_ragged, "vadi": extra_vadi,
               "block": extra_block, "split": extra_split, "adi_sizes": extra_adi_sizes}
        out["extras"] = {}
        for k in args.extras.split(","):
            k = k.strip()
            if k:
                print(f"[bench] extra {k}", file=sys.stderr, flush=True)
                out["extras"][k] = fns[k](torch, np, ta, not args.no_cpu)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

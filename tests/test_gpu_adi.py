"""GPU parity of the ADI driver (include/trid_adi.h).

* reference mode (a = c = -lambda/2, b = 1 + lambda, u = 1 on the boundary)
  against the output files of the reference's own gold ADI
  (apps/adi/src/adi_orig.c, compiled unmodified; tests/golden/adi_ref_*);
* both modes against the oracle's ADI at several grid shapes and at the
  benchmark size 256^3 (FP64 and FP32, 20 steps);
* each stage against the oracle (the right-hand side bit-exactly);
* every y / z sweep layout (tridAdiPlanSetSweep) against the oracle, and the
  layouts that do not fit a line length rejected with a status;
* HIP-graph replays against eager execution, several replays with eager
  work and host synchronisation in between.
Tolerance: normwise relative error 1e-12 (FP64), 1e-5 (FP32).
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

import oracle
import tridsolver_amd as ta
from tridsolver_amd.adi import AdiPlan

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def torch_dt(dt):
    import torch
    return torch.float64 if dt == np.float64 else torch.float32


def ref_init(nx, ny, nz, dt):
    u = np.zeros((nz, ny, nx), dt)
    u[0], u[-1], u[:, 0], u[:, -1], u[:, :, 0], u[:, :, -1] = 1, 1, 1, 1, 1, 1
    return u


def rand_init(nx, ny, nz, dt, seed):
    """interior U(0,1) from the seed, Dirichlet-zero boundary"""
    u = np.zeros((nz, ny, nx), dt)
    r = np.random.default_rng(seed).uniform(0, 1, (nz - 2, ny - 2, nx - 2)) if min(nx, ny, nz) > 2 \
        else np.zeros((0, 0, 0))
    if r.size:
        u[1:-1, 1:-1, 1:-1] = r.astype(dt)
    return u


def run_gpu(u0, steps, mode, lam=1.0):
    import torch
    nz, ny, nx = u0.shape
    plan = AdiPlan(nx, ny, nz, torch_dt(u0.dtype), lam, mode)
    u = torch.from_numpy(u0.copy()).to("cuda")
    du = torch.empty_like(u)
    plan.step(u, du, steps)
    return u.cpu().numpy(), plan


def tol(dt):
    return 1e-12 if dt == np.float64 else 1e-5


def gold_cases():
    with open(os.path.join(GOLD, "golden_meta.json")) as f:
        meta = json.load(f)
    return sorted(k for k in meta if k.startswith("adi_ref_"))


@pytest.mark.parametrize("name", gold_cases())
def test_adi_reference_mode_vs_reference_gold(name):
    with open(os.path.join(GOLD, "golden_meta.json")) as f:
        m = json.load(f)[name]
    dt = np.float64 if m["dtype"] == "f64" else np.float32
    ref = np.fromfile(os.path.join(GOLD, name), dtype=dt)
    got, plan = run_gpu(ref_init(m["nx"], m["ny"], m["nz"], dt), m["iter"], mode=0, lam=m["lambda"])
    err = oracle.rel_err(ref, got.ravel())
    print(f"{name}: rel err vs reference gold {err:.2e} [{plan.describe()}]")
    assert err <= tol(dt)


# y and z lines of at most 512 rows take the factor-reuse sweeps (precomputed
# constant-coefficient factors); (6, 600, 5) and (5, 6, 700) exercise the
# generic sweep kernels along y and z; short lines (3, 5, 9 rows) the
# single-chunk and identity-padded factor layouts
# 512-point lines along each axis: the 512^3 configuration's kernels
ADI_SHAPES = [(32, 32, 32), (33, 17, 9), (64, 40, 24), (8, 200, 6), (128, 64, 48),
              (6, 600, 5), (5, 6, 700), (16, 3, 5), (20, 9, 3),
              (512, 24, 20), (20, 512, 24), (24, 20, 512)]


@pytest.mark.parametrize("shape", ADI_SHAPES)
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("mode", [0, 1])
def test_adi_vs_oracle(shape, dt, mode):
    nx, ny, nz = shape
    u0 = rand_init(nx, ny, nz, dt, seed=nx + ny + nz) if mode == 1 else ref_init(nx, ny, nz, dt)
    got, plan = run_gpu(u0, 5, mode)
    exp = u0.ravel().copy()
    assert oracle.adi(exp, nx, ny, nz, 5, 1.0, mode) == 0
    err = oracle.rel_err(exp, got.ravel())
    assert err <= tol(dt), f"{shape} mode {mode}: {err:.2e} [{plan.describe()}]"


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_adi_benchmark_size(dt):
    """256^3, 20 steps, lambda = 1, a = c = -lambda, b = 1 + 2 lambda,
    Dirichlet-zero boundary, interior U(0,1) from seed 0."""
    n = 256
    u0 = rand_init(n, n, n, dt, seed=0)
    got, plan = run_gpu(u0, 20, 1)
    exp = u0.ravel().copy()
    assert oracle.adi(exp, n, n, n, 20, 1.0, 1) == 0
    err = oracle.rel_err(exp, got.ravel())
    print(f"ADI 256^3 {np.dtype(dt).name} 20 steps: rel err vs oracle {err:.2e} [{plan.describe()}]")
    assert err <= tol(dt)


# every y / z sweep layout: rows 0 (partitioned sweep) or 8/16/32 (precomputed
# factors) x two lines per lane x XCD order, on lines of 24..512 rows
SWEEP_LAYOUTS = [(0, 0, 0)] + [(r, p, x) for r in (8, 16, 32) for p in (0, 1) for x in (0, 1)]


@pytest.mark.parametrize("shape", [(64, 40, 24), (24, 256, 130), (16, 20, 512), (10, 9, 7)])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_adi_every_sweep_layout(shape, dt):
    import torch
    nx, ny, nz = shape
    u0 = rand_init(nx, ny, nz, dt, seed=nx * ny)
    exp = u0.ravel().copy()
    assert oracle.adi(exp, nx, ny, nz, 3, 1.0, 1) == 0
    ran = 0
    for direction in (1, 2):
        n = ny if direction == 1 else nz
        for rows, pair, xcd in SWEEP_LAYOUTS:
            plan = AdiPlan(nx, ny, nz, torch_dt(dt), 1.0, 1)
            cps = -(-n // rows) if rows else 1
            cps = 1 << (cps - 1).bit_length()
            fits = rows == 0 or (64 * cps <= 1024 and (not pair or (rows <= 32 and
                                                                  64 * cps <= (512 if rows == 32 else 1024))))
            if not fits:
                with pytest.raises(ta.TridError):
                    plan.set_sweep(direction, rows, pair, xcd)
                continue
            plan.set_sweep(direction, rows, pair, xcd)
            u = torch.from_numpy(u0.copy()).to("cuda")
            plan.step(u, torch.empty_like(u), 3)
            what = f"{shape} dir {direction} layout {(rows, pair, xcd)} [{plan.describe()}]"
            err = oracle.rel_err(exp, u.cpu().numpy().ravel())
            assert err <= tol(dt), f"{what}: {err:.2e}"
            ran += 1
    print(f"{shape} {dt.__name__}: {ran} layouts")


@pytest.mark.parametrize("shape", [(64, 40, 24), (256, 24, 30), (32, 8, 5), (128, 16, 6), (512, 6, 5)])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_adi_every_fused_x_layout(shape, dt):
    """The fused right-hand side + x stage pinned (tridAdiPlanSetSweep dir 0)
    to every chunk length, both block orders (the XCD-contiguous order
    included) and both forms of the x sweep (precomputed factors, pair = 1,
    or factors computed in the kernel), against the oracle; rows that do not
    fit are rejected.  nx = 128, 256 and 512 put one, two or four 64-lane
    rows of 16-byte pieces in a line, the sizes at which a wave with two or
    more lines takes its y-neighbour lines from its own span."""
    import torch
    nx, ny, nz = shape
    u0 = rand_init(nx, ny, nz, dt, seed=nx + ny)
    exp = u0.ravel().copy()
    assert oracle.adi(exp, nx, ny, nz, 3, 1.0, 1) == 0
    ran = 0
    for rows in (4, 8, 16):
        for xcd in (0, 1):
            for cc in (0, 1):
                plan = AdiPlan(nx, ny, nz, torch_dt(dt), 1.0, 1)
                assert "fused" in plan.describe()
                cps = nx // rows
                fits = nx % rows == 0 and cps <= 64 and ny % (64 // cps) == 0
                if not fits:
                    with pytest.raises(ta.TridError):
                        plan.set_sweep(0, rows, cc, xcd)
                    continue
                plan.set_sweep(0, rows, cc, xcd)
                desc = plan.describe().replace(",cached", "")  # the cache policy is not the layout
                tag = f"fused,R{rows}{',cc' if cc else ''}{',xcd' if xcd else ''})"
                assert tag in desc, plan.describe()
                u = torch.from_numpy(u0.copy()).to("cuda")
                plan.step(u, torch.empty_like(u), 3)
                err = oracle.rel_err(exp, u.cpu().numpy().ravel())
                assert err <= tol(dt), f"{shape} R{rows} xcd {xcd} cc {cc} [{plan.describe()}]: {err:.2e}"
                ran += 1
    assert ran >= 1


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_adi_step_graph_matches_single_steps(dt):
    """tridAdiStep with several steps launches a cached graph of up to 64
    steps from the second call with the same arrays and count on: bitwise
    equal to the same number of one-step calls (kernel by kernel) on the
    first call, the capture, a replay, other arrays, another count, past 64
    steps (64-step graph + remainder) and after re-pinning a sweep."""
    import torch
    nx, ny, nz = 40, 24, 16
    u0 = torch.from_numpy(rand_init(nx, ny, nz, dt, seed=12)).to("cuda")
    plan = AdiPlan(nx, ny, nz, torch_dt(dt), 1.0, 1)

    def single(n):
        u = u0.clone()
        du = torch.empty_like(u)
        for _ in range(n):
            plan.step(u, du, 1)
        return u.cpu().numpy()

    u_a, du_a = u0.clone(), torch.empty_like(u0)
    u_b, du_b = u0.clone(), torch.empty_like(u0)
    exp5 = single(5)

    def three_calls(u, du, n, exp):
        for _ in range(3):                       # eager, capture, replay
            u.copy_(u0)
            plan.step(u, du, n)
            np.testing.assert_array_equal(u.cpu().numpy(), exp)

    three_calls(u_a, du_a, 5, exp5)
    three_calls(u_b, du_b, 5, exp5)                    # other arrays
    three_calls(u_b, du_b, 3, single(3))               # other count
    three_calls(u_b, du_b, 70, single(70))             # 64-step graph + 6
    plan.step(u_b, du_b, 5)
    plan.step(u_b, du_b, 5)                      # a graph is held
    plan.set_sweep(1, 16, 0, 1)                  # drops it
    three_calls(u_b, du_b, 5, single(5))
    exp = u0.cpu().numpy().ravel().copy()
    assert oracle.adi(exp, nx, ny, nz, 5, 1.0, 1) == 0
    assert oracle.rel_err(exp, u_b.cpu().numpy().ravel()) <= tol(dt)


def test_adi_stages_vs_oracle():
    """RHS bit-exact; each sweep against the oracle's strided solve."""
    import torch
    nx, ny, nz = 40, 24, 16
    dt = np.float64
    u0 = rand_init(nx, ny, nz, dt, seed=3)
    plan = AdiPlan(nx, ny, nz, torch.float64, 1.0, 1)
    u = torch.from_numpy(u0.copy()).to("cuda")
    du = torch.empty_like(u)
    plan.stage(u, du, 0)
    rhs = du.cpu().numpy()
    # oracle RHS: one oracle ADI step with zero-coefficient-free sweeps is not
    # separable, so rebuild the RHS exactly as adi_orig.c does
    uu = u0
    exp = np.zeros_like(uu)
    c = uu[1:-1, 1:-1, 1:-1]
    s = uu[1:-1, 1:-1, :-2] + uu[1:-1, 1:-1, 2:]
    s = s + uu[1:-1, :-2, 1:-1]
    s = s + uu[1:-1, 2:, 1:-1]
    s = s + uu[:-2, 1:-1, 1:-1]
    s = s + uu[2:, 1:-1, 1:-1]
    s = s - 6.0 * c
    exp[1:-1, 1:-1, 1:-1] = 1.0 * s
    np.testing.assert_array_equal(rhs, exp)
    # sweeps: materialise the constant coefficients and use the oracle's batch solve
    dims = [nx, ny, nz]
    inter = np.zeros((nz, ny, nx), bool)
    inter[1:-1, 1:-1, 1:-1] = True
    a = np.where(inter, -1.0, 0.0).ravel()
    b = np.where(inter, 3.0, 1.0).ravel()
    cc = a.copy()
    d = rhs.ravel().copy()
    for stage, sd in ((1, 0), (2, 1)):
        plan.stage(u, du, stage)
        assert oracle.mtsv(a, b, cc, d, None, 3, sd, dims) == 0
        got = du.cpu().numpy().ravel()
        assert oracle.rel_err(d, got) < 1e-13, f"stage {stage}"
    uexp = u0.ravel().copy()
    assert oracle.mtsv(a, b, cc, d.copy(), uexp, 3, 2, dims, inc=True) == 0
    plan.stage(u, du, 3)
    assert oracle.rel_err(uexp, u.cpu().numpy().ravel()) < 1e-13


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(64, 40, 24), (256, 16, 12), (128, 6, 5)])
def test_adi_fused_rhs_x_equals_separate_stages(dt, shape):
    """The fused RHS + x sweep equals the RHS kernel followed by the x sweep
    (same operations and chunking; the compiler may fuse different
    multiply-adds in the two kernels, so equality is to a few ulp).  Every
    element is written (the output starts as NaN)."""
    import torch
    nx, ny, nz = shape
    u0 = rand_init(nx, ny, nz, dt, seed=11)
    plan = AdiPlan(nx, ny, nz, torch_dt(dt), 1.0, 1)
    if "fused" not in plan.describe():
        pytest.skip("no fused stage for this shape: " + plan.describe())
    u = torch.from_numpy(u0.copy()).to("cuda")
    d1 = torch.empty_like(u)
    d2 = torch.full_like(u, float("nan"))
    plan.stage(u, d1, 0)
    plan.stage(u, d1, 1)
    plan.stage(u, d2, 4)
    a, b = d1.cpu().numpy(), d2.cpu().numpy()
    assert not np.isnan(b).any()
    err = oracle.rel_err(a, b)
    assert err <= (1e-14 if dt == np.float64 else 1e-6), f"{err:.2e} [{plan.describe()}]"


def test_adi_graph_replays_match_eager():
    import torch
    nx, ny, nz = 64, 48, 40
    u0 = rand_init(nx, ny, nz, np.float64, seed=7)
    plan = AdiPlan(nx, ny, nz, torch.float64, 1.0, 1)
    ue = torch.from_numpy(u0.copy()).to("cuda")
    de = torch.empty_like(ue)
    ug = torch.from_numpy(u0.copy()).to("cuda")
    dg = torch.empty_like(ug)
    plan.capture(ug, dg, 3)
    for rep in range(4):
        plan.step(ue, de, 3)
        plan.replay()
        torch.cuda.synchronize()
        x = ue.sum()  # eager torch work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(x)
        np.testing.assert_array_equal(ug.cpu().numpy(), ue.cpu().numpy(), err_msg=f"replay {rep}")


@pytest.mark.parametrize("shape", [(64, 40, 24), (128, 64, 48), (6, 600, 5), (512, 24, 20),
                                   (24, 20, 512)])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_adi_autotuned_plan_vs_oracle(shape, dt):
    """With autotuning on, the plan times its kernel choices at its first use on
    the caller's u and du (saved and restored): the steps must still give the
    oracle's answer."""
    import tridsolver_amd as ta
    nx, ny, nz = shape
    u0 = rand_init(nx, ny, nz, dt, seed=11)
    ta.set_autotune(1)
    try:
        got, plan = run_gpu(u0, 5, 1)
    finally:
        ta.set_autotune(0)
    exp = u0.ravel().copy()
    assert oracle.adi(exp, nx, ny, nz, 5, 1.0, 1) == 0
    err = oracle.rel_err(exp, got.ravel())
    assert err <= tol(dt), f"{shape}: {err:.2e} [{plan.describe()}]"


def test_adi_autotuned_graph_capture_vs_oracle():
    """Tuning triggered by the capture call runs before the capture starts; the
    captured steps replay to the oracle's answer."""
    import torch
    import tridsolver_amd as ta
    nx, ny, nz = 64, 48, 40
    u0 = rand_init(nx, ny, nz, np.float64, seed=12)
    ta.set_autotune(1)
    try:
        plan = AdiPlan(nx, ny, nz, torch.float64, 1.0, 1)
        u = torch.from_numpy(u0.copy()).to("cuda")
        du = torch.empty_like(u)
        plan.capture(u, du, 2)
    finally:
        ta.set_autotune(0)
    plan.replay()
    plan.replay()
    exp = u0.ravel().copy()
    assert oracle.adi(exp, nx, ny, nz, 4, 1.0, 1) == 0
    err = oracle.rel_err(exp, u.cpu().numpy().ravel())
    assert err <= 1e-12, f"{err:.2e} [{plan.describe()}]"


def test_adi_invalid_arguments():
    import torch
    from tridsolver_amd import TridError
    with pytest.raises(TridError):
        AdiPlan(0, 4, 4, torch.float64)
    plan = AdiPlan(8, 8, 8, torch.float64)
    u = torch.zeros((8, 8, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(TridError):
        plan.step(u, u, 1)
    with pytest.raises(TridError):
        plan.replay()


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_adi_benchmark_size_512(dt):
    """The 512^3 grid of the ADI configuration: 20 steps, lambda = 1,
    a = c = -lambda, b = 1 + 2 lambda, Dirichlet-zero boundary, interior
    U(0,1) from seed 0, against the oracle's ADI over the same steps
    (apps/adi/src/adi_cuda.cu:285-330 is the reference driver)."""
    n = 512
    u0 = rand_init(n, n, n, dt, seed=0)
    got, plan = run_gpu(u0, 20, 1)
    exp = u0.ravel()
    del u0
    assert oracle.adi(exp, n, n, n, 20, 1.0, 1) == 0
    err = oracle.rel_err(exp, got.ravel())
    print(f"ADI 512^3 {np.dtype(dt).name} 20 steps: rel err vs oracle {err:.2e} [{plan.describe()}]")
    assert err <= tol(dt)


# y and z lines of more than 512 rows run the generic sweeps with the strided
# default layout (up to 8192 rows); 2100, 4100 and 8192 rows take 16-row
# chunks with 256 or 512 chunks per line
@pytest.mark.parametrize("shape", [(4, 2100, 5), (5, 4, 4100), (3, 4, 8192), (3, 1100, 4)])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("mode", [0, 1])
def test_adi_long_y_z_lines(shape, dt, mode):
    nx, ny, nz = shape
    u0 = rand_init(nx, ny, nz, dt, seed=nx * ny + nz) if mode == 1 else ref_init(nx, ny, nz, dt)
    got, plan = run_gpu(u0, 3, mode)
    exp = u0.ravel().copy()
    assert oracle.adi(exp, nx, ny, nz, 3, 1.0, mode) == 0
    err = oracle.rel_err(exp, got.ravel())
    assert err <= tol(dt), f"{shape} mode {mode}: {err:.2e} [{plan.describe()}]"

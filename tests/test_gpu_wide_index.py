"""GPU parity of the 64-bit offset forms of the kernels.

Every kernel is built twice: with 32-bit element offsets, used while every
array is under 4 GiB, and with 64-bit offsets and system-index arithmetic.
Two sets of checks reach the second form:

* tridForceWideIndex(1) on the small oracle cases: strided solves under each
  kernel configuration (including the split-line solve and the general wave
  staging), the ADI plan, the adjoint and the variable-diffusivity ADI, all
  against the oracle at the usual tolerance (1e-12 FP64, 1e-5 FP32 normwise);
* arrays past the limits themselves (automatic choice): strided batches of
  2.2e9 elements (FP32 8.8 GB and FP64 17.6 GB per array, element offsets
  past 2^31) checked system by system on sampled systems (first, last, those
  around element 2^31 and seeded random ones) against the oracle, and a
  FP64 ADI step on grids of 2.2e9 points checked through locality: the step's line solves
  (tridiag(-1, 3, -1), lambda = 1) damp an isolated bump by 0.382 per row, so
  a 16^3 bump placed 64 points from everything at the far end of the grid
  must give, in its 144^3 window, what the oracle gives for the same bump on a
  144^3 grid (1e-12 normwise), and less than 1e-20 everywhere else.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import tridsolver_amd as ta
from tests.helpers import check_close, oracle_solve
from tests.test_gpu_adi import rand_init, ref_init
from tests.test_gpu_scalar import run_case

pytestmark = pytest.mark.gpu


@pytest.fixture
def wide():
    ta.force_wide_index(True)
    try:
        yield
    finally:
        ta.force_wide_index(False)
        ta.force_config(0, 0, 0, 0)


# (kind, rows, cps, spb): automatic, the partitioned kernel at each chunk
# length, the one-thread fallback, the XCD-ordered variant, the split solve
CONFIGS = [(0, 0, 0, 0), (1, 4, 0, 0), (1, 8, 0, 16), (1, 16, 0, 0), (2, 0, 0, 0),
           (3, 8, 0, 16), (4, 4, 4, 0), (4, 8, 2, 0)]


@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("sd", [0, 1, 2])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_forced_wide_strided(wide, cfg, sd, dtype):
    ta.force_config(*cfg)
    for inc in (False, True):
        run_case([45, 38, 27], sd, dtype, inc, seed=5)
    run_case([70, 33, 21], sd, dtype, False, pads=[72, 35, 21], seed=6)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_forced_wide_staging_and_long(wide, dtype):
    """Wave staging (exact and general), multi-wave systems and the automatic
    split-line solve of long systems."""
    for dims, inc in (([256, 64], False), ([1000, 9], True), ([64 * 8 * 4, 3], False),
                      ([20000, 3], False), ([3, 4096], False)):
        run_case(dims, 0 if dims[0] > 3 else 1, dtype, inc, seed=dims[0])


@pytest.mark.parametrize("shape", [(33, 17, 9), (64, 40, 24), (512, 24, 20)])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_forced_wide_adi(wide, shape, dt):
    import torch
    from tridsolver_amd.adi import AdiPlan
    nx, ny, nz = shape
    for mode in (0, 1):
        u0 = rand_init(nx, ny, nz, dt, seed=nx) if mode else ref_init(nx, ny, nz, dt)
        plan = AdiPlan(nx, ny, nz, torch.float64 if dt == np.float64 else torch.float32, 1.0, mode)
        u = torch.from_numpy(u0.copy()).to("cuda")
        plan.step(u, torch.empty_like(u), 3)
        exp = u0.ravel().copy()
        assert oracle.adi(exp, nx, ny, nz, 3, 1.0, mode) == 0
        check_close(exp, u.cpu().numpy().ravel(), dt, f"ADI {shape} mode {mode}")


def test_forced_wide_adjoint_and_vadi(wide):
    from tests.test_gpu_adjoint import test_strided_adjoint_vs_oracle, \
        test_vadi_forward_and_grad_vs_oracle
    for sd in (0, 1, 2):
        for dtype in (np.float64, np.float32):
            test_strided_adjoint_vs_oracle([40, 30, 20], sd, dtype)
    test_vadi_forward_and_grad_vs_oracle(12, 5, 2)


def need_memory(nbytes):
    import torch
    free, _ = torch.cuda.mem_get_info()
    if free < nbytes * 1.1:
        pytest.skip(f"needs {nbytes / 2**30:.0f} GiB of device memory, {free / 2**30:.0f} free")


def fill_uniform(t, lo, hi, gen):
    """uniform_ over a flat tensor in slices of 2^28 elements"""
    flat = t.view(-1)
    step = 1 << 28
    for s in range(0, flat.numel(), step):
        flat[s:s + step].uniform_(lo, hi, generator=gen)


def system_index(dims, sd, q):
    """element offset of system q's first row and the row stride (unpadded)"""
    cum = [1, dims[0], dims[0] * dims[1]]
    other = [k for k in range(3) if k != sd]
    i0, i1 = q % dims[other[0]], q // dims[other[0]]
    return i0 * cum[other[0]] + i1 * cum[other[1]], cum[sd]


@pytest.mark.parametrize("dims,sd,dtype,inc", [
    ([256, 2048, 4200], 0, np.float32, False),
    ([256, 2048, 4200], 1, np.float32, True),
    ([256, 2048, 4200], 2, np.float32, False),
    ([1300, 1300, 1300], 0, np.float64, True),
    ([1300, 1300, 1300], 1, np.float64, False),
    ([1300, 1300, 1300], 2, np.float64, False),
])
def test_arrays_past_2g_elements(dims, sd, dtype, inc):
    import torch
    from tridsolver_amd import api
    n = dims[0] * dims[1] * dims[2]
    assert n >= 2**31
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    esz = 8 if dtype == np.float64 else 4
    need_memory((5 if inc else 4) * n * esz)
    gen = torch.Generator(device="cuda").manual_seed(sum(dims) + sd)
    a, b, c, d = (torch.empty(n, dtype=tdt, device="cuda") for _ in range(4))
    fill_uniform(a, -1.0, 1.0, gen)
    fill_uniform(c, -1.0, 1.0, gen)
    fill_uniform(b, 4.0, 5.0, gen)
    fill_uniform(d, -1.0, 1.0, gen)
    u = None
    if inc:
        u = torch.empty(n, dtype=tdt, device="cuda")
        fill_uniform(u, -1.0, 1.0, gen)
    N = dims[sd]
    nsys = n // N
    rng = np.random.default_rng(N)
    picks = {0, nsys - 1, *rng.integers(0, nsys, 12).tolist()}
    # the system holding element 2^31 and its neighbours
    x = [2**31 % dims[0], 2**31 // dims[0] % dims[1], 2**31 // (dims[0] * dims[1])]
    o0, o1 = [k for k in range(3) if k != sd]
    q31 = x[o0] + x[o1] * dims[o0]
    picks |= {q for q in range(q31 - 2, q31 + 3) if 0 <= q < nsys}
    picks = sorted(picks)

    def gather(t, q):
        off, st = system_index(dims, sd, q)
        idx = off + st * torch.arange(N, device="cuda", dtype=torch.int64)
        return t[idx].cpu().numpy()

    before = {q: [gather(t, q) for t in (a, b, c, d) + ((u,) if inc else ())] for q in picks}
    fn = {(True, False): api.tridDmtsvStridedBatch, (True, True): api.tridDmtsvStridedBatchInc,
          (False, False): api.tridSmtsvStridedBatch,
          (False, True): api.tridSmtsvStridedBatchInc}[(dtype == np.float64, inc)]
    st = fn(a, b, c, d, u, 3, sd, dims, dims, [0, 0, 0], 1)
    assert st == 0, ta.last_error()
    torch.cuda.synchronize()
    print(f"{dims} sd={sd}: {ta.last_config()}, {len(picks)} systems checked")
    out = u if inc else d
    for q in picks:
        arrs = before[q]
        exp = oracle_solve(arrs[0], arrs[1], arrs[2], arrs[3], arrs[4] if inc else None, [N], 0,
                           inc=inc)
        check_close(exp, gather(out, q), dtype, f"{dims} sd={sd} system {q}")
    del a, b, c, d, u, out
    torch.cuda.empty_cache()


@pytest.mark.parametrize("grid", [(1300, 1300, 1300), (8192, 512, 520)])
def test_adi_grid_past_2g_points(grid):
    """(8192, 512, 520) also runs the factor-reuse sweep along y (512 rows)."""
    import torch
    from tridsolver_amd.adi import AdiPlan
    nx, ny, nz = grid
    m, w = 64, 16                    # margin, bump side
    s = w + 2 * m                    # window side
    assert nx * ny * nz >= 2**31
    need_memory(2 * nx * ny * nz * 8)
    bump = np.random.default_rng(7).uniform(0, 1, (w, w, w))
    small = np.zeros((s, s, s))
    small[m:m + w, m:m + w, m:m + w] = bump
    exp = small.ravel().copy()
    assert oracle.adi(exp, s, s, s, 1, 1.0, 1) == 0

    # bump start and window start per axis (z, y, x), at the far end
    b0 = [n - 1 - m - w for n in (nz, ny, nx)]
    win = tuple(slice(b - m, b - m + s) for b in b0)
    u = torch.zeros((nz, ny, nx), dtype=torch.float64, device="cuda")
    u[tuple(slice(b, b + w) for b in b0)] = torch.from_numpy(bump).to("cuda")
    du = torch.empty_like(u)
    plan = AdiPlan(nx, ny, nz, torch.float64, 1.0, 1)
    plan.step(u, du, 1)
    torch.cuda.synchronize()
    got = u[win].cpu().numpy()
    check_close(exp, got.ravel(), np.float64, f"ADI {grid} window [{plan.describe()}]")
    u[win] = 0
    outside = u.abs_().amax().item()
    print(f"ADI {grid}: max |u| outside the window {outside:.2e}")
    assert outside < 1e-20
    del u, du, plan
    torch.cuda.empty_cache()

"""GPU parity of the strided-batch scalar solves (trid_hip.h) against the oracle.

Every case runs the same seeded inputs through the C ABI (ctypes, device
buffers from torch) and through the CPU oracle, and requires
  * normwise relative error <= 1e-12 (FP64) / 1e-5 (FP32), and
  * the reference tests' elementwise tolerance |e-a| <= abs + rel*min(|e|,|a|)
    (1e-11/1e-11 FP64, 1e-6/1e-5 FP32; scalar/test/utils.hpp:14-17),
  * untouched padding (NaN-filled padding must stay NaN).
FP32 results are also compared with the FP64 oracle on the same (rounded)
inputs and that error is printed.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import tridsolver_amd as ta
from tridsolver_amd import api
from tests.helpers import alloc, check_close, dd_system, oracle_solve, view

pytestmark = pytest.mark.gpu


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def gpu_solve(a, b, c, d, u, dims, sd, pads, inc, per_array=None):
    import torch
    ta_, tb, tc, td = (dev(x) for x in (a, b, c, d))
    tu = dev(u) if u is not None else None
    dbl = d.dtype == np.float64
    if per_array:
        fn = {(True, False): api.tridDmtsvStridedBatchPadded,
              (True, True): api.tridDmtsvStridedBatchPaddedInc,
              (False, False): api.tridSmtsvStridedBatchPadded,
              (False, True): api.tridSmtsvStridedBatchPaddedInc}[(dbl, inc)]
        st = fn(ta_, per_array[0], tb, per_array[1], tc, per_array[2], td, per_array[3], tu,
                per_array[4], len(dims), sd, dims, None, 1)
    else:
        fn = {(True, False): api.tridDmtsvStridedBatch, (True, True): api.tridDmtsvStridedBatchInc,
              (False, False): api.tridSmtsvStridedBatch,
              (False, True): api.tridSmtsvStridedBatchInc}[(dbl, inc)]
        st = fn(ta_, tb, tc, td, tu, len(dims), sd, dims, pads, [0] * len(dims), 1)
    assert st == 0, ta.last_error()
    torch.cuda.synchronize()
    if inc:
        # the Inc form leaves d untouched
        np.testing.assert_array_equal(td.cpu().numpy(), d)
    return (tu if inc else td).cpu().numpy()


def make_u(dims, pads, dtype, seed=99):
    u = alloc(dims, pads, dtype)
    view(u, dims, pads)[...] = np.random.default_rng(seed).uniform(-1, 1, tuple(reversed(dims)))
    return u


def run_case(dims, sd, dtype, inc, pads=None, seed=0, zero_ends=False):
    pads = pads or list(dims)
    a, b, c, d = dd_system(dims, sd, dtype, seed=seed, pads=pads, zero_ends=zero_ends)
    u = make_u(dims, pads, dtype) if inc else None
    got = gpu_solve(a, b, c, d, u, dims, sd, pads, inc)
    exp = oracle_solve(a, b, c, d, u, dims, sd, pads, inc)
    r = check_close(exp, got, dtype, f"dims={dims} sd={sd} inc={inc} [{ta.last_config()}]")
    if dtype == np.float32:
        exp64 = oracle_solve(*(x.astype(np.float64) for x in (a, b, c, d)),
                             None if u is None else u.astype(np.float64), dims, sd, pads, inc)
        r64 = oracle.rel_err(exp64[~np.isnan(exp64)], got[~np.isnan(got)])
        print(f"fp32 dims={dims} sd={sd}: vs fp32 oracle {r:.2e}, vs fp64 oracle {r64:.2e}")
    return r


SHAPES = [
    ([37, 29, 23], 0), ([37, 29, 23], 1), ([37, 29, 23], 2),
    ([64, 64, 64], 0), ([64, 64, 64], 1), ([64, 64, 64], 2),
    ([256, 16, 8], 0), ([8, 256, 9], 1), ([70, 5, 300], 2),
    ([1], 0), ([2], 0), ([3], 0), ([5], 0), ([200], 0), ([1000], 0), ([4096], 0),
    ([8, 8], 0), ([8, 8], 1), ([32, 32], 0), ([32, 32], 1), ([256, 256], 1),
    ([7, 5, 6, 4], 0), ([7, 5, 6, 4], 1), ([7, 5, 6, 4], 2), ([7, 5, 6, 4], 3),
    ([1, 40, 3], 1), ([40, 1, 3], 0), ([3, 2, 1], 2),
    # wave-staged solvedim-0 path: partial last work-group, odd batch counts
    ([256, 3], 0), ([128, 5, 3], 0), ([64, 33], 0), ([512, 7], 0), ([32, 9, 2], 0),
]


@pytest.mark.parametrize("dims,sd", SHAPES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("inc", [False, True])
def test_strided_matches_oracle(dims, sd, dtype, inc):
    run_case(dims, sd, dtype, inc, seed=len(dims) * 10 + sd)


@pytest.mark.parametrize("sd", [0, 1, 2])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("inc", [False, True])
def test_pads_differ_from_dims(sd, dtype, inc):
    run_case([33, 20, 17], sd, dtype, inc, pads=[40, 23, 19], seed=5)


@pytest.mark.parametrize("sd", [0, 1, 2])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_per_array_pads(sd, dtype):
    dims = [19, 13, 11]
    pa = [[19, 13, 11], [22, 13, 12], [20, 15, 11], [24, 14, 11], [21, 16, 13]]
    for inc in (False, True):
        a, b, c, d = dd_system(dims, sd, dtype, seed=3, pads_per_array=pa[:4])
        u = make_u(dims, pa[4], dtype) if inc else None
        got = gpu_solve(a, b, c, d, u, dims, sd, None, inc, per_array=pa)
        exp = oracle_solve(a, b, c, d, u, dims, sd, inc=inc, per_array=pa)
        check_close(exp, got, dtype, f"per-array pads sd={sd} inc={inc}")


@pytest.mark.parametrize("sd", [0, 1, 2])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_full_size_256cube(sd, dtype):
    """The benchmark configuration: 256^3, a_0 = c_{N-1} = 0, seed 0."""
    r = run_case([256, 256, 256], sd, dtype, False, seed=0, zero_ends=True)
    print(f"256^3 sd={sd} {np.dtype(dtype).name}: rel err vs oracle {r:.2e} [{ta.last_config()}]")


@pytest.mark.parametrize("sd", [0, 1, 2])
def test_a0_and_cN_are_never_read(sd):
    dims = [30, 20, 10]
    a, b, c, d = dd_system(dims, sd, np.float64, seed=1)
    r1 = gpu_solve(a, b, c, d, None, dims, sd, dims, False)
    a2, c2 = a.copy(), c.copy()
    ax = 2 - sd
    av, cv = view(a2, dims, dims), view(c2, dims, dims)
    i0 = [slice(None)] * 3
    i0[ax] = 0
    i1 = [slice(None)] * 3
    i1[ax] = -1
    av[tuple(i0)] = np.nan
    cv[tuple(i1)] = np.inf
    r2 = gpu_solve(a2, b, c2, d, None, dims, sd, dims, False)
    np.testing.assert_array_equal(r1, r2)


# (kind, rows, cps, spb); kind 3 = kind 1 with the XCD-contiguous block order
# (the [45, 38, 27] grid gives 65 blocks at spb 16: 64 remapped, 1 identity)
CONFIGS = [(1, 4, 0, 0), (1, 8, 0, 0), (1, 16, 0, 0), (1, 8, 0, 16), (1, 4, 0, 64),
           (1, 16, 0, 32), (2, 0, 0, 0), (3, 8, 0, 16), (3, 16, 0, 32), (3, 4, 0, 64)]


@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("sd", [0, 1, 2])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_every_kernel_configuration(cfg, sd, dtype):
    """Each schedule the dispatcher or autotuner may pick gives the oracle's answer."""
    ta.force_config(*cfg)
    try:
        for inc in (False, True):
            run_case([45, 38, 27], sd, dtype, inc, seed=2)
    finally:
        ta.force_config(0, 0, 0, 0)


@pytest.mark.parametrize("sd", [1, 2])
def test_fp32_wide_strided_groups(sd):
    """FP32 16-row chunks of strided systems in 1024-thread work-groups (64
    systems side by side); FP64 refuses that shape."""
    dims = [70, 200, 9] if sd == 1 else [70, 9, 200]  # 16 chunks x 64 systems
    ta.force_config(1, 16, 0, 64)
    try:
        for inc in (False, True):
            run_case(dims, sd, np.float32, inc, seed=8)
            assert "R16,cps16,spb64" in ta.last_config(), ta.last_config()
        run_case([70, 33, 21], sd, np.float32, False, seed=9)  # 4 chunks x 64 systems
        with pytest.raises(AssertionError, match="forced configuration"):
            run_case(dims, sd, np.float64, False)
    finally:
        ta.force_config(0, 0, 0, 0)


@pytest.mark.parametrize("R", [4, 8, 16])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_staged_multiwave_systems(R, dtype):
    """Wave-staged access for systems longer than 64 chunks (each wave a
    contiguous span of one system) with the two-level reduced solve."""
    ta.force_config(1, R, 0, 0)
    try:
        for dims, inc in (([64 * R * 2, 5], False), ([64 * R * 4, 3], True), ([64 * R * 8, 2], False)):
            run_case(dims, 0, dtype, inc, seed=dims[0])
            assert "staged" in ta.last_config(), ta.last_config()
    finally:
        ta.force_config(0, 0, 0, 0)


@pytest.mark.parametrize("R", [5, 6, 7, 12])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_general_staging_fill_rows(R, dtype):
    """Chunk lengths 5, 6, 7 and 12 (general staged layout only): several
    lengths, dense or padded, Inc, partial waves; solvedim 1 refuses them."""
    ta.force_config(1, R, 0, 0)
    try:
        cases = [([300, 21], None, False), ([77, 9], None, True), ([3000, 3], None, False),
                 ([5, 41], None, True), ([1500, 5], [1504, 5], False), ([6000, 2], None, True)]
        for dims, pads, inc in cases:
            if (dims[0] + R - 1) // R > 512:  # more than 512 chunks
                continue
            run_case(dims, 0, dtype, inc, pads=pads, seed=dims[0] + R)
            assert f"R{R}," in ta.last_config() and "gstaged" in ta.last_config(), \
                ta.last_config()
        with pytest.raises(AssertionError, match="general staged"):
            run_case([30, 40], 1, dtype, False)
    finally:
        ta.force_config(0, 0, 0, 0)


@pytest.mark.parametrize("dtype,N,cfg", [
    (np.float64, 300, "R12,cps32"), (np.float64, 200, "R7,cps32"), (np.float64, 3000, "R12,cps256"),
    (np.float64, 3500, "R7,cps512"), (np.float64, 500, "R8,cps64"),
    (np.float64, 600, "R12,cps64"), (np.float64, 1000, "R8,cps128"), (np.float64, 6000, "R12,cps512"),
    (np.float64, 7000, "R16,cps512"), (np.float32, 300, "R12,cps32"), (np.float32, 200, "R4,cps64"),
    (np.float32, 5000, "R12,cps512"), (np.float32, 8000, "R8,cps1024")])
def test_default_chunk_length(dtype, N, cfg):
    """Without tuning, solvedim-0 solves take 12-row (FP64 also 7-row) chunks
    with general staging where they pad the system to fewer rows than the
    8-row (FP64) or 4-row (FP32) chunks."""
    run_case([N, 7], 0, dtype, False, seed=N)
    assert cfg in ta.last_config(), ta.last_config()


@pytest.mark.parametrize("dtype,N,cfg", [
    (np.float64, 200, "R8,cps32,spb16>"), (np.float64, 300, "R16,cps32,spb16>"),
    (np.float64, 1500, "R16,cps128,spb4,xcd>"), (np.float32, 100, "R8,cps16,spb64>"),
    (np.float32, 200, "R8,cps32,spb32,xcd>"), (np.float32, 700, "R16,cps64,spb16,xcd>")])
@pytest.mark.parametrize("sd", [1, 2])
def test_default_strided_layout(dtype, N, cfg, sd):
    """Without tuning, strided batches (solvedim >= 1) take the longest chunks
    at the largest work-group the registers allow (most systems side by
    side)."""
    dims = [64, N, 3] if sd == 1 else [64, 3, N]
    run_case(dims, sd, dtype, False, seed=N)
    # (the cache policy, ",cached" for these small batches, is not the layout)
    assert ta.last_config().replace(",cached", "").endswith(cfg), ta.last_config()


@pytest.mark.parametrize("R", [4, 8, 16])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_general_staging_any_length(R, dtype):
    """Wave-staged access for any system length (spans that start and end
    anywhere, boundary pieces stored element by element): odd lengths,
    partial last waves, systems over several waves, Inc, and padded arrays
    with one or more waves per system."""
    ta.force_config(1, R, 0, 0)
    try:
        cases = [([100, 37], None, False), ([77, 9], None, True), ([301, 13], None, False),
                 ([999, 5], None, True), ([1500, 3], None, False), ([4000, 2], None, True),
                 ([5, 41], None, False), ([64 * R + 3, 3], [64 * R + 8, 3], False),
                 ([2000, 3], [2005, 4], True)]
        for dims, pads, inc in cases:
            run_case(dims, 0, dtype, inc, pads=pads, seed=dims[0])
            if 64 * R * (4 if R < 16 else 2) >= dims[0]:  # within the staged kernel's reach
                assert "gstaged" in ta.last_config() or "staged" in ta.last_config(), \
                    (dims, ta.last_config())
    finally:
        ta.force_config(0, 0, 0, 0)


@pytest.mark.parametrize("N", [8192, 20000])
def test_long_systems(N):
    for dtype in (np.float64, np.float32):
        run_case([N, 3], 0, dtype, False, seed=N)
        run_case([5, N], 1, dtype, True, seed=N)


def test_autotune_mode_matches_oracle():
    ta.set_autotune(1)
    try:
        for sd in (0, 1, 2):
            run_case([64, 48, 40], sd, np.float64, False, seed=8)
            run_case([64, 48, 40], sd, np.float64, True, seed=8)
    finally:
        ta.set_autotune(0)


def test_async_on_torch_stream():
    import torch
    dims, sd = [128, 64, 32], 1
    a, b, c, d = dd_system(dims, sd, np.float64, seed=4)
    exp = oracle_solve(a, b, c, d, None, dims, sd)
    s = torch.cuda.Stream()
    tas = [dev(x) for x in (a, b, c, d)]
    torch.cuda.synchronize()
    ta.set_stream(s)
    try:
        st = api.tridDmtsvStridedBatch(*tas, None, 3, sd, dims, dims, None, 0)
        assert st == 0
        s.synchronize()
    finally:
        ta.set_stream(None)
    check_close(exp, tas[3].cpu().numpy(), np.float64, "async stream")


def test_invalid_arguments_return_status():
    import torch
    x = torch.zeros(8, dtype=torch.float64, device="cuda")
    assert api.tridDmtsvStridedBatch(x, x, x, x, None, 1, 1, [8], [8]) == 3
    assert api.tridDmtsvStridedBatch(x, x, x, x, None, 1, 0, [8], [7]) == 3
    assert api.tridDmtsvStridedBatchInc(x, x, x, x, None, 1, 0, [8], [8]) == 3


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reciprocal_error_bound(dtype):
    """The pivot reciprocal (hardware estimate + Newton-Raphson refinement)
    stays within 1 ulp of the correctly rounded 1/x over a wide range."""
    import torch
    rng = np.random.default_rng(17)
    n = 1 << 20
    mant = rng.uniform(1.0, 2.0, n)
    lo, hi = (-300, 300) if dtype == np.float64 else (-120, 120)
    x = (mant * np.exp2(rng.integers(lo, hi, n)) * rng.choice([-1.0, 1.0], n)).astype(dtype)
    x[:8] = np.array([1.0, -1.0, 3.0, 0.1, 7.0, 1.5, 2.0, 0.5], dtype)
    tx = torch.from_numpy(x).to("cuda")
    ty = torch.empty_like(tx)
    st = ta.lib().tridDebugReciprocal(tx.data_ptr(), ty.data_ptr(), n, int(dtype == np.float64))
    assert st == 0
    y = ty.cpu().numpy()
    exact = (np.float64(1.0) / x.astype(np.float64)).astype(dtype)
    ulp = np.spacing(np.abs(exact).astype(dtype))
    err_ulp = np.abs(y.astype(np.float64) - exact.astype(np.float64)) / ulp.astype(np.float64)
    print(f"{np.dtype(dtype).name} reciprocal: max {err_ulp.max():.3f} ulp, "
          f"{np.mean(err_ulp > 0) * 100:.2f}% not correctly rounded")
    assert err_ulp.max() <= 1.0

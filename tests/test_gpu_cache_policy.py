"""GPU parity of the two cache policies of the streaming kernels.

The partitioned kernels, the factor-reuse sweeps and the ADI right-hand-side
kernels exist with non-temporal loads and stores (policy 0), with ordinary
ones (policy 1, which the automatic rule picks for the small arrays of the
other GPU tests) and, for coefficient arrays, with non-temporal a, b, c and
ordinary d, u (policy 2).  The policy changes no arithmetic, so both must give
bitwise equal results, each within the usual tolerance of the oracle (1e-12
FP64, 1e-5 FP32 normwise); the configuration text names the policy.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import tridsolver_amd as ta
from tests.helpers import check_close, oracle_solve
from tests.test_gpu_adi import rand_init
from tests.test_gpu_scalar import dd_system, gpu_solve, make_u

pytestmark = pytest.mark.gpu


@pytest.fixture
def policy_reset():
    try:
        yield
    finally:
        ta.force_cache_policy(-1)
        ta.force_config(0, 0, 0, 0)


def solve_under(policy, a, b, c, d, u, dims, sd, pads, inc):
    ta.force_cache_policy(policy)
    got = gpu_solve(a, b, c, d, u, dims, sd, pads, inc)
    return got, ta.last_config()


# (kind, rows, cps, spb): automatic, each chunk length, the XCD order
CONFIGS = [(0, 0, 0, 0), (1, 4, 0, 0), (1, 8, 0, 16), (1, 16, 0, 0), (3, 8, 0, 16)]


@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("sd", [0, 1, 2])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("inc", [False, True])
def test_strided_policies_bitwise_equal(policy_reset, cfg, sd, dtype, inc):
    dims, pads = [64, 40, 33], [66, 40, 35]
    a, b, c, d = dd_system(dims, sd, dtype, seed=sd + 3, pads=pads)
    u = make_u(dims, pads, dtype) if inc else None
    ta.force_config(*cfg)
    g0, c0 = solve_under(0, a, b, c, d, u, dims, sd, pads, inc)
    g1, c1 = solve_under(1, a, b, c, d, u, dims, sd, pads, inc)
    g2, c2 = solve_under(2, a, b, c, d, u, dims, sd, pads, inc)
    assert ",cached" not in c0 and ",mixed" not in c0
    if c1.startswith("part<"):
        assert ",cached" in c1 and ",mixed" in c2, (c1, c2)
    np.testing.assert_array_equal(g0, g1)
    np.testing.assert_array_equal(g0, g2)
    exp = oracle_solve(a, b, c, d, u, dims, sd, pads, inc)
    check_close(exp, g1, dtype, f"dims={dims} sd={sd} inc={inc} [{c1}]")


@pytest.mark.parametrize("dims,sd", [([256, 64, 8], 0), ([300, 70], 0), ([512, 33], 0),
                                     ([8, 256, 20], 1), ([40, 9, 256], 2)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_staged_policies_bitwise_equal(policy_reset, dims, sd, dtype):
    """wave-staged (exact and general) solvedim-0 layouts and long strided lines"""
    a, b, c, d = dd_system(dims, sd, dtype, seed=7, pads=dims)
    for inc in (False, True):
        u = make_u(dims, dims, dtype) if inc else None
        g0, _ = solve_under(0, a, b, c, d, u, dims, sd, dims, inc)
        g1, c1 = solve_under(1, a, b, c, d, u, dims, sd, dims, inc)
        g2, _ = solve_under(2, a, b, c, d, u, dims, sd, dims, inc)
        np.testing.assert_array_equal(g0, g1)
        np.testing.assert_array_equal(g0, g2)
        exp = oracle_solve(a, b, c, d, u, dims, sd, dims, inc)
        check_close(exp, g1, dtype, f"dims={dims} sd={sd} inc={inc} [{c1}]")


def test_automatic_policy_follows_footprint(policy_reset):
    """ordinary accesses while all arrays span at most 384 MiB, then
    non-temporal coefficients with ordinary d while d does, then
    non-temporal"""
    import torch
    ta.force_cache_policy(-1)
    dims = [64, 64, 64]
    a, b, c, d = dd_system(dims, 1, np.float64, seed=1, pads=dims)
    gpu_solve(a, b, c, d, None, dims, 1, dims, False)
    assert ",cached" in ta.last_config()
    for shape, tag in (((256, 256, 256), ",mixed"), ((512, 512, 256), "")):
        # 4 arrays of 128 MiB (d fits) and of 512 MiB (d does not)
        t = [torch.full(shape, v, dtype=torch.float64, device="cuda") for v in (-1.0, 4.0, -1.0, 1.0)]
        dims = list(reversed(shape))
        st = ta.api.tridDmtsvStridedBatch(*t, None, 3, 1, dims, dims, [0] * 3, 1)
        assert st == 0, ta.last_error()
        cfg = ta.last_config()
        assert ",cached" not in cfg and (tag in cfg if tag else ",mixed" not in cfg), cfg
        # the solution of tridiag(-1, 4, -1) x = 1 along y: symmetric, in (1/4, 1/2]
        x = t[3][0, :, 0].cpu().numpy()
        assert np.all(x > 0.25) and np.all(x <= 0.5 + 1e-14)
        np.testing.assert_allclose(x, x[::-1], rtol=0, atol=1e-14)
        del t
        torch.cuda.empty_cache()


@pytest.mark.parametrize("shape", [(64, 40, 24), (256, 24, 30), (33, 17, 9)])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_adi_policies_bitwise_equal(policy_reset, shape, dt):
    import torch
    from tridsolver_amd.adi import AdiPlan
    nx, ny, nz = shape
    u0 = rand_init(nx, ny, nz, dt, seed=nx + 1)
    tdt = torch.float64 if dt == np.float64 else torch.float32
    out = {}
    for policy in (0, 1):
        ta.force_cache_policy(policy)
        plan = AdiPlan(nx, ny, nz, tdt, 1.0, 1)
        assert ("cached" in plan.describe()) == (policy == 1), plan.describe()
        u = torch.from_numpy(u0.copy()).to("cuda")
        plan.step(u, torch.empty_like(u), 3)
        out[policy, "fused"] = u.cpu().numpy().ravel()
        # an unaligned du takes the unfused right-hand side and x sweep
        u2 = torch.from_numpy(u0.copy()).to("cuda")
        du2 = torch.empty(u2.numel() + 1, dtype=tdt, device="cuda")[1:].view_as(u2)
        plan.step(u2, du2, 3)
        out[policy, "unfused"] = u2.cpu().numpy().ravel()
    exp = u0.ravel().copy()
    assert oracle.adi(exp, nx, ny, nz, 3, 1.0, 1) == 0
    for path in ("fused", "unfused"):
        np.testing.assert_array_equal(out[0, path], out[1, path])
        check_close(exp, out[1, path], dt, f"ADI {shape} {path}")


@pytest.mark.parametrize("n,steps,ckpt", [(16, 6, 2), (32, 4, 4), (12, 5, 2)])
def test_vadi_policies_bitwise_equal(policy_reset, n, steps, ckpt):
    """differentiable ADI (staged x kernel for power-of-two grids, kappa
    sweeps): forward and gradients equal under both policies, and within the
    oracle's tolerance"""
    from tests.test_gpu_adjoint import dev, fourier_kappa
    from tridsolver_amd.autograd import vadi_forward, vadi_loss_grad
    lam = 0.8
    kap = fourier_kappa(n, seed=n)
    rng = np.random.default_rng(n + 1)
    u0 = rng.uniform(0, 1, (n, n, n))
    tgt = rng.uniform(0, 1, (n, n, n))
    res = []
    for policy in (0, 1):
        ta.force_cache_policy(policy)
        uT = vadi_forward(dev(kap), dev(u0), steps, lam).cpu().numpy()
        loss, gk, gu = vadi_loss_grad(dev(kap), dev(u0), dev(tgt), steps, lam, ckpt)
        res.append((uT, loss, gk.cpu().numpy(), gu.cpu().numpy()))
    for x0, x1 in zip(res[0], res[1]):
        np.testing.assert_array_equal(x0, x1)
    exp_uT = oracle.vadi_forward(kap.ravel(), u0.ravel(), n, steps, lam)
    assert oracle.rel_err(exp_uT, res[1][0].ravel()) <= 1e-12
    eloss, egk, egu = oracle.vadi_grad(kap.ravel(), u0.ravel(), tgt.ravel(), n, steps, lam)
    assert res[1][1] == pytest.approx(eloss, rel=1e-12)
    assert oracle.rel_err(egk, res[1][2].ravel()) <= 1e-10
    assert oracle.rel_err(egu, res[1][3].ravel()) <= 1e-10

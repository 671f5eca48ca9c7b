"""GPU parity of the adjoint kernels (trid_adjoint.h) and the differentiable
ADI.

* strided-batch adjoint against the oracle's adjoint (transposed Thomas +
  gradient formulas; pinned in test_oracle.py by the dot-product identity and
  finite differences), FP64 1e-12 / FP32 1e-5 normwise;
* torch.autograd.gradcheck of TridSolve (finite differences, FP64);
* variable-diffusivity ADI forward and gradient against the oracle's
  full-storage reverse mode on small grids, and the benchmark configuration
  (128^3, 16 steps, checkpoint every 4 steps, kappa = exp(0.3 g) with g a
  sum of 8 seeded Fourier modes, target U(0,1)) against the same CPU FP64
  gradient (<= 1e-8 relative, stated) and central finite differences
  (h = 1e-6, difference-of-squares form) at 64 seeded points.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from bench import fourier_kappa
from tests.helpers import dd_system, oracle_solve

pytestmark = pytest.mark.gpu


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


@pytest.mark.parametrize("dims,sd", [([40, 30, 20], 0), ([40, 30, 20], 1), ([40, 30, 20], 2),
                                     ([257], 0), ([5, 1000], 1), ([300, 7], 0)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_strided_adjoint_vs_oracle(dims, sd, dtype):
    import tridsolver_amd as ta
    from tridsolver_amd.api import _iv
    a, b, c, d = dd_system(dims, sd, np.float64, seed=3)
    x = oracle_solve(a, b, c, d, None, dims, sd)
    g = np.random.default_rng(1).uniform(-1, 1, x.shape)
    exp = oracle.adjoint(a, b, c, x, g, len(dims), sd, dims)
    t = [dev(v.astype(dtype)) for v in (a, b, c, x, g)]
    outs = [t[0].new_empty(t[0].shape) for _ in range(4)]
    fn = ta.lib().tridDmtsvStridedBatchAdjoint if dtype == np.float64 else \
        ta.lib().tridSmtsvStridedBatchAdjoint
    st = fn(*[v.data_ptr() for v in t], *[o.data_ptr() for o in outs], len(dims), sd, _iv(dims),
            _iv(dims), 1)
    assert st == 0, ta.last_error()
    tol = 1e-12 if dtype == np.float64 else 1e-5
    for name, e, o in zip("abcd", exp, outs):
        err = oracle.rel_err(e, o.cpu().numpy())
        assert err <= tol, f"grad_{name}: {err:.2e}"


def test_autograd_gradcheck():
    import torch
    from tridsolver_amd.autograd import TridSolve
    for sd, shape in ((0, (3, 2, 9)), (1, (3, 7, 4)), (2, (6, 3, 2))):
        dims = list(reversed(shape))
        a, b, c, d = dd_system(dims, sd, np.float64, seed=sd)
        ts = [dev(v).reshape(shape).requires_grad_(True) for v in (a, b, c, d)]
        assert torch.autograd.gradcheck(lambda a_, b_, c_, d_: TridSolve.apply(a_, b_, c_, d_, sd),
                                        ts, eps=1e-6, atol=1e-8, rtol=1e-6)


@pytest.mark.parametrize("n,steps,ckpt", [(12, 5, 2), (16, 8, 4), (10, 3, 1), (9, 6, 6)])
def test_vadi_forward_and_grad_vs_oracle(n, steps, ckpt):
    from tridsolver_amd.autograd import vadi_forward, vadi_loss_grad
    lam = 0.8
    kap = fourier_kappa(n, seed=n)
    rng = np.random.default_rng(n + 1)
    u0 = rng.uniform(0, 1, (n, n, n))
    tgt = rng.uniform(0, 1, (n, n, n))
    uT = vadi_forward(dev(kap), dev(u0), steps, lam).cpu().numpy()
    exp_uT = oracle.vadi_forward(kap.ravel(), u0.ravel(), n, steps, lam)
    assert oracle.rel_err(exp_uT, uT.ravel()) <= 1e-12
    loss, gk, gu = vadi_loss_grad(dev(kap), dev(u0), dev(tgt), steps, lam, ckpt)
    eloss, egk, egu = oracle.vadi_grad(kap.ravel(), u0.ravel(), tgt.ravel(), n, steps, lam)
    assert loss == pytest.approx(eloss, rel=1e-12)
    assert oracle.rel_err(egk, gk.cpu().numpy().ravel()) <= 1e-10
    assert oracle.rel_err(egu, gu.cpu().numpy().ravel()) <= 1e-10


def test_vadi_benchmark_configuration_fd_and_oracle():
    import torch
    from tridsolver_amd.autograd import vadi_forward, vadi_loss_grad
    n, steps, ckpt, lam, h = 128, 16, 4, 1.0, 1e-6
    kap = fourier_kappa(n, seed=0)
    rng = np.random.default_rng(0)
    u0 = rng.uniform(0, 1, (n, n, n))
    tgt = rng.uniform(0, 1, (n, n, n))
    tk, tu, tt = dev(kap), dev(u0), dev(tgt)
    loss, gk, gu = vadi_loss_grad(tk, tu, tt, steps, lam, ckpt)
    gk, gu = gk.cpu().numpy().ravel(), gu.cpu().numpy().ravel()
    eloss, egk, egu = oracle.vadi_grad(kap.ravel(), u0.ravel(), tgt.ravel(), n, steps, lam)
    ek, eu = oracle.rel_err(egk, gk), oracle.rel_err(egu, gu)
    print(f"vadi 128^3 x16: loss {loss:.6e}; grad rel err vs CPU FP64: kappa {ek:.2e}, u0 {eu:.2e}")
    assert loss == pytest.approx(eloss, rel=1e-12)
    assert ek <= 1e-8 and eu <= 1e-8

    def dL(kp, up, km, um):
        a = vadi_forward(kp, up, steps, lam)
        b = vadi_forward(km, um, steps, lam)
        return float(0.5 * torch.sum((a - b) * (a + b - 2 * tt)))

    pts = np.random.default_rng(64).choice(n ** 3, 64, replace=False)
    worst_k = worst_u = 0.0
    for p in pts:
        idx = np.unravel_index(p, (n, n, n))
        kp, km = tk.clone(), tk.clone()
        kp[idx] += h
        km[idx] -= h
        fd = dL(kp, tu, km, tu) / (2 * h)
        worst_k = max(worst_k, abs(fd - gk[p]) / max(abs(gk[p]), 1e-300))
        up, um = tu.clone(), tu.clone()
        up[idx] += h
        um[idx] -= h
        fd = dL(tk, up, tk, um) / (2 * h)
        worst_u = max(worst_u, abs(fd - gu[p]) / max(abs(gu[p]), 1e-300))
    print(f"vadi 128^3 x16: central FD (h=1e-6) worst relative deviation at 64 points: "
          f"kappa {worst_k:.2e}, u0 {worst_u:.2e}")
    assert worst_k <= 1e-4 and worst_u <= 1e-4

"""Differentiable wrappers over the adjoint kernels (include/trid_adjoint.h).

TridSolve: torch.autograd.Function for x = A^{-1} d along `solvedim` of
same-shaped tensors (a, b, c, d); backward runs the hand-written adjoint
kernel (transposed solve + fused coefficient gradients).  Used for the
gradient tests; it does not record anything for graph capture.

vadi_forward / vadi_loss_grad: the variable-diffusivity ADI forward pass and
its reverse-mode gradient (checkpoint every `ckpt` steps, recomputation).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _check, lib
from .api import _iv, dims_of, solve


class TridSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, b, c, d, solvedim: int):
        x = d.detach().clone().contiguous()
        solve(a.detach().contiguous(), b.detach().contiguous(), c.detach().contiguous(), x,
              solvedim)
        ctx.solvedim = solvedim
        ctx.save_for_backward(a, b, c, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        a, b, c, x = ctx.saved_tensors
        gx = gx.contiguous()
        ga, gb, gc, gd = (torch.empty_like(x) for _ in range(4))
        dims = dims_of(x)
        fn = (lib().tridDmtsvStridedBatchAdjoint if x.dtype == torch.float64
              else lib().tridSmtsvStridedBatchAdjoint)
        _check(fn(a.data_ptr(), b.data_ptr(), c.data_ptr(), x.data_ptr(), gx.data_ptr(),
                  ga.data_ptr(), gb.data_ptr(), gc.data_ptr(), gd.data_ptr(), len(dims),
                  ctx.solvedim, _iv(dims), _iv(dims), 1), "tridXmtsvStridedBatchAdjoint")
        return ga, gb, gc, gd, None


def trid_solve(a, b, c, d, solvedim: int):
    return TridSolve.apply(a, b, c, d, solvedim)


def _grid(t):
    nz, ny, nx = t.shape
    return nx, ny, nz


def vadi_forward(kappa, u0, steps: int, lam: float):
    u = u0.clone().contiguous()
    nx, ny, nz = _grid(u)
    _check(lib().tridVadiForward(kappa.data_ptr(), u.data_ptr(), nx, ny, nz, int(steps),
                                 float(lam), 1), "tridVadiForward")
    return u


def vadi_loss_grad(kappa, u0, target, steps: int, lam: float, ckpt: int = 4):
    """Returns (loss, grad_kappa, grad_u0)."""
    nx, ny, nz = _grid(u0)
    gk = torch.empty_like(kappa)
    gu = torch.empty_like(u0)
    loss = C.c_double(0.0)
    _check(lib().tridVadiLossGrad(kappa.data_ptr(), u0.data_ptr(), target.data_ptr(), nx, ny, nz,
                                  int(steps), float(lam), int(ckpt), C.byref(loss),
                                  gk.data_ptr(), gu.data_ptr()), "tridVadiLossGrad")
    return loss.value, gk, gu

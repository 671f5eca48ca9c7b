"""Python mirror of the tridsolver C API (reference: scalar/include/trid_cuda.h).

The functions keep the reference's names, argument order and meaning and
return its ``tridStatus_t`` codes (0 = success).  Array arguments are device
tensors (``torch.Tensor`` on a ROCm device) or raw device addresses (int);
``dims``/``pads``/``opts`` are sequences of ints.  Dimension 0 is the fastest
varying one, so a C-contiguous torch tensor of shape (nz, ny, nx) has
``dims = [nx, ny, nz]``.

:func:`solve` is a convenience wrapper that derives dims/pads from a tensor
and raises :class:`tridsolver_amd.TridError` on failure.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

from . import TridError, lib

__all__ = [
    "initTridMultiDimBatchSolve",
    "tridSmtsvStridedBatch", "tridDmtsvStridedBatch",
    "tridSmtsvStridedBatchInc", "tridDmtsvStridedBatchInc",
    "tridDmtsvStridedBatchPadded", "tridDmtsvStridedBatchPaddedInc",
    "tridSmtsvStridedBatchPadded", "tridSmtsvStridedBatchPaddedInc",
    "solve", "dims_of",
]


def _ptr(x) -> Optional[int]:
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return int(x.data_ptr())


def _iv(v: Optional[Sequence[int]]):
    if v is None:
        return None
    return (C.c_int * len(v))(*[int(t) for t in v])


def initTridMultiDimBatchSolve(ndim, dims, pads) -> None:
    lib().initTridMultiDimBatchSolve(int(ndim), _iv(dims), _iv(pads))


def _strided(name):
    def f(a, b, c, d, u, ndim, solvedim, dims, pads, opts=None, sync=1) -> int:
        fn = getattr(lib(), name)
        return fn(_ptr(a), _ptr(b), _ptr(c), _ptr(d), _ptr(u), int(ndim), int(solvedim),
                  _iv(dims), _iv(pads), _iv(opts), int(sync))
    f.__name__ = name
    f.__doc__ = f"{name}(a, b, c, d, u, ndim, solvedim, dims, pads, opts=None, sync=1) -> status"
    return f


def _padded(name):
    def f(a, a_pads, b, b_pads, c, c_pads, d, d_pads, u, u_pads, ndim, solvedim, dims,
          opts=None, sync=1) -> int:
        fn = getattr(lib(), name)
        return fn(_ptr(a), _iv(a_pads), _ptr(b), _iv(b_pads), _ptr(c), _iv(c_pads),
                  _ptr(d), _iv(d_pads), _ptr(u), _iv(u_pads), int(ndim), int(solvedim),
                  _iv(dims), _iv(opts), int(sync))
    f.__name__ = name
    return f


tridSmtsvStridedBatch = _strided("tridSmtsvStridedBatch")
tridDmtsvStridedBatch = _strided("tridDmtsvStridedBatch")
tridSmtsvStridedBatchInc = _strided("tridSmtsvStridedBatchInc")
tridDmtsvStridedBatchInc = _strided("tridDmtsvStridedBatchInc")
tridDmtsvStridedBatchPadded = _padded("tridDmtsvStridedBatchPadded")
tridDmtsvStridedBatchPaddedInc = _padded("tridDmtsvStridedBatchPaddedInc")
tridSmtsvStridedBatchPadded = _padded("tridSmtsvStridedBatchPadded")
tridSmtsvStridedBatchPaddedInc = _padded("tridSmtsvStridedBatchPaddedInc")


def dims_of(t) -> list:
    """dims (fastest first) of a C-contiguous tensor."""
    return [int(s) for s in reversed(t.shape)]


def solve(a, b, c, d, solvedim: int, u=None, inc: bool = False, sync: int = 1) -> None:
    """Solve along `solvedim` (0 = last torch axis) for same-shaped contiguous
    tensors; non-Inc overwrites d, Inc adds the solution into u."""
    import torch
    for t in (a, b, c, d) + ((u,) if inc else ()):
        if not t.is_contiguous() or t.shape != d.shape or t.dtype != d.dtype:
            raise ValueError("a, b, c, d (and u) must be contiguous, same shape and dtype")
    dims = dims_of(d)
    if d.dtype == torch.float64:
        fn = tridDmtsvStridedBatchInc if inc else tridDmtsvStridedBatch
    elif d.dtype == torch.float32:
        fn = tridSmtsvStridedBatchInc if inc else tridSmtsvStridedBatch
    else:
        raise TypeError(d.dtype)
    st = fn(a, b, c, d, u, len(dims), solvedim, dims, dims, None, sync)
    if st != 0:
        raise TridError(st, fn.__name__, lib().tridGetLastError().decode())

"""Block-tridiagonal solves (include/blktrid_hip.h) -- Python mirror.

Arrays are torch tensors in the reference GPU layout: A, B, C of shape
(N, P, M, M) (block (n, p) row-major), d and u of shape (N, P, M).
"""
from __future__ import annotations

import ctypes as C

from . import _check, lib


def _p(t):
    return None if t is None else t.data_ptr()


def dbtsv_gpu(A, B, Cm, d, Cstar, dstar, u, N, P, blkdim) -> None:
    """The reference's entry point (no status; errors via last_error())."""
    fn = lib().dbtsv_gpu if d.dtype.itemsize == 8 else lib().sbtsv_gpu
    fn(_p(A), _p(B), _p(Cm), _p(d), _p(Cstar), _p(dstar), _p(u), int(N), int(P), int(blkdim))


def btsv(A, B, Cm, d, u=None, Cstar=None, dstar=None, sync: int = 1):
    """Solve; returns u (allocated if not given)."""
    import torch
    N, P, M = d.shape
    if u is None:
        u = torch.empty_like(d)
    fn = lib().tridDbtsvStridedBatch if d.dtype == torch.float64 else lib().tridSbtsvStridedBatch
    _check(fn(_p(A), _p(B), _p(Cm), _p(d), _p(Cstar), _p(dstar), _p(u), N, P, M, int(sync)),
           "tridXbtsvStridedBatch")
    return u


class _Group(C.Structure):
    _fields_ = [("A", C.c_void_p), ("B", C.c_void_p), ("C", C.c_void_p), ("d", C.c_void_p),
                ("Cstar", C.c_void_p), ("dstar", C.c_void_p), ("u", C.c_void_p),
                ("N", C.c_int), ("P", C.c_int), ("M", C.c_int)]


def btsv_mixed(groups, sync: int = 1) -> None:
    """groups: list of dicts with A, B, C, d, Cstar, dstar, u tensors (FP64)."""
    arr = (_Group * len(groups))()
    for i, g in enumerate(groups):
        N, P, M = g["d"].shape
        arr[i] = _Group(_p(g["A"]), _p(g["B"]), _p(g["C"]), _p(g["d"]), _p(g["Cstar"]),
                        _p(g["dstar"]), _p(g["u"]), N, P, M)
    _check(lib().tridDbtsvMixedBatch(C.cast(arr, C.c_void_p), len(groups), int(sync)),
           "tridDbtsvMixedBatch")


def btsv_adjoint(A, B, Cm, u, g, sync: int = 1):
    """Returns (grad_A, grad_B, grad_C, grad_d) for dL/du = g."""
    import torch
    N, P, M = u.shape
    gA, gB, gC = (torch.empty_like(A) for _ in range(3))
    gd = torch.empty_like(u)
    _check(lib().tridDbtsvAdjoint(_p(A), _p(B), _p(Cm), _p(u), _p(g), _p(gA), _p(gB), _p(gC),
                                  _p(gd), N, P, M, int(sync)), "tridDbtsvAdjoint")
    return gA, gB, gC, gd


def force_lanes(lanes: int) -> None:
    _check(lib().tridBlkForceLanes(int(lanes)), "tridBlkForceLanes")

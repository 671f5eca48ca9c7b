"""Ragged batches in one persistent launch (include/trid_ragged.h).

    plan = RaggedPlan(offsets, torch.float64)   # offsets: int64 device tensor, nsys+1
    plan.solve(a, b, c, d)                      # d <- solution, per system
    plan.solve(a, b, c, d, u=u, inc=True)       # u += solution
"""
from __future__ import annotations

import ctypes as C

from . import _check, lib


def check_arrays(extent: int, dtype, a, b, c, d, u=None, inc: bool = False) -> None:
    """Raises ValueError unless a, b, c, d (and u for inc) are contiguous
    tensors of `dtype` holding at least `extent` (= offsets[nsys]) elements:
    the kernels access every element the offsets name."""
    named = [("a", a), ("b", b), ("c", c), ("d", d)]
    if inc:
        if u is None:
            raise ValueError("u is required with inc=True")
        named.append(("u", u))
    for name, t in named:
        if t.dtype != dtype:
            raise ValueError(f"{name}: dtype {t.dtype}, the plan's is {dtype}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        if t.numel() < extent:
            raise ValueError(f"{name} holds {t.numel()} elements; the offsets name {extent}")


def force_choice(family: int = -1, red: int = -1) -> None:
    """Plans created afterwards use row family `family` (0: chunks of 5..8
    rows, 1: 9..16) and reduced solve `red` past 64 chunks (0: PCR through
    LDS, 1: per-wave PCR and a sweep across waves); -1 = automatic."""
    _check(lib().tridRaggedForceChoice(int(family), int(red)), "tridRaggedForceChoice")


class RaggedPlan:
    def __init__(self, offsets, dtype):
        import torch
        if offsets.dtype != torch.int64 or not offsets.is_contiguous():
            raise ValueError("offsets must be a contiguous int64 device tensor")
        self.offsets = offsets  # the plan refers to this device array
        self.dtype = dtype
        h = C.c_void_p()
        _check(lib().tridRaggedPlanCreate(C.byref(h), offsets.data_ptr(), offsets.numel() - 1,
                                          int(dtype == torch.float64)), "tridRaggedPlanCreate")
        self._h = h
        ext = C.c_longlong()
        _check(lib().tridRaggedPlanExtent(h, C.byref(ext)), "tridRaggedPlanExtent")
        self.extent = int(ext.value)

    def solve(self, a, b, c, d, u=None, inc: bool = False, sync: int = 1) -> None:
        import torch
        check_arrays(self.extent, self.dtype, a, b, c, d, u, inc)
        fn = lib().tridDgtsvRaggedBatch if self.dtype == torch.float64 else lib().tridSgtsvRaggedBatch
        _check(fn(self._h, a.data_ptr(), b.data_ptr(), c.data_ptr(), d.data_ptr(),
                  None if u is None else u.data_ptr(), int(bool(inc)), int(sync)),
               "tridXgtsvRaggedBatch")

    def describe(self) -> str:
        return lib().tridRaggedPlanDescribe(self._h).decode()

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().tridRaggedPlanDestroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

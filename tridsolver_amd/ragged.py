"""Ragged batches in one persistent launch (include/trid_ragged.h).

    plan = RaggedPlan(offsets, torch.float64)   # offsets: int64 device tensor, nsys+1
    plan.solve(a, b, c, d)                      # d <- solution, per system
    plan.solve(a, b, c, d, u=u, inc=True)       # u += solution
"""
from __future__ import annotations

import ctypes as C

from . import _check, lib


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

    def solve(self, a, b, c, d, u=None, inc: bool = False, sync: int = 1) -> None:
        import torch
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

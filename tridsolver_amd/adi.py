"""ADI heat-diffusion time stepping (include/trid_adi.h) -- Python mirror.

    plan = AdiPlan(nx, ny, nz, torch.float64, lam=1.0, mode=1)
    plan.step(u, du, nsteps)          # u updated in place, du scratch
    plan.capture(u, du, nsteps); plan.replay()   # HIP graph

mode 0: the reference's coefficients a = c = -lambda/2, b = 1 + lambda
(apps/adi/src/preproc.cu:70-72); mode 1: a = c = -lambda, b = 1 + 2 lambda.
"""
from __future__ import annotations

import ctypes as C

from . import _check, lib


class AdiPlan:
    def __init__(self, nx: int, ny: int, nz: int, dtype, lam: float = 1.0, mode: int = 1):
        import torch
        self.shape = (nz, ny, nx)
        self.dtype = dtype
        h = C.c_void_p()
        _check(lib().tridAdiPlanCreate(C.byref(h), nx, ny, nz, int(dtype == torch.float64),
                                       float(lam), int(mode)), "tridAdiPlanCreate")
        self._h = h

    def _ok(self, *ts):
        for t in ts:
            if tuple(t.shape) != self.shape or t.dtype != self.dtype or not t.is_contiguous():
                raise ValueError(f"expected contiguous {self.dtype} tensors of shape {self.shape}")

    def step(self, u, du, nsteps: int = 1, sync: int = 1) -> None:
        self._ok(u, du)
        _check(lib().tridAdiStep(self._h, u.data_ptr(), du.data_ptr(), int(nsteps), int(sync)),
               "tridAdiStep")

    def stage(self, u, du, stage: int) -> None:
        self._ok(u, du)
        _check(lib().tridAdiStage(self._h, u.data_ptr(), du.data_ptr(), int(stage)), "tridAdiStage")

    def capture(self, u, du, nsteps: int) -> None:
        self._ok(u, du)
        _check(lib().tridAdiGraphCapture(self._h, u.data_ptr(), du.data_ptr(), int(nsteps)),
               "tridAdiGraphCapture")

    def replay(self, sync: int = 1) -> None:
        _check(lib().tridAdiGraphLaunch(self._h, int(sync)), "tridAdiGraphLaunch")

    def describe(self) -> str:
        return lib().tridAdiPlanDescribe(self._h).decode()

    def set_sweep(self, direction: int, rows: int, pair: int = 0, xcd: int = 0) -> None:
        """Pins the y (1) or z (2) sweep: rows 8/16/32 = precomputed factors
        with that chunk length, 0 = the partitioned sweep.  Direction 0 pins
        the fused right-hand side + x stage: rows 4/8/16, pair = 1 for
        precomputed factors of the x sweep."""
        _check(lib().tridAdiPlanSetSweep(self._h, int(direction), int(rows), int(pair), int(xcd)),
               "tridAdiPlanSetSweep")

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().tridAdiPlanDestroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

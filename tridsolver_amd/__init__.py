"""tridsolver_amd -- MI355X-native batched tridiagonal solvers.

Python mirror of the C ABI in ``include/*.h`` (the drop-in replacement for
TobyFlynn/tridsolver's ``trid_cuda.h`` and ``blktrid_gpu.h``).  Every call
goes through ctypes into ``tridsolver_amd/libtridhip.so``; there is no other
implementation behind these functions.  If the library is missing the import
of any entry point raises: build it with ``python -m tridsolver_amd.build``.

Device memory and streams come from PyTorch (ROCm): tensors are passed by
``data_ptr()``; the library enqueues on the stream set with
:func:`set_stream` (default: the null stream, which is PyTorch's default
stream too).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtridhip.so")

TRID_STATUS = {
    0: "TRID_STATUS_SUCCESS",
    1: "TRID_STATUS_NOT_INITIALIZED",
    2: "TRID_STATUS_ALLOC_FAILED",
    3: "TRID_STATUS_INVALID_VALUE",
    4: "TRID_STATUS_ARCH_MISMATCH",
    5: "TRID_STATUS_MAPPING_ERROR",
    6: "TRID_STATUS_EXECUTION_FAILED",
    7: "TRID_STATUS_INTERNAL_ERROR",
    8: "TRID_STATUS_MATRIX_TYPE_NOT_SUPPORTED",
    9: "TRID_STATUS_ZERO_PIVOT",
}

_lib = None


class TridError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str = ""):
        self.status = status
        super().__init__(f"{where}: {TRID_STATUS.get(status, status)}" + (f" ({msg})" if msg else ""))


def lib() -> C.CDLL:
    """The loaded libtridhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m tridsolver_amd.build` "
                "(there is no non-HIP fallback)")
        _lib = C.CDLL(LIB_PATH)
        from . import _abi
        _abi.declare(_lib)
    return _lib


def _check(st: int, where: str) -> None:
    if st != 0:
        msg = lib().tridGetLastError().decode(errors="replace")
        raise TridError(st, where, msg)


def set_stream(stream) -> None:
    """Stream (torch.cuda.Stream, raw handle int, or None) for all launches."""
    h = 0
    if stream is not None:
        h = int(getattr(stream, "cuda_stream", stream))
    _check(lib().tridSetStream(C.c_void_p(h)), "tridSetStream")


def set_autotune(mode: int) -> None:
    _check(lib().tridSetAutotune(int(mode)), "tridSetAutotune")


def force_config(kind: int = 0, rows: int = 0, cps: int = 0, spb: int = 0) -> None:
    _check(lib().tridForceConfig(kind, rows, cps, spb), "tridForceConfig")


def force_wide_index(on: bool) -> None:
    """Makes later calls use the 64-bit offset forms of the kernels (taken
    automatically for arrays of 4 GiB or more) whatever the sizes."""
    _check(lib().tridForceWideIndex(1 if on else 0), "tridForceWideIndex")


def force_cache_policy(policy: int) -> None:
    """Cache policy of the solvers' streaming loads and stores: 0 non-temporal,
    1 ordinary (arrays that fit the 256 MB last-level cache stay there for
    the next kernel or call), 2 non-temporal coefficient arrays with ordinary
    d and u, -1 automatic (ordinary up to 384 MiB of arrays per call or ADI
    step, else 2 while d and u fit in 384 MiB, else 0).  All give bitwise
    equal results."""
    _check(lib().tridForceCachePolicy(int(policy)), "tridForceCachePolicy")


def last_config() -> str:
    return lib().tridGetLastConfig().decode()


def last_error() -> str:
    return lib().tridGetLastError().decode()


from .api import *  # noqa: E402,F401,F403

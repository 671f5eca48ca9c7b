"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's ``cpu_baseline`` leg, always as the checker or as the timed
CPU baseline -- never by the product path (the ``tridsolver_amd`` package).

Arrays are numpy arrays; every wrapper works in place like the C routine it
calls.  The C routines cite the reference lines they restate.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and, with ref=True, the reference binaries)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
    if ref:
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_P = C.c_void_p
_I = C.c_int
_L = C.c_long


def _declare(L: C.CDLL) -> None:
    for suf in ("d", "s"):
        f = getattr(L, f"oracle_mtsv_{suf}")
        f.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _P, _I]
        f.restype = _I
        f = getattr(L, f"oracle_ragged_{suf}")
        f.argtypes = [_P, _P, _P, _P, _P, _P, _L, _I]
        f.restype = _I
        f = getattr(L, f"oracle_trid_scalar_{suf}")
        f.argtypes = [_P, _P, _P, _P, _P, _I, _I, _I]
        f.restype = None
    L.oracle_adi_d.argtypes = [_I, C.c_double, _P, _I, _I, _I, _I]
    L.oracle_adi_d.restype = _I
    L.oracle_adi_s.argtypes = [_I, C.c_float, _P, _I, _I, _I, _I]
    L.oracle_adi_s.restype = _I
    L.oracle_adi_ws_d.argtypes = [_I, C.c_double, _P, _I, _I, _I, _I, _P]
    L.oracle_adi_ws_d.restype = _I
    L.oracle_adi_ws_s.argtypes = [_I, C.c_float, _P, _I, _I, _I, _I, _P]
    L.oracle_adi_ws_s.restype = _I
    L.oracle_adjoint_d.argtypes = [_P] * 9 + [_I, _I, _P, _P]
    L.oracle_adjoint_d.restype = _I
    L.oracle_btsv_d.argtypes = [_P, _P, _P, _P, _P, _I, _I, _I, _I]
    L.oracle_btsv_d.restype = _I
    L.oracle_vadi_forward.argtypes = [_P, _P, _P, _I, _I, C.c_double]
    L.oracle_vadi_forward.restype = _I
    L.oracle_vadi_grad.argtypes = [_P, _P, _P, _I, _I, C.c_double, _P, _P, _P]
    L.oracle_vadi_grad.restype = _I
    L.oracle_num_threads.restype = _I
    L.oracle_set_num_threads.argtypes = [_I]


def _ptr(x):
    if x is None:
        return None
    assert isinstance(x, np.ndarray) and x.flags.c_contiguous, "contiguous numpy array expected"
    return x.ctypes.data


def _ipt(v):
    if v is None:
        return None
    arr = (C.c_int * len(v))(*[int(t) for t in v])
    return arr


def _suf(x: np.ndarray) -> str:
    if x.dtype == np.float64:
        return "d"
    if x.dtype == np.float32:
        return "s"
    raise TypeError(x.dtype)


def num_threads() -> int:
    return lib().oracle_num_threads()


def set_num_threads(n: int) -> None:
    lib().oracle_set_num_threads(n)


def mtsv(a, b, c, d, u, ndim, solvedim, dims, pads=None, inc=False,
         a_pads=None, b_pads=None, c_pads=None, d_pads=None, u_pads=None) -> int:
    """Strided-batch solve on flat buffers (in place: d, or u when inc)."""
    pads = list(pads) if pads is not None else list(dims)
    ap = _ipt(a_pads or pads)
    bp = _ipt(b_pads or pads)
    cp = _ipt(c_pads or pads)
    dp = _ipt(d_pads or pads)
    up = _ipt(u_pads or pads)
    f = getattr(lib(), f"oracle_mtsv_{_suf(d)}")
    return f(_ptr(a), ap, _ptr(b), bp, _ptr(c), cp, _ptr(d), dp, _ptr(u), up,
             ndim, solvedim, _ipt(dims), int(bool(inc)))


def trid_scalar(a, b, c, d, u, N, stride, inc=False) -> None:
    f = getattr(lib(), f"oracle_trid_scalar_{_suf(d)}")
    f(_ptr(a), _ptr(b), _ptr(c), _ptr(d), _ptr(u), N, stride, int(bool(inc)))


def ragged(a, b, c, d, u, offsets, inc=False) -> int:
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    f = getattr(lib(), f"oracle_ragged_{_suf(d)}")
    return f(_ptr(a), _ptr(b), _ptr(c), _ptr(d), _ptr(u), offsets.ctypes.data,
             len(offsets) - 1, int(bool(inc)))


def adi(u, nx, ny, nz, iters, lam=1.0, mode=0, ws=None) -> int:
    """ADI time steps on u in place; ws (optional): work buffer of
    4*nx*ny*nz values of u's dtype, reused across calls."""
    if ws is not None:
        assert ws.dtype == u.dtype and ws.size >= 4 * nx * ny * nz
        f = lib().oracle_adi_ws_d if u.dtype == np.float64 else lib().oracle_adi_ws_s
        return f(mode, lam, _ptr(u), nx, ny, nz, iters, _ptr(ws))
    if u.dtype == np.float64:
        return lib().oracle_adi_d(mode, lam, _ptr(u), nx, ny, nz, iters)
    return lib().oracle_adi_s(mode, lam, _ptr(u), nx, ny, nz, iters)


def adjoint(a, b, c, x, g, ndim, solvedim, dims, pads=None):
    """Returns (grad_a, grad_b, grad_c, grad_d) of L with dL/dx = g."""
    pads = list(pads) if pads is not None else list(dims)
    ga, gb, gc, gd = (np.zeros_like(a) for _ in range(4))
    st = lib().oracle_adjoint_d(_ptr(a), _ptr(b), _ptr(c), _ptr(x), _ptr(g), _ptr(ga),
                                _ptr(gb), _ptr(gc), _ptr(gd), ndim, solvedim,
                                _ipt(dims), _ipt(pads))
    if st:
        raise RuntimeError(f"oracle_adjoint_d status {st}")
    return ga, gb, gc, gd


def btsv(A, B, C_, d, N, P, M, layout=0):
    """Block Thomas; returns u.  layout 0 = GPU row-interleaved, 1 = system-major."""
    u = np.zeros_like(d)
    st = lib().oracle_btsv_d(_ptr(A), _ptr(B), _ptr(C_), _ptr(d), _ptr(u), N, P, M, layout)
    if st:
        raise RuntimeError(f"oracle_btsv_d status {st}")
    return u


def vadi_forward(kappa, u0, n, steps, lam):
    uT = np.empty_like(u0)
    st = lib().oracle_vadi_forward(_ptr(kappa), _ptr(u0), _ptr(uT), n, steps, lam)
    if st:
        raise RuntimeError("oracle_vadi_forward failed")
    return uT


def vadi_grad(kappa, u0, target, n, steps, lam):
    """Returns (loss, grad_kappa, grad_u0)."""
    gk = np.empty_like(kappa)
    gu = np.empty_like(u0)
    loss = C.c_double(0.0)
    st = lib().oracle_vadi_grad(_ptr(kappa), _ptr(u0), _ptr(target), n, steps, lam,
                                C.addressof(loss), _ptr(gk), _ptr(gu))
    if st:
        raise RuntimeError("oracle_vadi_grad failed")
    return loss.value, gk, gu


# ---------------------------------------------------------------------------
# tolerance of the reference's tests (scalar/test/utils.hpp:14-17 and the
# mixed abs+rel check of require_allclose, scalar/test/test_cpu.cpp:9-32)
ABS_TOL = {np.float64: 1e-11, np.float32: 1e-6}
REL_TOL = {np.float64: 1e-11, np.float32: 1e-5}


def allclose_ref(expected, actual, abs_tol=None, rel_tol=None):
    """|e-a| <= abs + rel*min(|e|,|a|) elementwise (reference test semantics).

    Returns (ok, worst_index, worst_excess)."""
    e = np.asarray(expected, dtype=np.float64)
    a = np.asarray(actual, dtype=np.float64)
    dt = np.float32 if np.asarray(actual).dtype == np.float32 else np.float64
    at = ABS_TOL[dt] if abs_tol is None else abs_tol
    rt = REL_TOL[dt] if rel_tol is None else rel_tol
    tol = at + rt * np.minimum(np.abs(e), np.abs(a))
    diff = np.abs(e - a)
    bad = ~(diff <= tol)
    if not bad.any():
        return True, -1, 0.0
    excess = np.where(bad, diff - tol, -np.inf)
    i = int(np.argmax(excess))
    return False, i, float(diff.flat[i])


def rel_err(expected, actual) -> float:
    """Normwise relative error max|e-a| / max|e|."""
    e = np.asarray(expected, dtype=np.float64)
    a = np.asarray(actual, dtype=np.float64)
    den = float(np.max(np.abs(e))) if e.size else 1.0
    return float(np.max(np.abs(e - a))) / (den if den > 0 else 1.0) if e.size else 0.0

"""Shared generators and checks of the test-suite."""
from __future__ import annotations

import numpy as np

import oracle


def cum(pads):
    out, c = [], 1
    for p in pads:
        out.append(c)
        c *= p
    return out, c


def extent(dims, pads):
    if any(d == 0 for d in dims):
        return 0
    cp, _ = cum(pads)
    return sum((d - 1) * c for d, c in zip(dims, cp)) + 1


def alloc(dims, pads, dtype, fill=np.nan):
    """Flat buffer holding an array of `dims` with allocation `pads`."""
    _, tot = cum(pads)
    return np.full(tot, fill, dtype=dtype)


def view(buf, dims, pads):
    """numpy view (axis order reversed: last axis = dim 0) of the valid region."""
    shape = tuple(reversed(pads))
    v = buf.reshape(shape)
    sl = tuple(slice(0, d) for d in reversed(dims))
    return v[sl]


def dd_system(dims, solvedim, dtype, seed, pads=None, zero_ends=False, pad_fill=np.nan,
              pads_per_array=None):
    """Diagonally dominant batch of the benchmark distribution:
    a, c ~ U(-1,-0.1), b = |a|+|c|+U(0.5,1.5), d ~ U(-1,1)  (numpy PCG64,
    seed `seed`, drawn in the order a, b, c, d over the dims-shaped grid).
    With zero_ends, a_0 = c_{N-1} = 0 along solvedim.  Returns flat buffers
    (a, b, c, d) laid out with `pads` (or per-array pads), padding = pad_fill."""
    rng = np.random.default_rng(seed)
    shape = tuple(reversed(dims))
    a = rng.uniform(-1.0, -0.1, shape)
    c = rng.uniform(-1.0, -0.1, shape)
    b = np.abs(a) + np.abs(c) + rng.uniform(0.5, 1.5, shape)
    d = rng.uniform(-1.0, 1.0, shape)
    if zero_ends:
        ax = len(dims) - 1 - solvedim
        idx0 = [slice(None)] * len(dims)
        idx0[ax] = 0
        a[tuple(idx0)] = 0.0
        idx1 = [slice(None)] * len(dims)
        idx1[ax] = -1
        c[tuple(idx1)] = 0.0
    out = []
    for i, x in enumerate((a, b, c, d)):
        p = pads_per_array[i] if pads_per_array else (pads or dims)
        buf = alloc(dims, p, dtype, pad_fill)
        view(buf, dims, p)[...] = x.astype(dtype)
        out.append(buf)
    return tuple(out)


def oracle_solve(a, b, c, d, u, dims, solvedim, pads=None, inc=False, per_array=None):
    """Runs the oracle on copies; returns the written array (d, or u if inc)."""
    a, b, c, d = (x.copy() for x in (a, b, c, d))
    u = None if u is None else u.copy()
    kw = {}
    if per_array:
        kw = dict(a_pads=per_array[0], b_pads=per_array[1], c_pads=per_array[2],
                  d_pads=per_array[3], u_pads=per_array[4])
    st = oracle.mtsv(a, b, c, d, u, len(dims), solvedim, dims, pads, inc=inc, **kw)
    assert st == 0
    return u if inc else d


def check_close(expected, actual, dtype, what="", rel=None):
    """Normwise relative error within the stated tolerance (1e-12 FP64, 1e-5
    FP32) AND the reference tests' elementwise mixed tolerance."""
    exp = np.asarray(expected, dtype=np.float64)
    act = np.asarray(actual, dtype=np.float64)
    mask = ~np.isnan(exp)
    assert np.array_equal(mask, ~np.isnan(act)), f"{what}: NaN pattern differs (padding written?)"
    e, a = exp[mask], act[mask]
    r = oracle.rel_err(e, a)
    lim = rel if rel is not None else (1e-12 if dtype == np.float64 else 1e-5)
    assert r <= lim, f"{what}: normwise relative error {r:.3e} > {lim:.0e}"
    dt = np.float32 if dtype == np.float32 else np.float64
    ok, i, diff = oracle.allclose_ref(e, a, abs_tol=oracle.ABS_TOL[dt], rel_tol=oracle.REL_TOL[dt])
    assert ok, f"{what}: element {i} differs by {diff:.3e} (reference tolerance)"
    return r

"""Pins the CPU oracle (oracle/*.c) before any GPU result is compared with it.

Which check fixes which oracle function:
  oracle_mtsv_{d,s} + thomas (restating scalar/src/cpu/trid_cpu.cpp:360-531)
      - bit-exact against the reference's own gold ADI (apps/adi/src/adi_orig.c,
        compiled unmodified; fixtures in tests/golden/) whose steps call the same
        Thomas recurrence along x, y (into du) and z (then u += du = the Inc form);
      - dense LU solves (numpy.linalg.solve) of every system for ndim 1..4, every
        solvedim, pads != dims and per-array pads, Inc and non-Inc;
      - a closed form: constant-coefficient systems solved by the type-I sine
        transform (scipy.fft.dst), and a manufactured solution d = A x.
  oracle_ragged_{d,s}: per-system dense solves.
  oracle_btsv_d (restating block/src/blktrid_cpu.cpp:388-504, index bug fixed):
      - dense LU of the assembled (N*M)^2 matrix, both layouts, M in 1..8;
      - the reference's own CPU block solver (compiled unmodified) on diagonal
        blocks, the inputs its gold_MMmultiply index error cannot affect
        (fixtures blk_ref_diag_M*.npz).
  oracle_adjoint_d: the dot-product identity <A^-1 d, g> = <d, A^-T g> and
      central finite differences of every coefficient gradient.
  oracle_vadi_forward/grad: the forward step against an independent dense
      numpy restatement on a tiny grid, and central finite differences of the
      loss at random points (h = 1e-6).
  oracle_adi_{d,s}: the gold ADI fixtures above (reference mode).
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest
import scipy.fft

import oracle
from tests.helpers import alloc, dd_system, oracle_solve, view

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def dense_solve_batch(a, b, c, d, dims, solvedim, pads, u=None):
    """Independent restatement: build each system's dense matrix and LU-solve."""
    av, bv, cv, dv = (view(x, dims, pads).astype(np.float64) for x in (a, b, c, d))
    ax = len(dims) - 1 - solvedim
    av, bv, cv, dv = (np.moveaxis(x, ax, -1) for x in (av, bv, cv, dv))
    N = dims[solvedim]
    out = np.empty_like(dv)
    for idx in np.ndindex(av.shape[:-1]):
        A = np.diag(bv[idx])
        if N > 1:
            A += np.diag(av[idx][1:], -1) + np.diag(cv[idx][:-1], 1)
        out[idx] = np.linalg.solve(A, dv[idx])
    return np.moveaxis(out, -1, ax)


# ---------------------------------------------------------------------------
# gold ADI (reference apps/adi/src/adi_orig.c) -- bit-exact


def adi_cases():
    with open(os.path.join(GOLD, "golden_meta.json")) as f:
        meta = json.load(f)
    return sorted(k for k in meta if k.startswith("adi_ref_"))


@pytest.mark.parametrize("name", adi_cases())
def test_oracle_adi_bitexact_vs_reference_gold(name):
    with open(os.path.join(GOLD, "golden_meta.json")) as f:
        m = json.load(f)[name]
    dt = np.float64 if m["dtype"] == "f64" else np.float32
    nx, ny, nz, it = m["nx"], m["ny"], m["nz"], m["iter"]
    ref = np.fromfile(os.path.join(GOLD, name), dtype=dt)
    u = np.zeros((nz, ny, nx), dt)
    u[0], u[-1], u[:, 0], u[:, -1], u[:, :, 0], u[:, :, -1] = 1, 1, 1, 1, 1, 1
    u = u.ravel().copy()
    assert oracle.adi(u, nx, ny, nz, it, 1.0, 0) == 0
    assert np.array_equal(u, ref), f"max diff {np.max(np.abs(u - ref))}"
    # the caller-workspace form, one step per call on a reused (dirty)
    # buffer, reproduces the gold run bit for bit
    u2 = np.zeros((nz, ny, nx), dt)
    u2[0], u2[-1], u2[:, 0], u2[:, -1], u2[:, :, 0], u2[:, :, -1] = 1, 1, 1, 1, 1, 1
    u2 = u2.ravel().copy()
    ws = np.full(4 * u2.size, np.nan, dt)
    for _ in range(it):
        assert oracle.adi(u2, nx, ny, nz, 1, 1.0, 0, ws=ws) == 0
    assert np.array_equal(u2, ref)


# ---------------------------------------------------------------------------
# strided batch vs dense LU

SHAPES = [
    ([7], 0), ([1], 0), ([2], 0), ([9, 5], 0), ([9, 5], 1), ([6, 5, 4], 0), ([6, 5, 4], 1),
    ([6, 5, 4], 2), ([3, 4, 2, 5], 0), ([3, 4, 2, 5], 1), ([3, 4, 2, 5], 2), ([3, 4, 2, 5], 3),
]


@pytest.mark.parametrize("dims,sd", SHAPES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("inc", [False, True])
def test_oracle_mtsv_vs_dense(dims, sd, dtype, inc):
    pads = [x + (k % 2) * 2 for k, x in enumerate(dims)]  # pads != dims on odd dims
    a, b, c, d = dd_system(dims, sd, dtype, seed=11 + sd, pads=pads)
    u = None
    if inc:
        u = alloc(dims, pads, dtype)
        view(u, dims, pads)[...] = np.random.default_rng(3).uniform(-1, 1, tuple(reversed(dims)))
    got = oracle_solve(a, b, c, d, u, dims, sd, pads, inc)
    exp = dense_solve_batch(a, b, c, d, dims, sd, pads)
    if inc:
        exp = exp + view(u, dims, pads).astype(np.float64)
    tol = 1e-13 if dtype == np.float64 else 2e-6
    gv = view(got, dims, pads).astype(np.float64)
    assert np.max(np.abs(gv - exp)) <= tol * max(1.0, np.max(np.abs(exp)))
    # the padding is never written
    mask = np.ones(got.shape, bool)
    view(mask, dims, pads)[...] = False
    assert np.all(np.isnan(got[mask]))


def test_oracle_per_array_pads_vs_dense():
    dims, sd = [5, 4, 3], 1
    pa = [[5, 4, 3], [7, 4, 3], [5, 6, 4], [8, 5, 3], [6, 4, 5]]
    a, b, c, d = dd_system(dims, sd, np.float64, seed=5, pads_per_array=pa[:4])
    u = alloc(dims, pa[4], np.float64)
    view(u, dims, pa[4])[...] = 0.5
    got = oracle_solve(a, b, c, d, u, dims, sd, inc=True, per_array=pa)
    ref = np.zeros(tuple(reversed(dims)))
    # solve on compacted copies
    comp = [np.ascontiguousarray(view(x, dims, p)).ravel() for x, p in zip((a, b, c, d), pa)]
    ref = dense_solve_batch(*comp, dims, sd, dims) + 0.5
    np.testing.assert_allclose(view(got, dims, pa[4]), ref, rtol=1e-13, atol=1e-14)


def test_oracle_ignores_a0_and_cN():
    dims, sd = [6, 4], 0
    a, b, c, d = dd_system(dims, sd, np.float64, seed=2)
    r1 = oracle_solve(a, b, c, d, None, dims, sd)
    a2, c2 = a.copy(), c.copy()
    view(a2, dims, dims)[:, 0] = 1e30
    view(c2, dims, dims)[:, -1] = -7e20
    r2 = oracle_solve(a2, b, c2, d, None, dims, sd)
    np.testing.assert_array_equal(r1, r2)


@pytest.mark.parametrize("N", [1, 2, 3, 17, 256, 1000])
def test_oracle_closed_form_sine_transform(N):
    """tridiag(alpha, beta, alpha) x = f has x = S diag(1/lambda_k) S f with the
    orthogonal DST-I basis S and lambda_k = beta + 2 alpha cos(k pi/(N+1))."""
    alpha, beta = -0.7, 2.3
    f = np.random.default_rng(N).uniform(-1, 1, N)
    k = np.arange(1, N + 1)
    lam = beta + 2 * alpha * np.cos(k * np.pi / (N + 1))
    x_exact = scipy.fft.idst(scipy.fft.dst(f, type=1) / lam, type=1)
    a = np.full(N, alpha)
    b = np.full(N, beta)
    c = np.full(N, alpha)
    d = f.copy()
    oracle.trid_scalar(a, b, c, d, None, N, 1)
    np.testing.assert_allclose(d, x_exact, rtol=0, atol=1e-13)


def test_oracle_manufactured_solution_f32_vs_f64():
    dims, sd = [64, 33], 1
    rng = np.random.default_rng(9)
    a, b, c, _ = dd_system(dims, sd, np.float64, seed=4)
    x = rng.uniform(-1, 1, dims[::-1])
    av, bv, cv = (view(t, dims, dims) for t in (a, b, c))
    d = b.copy()
    dv = view(d, dims, dims)
    dv[...] = bv * x
    dv[1:, :] += av[1:, :] * x[:-1, :]
    dv[:-1, :] += cv[:-1, :] * x[1:, :]
    got = oracle_solve(a, b, c, d, None, dims, sd)
    np.testing.assert_allclose(view(got, dims, dims), x, atol=1e-14)
    got32 = oracle_solve(*(t.astype(np.float32) for t in (a, b, c, d)), None, dims, sd)
    assert np.max(np.abs(view(got32, dims, dims) - x)) < 1e-5


# ---------------------------------------------------------------------------
# ragged


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_oracle_ragged_vs_dense(dtype):
    rng = np.random.default_rng(7)
    lens = np.array([1, 2, 5, 17, 3, 64, 8, 1, 130])
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(off[-1])
    a = rng.uniform(-1, -0.1, n).astype(dtype)
    c = rng.uniform(-1, -0.1, n).astype(dtype)
    b = (np.abs(a) + np.abs(c) + rng.uniform(0.5, 1.5, n)).astype(dtype)
    d = rng.uniform(-1, 1, n).astype(dtype)
    u = rng.uniform(-1, 1, n).astype(dtype)
    dd, uu = d.copy(), u.copy()
    assert oracle.ragged(a, b, c, dd, None, off) == 0
    assert oracle.ragged(a, b, c, d.copy(), uu, off, inc=True) == 0
    for k in range(len(lens)):
        s, e = off[k], off[k + 1]
        L = e - s
        A = np.diag(b[s:e].astype(np.float64))
        if L > 1:
            A += np.diag(a[s + 1:e], -1) + np.diag(c[s:e - 1], 1)
        x = np.linalg.solve(A, d[s:e].astype(np.float64))
        tol = 1e-13 if dtype == np.float64 else 3e-6
        np.testing.assert_allclose(dd[s:e], x, atol=tol)
        np.testing.assert_allclose(uu[s:e], u[s:e] + x, atol=tol)


# ---------------------------------------------------------------------------
# block


def blk_system(M, N, P, seed, layout=0, full=True):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, (N, P, M, M)) / M
    C = rng.uniform(-1, 1, (N, P, M, M)) / M
    B = rng.uniform(-1, 1, (N, P, M, M)) / M
    rs = np.abs(A).sum(-1) + np.abs(C).sum(-1)
    B += (rs + 1.0)[..., None] * np.eye(M)
    A[0] = 0.0
    C[-1] = 0.0
    d = rng.uniform(-1, 1, (N, P, M))
    if layout == 1:
        A, B, C = (np.ascontiguousarray(x.transpose(1, 0, 2, 3)) for x in (A, B, C))
        d = np.ascontiguousarray(d.transpose(1, 0, 2))
    return A, B, C, d


def blk_dense(A, B, C, d, layout):
    if layout == 1:
        A, B, C = (x.transpose(1, 0, 2, 3) for x in (A, B, C))
        d = d.transpose(1, 0, 2)
    N, P, M, _ = A.shape
    out = np.empty((N, P, M))
    for p in range(P):
        K = np.zeros((N * M, N * M))
        for n in range(N):
            K[n * M:(n + 1) * M, n * M:(n + 1) * M] = B[n, p]
            if n > 0:
                K[n * M:(n + 1) * M, (n - 1) * M:n * M] = A[n, p]
            if n < N - 1:
                K[n * M:(n + 1) * M, (n + 1) * M:(n + 2) * M] = C[n, p]
        out[:, p] = np.linalg.solve(K, d[:, p].ravel()).reshape(N, M)
    if layout == 1:
        out = out.transpose(1, 0, 2)
    return out


@pytest.mark.parametrize("M", [1, 2, 3, 4, 5, 8])
@pytest.mark.parametrize("layout", [0, 1])
def test_oracle_block_vs_dense(M, layout):
    N, P = 9, 4
    A, B, C, d = blk_system(M, N, P, seed=M, layout=layout)
    u = oracle.btsv(*(np.ascontiguousarray(x) for x in (A, B, C, d)), N, P, M, layout)
    np.testing.assert_allclose(u, blk_dense(A, B, C, d, layout), rtol=0, atol=1e-13)


@pytest.mark.parametrize("M", [2, 3, 4, 5, 8])
def test_oracle_block_vs_reference_cpu_diag(M):
    z = np.load(os.path.join(GOLD, f"blk_ref_diag_M{M}.npz"))
    N, P = int(z["N"]), int(z["P"])
    u = oracle.btsv(z["A"], z["B"], z["C"], z["d"], N, P, M, layout=1)
    np.testing.assert_allclose(u, z["u"], rtol=0, atol=1e-14)


# ---------------------------------------------------------------------------
# adjoint


def test_oracle_adjoint_dot_identity_and_fd():
    dims, sd = [7, 3], 0
    a, b, c, d = dd_system(dims, sd, np.float64, seed=8)
    g = np.random.default_rng(1).uniform(-1, 1, a.shape)
    x = oracle_solve(a, b, c, d, None, dims, sd)
    ga, gb, gc, gd = oracle.adjoint(a, b, c, x, g, 2, sd, dims)
    # <A^-1 d, g> = <d, A^-T g>
    assert abs(np.dot(x, g) - np.dot(d, gd)) < 1e-13
    # central differences of L = <x(a,b,c,d), g>
    h = 1e-6
    for name, arr, grad in (("a", a, ga), ("b", b, gb), ("c", c, gc), ("d", d, gd)):
        for flat in [1, 5, 8, 12, 20]:
            r, col = divmod(flat, dims[0])
            p, m = arr.copy(), arr.copy()
            p[flat] += h
            m[flat] -= h
            args_p = [p if t is arr else t for t in (a, b, c, d)]
            args_m = [m if t is arr else t for t in (a, b, c, d)]
            lp = np.dot(oracle_solve(*args_p, None, dims, sd), g)
            lm = np.dot(oracle_solve(*args_m, None, dims, sd), g)
            fd = (lp - lm) / (2 * h)
            if (name == "a" and col == 0) or (name == "c" and col == dims[0] - 1):
                # a_0 and c_{N-1} are never read: zero gradient by definition
                assert grad[flat] == 0.0 and fd == 0.0, (name, flat)
                continue
            assert grad[flat] == pytest.approx(fd, rel=1e-6, abs=1e-9), (name, flat)


# ---------------------------------------------------------------------------
# variable-coefficient ADI (forward + reverse)


def vadi_forward_numpy(kap, u0, n, steps, lam):
    """Independent dense restatement of one forward pass on an n^3 grid."""
    u = u0.reshape(n, n, n).copy()  # [k, j, i]
    K = kap.reshape(n, n, n)
    inter = np.zeros((n, n, n), bool)
    inter[1:-1, 1:-1, 1:-1] = True
    for _ in range(steps):
        r = np.zeros_like(u)
        for ax in range(3):
            up = np.roll(u, -1, axis=ax)
            um = np.roll(u, 1, axis=ax)
            kp = 0.5 * (K + np.roll(K, -1, axis=ax))
            km = 0.5 * (K + np.roll(K, 1, axis=ax))
            r += kp * (up - u) - km * (u - um)
        r = np.where(inter, lam * r, 0.0)
        for ax in (2, 1, 0):  # x (i, last numpy axis), y, z
            rr = np.moveaxis(r, ax, -1).copy()
            KK = np.moveaxis(K, ax, -1)
            II = np.moveaxis(inter, ax, -1)
            for idx in np.ndindex(rr.shape[:-1]):
                kk = KK[idx]
                it = II[idx]
                A = np.eye(n)
                for t in range(n):
                    if it[t]:
                        am = -lam * 0.5 * (kk[t] + kk[t - 1])
                        cp = -lam * 0.5 * (kk[t] + kk[t + 1])
                        A[t, t - 1] = am
                        A[t, t + 1] = cp
                        A[t, t] = 1 - am - cp
                rr[idx] = np.linalg.solve(A, rr[idx])
            r = np.moveaxis(rr, -1, ax)
        u = u + r
    return u.ravel()


def test_oracle_vadi_forward_vs_numpy():
    n, steps, lam = 6, 3, 0.7
    rng = np.random.default_rng(3)
    kap = np.exp(0.3 * rng.uniform(-1, 1, n ** 3))
    u0 = rng.uniform(0, 1, n ** 3)
    got = oracle.vadi_forward(kap, u0, n, steps, lam)
    np.testing.assert_allclose(got, vadi_forward_numpy(kap, u0, n, steps, lam), atol=1e-13)


def test_oracle_vadi_grad_fd():
    n, steps, lam = 6, 3, 0.8
    rng = np.random.default_rng(5)
    kap = np.exp(0.3 * rng.uniform(-1, 1, n ** 3))
    u0 = rng.uniform(0, 1, n ** 3)
    tgt = rng.uniform(0, 1, n ** 3)
    loss, gk, gu = oracle.vadi_grad(kap, u0, tgt, n, steps, lam)

    def L(k, u):
        uT = oracle.vadi_forward(k, u, n, steps, lam)
        return 0.5 * np.sum((uT - tgt) ** 2)

    def dL(kp, up, km, um):
        # L(+) - L(-) as a difference of squares: the two forward results are
        # subtracted before the sum, so the central difference keeps full
        # precision at h = 1e-6
        a = oracle.vadi_forward(kp, up, n, steps, lam)
        b = oracle.vadi_forward(km, um, n, steps, lam)
        return 0.5 * np.sum((a - b) * (a + b - 2 * tgt))

    assert loss == pytest.approx(L(kap, u0), rel=1e-14)
    h = 1e-6
    pts = rng.choice(n ** 3, 12, replace=False)
    for p in pts:
        kp, km = kap.copy(), kap.copy()
        kp[p] += h
        km[p] -= h
        fd = dL(kp, u0, km, u0) / (2 * h)
        assert gk[p] == pytest.approx(fd, rel=1e-6, abs=1e-12), ("kappa", p)
        up, um = u0.copy(), u0.copy()
        up[p] += h
        um[p] -= h
        fd = dL(kap, up, kap, um) / (2 * h)
        assert gu[p] == pytest.approx(fd, rel=1e-6, abs=1e-12), ("u0", p)

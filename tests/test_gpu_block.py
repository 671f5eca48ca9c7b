"""GPU parity of the block-tridiagonal solver (blktrid_hip.h).

Inputs follow the benchmark distribution: off-diagonal blocks U(-1,1)/M,
diagonal blocks U(-1,1)/M + (row-sum |off-diagonals| + 1) I (block
diagonally dominant), rhs U(-1,1), A_0 = C_{N-1} = 0.  Each solve is
compared with the oracle's block Thomas (restating block/src/blktrid_cpu.cpp
with its product index fixed; pinned against dense LU in test_oracle.py):
normwise relative error <= 1e-12 (FP64), 1e-5 (FP32).  At the benchmark
size (P = 65536, N = 128, every M of the block configuration) the inputs
are generated on the device; every system's residual ||K u - d|| / ||d||
is checked there, and a sample of systems against the oracle.  The adjoint
is checked against dense transposed solves and the dot-product identity,
and at the configuration's size through the transposed residual, the dot
identity, its outer-product gradients and a sample against the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tridsolver_amd import block as blk
from tests.test_oracle import blk_system

pytestmark = pytest.mark.gpu


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


@pytest.mark.parametrize("M", list(range(1, 11)))
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_block_matches_oracle(M, dt):
    N, P = 17, 300
    A, B, C, d = blk_system(M, N, P, seed=M)
    exp = oracle.btsv(A, B, C, d, N, P, M, layout=0)
    tA, tB, tC, td = (dev(x.astype(dt)) for x in (A, B, C, d))
    u = blk.btsv(tA, tB, tC, td)
    err = oracle.rel_err(exp, u.cpu().numpy())
    assert err <= (1e-12 if dt == np.float64 else 1e-5), f"M={M}: {err:.2e}"


@pytest.mark.parametrize("M,L", [(2, 1), (3, 1), (4, 1), (4, 2), (6, 2), (6, 3), (8, 2), (8, 4),
                                 (10, 5), (5, 1)])
def test_block_every_lane_mapping(M, L):
    N, P = 9, 130
    A, B, C, d = blk_system(M, N, P, seed=10 + M)
    exp = oracle.btsv(A, B, C, d, N, P, M, layout=0)
    blk.force_lanes(L)
    try:
        u = blk.btsv(*(dev(x) for x in (A, B, C, d)))
    finally:
        blk.force_lanes(0)
    assert oracle.rel_err(exp, u.cpu().numpy()) <= 1e-12


@pytest.mark.parametrize("M,L", [(1, 1), (2, 1), (2, 2), (3, 1), (4, 1), (4, 2), (4, 4), (5, 1),
                                 (6, 2), (8, 2), (8, 4), (10, 2)])
@pytest.mark.parametrize("N", [1, 2, 3, 8, 17])
def test_block_two_ended_layouts(M, L, N):
    """The two-ended elimination (force_lanes(100 + L): L lanes per half;
    the upper half eliminates rows 0..N/2-1 downward, the lower half the rest
    upward, both meet at row N/2), forward and transposed, against the
    oracle; odd and even N, and N = 1, 2, 3 where a half is empty."""
    import torch
    P = 130
    A, B, C, d = blk_system(M, N, P, seed=30 + M + N)
    exp = oracle.btsv(A, B, C, d, N, P, M, layout=0)
    t = [dev(x) for x in (A, B, C, d)]
    blk.force_lanes(100 + L)
    try:
        u = blk.btsv(*t)
        g = torch.ones_like(u)
        lam = blk.btsv_adjoint(t[0], t[1], t[2], u, g)[3]
    finally:
        blk.force_lanes(0)
    err = oracle.rel_err(exp, u.cpu().numpy())
    assert err <= 1e-12, f"M={M} L={L} N={N}: {err:.2e}"
    # the transposed solve through the dot identity <A^{-1} d, 1> = <d, A^{-T} 1>
    lhs, rhs = float(u.sum()), float((t[3] * lam).sum())
    assert abs(lhs - rhs) <= 1e-10 * max(1.0, abs(lhs)), (M, L, N, lhs, rhs)


@pytest.mark.parametrize("M", [3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("N", [1, 2, 3, 8, 17])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_block_two_ended_tiled(M, N, dt):
    """The two-ended elimination on the tiled kernel (force_lanes(200): two
    groups of M lanes per problem, each with its own LDS tile), forward and
    transposed, against the oracle; P = 131 leaves a partial last wave."""
    import torch
    P = 131
    A, B, C, d = blk_system(M, N, P, seed=50 + M + N)
    exp = oracle.btsv(A, B, C, d, N, P, M, layout=0)
    t = [dev(x.astype(dt)) for x in (A, B, C, d)]
    blk.force_lanes(200)
    try:
        u = blk.btsv(*t)
        lam = None
        if dt == np.float64:
            lam = blk.btsv_adjoint(t[0], t[1], t[2], u, torch.ones_like(u))[3]
    finally:
        blk.force_lanes(0)
    err = oracle.rel_err(exp, u.cpu().numpy())
    assert err <= (1e-12 if dt == np.float64 else 1e-5), f"M={M} N={N}: {err:.2e}"
    if lam is not None:
        # the transposed solve through <A^{-1} d, 1> = <d, A^{-T} 1>
        lhs, rhs = float(u.sum()), float((t[3] * lam).sum())
        assert abs(lhs - rhs) <= 1e-10 * max(1.0, abs(lhs)), (M, N, lhs, rhs)


def test_block_force_lanes_rejects():
    import tridsolver_amd as ta
    for bad in (11, 100, 103, 105, -1, 199, 201):
        with pytest.raises(ta.TridError):
            blk.force_lanes(bad)
    M, N, P = 3, 4, 10
    A, B, C, d = blk_system(M, N, P, seed=1)
    blk.force_lanes(102)  # 2 lanes per half do not divide M = 3
    try:
        with pytest.raises(ta.TridError):
            blk.btsv(*(dev(x) for x in (A, B, C, d)))
    finally:
        blk.force_lanes(0)
    M = 2
    A, B, C, d = blk_system(M, N, P, seed=2)
    blk.force_lanes(200)  # the two-ended tiled kernel takes M = 3..8
    try:
        with pytest.raises(ta.TridError):
            blk.btsv(*(dev(x) for x in (A, B, C, d)))
    finally:
        blk.force_lanes(0)


@pytest.mark.parametrize("M", [2, 3, 4, 8])
def test_block_tuned_and_default_layouts(M):
    """With autotuning on, the first call of a shape times every layout (the
    tiled kernel, each lane count that divides M and the two-ended forms) and
    keeps the fastest; the untuned default comes from a table per precision,
    block size and batch band (two-ended for most FP64 entries at these
    sizes).  Every layout solves as the oracle does, forward and transposed
    (adjoint)."""
    import tridsolver_amd as ta
    import torch
    for N, P in ((6, 3000), (3, 65536)):
        A, B, C, d = blk_system(M, N, P, seed=20 + M)
        exp = oracle.btsv(A, B, C, d, N, P, M, layout=0)
        t = [dev(x) for x in (A, B, C, d)]
        for tune in (0, 1):
            ta.set_autotune(tune)
            try:
                u = blk.btsv(*t)
                u2 = blk.btsv(*t)  # the tuned layout from the cache
                g = torch.ones_like(u)
                lam = blk.btsv_adjoint(t[0], t[1], t[2], u, g)[3]
            finally:
                ta.set_autotune(0)
            err = oracle.rel_err(exp, u.cpu().numpy())
            assert err <= 1e-12, f"M={M} N={N} P={P} tune={tune}: {err:.2e}"
            assert torch.equal(u, u2)
            # the transposed solve: lambda = A^{-T} 1 satisfies A^T lambda = 1;
            # check sum_n,p <lambda, A d-direction> via the dot identity
            # <A^{-1} d, 1> = <d, A^{-T} 1>
            lhs = float(u.sum())
            rhs = float((t[3] * lam).sum())
            assert abs(lhs - rhs) <= 1e-10 * max(1.0, abs(lhs)), (M, N, P, tune, lhs, rhs)


def test_block_reference_entry_point_and_scratch():
    """dbtsv_gpu with caller-provided Cstar/dstar (the reference interface);
    Cstar/dstar hold C'_n and d'_n of the forward sweep."""
    import torch
    M, N, P = 3, 12, 77
    A, B, C, d = blk_system(M, N, P, seed=4)
    t = [dev(x) for x in (A, B, C, d)]
    Cs = torch.zeros((N, P, M, M), dtype=torch.float64, device="cuda")
    ds = torch.zeros((N, P, M), dtype=torch.float64, device="cuda")
    u = torch.zeros((N, P, M), dtype=torch.float64, device="cuda")
    blk.dbtsv_gpu(t[0], t[1], t[2], t[3], Cs, ds, u, N, P, M)
    torch.cuda.synchronize()
    assert oracle.rel_err(oracle.btsv(A, B, C, d, N, P, M), u.cpu().numpy()) <= 1e-12
    # d'_0 = B_0^{-1} d_0
    np.testing.assert_allclose(ds.cpu().numpy()[0], np.linalg.solve(B[0], d[0][..., None])[..., 0],
                               rtol=1e-12, atol=1e-14)


def test_block_mixed_single_launch():
    import torch
    groups, exps = [], []
    for M in (2, 3, 4, 5, 8):
        N, P = 11, 97 + M
        A, B, C, d = blk_system(M, N, P, seed=100 + M)
        exps.append(oracle.btsv(A, B, C, d, N, P, M))
        g = {k: dev(v) for k, v in zip("ABCd", (A, B, C, d))}
        g["Cstar"] = torch.empty((N, P, M, M), dtype=torch.float64, device="cuda")
        g["dstar"] = torch.empty((N, P, M), dtype=torch.float64, device="cuda")
        g["u"] = torch.empty((N, P, M), dtype=torch.float64, device="cuda")
        groups.append(g)
    blk.btsv_mixed(groups)
    for g, e in zip(groups, exps):
        assert oracle.rel_err(e, g["u"].cpu().numpy()) <= 1e-12


def dense_adjoint(A, B, C, u, g):
    N, P, M, _ = A.shape
    gA, gB, gC = (np.zeros_like(A) for _ in range(3))
    gd = np.zeros_like(u)
    for p in range(P):
        K = np.zeros((N * M, N * M))
        for n in range(N):
            K[n * M:(n + 1) * M, n * M:(n + 1) * M] = B[n, p]
            if n > 0:
                K[n * M:(n + 1) * M, (n - 1) * M:n * M] = A[n, p]
            if n < N - 1:
                K[n * M:(n + 1) * M, (n + 1) * M:(n + 2) * M] = C[n, p]
        lam = np.linalg.solve(K.T, g[:, p].ravel()).reshape(N, M)
        gd[:, p] = lam
        for n in range(N):
            gB[n, p] = -np.outer(lam[n], u[n, p])
            if n > 0:
                gA[n, p] = -np.outer(lam[n], u[n - 1, p])
            if n < N - 1:
                gC[n, p] = -np.outer(lam[n], u[n + 1, p])
    return gA, gB, gC, gd


@pytest.mark.parametrize("M", [1, 2, 3, 4, 5, 8])
def test_block_adjoint(M):
    N, P = 7, 9
    A, B, C, d = blk_system(M, N, P, seed=30 + M)
    u = oracle.btsv(A, B, C, d, N, P, M)
    g = np.random.default_rng(M).uniform(-1, 1, u.shape)
    got = [x.cpu().numpy() for x in blk.btsv_adjoint(*(dev(x) for x in (A, B, C, u, g)))]
    exp = dense_adjoint(A, B, C, u, g)
    for name, e, a in zip(("A", "B", "C", "d"), exp, got):
        assert oracle.rel_err(e, a) <= 1e-12, name
    # <A^{-1} d, g> = <d, A^{-T} g>
    assert abs(np.sum(u * g) - np.sum(d * got[3])) <= 1e-12 * np.sum(np.abs(u * g))


def bench_inputs(N, P, M, seed):
    """The block configuration's inputs, generated on the device (torch
    Philox, `seed`): off-diagonal blocks U(-1,1)/M, B = U(-1,1)/M +
    (row-sum |A| + |C| + 1) I, d ~ U(-1,1), A_0 = C_{N-1} = 0, layout
    (N, P, M, M)."""
    import torch
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)

    def U(*shape):
        return torch.rand(*shape, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
    A = U(N, P, M, M) / M
    Cm = U(N, P, M, M) / M
    A[0].zero_()
    Cm[N - 1].zero_()
    B = U(N, P, M, M) / M
    B += torch.diag_embed(A.abs().sum(-1) + Cm.abs().sum(-1) + 1.0)
    d = U(N, P, M)
    return A, B, Cm, d


def apply_blocks(A, B, Cm, u, transpose=False):
    """K u for the block-tridiagonal K of (A, B, C), or K^T u: elementwise
    products summed on the device (no library GEMM)."""
    if transpose:
        r = (B * u.unsqueeze(-1)).sum(-2)
        r[1:] += (Cm[:-1] * u[:-1].unsqueeze(-1)).sum(-2)
        r[:-1] += (A[1:] * u[1:].unsqueeze(-1)).sum(-2)
    else:
        r = (B * u.unsqueeze(-2)).sum(-1)
        r[1:] += (A[1:] * u[:-1].unsqueeze(-2)).sum(-1)
        r[:-1] += (Cm[:-1] * u[1:].unsqueeze(-2)).sum(-1)
    return r


SAMPLE = [0, 1, 2, 31, 32, 63, 64, 1000, 32767, 32768, 65534, 65535]


@pytest.mark.parametrize("M", [2, 3, 4, 5, 8])
def test_block_benchmark_size(M):
    """The block configuration at full size: N = 128 block rows, P = 65536
    systems, FP64 (reference solve block/src/blktrid_gpu.cu:57-179).  Every
    system: residual ||K u - d||_inf / ||d||_inf <= 1e-12 on the device; a
    sample of systems (both ends of the batch, wave and group boundaries):
    the oracle's block Thomas, normwise relative error <= 1e-12."""
    N, P = 128, 65536
    A, B, Cm, d = bench_inputs(N, P, M, seed=M)
    u = blk.btsv(A, B, Cm, d)
    rel = float((apply_blocks(A, B, Cm, u) - d).abs().max() / d.abs().max())
    print(f"block M={M} P={P} N={N}: residual {rel:.2e}")
    assert rel <= 1e-12
    h = [x[:, SAMPLE].contiguous().cpu().numpy() for x in (A, B, Cm, d, u)]
    exp = oracle.btsv(h[0], h[1], h[2], h[3], N, len(SAMPLE), M, layout=0)
    err = oracle.rel_err(exp, h[4])
    print(f"block M={M}: sampled systems vs oracle {err:.2e}")
    assert err <= 1e-12


@pytest.mark.parametrize("M", [2, 3, 4, 5, 8])
def test_block_adjoint_benchmark_size(M):
    """The block adjoint at the configuration's size (N = 128, P = 65536):
    lambda = grad_d solves K^T lambda = g (device residual <= 1e-12), the
    dot identity <K^{-1} d, g> = <d, K^{-T} g>, the outer-product gradients
    grad_B_n = -lambda_n u_n^T, grad_A_n = -lambda_n u_{n-1}^T, grad_C_n =
    -lambda_n u_{n+1}^T, and a sample of systems against the oracle's block
    Thomas on the transposed blocks (C_{n-1}^T, B_n^T, A_{n+1}^T)."""
    import torch
    N, P = 128, 65536
    A, B, Cm, d = bench_inputs(N, P, M, seed=50 + M)
    u = blk.btsv(A, B, Cm, d)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    g = torch.rand(N, P, M, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
    gA, gB, gC, lam = blk.btsv_adjoint(A, B, Cm, u, g)
    rel = float((apply_blocks(A, B, Cm, lam, transpose=True) - g).abs().max() / g.abs().max())
    print(f"block adjoint M={M}: transposed residual {rel:.2e}")
    assert rel <= 1e-12
    lhs, rhs = float((u * g).sum()), float((d * lam).sum())
    assert abs(lhs - rhs) <= 1e-12 * float((u * g).abs().sum())
    outer = -lam.unsqueeze(-1) * u.unsqueeze(-2)
    assert float((gB - outer).abs().max()) <= 1e-15 * float(outer.abs().max())
    assert float(gA[0].abs().max()) == 0.0 and float(gC[N - 1].abs().max()) == 0.0
    assert torch.equal(gA[1:], -lam[1:].unsqueeze(-1) * u[:-1].unsqueeze(-2))
    assert torch.equal(gC[:-1], -lam[:-1].unsqueeze(-1) * u[1:].unsqueeze(-2))
    h = [x[:, SAMPLE].contiguous().cpu().numpy() for x in (A, B, Cm, g, lam)]
    At = np.zeros_like(h[0])
    Ct = np.zeros_like(h[0])
    At[1:] = np.swapaxes(h[2][:-1], -1, -2)
    Ct[:-1] = np.swapaxes(h[0][1:], -1, -2)
    Bt = np.ascontiguousarray(np.swapaxes(h[1], -1, -2))
    exp = oracle.btsv(At, Bt, Ct, h[3], N, len(SAMPLE), M, layout=0)
    err = oracle.rel_err(exp, h[4])
    print(f"block adjoint M={M}: sampled lambda vs oracle {err:.2e}")
    assert err <= 1e-12


def test_block_mixed_benchmark_size():
    """All five block sizes of the configuration (M = 2, 3, 4, 5, 8; N = 128,
    P = 65536 each) in one mixed launch; every group agrees with its own
    separate solve within 1e-12 and its residual is <= 1e-12."""
    import torch
    N, P = 128, 65536
    groups = []
    for M in (2, 3, 4, 5, 8):
        A, B, Cm, d = bench_inputs(N, P, M, seed=90 + M)
        g = dict(A=A, B=B, C=Cm, d=d, Cstar=torch.empty_like(A), dstar=torch.empty_like(d),
                 u=torch.empty_like(d))
        groups.append(g)
    blk.btsv_mixed(groups)
    for g in groups:
        M = g["d"].shape[-1]
        ref = blk.btsv(g["A"], g["B"], g["C"], g["d"])
        diff = float((ref - g["u"]).abs().max() / ref.abs().max())
        rel = float((apply_blocks(g["A"], g["B"], g["C"], g["u"]) - g["d"]).abs().max()
                    / g["d"].abs().max())
        print(f"mixed M={M}: vs separate solve {diff:.2e}, residual {rel:.2e}")
        assert diff <= 1e-12 and rel <= 1e-12
        del ref

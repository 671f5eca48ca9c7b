"""GPU parity of the block-tridiagonal solver (blktrid_hip.h).

Inputs follow the benchmark distribution: off-diagonal blocks U(-1,1)/M,
diagonal blocks U(-1,1)/M + (row-sum |off-diagonals| + 1) I (block
diagonally dominant), rhs U(-1,1), A_0 = C_{N-1} = 0.  Each solve is
compared with the oracle's block Thomas (restating block/src/blktrid_cpu.cpp
with its product index fixed; pinned against dense LU in test_oracle.py):
normwise relative error <= 1e-12 (FP64), 1e-5 (FP32).  At the benchmark
size (P = 65536, N = 128) the residual ||A u - d|| / ||d|| is checked
instead.  The adjoint is checked against dense transposed solves and the
dot-product identity.
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


@pytest.mark.parametrize("M", [2, 3, 4, 5, 8])
def test_block_benchmark_size_residual(M):
    """N = 128 block rows, P = 65536 systems (fewer for M >= 5 to keep the
    host-side generation of the inputs within seconds; the benchmark runs
    the full batch): residual ||A u - d|| / ||d||."""
    import torch
    N, P = 128, {2: 65536, 3: 65536, 4: 65536, 5: 32768, 8: 8192}[M]
    A, B, C, d = blk_system(M, N, P, seed=M)
    tA, tB, tC, td = (dev(x) for x in (A, B, C, d))
    del A, B, C
    u = blk.btsv(tA, tB, tC, td)
    r = torch.matmul(tB, u.unsqueeze(-1)).squeeze(-1)
    r[1:] += torch.matmul(tA[1:], u[:-1].unsqueeze(-1)).squeeze(-1)
    r[:-1] += torch.matmul(tC[:-1], u[1:].unsqueeze(-1)).squeeze(-1)
    rel = float((r - td).abs().max() / td.abs().max())
    print(f"block M={M} P={P} N={N}: residual {rel:.2e}")
    assert rel <= 1e-12

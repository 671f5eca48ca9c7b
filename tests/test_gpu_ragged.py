"""GPU parity of ragged batches (trid_ragged.h) against the oracle.

Each system k of [offsets[k], offsets[k+1]) must equal the reference's 1-D
solve of that system (oracle_ragged restates trid_scalarD per system),
within 1e-12 (FP64) / 1e-5 (FP32) normwise, plus the reference tests'
elementwise tolerance.  Covers every length bin (1..8, wave-level 2..64
chunks, multi-wave up to 4096, 512 chunks of 9..16 rows up to 8192, and
beyond: per-system strided solves for the longest systems, the one-thread
kernel for the rest), empty systems,
Inc, repeated solves with one plan (the persistent kernel's self-reset
counter) and the benchmark configuration (200,000 systems, lengths
log-uniform in [8, 4096], seed 0).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tridsolver_amd.ragged import RaggedPlan
from bench import ragged_lengths
from tests.helpers import check_close

pytestmark = pytest.mark.gpu


def ragged_inputs(lengths, dtype, seed):
    rng = np.random.default_rng(seed)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    n = int(off[-1])
    a = rng.uniform(-1.0, -0.1, n)
    c = rng.uniform(-1.0, -0.1, n)
    b = np.abs(a) + np.abs(c) + rng.uniform(0.5, 1.5, n)
    d = rng.uniform(-1.0, 1.0, n)
    return off, [x.astype(dtype) for x in (a, b, c, d)]


def torch_dt(dt):
    import torch
    return torch.float64 if dt == np.float64 else torch.float32


def run(off, arrs, dtype, inc=False, u=None, reps=1):
    import torch
    t_off = torch.from_numpy(off).to("cuda")
    plan = RaggedPlan(t_off, torch_dt(dtype))
    ta, tb, tc, td = (torch.from_numpy(x).to("cuda") for x in arrs)
    d0 = td.clone()
    tu = torch.from_numpy(u).to("cuda") if u is not None else None
    u0 = tu.clone() if tu is not None else None
    outs = []
    for _ in range(reps):
        if inc:
            tu.copy_(u0)
        else:
            td.copy_(d0)
        plan.solve(ta, tb, tc, td, tu, inc)
        outs.append((tu if inc else td).cpu().numpy())
    return outs, plan


EDGE = [1, 2, 3, 7, 8, 9, 15, 16, 17, 31, 33, 63, 64, 65, 100, 255, 256, 257, 511, 512, 513,
        1000, 1023, 1024, 1025, 2047, 2048, 2049, 4095, 4096, 4097, 5000, 0, 1, 12]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("inc", [False, True])
def test_ragged_every_bin(dtype, inc):
    rng = np.random.default_rng(5)
    lengths = np.array(EDGE * 3 + list(rng.integers(1, 700, 500)), dtype=np.int64)
    rng.shuffle(lengths)
    off, arrs = ragged_inputs(lengths, dtype, seed=1)
    u = rng.uniform(-1, 1, int(off[-1])).astype(dtype) if inc else None
    (got,), plan = run(off, arrs, dtype, inc, u)
    exp_d = arrs[3].copy()
    exp_u = u.copy() if inc else None
    assert oracle.ragged(arrs[0], arrs[1], arrs[2], exp_d, exp_u, off, inc=inc) == 0
    check_close(exp_u if inc else exp_d, got, dtype, f"ragged [{plan.describe()}]")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ragged_every_chunk_length(dtype):
    """Every bin at both of its edges, cps*(R-1)+1 and cps*R rows: FP64
    chunks of 5..8 rows (1..512 chunks per system, up to 4096 rows) and 512
    chunks of 9..16 rows (the second launch, up to 8192 rows); FP32 chunks
    of 9..16 rows (1..512 chunks, up to 8192 rows)."""
    rmin, rmax = (5, 8) if dtype == np.float64 else (9, 16)
    lengths = []
    for cl in range(10):
        cps = 1 << cl
        for R in range(rmin, rmax + 1):
            lengths += [cps * (R - 1) + 1, cps * R]
    if dtype == np.float64:
        lengths += [L for R in range(9, 17) for L in (512 * (R - 1) + 1, 512 * R)]
    lengths = np.array(lengths * 2, dtype=np.int64)
    np.random.default_rng(9).shuffle(lengths)
    off, arrs = ragged_inputs(lengths, dtype, seed=3)
    (got,), plan = run(off, arrs, dtype)
    exp = arrs[3].copy()
    assert oracle.ragged(arrs[0], arrs[1], arrs[2], exp, None, off) == 0
    check_close(exp, got, dtype, f"ragged [{plan.describe()}]")
    for R in range(rmin, rmax + 1):
        assert f"R{R}/cps512:4" in plan.describe(), plan.describe()
    if dtype == np.float64:
        for R in range(9, 17):
            assert f"R{R}/cps512:4" in plan.describe(), plan.describe()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("inc", [False, True])
def test_ragged_long_systems_both_routes(dtype, inc):
    """Systems of 4097..8192 rows in the second persistent launch; past 8192
    rows the longest solved one by one as strided batches (split-line), the
    rest by the one-thread kernel."""
    rng = np.random.default_rng(11)
    lengths = np.array([150000, 40000, 20000] + list(rng.integers(8193, 8300, 300)) +
                       list(rng.integers(4097, 8193, 20)) + list(rng.integers(1, 3000, 300)),
                       dtype=np.int64)
    rng.shuffle(lengths)
    off, arrs = ragged_inputs(lengths, dtype, seed=4)
    u = rng.uniform(-1, 1, int(off[-1])).astype(dtype) if inc else None
    (got,), plan = run(off, arrs, dtype, inc, u)
    desc = plan.describe()
    assert "solo:3" in desc and "long(>8192):300" in desc and "cps512" in desc, desc
    exp_d = arrs[3].copy()
    exp_u = u.copy() if inc else None
    assert oracle.ragged(arrs[0], arrs[1], arrs[2], exp_d, exp_u, off, inc=inc) == 0
    check_close(exp_u if inc else exp_d, got, dtype, f"ragged [{desc}]")


@pytest.mark.parametrize("fam,red", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ragged_forced_choices(fam, red, dtype):
    """Every schedule the autotuner chooses from: both row families (5..8
    and 9..16 rows per chunk) and both reduced solves past 64 chunks (PCR
    through LDS, per-wave PCR with a sweep across waves), at the edges of
    every chunk count."""
    from tridsolver_amd.ragged import force_choice
    rng = np.random.default_rng(20 + fam * 2 + red)
    edges = [L for cl in range(10) for L in ((8 << cl) - 1, 8 << cl, (8 << cl) + 1)]
    lengths = np.array(EDGE + edges + list(rng.integers(1, 8193, 60)), dtype=np.int64)
    lengths = lengths[lengths <= 8192]
    rng.shuffle(lengths)
    off, arrs = ragged_inputs(lengths, dtype, seed=6)
    force_choice(fam, red)
    try:
        (got,), plan = run(off, arrs, dtype)
    finally:
        force_choice(-1, -1)
    desc = plan.describe()
    for c in range(10):
        assert f"c{c}={fam}{'w' if red else 'p'}" in desc, desc
    assert f"c10=1{'w' if red else 'p'}" in desc, desc
    exp = arrs[3].copy()
    assert oracle.ragged(arrs[0], arrs[1], arrs[2], exp, None, off) == 0
    check_close(exp, got, dtype, f"ragged [{desc}]")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ragged_tuned_plan(dtype):
    """A plan created with the autotuner on times every class's candidates
    (and the long-system cost constants) and still solves every system as
    the oracle does, with and without Inc."""
    import tridsolver_amd as ta
    rng = np.random.default_rng(31)
    lengths = np.array([30000, 9000] + list(rng.integers(8193, 8400, 5)) +
                       [int(x) for x in np.exp(rng.uniform(0, np.log(8192), 3000))],
                       dtype=np.int64)
    rng.shuffle(lengths)
    off, arrs = ragged_inputs(lengths, dtype, seed=8)
    u = rng.uniform(-1, 1, int(off[-1])).astype(dtype)
    ta.set_autotune(1)
    try:
        (got,), plan = run(off, arrs, dtype)
        (got_u,), _ = run(off, arrs, dtype, inc=True, u=u)
    finally:
        ta.set_autotune(0)
    desc = plan.describe()
    assert "tuned:" in desc and "long:solo=" in desc, desc
    print(f"tuned ragged plan {np.dtype(dtype).name}: {desc}")
    exp = arrs[3].copy()
    assert oracle.ragged(arrs[0], arrs[1], arrs[2], exp, None, off) == 0
    check_close(exp, got, dtype, f"ragged [{desc}]")
    exp_d = arrs[3].copy()
    exp_u = u.copy()
    assert oracle.ragged(arrs[0], arrs[1], arrs[2], exp_d, exp_u, off, inc=True) == 0
    check_close(exp_u, got_u, dtype, f"ragged inc [{desc}]")


def test_ragged_repeated_solves_one_plan():
    lengths = np.array([5, 300, 1200, 17, 4000, 64, 9000] * 40, dtype=np.int64)
    off, arrs = ragged_inputs(lengths, np.float64, seed=2)
    outs, _ = run(off, arrs, np.float64, reps=4)
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


def test_ragged_empty_and_all_empty():
    import torch
    off = torch.zeros(5, dtype=torch.int64, device="cuda")
    plan = RaggedPlan(off, torch.float64)
    x = torch.zeros(1, dtype=torch.float64, device="cuda")
    plan.solve(x, x, x, x)


def test_ragged_rejects_bad_offsets():
    import torch
    from tridsolver_amd import TridError
    off = torch.tensor([0, 5, 3], dtype=torch.int64, device="cuda")
    with pytest.raises(TridError):
        RaggedPlan(off, torch.float64)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ragged_benchmark_configuration(dtype):
    lengths = ragged_lengths()
    off, arrs = ragged_inputs(lengths, dtype, seed=0)
    (got,), plan = run(off, arrs, dtype)
    exp = arrs[3].copy()
    assert oracle.ragged(arrs[0], arrs[1], arrs[2], exp, None, off) == 0
    r = check_close(exp, got, dtype, f"ragged benchmark [{plan.describe()}]")
    print(f"ragged 200k systems, {int(off[-1])} unknowns, {np.dtype(dtype).name}: rel err {r:.2e} "
          f"[{plan.describe()}]")

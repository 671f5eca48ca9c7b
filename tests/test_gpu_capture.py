"""Calls made inside a caller's stream capture (torch.cuda.graph).

With autotuning on, a shape first seen under capture must not be tuned
(tuning synchronises and times launches, which invalidates a capture): the
call takes the untuned schedule, the capture succeeds and every replay
matches the oracle.  A split-line solve whose workspace does not exist yet
on the capturing stream returns TRID_STATUS_ALLOC_FAILED instead of
allocating; after one eager call on that stream it captures, and the
captured buffer outlives a later regrowth of the stream's workspace.  An
ADI plan first used under capture stays untuned.

Each graph is replayed several times with eager work and host
synchronisation in between, and every replay is checked against the
oracle: normwise relative error 1e-12 (FP64).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import tridsolver_amd as ta
from tridsolver_amd import api
from tridsolver_amd.adi import AdiPlan
from tests.helpers import check_close, dd_system, oracle_solve

pytestmark = pytest.mark.gpu


@pytest.fixture
def autotune_on():
    ta.set_autotune(1)
    try:
        yield
    finally:
        ta.set_autotune(0)
        ta.set_stream(None)


def _dev(*xs):
    import torch
    return [torch.from_numpy(x).to("cuda") for x in xs]


def _capture(fn, stream):
    """Captures fn() on `stream` (the library's stream during the capture);
    a small torch op keeps the graph non-empty.  Returns (graph, fn's value)."""
    import torch
    g = torch.cuda.CUDAGraph()
    dummy = torch.zeros(16, device="cuda")
    torch.cuda.synchronize()
    out = None
    try:
        with torch.cuda.graph(g, stream=stream):
            ta.set_stream(stream)
            out = fn()
            dummy.add_(1.0)
    finally:
        ta.set_stream(None)
    return g, out


def _replays(g, reset, check, n=3):
    import torch
    for _ in range(n):
        reset()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        check()


@pytest.mark.parametrize("sd", [0, 1, 2])
def test_strided_shape_first_seen_under_capture(autotune_on, sd):
    import torch
    dims = [72 + sd, 88, 40]  # shapes no other test tunes
    a, b, c, d = dd_system(dims, sd, np.float64, seed=40 + sd)
    exp = oracle_solve(a, b, c, d, None, dims, sd)
    ta_, tb, tc, td = _dev(a, b, c, d)
    d0 = td.clone()
    s = torch.cuda.Stream()
    g, st = _capture(lambda: api.tridDmtsvStridedBatch(ta_, tb, tc, td, None, 3, sd, dims, dims,
                                                        None, 0), s)
    assert st == 0, ta.last_error()
    _replays(g, lambda: td.copy_(d0),
             lambda: check_close(exp, td.cpu().numpy(), np.float64, f"sd{sd} replay"))
    # the first eager call tunes the shape; results are unchanged
    td.copy_(d0)
    assert api.tridDmtsvStridedBatch(ta_, tb, tc, td, None, 3, sd, dims, dims, None, 1) == 0
    check_close(exp, td.cpu().numpy(), np.float64, f"sd{sd} eager after capture")
    _replays(g, lambda: td.copy_(d0),
             lambda: check_close(exp, td.cpu().numpy(), np.float64, f"sd{sd} replay 2"), n=1)


def test_split_line_workspace_under_capture(autotune_on):
    """Strided systems of 2048+ rows take the split-line solve, whose
    workspace is per stream."""
    import torch
    dims = [64, 2048, 2]
    a, b, c, d = dd_system(dims, 1, np.float64, seed=7)
    exp = oracle_solve(a, b, c, d, None, dims, 1)
    ta_, tb, tc, td = _dev(a, b, c, d)
    d0 = td.clone()
    s = torch.cuda.Stream()  # no workspace on this stream yet

    def call(sync=0):
        return api.tridDmtsvStridedBatch(ta_, tb, tc, td, None, 3, 1, dims, dims, None, sync)
    _, st = _capture(call, s)
    assert st == 2 and "capture" in ta.last_error(), (st, ta.last_error())
    torch.cuda.synchronize()
    # one eager call on the stream allocates its workspace
    ta.set_stream(s)
    try:
        assert call(1) == 0, ta.last_error()
    finally:
        ta.set_stream(None)
    check_close(exp, td.cpu().numpy(), np.float64, "eager split")
    g, st = _capture(call, s)
    assert st == 0, ta.last_error()
    _replays(g, lambda: td.copy_(d0),
             lambda: check_close(exp, td.cpu().numpy(), np.float64, "split replay"))
    # a larger split-line solve on the same stream grows its workspace; the
    # buffer the graph holds is kept
    big = [256, 4096, 4]
    a2, b2, c2, d2 = dd_system(big, 1, np.float64, seed=8)
    exp2 = oracle_solve(a2, b2, c2, d2, None, big, 1)
    t2 = _dev(a2, b2, c2, d2)
    ta.set_stream(s)
    try:
        assert api.tridDmtsvStridedBatch(*t2, None, 3, 1, big, big, None, 1) == 0
    finally:
        ta.set_stream(None)
    check_close(exp2, t2[3].cpu().numpy(), np.float64, "larger split")
    _replays(g, lambda: td.copy_(d0),
             lambda: check_close(exp, td.cpu().numpy(), np.float64, "split replay after growth"))


def test_adi_plan_first_used_under_capture(autotune_on):
    import torch
    from tests.test_gpu_adi import rand_init
    n = 48
    u0 = rand_init(n, n, n, np.float64, seed=3)
    exp = u0.ravel().copy()
    assert oracle.adi(exp, n, n, n, 2, 1.0, 1) == 0
    plan = AdiPlan(n, n, n, torch.float64, 1.0, 1)
    before = plan.describe()
    u = torch.from_numpy(u0).to("cuda")
    du = torch.empty_like(u)
    u_init = u.clone()
    s = torch.cuda.Stream()
    g, _ = _capture(lambda: plan.step(u, du, 2, sync=0), s)
    _replays(g, lambda: u.copy_(u_init),
             lambda: check_close(exp, u.cpu().numpy().ravel(), np.float64, "ADI replay"))
    # later eager use does not retune the plan (the graph holds its tables)
    u.copy_(u_init)
    plan.step(u, du, 2)
    check_close(exp, u.cpu().numpy().ravel(), np.float64, "ADI eager after capture")
    assert plan.describe() == before
    _replays(g, lambda: u.copy_(u_init),
             lambda: check_close(exp, u.cpu().numpy().ravel(), np.float64, "ADI replay 2"), n=1)
    plan.close()

"""GPU tests of tridDmtsvStridedBatchNoAlloc, the no-allocation form of
tridDmtsvStridedBatch (include/trid_hip_ext.h).

It runs the untuned schedule of tridDmtsvStridedBatch with the split-line
scratch taken from a static device array, so with autotuning off both entry
points launch the same kernels: results must agree bit for bit.  Two calls in
a row (the second on a smaller input, with the scratch still holding the
first call's values) are each compared with tridDmtsvStridedBatch and with
the oracle (normwise relative error <= 1e-12, the FP64 bar of the scalar
tests).  A solve whose scratch exceeds the array's size is rejected before
anything is launched.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

import oracle
import tridsolver_amd as ta
from tests.helpers import dd_system

pytestmark = pytest.mark.gpu

LIMIT = 64 << 20  # TRID_NOALLOC_SCRATCH_BYTES


def _iv(v):
    return (C.c_int * len(v))(*v)


def _solve(name, dev, dims, sd):
    fn = getattr(ta.lib(), name)
    return fn(*(t.data_ptr() for t in dev), None, len(dims), sd, _iv(dims), _iv(dims), None, 1)


@pytest.fixture(autouse=True)
def _untuned():
    ta.set_stream(None)
    ta.set_autotune(0)
    ta.force_config(0, 0, 0, 0)
    yield
    ta.force_config(0, 0, 0, 0)


def _case(dims, sd, seed, check_oracle=True):
    a, b, c, d = dd_system(dims, sd, np.float64, seed=seed)
    print(f"no-alloc case dims={dims} sd={sd}", flush=True)
    dev = [torch.from_numpy(x).to("cuda") for x in (a, b, c, d)]
    dev2 = [t.clone() for t in dev]
    st = _solve("tridDmtsvStridedBatchNoAlloc", dev, dims, sd)
    assert st == 0, ta.last_error()
    cfg = ta.last_config()
    st = _solve("tridDmtsvStridedBatch", dev2, dims, sd)
    assert st == 0, ta.last_error()
    assert ta.last_config() == cfg
    got = dev[3].cpu().numpy()
    assert np.array_equal(got, dev2[3].cpu().numpy()), f"dims={dims}: differs from the allocating form"
    if check_oracle:
        ref = d.copy()
        assert oracle.mtsv(a, b, c, ref, None, len(dims), sd, dims) == 0
        err = oracle.rel_err(ref, got)
        print(f"  [{cfg}] rel err vs oracle {err:.2e}", flush=True)
        assert err <= 1e-12
    return cfg


def test_two_calls_second_smaller():
    # 2^25 unknowns in 64 systems of 2^19 rows along solvedim 1: a split-line
    # solve whose reduced systems take about 33 MiB of the scratch
    cfg = _case([64, 1 << 19], 1, seed=1)
    assert cfg.startswith("split<"), cfg
    # then a smaller split-line solve along solvedim 0 (systems of 20000 rows)
    cfg = _case([20000, 3], 0, seed=2)
    assert cfg.startswith("split<"), cfg
    # and a solve that needs no scratch at all
    _case([64, 64, 64], 2, seed=3)


def test_strided_long_lines():
    # solvedim >= 1 from 2048 rows takes the split-line path too
    _case([300, 4096], 1, seed=4)
    _case([5, 3, 2500], 2, seed=5)


def test_scratch_limit_rejected():
    # the forced one-thread-per-system kernel needs 2 * N * nsys doubles:
    # 128 MiB for 1024 systems of 8192 rows, over the 64 MiB array
    dims = [1024, 8192]
    assert 2 * 8 * dims[0] * dims[1] > LIMIT
    d = torch.rand(dims[0] * dims[1], dtype=torch.float64, device="cuda")
    b = torch.full_like(d, 4.0)
    a = torch.full_like(d, -1.0)
    c = torch.full_like(d, -1.0)
    d0 = d.clone()
    ta.force_config(2, 0, 0, 0)
    st = _solve("tridDmtsvStridedBatchNoAlloc", [a, b, c, d], dims, 1)
    assert st == 3, st  # TRID_STATUS_INVALID_VALUE
    assert "no-allocation" in ta.last_error()
    assert torch.equal(d, d0)

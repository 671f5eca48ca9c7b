"""GPU parity of the split-line solves (systems longer than one work-group)
against the oracle.

Systems longer than 8192 rows are cut into segments that are reduced, coupled
through a reduced system of two rows per segment, and solved as one batch
(recursively when the reduced system is itself long).  Forcing the split
configuration (tridForceConfig kind 4) runs the same path on short systems
with tiny segments, which covers tails, single segments and recursion at
sizes the oracle solves in milliseconds.  Bars as in test_gpu_scalar:
normwise relative error <= 1e-12 (FP64) / 1e-5 (FP32) plus the reference
tests' elementwise tolerance, untouched padding, and d unchanged by Inc.
"""
from __future__ import annotations

import numpy as np
import pytest

import tridsolver_amd as ta
from tests.test_gpu_scalar import run_case

pytestmark = pytest.mark.gpu


def forced(kind, rows, cps, spb, cases):
    ta.force_config(kind, rows, cps, spb)
    try:
        for args in cases:
            run_case(*args)
            assert ta.last_config().startswith("split<"), ta.last_config()
    finally:
        ta.force_config(0, 0, 0, 0)


LONG = [([20000, 3], 0), ([100000], 0), ([3, 40000], 1), ([2, 3, 30000], 2),
        ([64 * 256, 4], 0), ([12345, 2], 0), ([70, 9000], 1)]


@pytest.mark.parametrize("dims,sd", LONG)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("inc", [False, True])
def test_long_systems_split(dims, sd, dtype, inc):
    run_case(dims, sd, dtype, inc, seed=sum(dims) + sd)
    assert ta.last_config().startswith("split<"), ta.last_config()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_long_padded(dtype):
    for sd, dims, pads in ((0, [20001, 3], [20008, 4]), (1, [5, 17000], [7, 17003])):
        for inc in (False, True):
            run_case(dims, sd, dtype, inc, pads=pads, seed=11)


# forced segment shapes (rows, cps, spb): segments of 32, 16, 8 and 128 rows
FORCED = [(8, 4, 0), (4, 4, 16), (4, 2, 0), (16, 8, 8)]
SHORT = [([1000, 7], 0), ([777, 5], 0), ([31, 3], 0), ([32, 2], 0), ([33, 2], 0),
         ([7, 1000], 1), ([5, 777], 1), ([5, 3, 300], 2), ([4, 6, 5, 97], 3)]


@pytest.mark.parametrize("cfg", FORCED)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_forced_split_short(cfg, dtype):
    cases = [(dims, sd, dtype, inc, None, 7) for dims, sd in SHORT for inc in (False, True)]
    forced(4, *cfg, cases)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_forced_split_recursive(dtype):
    """Segments of 8 rows: 40000 rows give a reduced system of 10000 rows,
    itself split (segments of the default length)."""
    cases = [([40000, 2], 0, dtype, False, None, 3), ([2, 40000], 1, dtype, True, None, 3)]
    forced(4, 4, 2, 0, cases)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_one_very_long_system(dtype):
    """One system of 2^22 rows: 16384 segments, a reduced system of 32768
    rows that is split again."""
    run_case([1 << 22], 0, dtype, False, seed=5, zero_ends=True)
    run_case([1 << 22], 0, dtype, True, seed=6)

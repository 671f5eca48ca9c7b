"""The command-line tools of tridsolver_amd/apps, CPU side.

compare_{d,s} (SumOfDiff of two raw ADI outputs) are checked
  * against the reference's own comparison tool (apps/adi/src/compare.c,
    built unmodified into oracle/_ref/ where the reference tree is present):
    identical standard output on the same files, byte for byte;
  * against the sequential sum of the differences in the file's precision.
adi_hip_{d,s} and blktrid_hip_{d,s} need a GPU for a run
(tests/test_gpu_apps.py); here only their option handling is checked.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tridsolver_amd", "apps", "bin")
REF = os.path.join(ROOT, "oracle", "_ref")
GOLD = os.path.join(ROOT, "tests", "golden")
PRECS = [("d", np.float64), ("s", np.float32)]


def run(exe, *args):
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=60)


def test_tools_are_built():
    for name in ("adi_hip_d", "adi_hip_s", "compare_d", "compare_s", "blktrid_hip_d",
                 "blktrid_hip_s"):
        assert os.access(os.path.join(BIN, name), os.X_OK), name


@pytest.mark.parametrize("tag", ["d", "s"])
def test_blktrid_tool_options(tag):
    """the reference block app's usage text for -help, and bad sizes refused
    before any device work"""
    exe = os.path.join(BIN, f"blktrid_hip_{tag}")
    out = run(exe, "-help")
    assert out.returncode == 0 and "-blkdim=BLOCKDIM" in out.stdout, out.stdout
    for bad in (["-n=0"], ["-blkdim=11"], ["-p=-3"], ["-iter=0"]):
        out = run(exe, *bad)
        assert out.returncode == 1 and "invalid arguments" in out.stderr, (bad, out.stderr)


def files(tmp_path, dt, nan_at=None, seed=0, n=16 * 12 * 10):
    rng = np.random.default_rng(seed)
    u1 = rng.uniform(0, 1, n).astype(dt)
    u2 = (u1 + rng.normal(0, 1e-3, n)).astype(dt)
    if nan_at is not None:
        u2[nan_at] = np.nan
    p1, p2 = str(tmp_path / "f1.dat"), str(tmp_path / "f2.dat")
    u1.tofile(p1)
    u2.tofile(p2)
    return p1, p2, u1, u2


@pytest.mark.parametrize("tag,dt", PRECS)
def test_compare_sum_of_diff(tmp_path, tag, dt):
    p1, p2, u1, u2 = files(tmp_path, dt)
    out = run(os.path.join(BIN, f"compare_{tag}"), p1, p2, "-nx", "16", "-ny", "12", "-nz", "10")
    assert out.returncode == 0
    s = np.cumsum(u2 - u1, dtype=dt)[-1]  # sequential, in the file's precision
    exp = "SumOfDiff = %e; Normalized SumOfDiff = %e" % (s, s / dt(16 * 12 * 10))
    assert exp in out.stdout, out.stdout


@pytest.mark.parametrize("tag,dt", PRECS)
def test_compare_reports_nan_points(tmp_path, tag, dt):
    p1, p2, _, _ = files(tmp_path, dt, nan_at=5 + 16 * 3 + 16 * 12 * 2)
    out = run(os.path.join(BIN, f"compare_{tag}"), p1, p2, "-nx=16", "-ny=12", "-nz=10")
    assert "\n5 3 2\n" in out.stdout
    assert "SumOfDiff = nan" in out.stdout


@pytest.mark.parametrize("tag,dt", PRECS)
def test_compare_matches_reference_tool(tmp_path, tag, dt):
    ref = os.path.join(REF, f"compare_{tag}")
    if not os.access(ref, os.X_OK):
        pytest.skip("reference compare tool not built (no reference tree)")
    cases = [files(tmp_path, dt, seed=1)[:2], files(tmp_path, dt, nan_at=77, seed=2)[:2]]
    gold = os.path.join(GOLD, f"adi_ref_f{64 if tag == 'd' else 32}_16x12x10_it5.bin")
    cases.append((gold, gold))
    for p1, p2 in cases:
        args = [p1, p2, "-nx", "16", "-ny", "12", "-nz", "10"]
        ours, theirs = run(os.path.join(BIN, f"compare_{tag}"), *args), run(ref, *args)
        assert ours.stdout == theirs.stdout


def test_compare_missing_file(tmp_path):
    out = run(os.path.join(BIN, "compare_d"), str(tmp_path / "nope1"), str(tmp_path / "nope2"))
    assert out.returncode == 2


def test_adi_app_options():
    exe = os.path.join(BIN, "adi_hip_d")
    out = run(exe, "-help")
    assert out.returncode == 0 and "Usage" in out.stdout
    assert run(exe, "-bogus", "1").returncode == 1
    assert run(exe, "-nx").returncode == 1

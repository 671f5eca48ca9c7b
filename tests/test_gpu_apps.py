"""The ADI application adi_hip_{d,s} on the GPU against the reference's gold
ADI (apps/adi/src/adi_orig.c, built unmodified; tests/golden/adi_ref_*):

  * the raw output <argv[0]>.dat matches the gold output (1e-12 FP64 /
    1e-5 FP32 normwise relative), read back through compare_{d,s};
  * the corner tables it prints equal the gold run's printed tables
    (apps/adi/src/print_array.c format; FP64 text identical, FP32 values
    within 1e-4 after their 5-digit rounding);
  * the "end h_u sum" checksum line matches the sum of the gold output;
both for the staged step (-prof 1, one launch per stage) and the fused
step (-prof 0).
"""
from __future__ import annotations

import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tridsolver_amd", "apps", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")

pytestmark = pytest.mark.gpu


def gold_cases():
    with open(os.path.join(GOLD, "golden_meta.json")) as f:
        meta = json.load(f)
    return sorted((k, v) for k, v in meta.items() if k.startswith("adi_ref_"))


def corners_block(text):
    lines = text.splitlines(keepends=True)
    first = next(i for i, ln in enumerate(lines) if ln.startswith("k = "))
    last = next(i for i, ln in enumerate(lines) if ln.startswith("Done."))
    return "".join(lines[first:last])


@pytest.mark.parametrize("prof", [1, 0])
@pytest.mark.parametrize("name,meta", gold_cases(), ids=[c[0] for c in gold_cases()])
def test_adi_app_vs_reference_gold(tmp_path, name, meta, prof):
    tag = "d" if meta["dtype"] == "f64" else "s"
    dt = np.float64 if tag == "d" else np.float32
    nx, ny, nz, it = meta["nx"], meta["ny"], meta["nz"], meta["iter"]
    argv0 = str(tmp_path / f"adi_hip_{tag}")
    print(f"adi_hip_{tag} {nx}x{ny}x{nz} iter {it} prof {prof}", flush=True)
    out = subprocess.run([argv0, "-nx", str(nx), "-ny", str(ny), "-nz", str(nz),
                          "-iter", str(it), "-prof", str(prof)],
                         executable=os.path.join(BIN, f"adi_hip_{tag}"),
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    gold_path = os.path.join(GOLD, name)
    gold = np.fromfile(gold_path, dtype=dt)
    got = np.fromfile(argv0 + ".dat", dtype=dt)
    assert got.size == nx * ny * nz
    tol = 1e-12 if tag == "d" else 1e-5
    err = np.abs(got.astype(np.float64) - gold).max() / np.abs(gold).max()
    assert err <= tol, err

    # SumOfDiff through the compare tool
    cmp = subprocess.run([os.path.join(BIN, f"compare_{tag}"), gold_path, argv0 + ".dat",
                          "-nx", str(nx), "-ny", str(ny), "-nz", str(nz)],
                         capture_output=True, text=True, timeout=60)
    m = re.search(r"Normalized SumOfDiff = (\S+)", cmp.stdout)
    assert m and abs(float(m.group(1))) <= tol, cmp.stdout

    # checksum line
    m = re.search(r"end h_u sum = (\S+)", out.stdout)
    ref_sum = float(gold.astype(np.float64).sum())
    assert m and abs(float(m.group(1)) - ref_sum) <= 1e-5 * abs(ref_sum), out.stdout

    # corner tables
    with open(gold_path[:-4] + "_corners.txt") as f:
        want = f.read()
    have = corners_block(out.stdout)
    if tag == "d":
        assert have == want
    else:
        w = np.array(want.split(), dtype=object)
        h = np.array(have.split(), dtype=object)
        assert w.shape == h.shape
        num = np.array([not (s in ("k", "=") ) for s in w])
        wn, hn = w[num].astype(np.float64), h[num].astype(np.float64)
        assert np.allclose(hn, wn, rtol=1e-4, atol=1e-4)


def splitmix64(z):
    """S. Vigna's SplitMix64 on a numpy uint64 array"""
    z = z + np.uint64(0x9e3779b97f4a7c15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
    return z ^ (z >> np.uint64(31))


def blktrid_system(n, p, m):
    """the tool's generated system in the GPU layout (FP64)"""
    A = np.zeros((n, p, m, m))
    C = np.zeros((n, p, m, m))
    B = np.zeros((n, p, m, m))
    idx = np.arange(m)
    A[:, :, idx, idx] = 0.01
    C[:, :, idx, idx] = 0.01
    B[:, :, idx[1:], idx[:-1]] = 0.01
    B[:, :, idx[:-1], idx[1:]] = 0.01
    k = (np.arange(n * p, dtype=np.uint64)[:, None] * np.uint64(m) + idx.astype(np.uint64)[None, :])
    with np.errstate(over="ignore"):
        r = splitmix64(np.uint64(10) + k)
    B[:, :, idx, idx] = (4.0 + 0.01 * ((r >> np.uint64(11)).astype(np.float64) * 2.0 ** -53)).reshape(n, p, m)
    d = np.ones((n, p, m))
    return A.ravel(), B.ravel(), C.ravel(), d.ravel()


@pytest.mark.parametrize("tag", ["d", "s"])
@pytest.mark.parametrize("m", [2, 3, 5, 8])
def test_blktrid_tool(tmp_path, tag, m):
    """blktrid_hip_{d,s}: the reference block app's report, a relative
    residual at the precision's level, and results_binary (the solution in
    the GPU layout) against the oracle's block Thomas (FP64) on the same
    system, rebuilt here from the tool's generator (1e-12 FP64, 1e-5 FP32
    normwise relative)."""
    import oracle
    n, p = 12, 300
    exe = os.path.join(BIN, f"blktrid_hip_{tag}")
    print(f"blktrid_hip_{tag} n={n} p={p} blkdim={m}", flush=True)
    out = subprocess.run([exe, f"-n={n}", f"-p={p}", f"-blkdim={m}", "-iter=2"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert f"N:\t{n}\nP:\t{p}\nblkdim:\t{m}" in out.stdout
    assert re.search(r"GPU \n  duration \[ms\]: [0-9.]+\n  GFLOPS: +[0-9.]+\n  GB/sec: +[0-9.]+",
                     out.stdout), out.stdout
    res = float(re.search(r"Relative residual \|\|A u - d\|\| / \|\|d\|\|: (\S+)",
                          out.stdout).group(1))
    assert res <= (1e-13 if tag == "d" else 1e-6), res
    dt = np.float64 if tag == "d" else np.float32
    u = np.fromfile(tmp_path / "results_binary", dtype=dt)
    assert u.size == n * p * m
    A, B, C, d = blktrid_system(n, p, m)
    exp = oracle.btsv(A, B, C, d, n, p, m, layout=0)
    err = oracle.rel_err(exp, u.astype(np.float64))
    assert err <= (1e-12 if tag == "d" else 1e-5), err

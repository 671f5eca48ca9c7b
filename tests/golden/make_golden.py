"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Run in a container where the reference tree is mounted (it builds the
reference's buildable CPU sources into oracle/_ref via oracle/Makefile):

    python tests/golden/make_golden.py

Fixtures written (inputs and expected outputs only -- data, not code):
  adi_ref_{f64,f32}_<nx>x<ny>x<nz>_it<n>.bin
      final u of the reference's gold ADI (apps/adi/src/adi_orig.c, built
      unmodified from source), raw little-endian floats, x fastest; its own
      initial condition (1 on the boundary, 0 inside) and lambda = 1.
  adi_ref_{f64,f32}_<nx>x<ny>x<nz>_it<n>_corners.txt
      the corner tables the same run prints (apps/adi/src/print_array.c),
      i.e. its standard output between the timing line and "Done.".
  blk_ref_diag_M<m>.npz
      inputs and output of the reference's CPU block solver
      (block/src/blktrid_cpu.cpp dbtsv_cpu, built unmodified) on block
      systems whose A, B, C blocks are diagonal -- the inputs on which its
      gold_MMmultiply index error (blktrid_cpu.cpp:310-317) has no effect.
      Layout is the reference CPU layout, system-major ((p*N + n)*M*M).
"""
from __future__ import annotations

import ctypes as C
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")

ADI_CASES = [(16, 12, 10, 5), (33, 7, 9, 3)]
BLK_CASES = [(2, 9, 5), (3, 7, 4), (4, 6, 3), (5, 5, 3), (8, 4, 2)]  # (M, N, P)


def build_ref() -> None:
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])


def adi_fixtures(meta: dict) -> None:
    for exe, tag in (("adi_orig_d", "f64"), ("adi_orig_s", "f32")):
        for nx, ny, nz, it in ADI_CASES:
            with tempfile.TemporaryDirectory() as td:
                dst = os.path.join(td, exe)
                shutil.copy(os.path.join(REF, exe), dst)
                out = subprocess.run([dst, "-nx", str(nx), "-ny", str(ny), "-nz", str(nz),
                                      "-iter", str(it), "-prof", "0"],
                                     stdout=subprocess.PIPE, check=True, text=True).stdout
                name = f"adi_ref_{tag}_{nx}x{ny}x{nz}_it{it}.bin"
                shutil.copy(dst + ".dat", os.path.join(HERE, name))
                lines = out.splitlines(keepends=True)
                first = next(i for i, ln in enumerate(lines) if ln.startswith("k = "))
                last = next(i for i, ln in enumerate(lines) if ln.startswith("Done."))
                with open(os.path.join(HERE, name[:-4] + "_corners.txt"), "w") as f:
                    f.write("".join(lines[first:last]))
                meta[name] = {"source": "apps/adi/src/adi_orig.c", "nx": nx, "ny": ny,
                              "nz": nz, "iter": it, "dtype": tag, "lambda": 1.0}


def blk_fixtures(meta: dict) -> None:
    lib = C.CDLL(os.path.join(REF, "libblktrid_ref.so"))
    fn = getattr(lib, "_Z9dbtsv_cpuPdS_S_S_S_S_S_iii")  # dbtsv_cpu(double*...,int,int,int)
    fn.argtypes = [C.c_void_p] * 7 + [C.c_int] * 3
    fn.restype = None
    rng = np.random.default_rng(20240517)
    for M, N, P in BLK_CASES:
        shape = (P, N, M, M)
        A = np.zeros(shape)
        B = np.zeros(shape)
        Cc = np.zeros(shape)
        idx = np.arange(M)
        A[..., idx, idx] = rng.uniform(-1, 1, (P, N, M)) / M
        Cc[..., idx, idx] = rng.uniform(-1, 1, (P, N, M)) / M
        B[..., idx, idx] = (np.abs(A[..., idx, idx]) + np.abs(Cc[..., idx, idx]) + 1.0
                            + rng.uniform(0, 0.5, (P, N, M)))
        A[:, 0] = 0.0
        Cc[:, -1] = 0.0
        d = rng.uniform(-1, 1, (P, N, M))
        Cs = np.zeros(shape)
        ds = np.zeros((P, N, M))
        u = np.zeros((P, N, M))
        arrs = [np.ascontiguousarray(x) for x in (A, B, Cc, Cs, ds, d, u)]
        fn(*[x.ctypes.data for x in arrs], N, P, M)
        name = f"blk_ref_diag_M{M}.npz"
        np.savez_compressed(os.path.join(HERE, name), A=arrs[0], B=arrs[1], C=arrs[2],
                            d=arrs[5], u=arrs[6], N=N, P=P, M=M)
        meta[name] = {"source": "block/src/blktrid_cpu.cpp (dbtsv_cpu)", "M": M, "N": N,
                      "P": P, "layout": "system-major (p*N+n)*M*M", "blocks": "diagonal"}


def main() -> int:
    if not os.path.isdir("/root/reference"):
        print("reference tree not mounted; fixtures are committed already", file=sys.stderr)
        return 1
    build_ref()
    meta: dict = {}
    adi_fixtures(meta)
    blk_fixtures(meta)
    with open(os.path.join(HERE, "golden_meta.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print(f"wrote {len(meta)} fixtures")
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Builds libtridhip.so (all HIP sources of tridsolver_amd/csrc) for gfx950,
then the command-line tools of tridsolver_amd/apps against it.

    python -m tridsolver_amd.build          # or: python tridsolver_amd/build.py

The library is built in-tree (tridsolver_amd/libtridhip.so) so that it
travels with the source tree; hipcc cross-compiles without a GPU present.
Tools (tridsolver_amd/apps/bin/):
    adi_hip_d, adi_hip_s   ADI application (FP64 / FP32) on the library
    compare_d, compare_s   SumOfDiff comparison of two raw ADI outputs
    blktrid_hip_d, _s      block-tridiagonal benchmark driver (FP64 / FP32)
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtridhip.so")
APPS = os.path.join(HERE, "apps")
BIN = os.path.join(APPS, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TRIDHIP_ARCH", "gfx950")

FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-shared",
    # no packed-FP32 instructions (v_pk_*_f32) in any device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-Wall",
    "-Wno-unknown-pragmas",
]


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def build(verbose: bool = True) -> str:
    """Compile every source into LIB (always recompiles; parallel per file)."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    t0 = time.time()
    procs = []
    objs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [HIPCC] + [f for f in FLAGS if f != "-shared"] + ["-c"] + lang + [src, "-o", obj]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for src, p in procs:
        out, _ = p.communicate()
        txt = out.decode(errors="replace")
        # the host pass does not know the packed-fp32 feature: harmless
        lines = [ln for ln in txt.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"[build] FAILED {os.path.basename(src)}\n" + "\n".join(lines) + "\n")
        elif verbose and lines:
            sys.stderr.write(f"[build] {os.path.basename(src)}:\n" + "\n".join(lines[:40]) + "\n")
    if failed:
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    if verbose:
        print(f"[build] {LIB} ({time.time() - t0:.1f}s)")
    build_apps(verbose)
    return LIB


def build_apps(verbose: bool = True) -> None:
    """Compile the tools of apps/ (needs LIB) into apps/bin/."""
    os.makedirs(BIN, exist_ok=True)
    jobs = []
    for prec, tag in ((1, "d"), (0, "s")):
        jobs.append([HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}",
                     "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                     "-Wno-unused-command-line-argument", f"-DFPPREC={prec}",
                     os.path.join(APPS, "adi_hip.cpp"), "-L" + HERE, "-ltridhip",
                     "-Wl,-rpath,$ORIGIN/../..", "-o", os.path.join(BIN, f"adi_hip_{tag}")])
        jobs.append(["g++", "-O2", "-std=c++17", "-Wall", f"-DFPPREC={prec}",
                     os.path.join(APPS, "compare.cpp"), "-o", os.path.join(BIN, f"compare_{tag}")])
        jobs.append([HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}",
                     "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                     "-Wno-unused-command-line-argument", f"-DFPPREC={prec}",
                     os.path.join(APPS, "blktrid_hip.cpp"), "-L" + HERE, "-ltridhip",
                     "-Wl,-rpath,$ORIGIN/../..", "-o", os.path.join(BIN, f"blktrid_hip_{tag}")])
    procs = [(j, subprocess.Popen(j, stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
             for j in jobs]
    for j, p in procs:
        out, _ = p.communicate()
        lines = [ln for ln in out.decode(errors="replace").splitlines()
                 if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            raise RuntimeError(f"[build] FAILED {os.path.basename(j[-1])}\n" + "\n".join(lines))
    if verbose:
        print(f"[build] tools in {BIN}")


if __name__ == "__main__":
    build()

"""CPU checks of the C-ABI boundary: the library loads, exports every symbol
declared in include/*.h, and the Python mirror declares exactly those.
No compute call is made (no GPU here)."""
from __future__ import annotations

import ctypes as C
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tridsolver_amd", "libtridhip.so")


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"EXTERN_C\s+[\w\s\*]+?\b(\w+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_headers_declare_the_reference_api():
    names = declared_symbols()
    # scalar/include/trid_cuda.h:43-84 and block/include/blktrid_gpu.h:38-40
    for ref in ["initTridMultiDimBatchSolve", "tridSmtsvStridedBatch", "tridDmtsvStridedBatch",
                "tridDmtsvStridedBatchPadded", "tridSmtsvStridedBatchInc",
                "tridDmtsvStridedBatchInc", "tridDmtsvStridedBatchPaddedInc"]:
        assert ref in names, ref


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libtridhip.so not built (python -m tridsolver_amd.build)"
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB]).decode()
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = declared_symbols() - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"


def test_python_mirror_matches_headers():
    from tridsolver_amd import _abi
    assert set(_abi.SIGNATURES) == declared_symbols()


def test_library_loads_without_gpu():
    import tridsolver_amd
    L = tridsolver_amd.lib()
    assert L.tridGetStatusString(3) == b"TRID_STATUS_INVALID_VALUE"
    # argument validation happens before any device work
    dims = (C.c_int * 3)(4, 4, 4)
    st = L.tridDmtsvStridedBatch(None, None, None, None, None, 3, 5, dims, dims, None, 1)
    assert st == 3 and b"solvedim" in L.tridGetLastError()
    st = L.tridDmtsvStridedBatch(None, None, None, None, None, 9, 0, dims, dims, None, 1)
    assert st == 3
    bad = (C.c_int * 3)(4, 3, 4)
    st = L.tridDmtsvStridedBatch(None, None, None, None, None, 3, 0, dims, bad, None, 1)
    assert st == 3 and b"pads" in L.tridGetLastError()
    zero = (C.c_int * 3)(4, 0, 4)
    assert L.tridDmtsvStridedBatch(None, None, None, None, None, 3, 0, zero, zero, None, 1) == 0
    st = L.tridDmtsvStridedBatchInc(None, None, None, None, None, 3, 0, dims, dims, None, 1)
    assert st == 3


def test_cache_policy_switch_validates():
    import tridsolver_amd as ta
    for bad in (3, -2):
        with pytest.raises(ta.TridError):
            ta.force_cache_policy(bad)
    for ok in (0, 1, 2, -1):
        ta.force_cache_policy(ok)


def test_status_values_match_reference_enum():
    # scalar/include/trid_common.h:65-76
    import tridsolver_amd
    assert tridsolver_amd.TRID_STATUS[0] == "TRID_STATUS_SUCCESS"
    assert tridsolver_amd.TRID_STATUS[9] == "TRID_STATUS_ZERO_PIVOT"
    hdr = open(os.path.join(ROOT, "include", "trid_common.h")).read()
    for v, name in tridsolver_amd.TRID_STATUS.items():
        assert re.search(rf"{name}\s*=\s*{v}\b", hdr), name


def test_no_oracle_in_product_path():
    """The product package never imports or links the oracle."""
    for f in glob.glob(os.path.join(ROOT, "tridsolver_amd", "**", "*"), recursive=True):
        if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
            txt = open(f).read()
            assert "import oracle" not in txt and "liboracle" not in txt, f


@pytest.mark.parametrize("fn", ["tridSmtsvStridedBatchPadded", "tridSmtsvStridedBatchPaddedInc"])
def test_float_padded_forms_exported(fn):
    import tridsolver_amd
    assert hasattr(tridsolver_amd.lib(), fn)


def test_block_cxx_linkage_exports(tmp_path):
    """An object compiled against declarations with the reference's C++
    linkage (block/include/blktrid_gpu.h:38-40) references mangled names;
    the library exports exactly those."""
    src = tmp_path / "caller.cpp"
    src.write_text(
        "void sbtsv_gpu(float *A, float *B, float *C, float *d, float *Cstar, float *dstar,\n"
        "               float *u, const int N, const int P, const int blkdim);\n"
        "void dbtsv_gpu(double *A, double *B, double *C, double *d, double *Cstar,\n"
        "               double *dstar, double *u, const int N, const int P, const int blkdim);\n"
        "void call(float *f, double *g) {\n"
        "  sbtsv_gpu(f, f, f, f, f, f, f, 1, 1, 1);\n"
        "  dbtsv_gpu(g, g, g, g, g, g, g, 1, 1, 1);\n"
        "}\n")
    obj = tmp_path / "caller.o"
    subprocess.check_call(["g++", "-c", str(src), "-o", str(obj)])
    undef = {ln.split()[-1] for ln in
             subprocess.check_output(["nm", "-u", str(obj)]).decode().splitlines() if ln.strip()}
    wanted = {s for s in undef if "btsv_gpu" in s}
    assert len(wanted) == 2 and all(s.startswith("_Z") for s in wanted), wanted
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB]).decode()
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    assert wanted <= exported, sorted(wanted - exported)
    assert {"sbtsv_gpu", "dbtsv_gpu"} <= exported  # the extern "C" forms stay


def test_ragged_array_size_check():
    """RaggedPlan.solve refuses arrays shorter than offsets[nsys] (the
    kernels access every element the offsets name), wrong dtypes and
    non-contiguous arrays, before any device call."""
    import torch
    from tridsolver_amd.ragged import check_arrays
    f64 = torch.float64
    ok = [torch.zeros(100, dtype=f64) for _ in range(4)]
    check_arrays(100, f64, *ok)
    check_arrays(0, f64, *ok)
    short = [torch.zeros(100, dtype=f64)] * 3 + [torch.zeros(99, dtype=f64)]
    with pytest.raises(ValueError, match="d holds 99"):
        check_arrays(100, f64, *short)
    with pytest.raises(ValueError, match="dtype"):
        check_arrays(10, f64, *ok[:3], torch.zeros(100, dtype=torch.float32))
    with pytest.raises(ValueError, match="contiguous"):
        check_arrays(10, f64, *ok[:3], torch.zeros(200, dtype=f64)[::2])
    with pytest.raises(ValueError, match="u is required"):
        check_arrays(10, f64, *ok, inc=True)
    with pytest.raises(ValueError, match="u holds 50"):
        check_arrays(100, f64, *ok, u=torch.zeros(50, dtype=f64), inc=True)

"""ctypes signatures of every symbol declared in include/*.h."""
from __future__ import annotations

import ctypes as C

P = C.c_void_p
I = C.c_int
IP = C.POINTER(C.c_int)
ST = C.c_int  # tridStatus_t

# name -> (restype, argtypes); kept in step with include/*.h (a CPU test
# checks that every declared symbol is exported and listed here)
SIGNATURES = {
    # trid_hip.h
    "initTridMultiDimBatchSolve": (None, [I, IP, IP]),
    "tridSmtsvStridedBatch": (ST, [P, P, P, P, P, I, I, IP, IP, IP, I]),
    "tridDmtsvStridedBatch": (ST, [P, P, P, P, P, I, I, IP, IP, IP, I]),
    "tridSmtsvStridedBatchInc": (ST, [P, P, P, P, P, I, I, IP, IP, IP, I]),
    "tridDmtsvStridedBatchInc": (ST, [P, P, P, P, P, I, I, IP, IP, IP, I]),
    "tridDmtsvStridedBatchPadded": (ST, [P, IP, P, IP, P, IP, P, IP, P, IP, I, I, IP, IP, I]),
    "tridDmtsvStridedBatchPaddedInc": (ST, [P, IP, P, IP, P, IP, P, IP, P, IP, I, I, IP, IP, I]),
    "tridSmtsvStridedBatchPadded": (ST, [P, IP, P, IP, P, IP, P, IP, P, IP, I, I, IP, IP, I]),
    "tridSmtsvStridedBatchPaddedInc": (ST, [P, IP, P, IP, P, IP, P, IP, P, IP, I, I, IP, IP, I]),
    # trid_hip_ext.h
    "tridSetStream": (ST, [P]),
    "tridSetLaunchEvents": (ST, [P, P]),
    "tridGetStream": (P, []),
    "tridSetAutotune": (ST, [I]),
    "tridGetLastError": (C.c_char_p, []),
    "tridGetStatusString": (C.c_char_p, [I]),
    "tridGetLastConfig": (C.c_char_p, []),
    "tridForceConfig": (ST, [I, I, I, I]),
    "tridForceWideIndex": (ST, [I]),
    "tridForceCachePolicy": (ST, [I]),
    "tridDebugReciprocal": (ST, [P, P, C.c_longlong, I]),
    "tridDmtsvStridedBatchNoAlloc": (ST, [P, P, P, P, P, I, I, IP, IP, IP, I]),
    # trid_adi.h
    "tridAdiPlanCreate": (ST, [C.POINTER(P), I, I, I, I, C.c_double, I]),
    "tridAdiPlanDestroy": (ST, [P]),
    "tridAdiStep": (ST, [P, P, P, I, I]),
    "tridAdiStage": (ST, [P, P, P, I]),
    "tridAdiGraphCapture": (ST, [P, P, P, I]),
    "tridAdiGraphLaunch": (ST, [P, I]),
    "tridAdiPlanDescribe": (C.c_char_p, [P]),
    "tridAdiPlanSetSweep": (ST, [P, I, I, I, I]),
    # blktrid_hip.h
    "sbtsv_gpu": (None, [P, P, P, P, P, P, P, I, I, I]),
    "dbtsv_gpu": (None, [P, P, P, P, P, P, P, I, I, I]),
    "tridSbtsvStridedBatch": (ST, [P, P, P, P, P, P, P, I, I, I, I]),
    "tridDbtsvStridedBatch": (ST, [P, P, P, P, P, P, P, I, I, I, I]),
    "tridDbtsvMixedBatch": (ST, [P, I, I]),
    "tridDbtsvAdjoint": (ST, [P, P, P, P, P, P, P, P, P, I, I, I, I]),
    "tridBlkForceLanes": (ST, [I]),
    # trid_ragged.h
    "tridRaggedPlanCreate": (ST, [C.POINTER(P), P, C.c_longlong, I]),
    "tridRaggedPlanDestroy": (ST, [P]),
    "tridDgtsvRaggedBatch": (ST, [P, P, P, P, P, P, I, I]),
    "tridSgtsvRaggedBatch": (ST, [P, P, P, P, P, P, I, I]),
    "tridRaggedPlanDescribe": (C.c_char_p, [P]),
    "tridRaggedForceChoice": (ST, [I, I]),
    "tridRaggedPlanExtent": (ST, [P, C.POINTER(C.c_longlong)]),
    # trid_adjoint.h
    "tridDmtsvStridedBatchAdjoint": (ST, [P] * 9 + [I, I, IP, IP, I]),
    "tridSmtsvStridedBatchAdjoint": (ST, [P] * 9 + [I, I, IP, IP, I]),
    "tridVadiForward": (ST, [P, P, I, I, I, I, C.c_double, I]),
    "tridVadiLossGrad": (ST, [P, P, P, I, I, I, I, C.c_double, I, C.POINTER(C.c_double), P, P]),
    "tridVadiRhs": (ST, [P, P, P, I, I, I, C.c_double]),
    "tridVadiSweep": (ST, [P, P, P, I, I, I, I, C.c_double, I, I]),
}


def declare(lib: C.CDLL) -> None:
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args

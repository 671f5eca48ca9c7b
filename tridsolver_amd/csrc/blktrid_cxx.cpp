// blktrid_cxx.cpp -- the block solves with C++ linkage.
//
// block/include/blktrid_gpu.h:38-40 declares sbtsv_gpu and dbtsv_gpu with
// C++ linkage, so an object compiled against that header references the
// mangled names.  This file defines them (forwarding to the status-returning
// entry points of blktrid_hip.h) beside the extern "C" forms that
// blktrid.hip exports; it must not include blktrid_hip.h, whose extern "C"
// declarations of the same names would conflict with these.
#include "../../include/trid_common.h"

EXTERN_C tridStatus_t tridSbtsvStridedBatch(const float *A, const float *B, const float *C,
                                            const float *d, float *Cstar, float *dstar,
                                            float *u, int N, int P, int blkdim, int sync);
EXTERN_C tridStatus_t tridDbtsvStridedBatch(const double *A, const double *B, const double *C,
                                            const double *d, double *Cstar, double *dstar,
                                            double *u, int N, int P, int blkdim, int sync);

// block/include/blktrid_gpu.h:38; waits for completion, as the reference
void sbtsv_gpu(float *A, float *B, float *C, float *d, float *Cstar, float *dstar, float *u,
               const int N, const int P, const int blkdim) {
  (void)tridSbtsvStridedBatch(A, B, C, d, Cstar, dstar, u, N, P, blkdim, 1);
}

// block/include/blktrid_gpu.h:40; enqueues without waiting, as the reference
void dbtsv_gpu(double *A, double *B, double *C, double *d, double *Cstar, double *dstar,
               double *u, const int N, const int P, const int blkdim) {
  (void)tridDbtsvStridedBatch(A, B, C, d, Cstar, dstar, u, N, P, blkdim, 0);
}

/*
 * blktrid_hip.h -- batched block-tridiagonal solves on one MI355X.
 *
 * Drop-in replacement for block/include/blktrid_gpu.h:38-40 of
 * TobyFlynn/tridsolver.  P independent systems of N block rows with M x M
 * blocks (M = blkdim, 1..10):
 *     A_n u_{n-1} + B_n u_n + C_n u_{n+1} = d_n,   n = 0..N-1
 * (A_0 and C_{N-1} are never read).  Layout as the reference GPU solver
 * (block/src/blktrid_gpu.cuh:265-431): block (n, p) is row-major at
 * (n*P + p)*M*M, vector entry (n, p) at (n*P + p)*M.  Cstar (N*P*M*M) and
 * dstar (N*P*M) are caller-provided scratch that receives the elimination's
 * intermediates: in the one-ended layouts the downward sweep's
 * C'_n = X_n^{-1} C_n (n < N-1) and d'_n (every n); in the two-ended ones
 * (the default for most block sizes, see tridBlkForceLanes) C'_n and d'_n
 * at rows 0 .. N/2-1 and the upward sweep's A''_n = X_n^{-1} A_n and d''_n
 * at rows N/2+1 .. N-1.  The solution is u either way.  All pointers are
 * device pointers.
 *
 * The reference declares these with C++ linkage and no status.  Here they
 * are declared extern "C"; the library exports them with C++ linkage too
 * (the mangled names), so that objects compiled against the reference's
 * own blktrid_gpu.h link unchanged.  dbtsv_gpu enqueues without waiting
 * (as the reference),
 * sbtsv_gpu waits for completion (as the reference, without its timing
 * printout).  Errors are reported through tridGetLastError(); the
 * tridXbtsvStridedBatch forms return the status instead.
 */
#ifndef TRIDSOLVER_AMD_BLKTRID_HIP_H
#define TRIDSOLVER_AMD_BLKTRID_HIP_H

#include "trid_common.h"

/* block/include/blktrid_gpu.h:38 */
EXTERN_C void sbtsv_gpu(float *A, float *B, float *C, float *d, float *Cstar,
                        float *dstar, float *u, const int N, const int P,
                        const int blkdim);
/* block/include/blktrid_gpu.h:40 */
EXTERN_C void dbtsv_gpu(double *A, double *B, double *C, double *d,
                        double *Cstar, double *dstar, double *u, const int N,
                        const int P, const int blkdim);

/* No reference counterpart: the same solves returning a status; sync = 1
 * waits for completion.  Cstar/dstar may be NULL, in which case the library
 * uses its own workspace. */
EXTERN_C tridStatus_t tridSbtsvStridedBatch(const float *A, const float *B,
                                            const float *C, const float *d,
                                            float *Cstar, float *dstar,
                                            float *u, int N, int P, int blkdim,
                                            int sync);
EXTERN_C tridStatus_t tridDbtsvStridedBatch(const double *A, const double *B,
                                            const double *C, const double *d,
                                            double *Cstar, double *dstar,
                                            double *u, int N, int P,
                                            int blkdim, int sync);

/* No reference counterpart: several block batches of different block sizes
 * (blkdim 1..8, at most 16 groups) in ONE kernel launch.  Each group
 * describes a batch exactly as the arguments of tridDbtsvStridedBatch
 * (Cstar/dstar required).  The group table itself is read on the host. */
typedef struct {
  const double *A, *B, *C, *d;
  double *Cstar, *dstar, *u;
  int N, P, M;
} tridDbtsvGroup_t;
EXTERN_C tridStatus_t tridDbtsvMixedBatch(const tridDbtsvGroup_t *groups,
                                          int ngroups, int sync);

/* No reference counterpart: adjoint of tridDbtsvStridedBatch.  Given the
 * forward solution u and the upstream gradient g = dL/du, solves the
 * transposed block system for lambda = A^{-T} g (block row n of A^T has
 * C_{n-1}^T, B_n^T, A_{n+1}^T) and writes
 *   grad_d = lambda,  grad_A_n = -lambda_n u_{n-1}^T,
 *   grad_B_n = -lambda_n u_n^T,  grad_C_n = -lambda_n u_{n+1}^T
 * (outer products; grad_A_0 = grad_C_{N-1} = 0).  Layouts as above; the
 * gradient arrays may alias nothing else.  Workspace is internal. */
EXTERN_C tridStatus_t tridDbtsvAdjoint(const double *A, const double *B,
                                       const double *C, const double *u,
                                       const double *g, double *grad_A,
                                       double *grad_B, double *grad_C,
                                       double *grad_d, int N, int P,
                                       int blkdim, int sync);

/* Forces the lanes-per-problem of the block kernels (0 = automatic; else a
 * divisor of blkdim; 100 + L, L = 1, 2 or 4 dividing blkdim, for the
 * two-ended elimination with L lanes per half: one half eliminates
 * downward to row N/2, the other upward, two dependency chains of N/2 rows
 * per problem; 200 for the two-ended elimination on the tiled kernel, M
 * lanes per half, blkdim 3..8) -- for tests. */
EXTERN_C tridStatus_t tridBlkForceLanes(int lanes);

#endif

/*
 * trid_hip.h -- batched scalar tridiagonal solves on one MI355X (gfx950).
 *
 * Drop-in replacement for the GPU API of TobyFlynn/tridsolver,
 * scalar/include/trid_cuda.h.  Names, argument order and argument meaning are
 * the reference's; all pointers are device pointers.
 *
 *   a, b, c   sub-, main- and super-diagonal, d the right-hand side;
 *   ndim      number of dimensions (1..MAXDIM); dims[ndim] the extents,
 *             dimension 0 fastest; pads[ndim] the allocated extents, so that
 *             element (i_0..i_{n-1}) lives at sum_k i_k * prod_{j<k} pads[j];
 *   solvedim  the dimension along which each system runs: one independent
 *             system of length dims[solvedim] for every index of the others;
 *   a_0 and c_{N-1} of every system are never read (zero is assumed);
 *   non-Inc   the solution overwrites d, u is ignored (may be NULL);
 *   Inc       the solution is added into u, d is left unchanged;
 *   opts      per-dimension algorithm hint (reference meaning: opts[solvedim]
 *             picks a kernel).  Any value is accepted; 0 means "automatic".
 *             opts may be NULL.
 *   sync      1: return after the solve has finished (the reference always
 *             behaves this way); 0: return once it is queued on the stream
 *             set by tridSetStream (default: the null stream).
 *
 * Differences from the reference, all additive: invalid arguments return
 * TRID_STATUS_INVALID_VALUE instead of exit(-1) (trid_cuda.cu:164-166,
 * 214-219); pads[i] != dims[i] is supported along every solvedim (the
 * reference exits for solvedim 0, trid_cuda.cu:207-219); there is no N_MAX
 * cap on the system length; float forms of the Padded entry points exist.
 */
#ifndef TRIDSOLVER_AMD_TRID_HIP_H
#define TRIDSOLVER_AMD_TRID_HIP_H

#include "trid_common.h"

/* scalar/include/trid_cuda.h:43 (no-op in the reference, trid_cuda.cu:612).
 * Here: validates nothing, allocates nothing; kept for source compatibility. */
EXTERN_C void initTridMultiDimBatchSolve(int ndim, int *dims, int *pads);

/* scalar/include/trid_cuda.h:56-57 (trid_cuda.cu:515-522) */
EXTERN_C tridStatus_t tridSmtsvStridedBatch(const float *a, const float *b,
                                            const float *c, float *d, float *u,
                                            int ndim, int solvedim, int *dims,
                                            int *pads, int *opts, int sync);
/* scalar/include/trid_cuda.h:58-59 (trid_cuda.cu:525-532) */
EXTERN_C tridStatus_t tridDmtsvStridedBatch(const double *a, const double *b,
                                            const double *c, double *d,
                                            double *u, int ndim, int solvedim,
                                            int *dims, int *pads, int *opts,
                                            int sync);

/* scalar/include/trid_cuda.h:60-67 (trid_cuda.cu:534-544): one pads vector per
 * array. */
EXTERN_C tridStatus_t tridDmtsvStridedBatchPadded(
    const double *a, const int *a_pads, const double *b, const int *b_pads,
    const double *c, const int *c_pads, double *d, const int *d_pads,
    double *u, const int *u_pads, int ndim, int solvedim, int *dims, int *opts,
    int sync);

/* scalar/include/trid_cuda.h:73-74 (trid_cuda.cu:563-570) */
EXTERN_C tridStatus_t tridSmtsvStridedBatchInc(const float *a, const float *b,
                                               const float *c, float *d,
                                               float *u, int ndim,
                                               int solvedim, int *dims,
                                               int *pads, int *opts, int sync);
/* scalar/include/trid_cuda.h:75-76 (trid_cuda.cu:573-580) */
EXTERN_C tridStatus_t tridDmtsvStridedBatchInc(const double *a,
                                               const double *b,
                                               const double *c, double *d,
                                               double *u, int ndim,
                                               int solvedim, int *dims,
                                               int *pads, int *opts, int sync);

/* scalar/include/trid_cuda.h:77-84 (trid_cuda.cu:582-592) */
EXTERN_C tridStatus_t tridDmtsvStridedBatchPaddedInc(
    const double *a, const int *a_pads, const double *b, const int *b_pads,
    const double *c, const int *c_pads, double *d, const int *d_pads,
    double *u, const int *u_pads, int ndim, int solvedim, int *dims, int *opts,
    int sync);

/* No reference counterpart: float forms of the two Padded entry points
 * (the reference declares the double forms only). */
EXTERN_C tridStatus_t tridSmtsvStridedBatchPadded(
    const float *a, const int *a_pads, const float *b, const int *b_pads,
    const float *c, const int *c_pads, float *d, const int *d_pads, float *u,
    const int *u_pads, int ndim, int solvedim, int *dims, int *opts, int sync);
EXTERN_C tridStatus_t tridSmtsvStridedBatchPaddedInc(
    const float *a, const int *a_pads, const float *b, const int *b_pads,
    const float *c, const int *c_pads, float *d, const int *d_pads, float *u,
    const int *u_pads, int ndim, int solvedim, int *dims, int *opts, int sync);

#endif

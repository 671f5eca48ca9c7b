/*
 * trid_adjoint.h -- adjoints of the batched solves and a differentiable,
 * variable-diffusivity ADI (gradients by reverse-mode with checkpointing).
 *
 * No reference counterpart: TobyFlynn/tridsolver has no derivatives.
 *
 * Strided-batch adjoint.  For the forward solve x = A^{-1} d of
 * tridXmtsvStridedBatch (same a, b, c, ndim, solvedim, dims, pads; one pads
 * vector for every array) and the upstream gradient g = dL/dx:
 *   lambda = A^{-T} g   (transposed system: sub-diagonal c_{i-1}, diagonal
 *                        b_i, super-diagonal a_{i+1})
 *   grad_d = lambda, grad_a_i = -lambda_i x_{i-1}, grad_b_i = -lambda_i x_i,
 *   grad_c_i = -lambda_i x_{i+1}   (zero where the neighbour does not exist:
 *                                   grad_a_0 = grad_c_{N-1} = 0).
 *
 * Variable-diffusivity ADI ("vadi") on a dense nx*ny*nz grid, x fastest,
 * diffusivity field kappa (per point), face values
 * kf_{i+1/2} = (kappa_i + kappa_{i+1})/2 along each axis.  One step:
 *   r  = lam * sum_axes [kf_{i+1/2}(u_{i+1} - u_i) - kf_{i-1/2}(u_i - u_{i-1})]
 *        at interior points, 0 on the boundary;
 *   r <- A_x^{-1} r,  r <- A_y^{-1} r,  u <- u + A_z^{-1} r
 * with, at interior points of each line, a_i = -lam kf_{i-1/2},
 * c_i = -lam kf_{i+1/2}, b_i = 1 - a_i - c_i, identity rows on the boundary
 * (the reference ADI of apps/adi/src/adi_orig.c:147-243 with per-point
 * coefficients).  Loss = 0.5 * ||u_T - target||^2.
 */
#ifndef TRIDSOLVER_AMD_TRID_ADJOINT_H
#define TRIDSOLVER_AMD_TRID_ADJOINT_H

#include "trid_common.h"

EXTERN_C tridStatus_t tridDmtsvStridedBatchAdjoint(
    const double *a, const double *b, const double *c, const double *x,
    const double *g, double *grad_a, double *grad_b, double *grad_c,
    double *grad_d, int ndim, int solvedim, int *dims, int *pads, int sync);
EXTERN_C tridStatus_t tridSmtsvStridedBatchAdjoint(
    const float *a, const float *b, const float *c, const float *x,
    const float *g, float *grad_a, float *grad_b, float *grad_c, float *grad_d,
    int ndim, int solvedim, int *dims, int *pads, int sync);

/* u <- u_T after `steps` vadi steps from u (in place).  FP64. */
EXTERN_C tridStatus_t tridVadiForward(const double *kappa, double *u, int nx,
                                      int ny, int nz, int steps, double lam,
                                      int sync);

/* loss, d loss / d kappa and d loss / d u0 by reverse mode: the forward pass
 * keeps a checkpoint of u every `ckpt` steps; the backward pass recomputes
 * each segment from its checkpoint (storing the segment's intermediates)
 * and runs the adjoint sweeps.  *loss is written on the host (the call
 * waits for completion).  Workspace is internal. */
EXTERN_C tridStatus_t tridVadiLossGrad(const double *kappa, const double *u0,
                                       const double *target, int nx, int ny,
                                       int nz, int steps, double lam, int ckpt,
                                       double *loss, double *grad_kappa,
                                       double *grad_u0);

/* Single stages (for tests and timing): the right-hand side r = D(kappa) u;
 * one sweep d <- A_axis^{-1} d (transpose: A^{-T}; inc: u += A^{-1} d and d
 * unchanged). */
EXTERN_C tridStatus_t tridVadiRhs(const double *kappa, const double *u,
                                  double *r, int nx, int ny, int nz,
                                  double lam);
EXTERN_C tridStatus_t tridVadiSweep(const double *kappa, double *d, double *u,
                                    int nx, int ny, int nz, int axis,
                                    double lam, int transpose, int inc);

#endif

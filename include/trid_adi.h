/*
 * trid_adi.h -- ADI heat-diffusion time stepping on one MI355X.
 *
 * Replaces the time loop of the reference's ADI application
 * (apps/adi/src/adi_cuda.cu:285-330 with the preproc kernel of
 * apps/adi/src/preproc.cu:38-90).  The reference materialises nine
 * coefficient arrays every step and calls tridXmtsvStridedBatch along x, y
 * and (Inc) z; this driver keeps the constant coefficients in registers and
 * moves only u and du: per step
 *     du = lambda * (7-point Laplacian of u)  at interior points, 0 on the
 *          boundary (Dirichlet rows a = c = 0, b = 1)
 *     du <- A_x^{-1} du,  du <- A_y^{-1} du,  u <- u + A_z^{-1} du
 * with interior rows a = c = -lambda/2, b = 1 + lambda (mode 0, the
 * reference's coefficients, apps/adi/src/preproc.cu:70-72) or
 * a = c = -lambda, b = 1 + 2 lambda (mode 1).
 *
 * Arrays are dense nx*ny*nz device arrays, x fastest; u is updated in place,
 * du is caller-provided scratch of the same size.  All work is enqueued on
 * the stream set by tridSetStream.  This header has no counterpart in the
 * reference library (the reference ADI is an application, not an API).
 */
#ifndef TRIDSOLVER_AMD_TRID_ADI_H
#define TRIDSOLVER_AMD_TRID_ADI_H

#include "trid_common.h"

typedef struct tridAdiPlan_st *tridAdiPlan_t;

/* is_double: 1 = FP64, 0 = FP32.  1 <= nx, ny, nz <= 8192.  The y and z
 * sweeps of lines of at most 512 rows use precomputed factors of the
 * constant-coefficient matrix (computed here, once).  With autotuning on
 * (tridSetAutotune), the plan times its kernel choices at its first use, on
 * the arrays of that call, inside the time step, and keeps the fastest.  A
 * plan whose first use is inside a caller's stream capture is not tuned (a
 * capture cannot hold the timing; later tuning would replace the factor
 * tables that the caller's graph reads).  tridAdiPlanSetSweep and
 * tridAdiPlanDestroy free factor tables: a caller's graph that holds steps
 * of the plan must not be launched after them. */
EXTERN_C tridStatus_t tridAdiPlanCreate(tridAdiPlan_t *plan, int nx, int ny,
                                        int nz, int is_double, double lambda,
                                        int mode);
EXTERN_C tridStatus_t tridAdiPlanDestroy(tridAdiPlan_t plan);

/* Enqueues nsteps time steps; sync = 1 waits for them to finish.  From the
 * second consecutive call with the same arrays and nsteps >= 2 on, the steps
 * run as launches of a HIP graph of min(nsteps, 64) steps, captured at that
 * call (the plan keeps one such graph; the kernels are those of
 * tridAdiStage, so results are bitwise the same). */
EXTERN_C tridStatus_t tridAdiStep(tridAdiPlan_t plan, void *u, void *du,
                                  int nsteps, int sync);

/* Enqueues one stage of a time step: 0 = right-hand side, 1 = x sweep,
 * 2 = y sweep, 3 = z sweep with u += , 4 = right-hand side fused with the x
 * sweep (for tests and per-kernel timing).  tridAdiStep uses 4, 2, 3 when
 * the plan supports the fused stage (tridAdiPlanDescribe says so) and u, du
 * are 16-byte aligned, else 0, 1, 2, 3. */
EXTERN_C tridStatus_t tridAdiStage(tridAdiPlan_t plan, void *u, void *du,
                                   int stage);

/* Captures nsteps time steps on (u, du) into a HIP graph held by the plan
 * (replacing any earlier one); tridAdiGraphLaunch replays it. */
EXTERN_C tridStatus_t tridAdiGraphCapture(tridAdiPlan_t plan, void *u,
                                          void *du, int nsteps);
EXTERN_C tridStatus_t tridAdiGraphLaunch(tridAdiPlan_t plan, int sync);

/* Describes the kernel configurations of the plan's four stages. */
EXTERN_C const char *tridAdiPlanDescribe(tridAdiPlan_t plan);

/* Pins the y (dir = 1) or z (dir = 2) sweep of the plan to one kernel
 * layout, for tests and reproducible timings.  rows = 8, 16 or 32 selects
 * the sweep with precomputed factors and chunks of that many rows (pair = 1:
 * two adjacent lines per lane where the arrays allow it; xcd = 1: every XCD
 * takes one contiguous range of lines); rows = 0 the partitioned sweep
 * without precomputed factors.  dir = 0 pins the fused right-hand side + x
 * stage of a plan that has one: rows = 4, 8 or 16 (chunk length), pair = 1
 * for the x sweep with precomputed factors (0: factors computed in the
 * kernel), xcd = 1 for the XCD-contiguous block order (each XCD sweeps one
 * range of planes, so the stencil's neighbour planes hit its own L2).  Returns TRID_STATUS_INVALID_VALUE when the
 * layout does not fit the line length.  Pending autotuning of the plan is
 * cancelled; a graph captured earlier keeps the kernels it was captured
 * with. */
EXTERN_C tridStatus_t tridAdiPlanSetSweep(tridAdiPlan_t plan, int dir, int rows, int pair,
                                          int xcd);

#endif

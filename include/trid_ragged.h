/*
 * trid_ragged.h -- many independent tridiagonal systems of uneven length in
 * one persistent kernel launch.
 *
 * No reference counterpart as an API: the reference solves one length per
 * call (tridXmtsvStridedBatch) or one system per call (trid_scalarD,
 * scalar/include/trid_cpu.h:84-98).  Here system k occupies elements
 * [offsets[k], offsets[k+1]) of a, b, c, d (and u), contiguous, and is
 * solved exactly as trid_scalarD would solve it: a of its first row and c
 * of its last row are never read; non-Inc overwrites d, Inc adds into u.
 *
 * The arrays must hold at least offsets[nsys] elements: the library sees
 * only the offsets and accesses every element they name.
 *
 * tridRaggedPlanCreate reads the offsets once (device array of nsys+1
 * int64 values, non-decreasing), bins the systems by length into chunked
 * schedules (a power-of-two number of chunks of 5..8 rows, FP32 9..16
 * rows: one lane per system up to 8 rows (FP32 16), 2..64 lanes per system
 * with an in-wave reduced solve, 128..512 lanes per system across the waves
 * of a work-group through LDS) and builds a work list ordered from the most
 * to the least expensive item.  Each solve is one
 * launch of a persistent kernel whose work-groups pull items from an atomic
 * counter (reset by the kernel itself).  FP64 systems of 4097..8192 rows
 * go to a second launch of the same kernel (512 chunks of 9..16 rows).  Systems
 * longer than 8192 rows are solved after it: the longest of them one by
 * one, each as a strided batch
 * of one system (tridXmtsvStridedBatch, split-line solve past 8192 rows),
 * the rest by a one-thread-per-system launch; the plan picks the split by a
 * cost model (a one-system solve ~25 us, the one-thread launch ~0.5 us per
 * row of its longest system).
 * A plan may be used for any arrays with the same offsets; calls that
 * overlap on one plan race on its counter and must be stream-ordered.
 */
#ifndef TRIDSOLVER_AMD_TRID_RAGGED_H
#define TRIDSOLVER_AMD_TRID_RAGGED_H

#include <stdint.h>

#include "trid_common.h"

typedef struct tridRaggedPlan_st *tridRaggedPlan_t;

/* is_double: 1 = FP64, 0 = FP32 (the plan is bound to one precision). */
EXTERN_C tridStatus_t tridRaggedPlanCreate(tridRaggedPlan_t *plan,
                                           const int64_t *offsets,
                                           long long nsys, int is_double);
EXTERN_C tridStatus_t tridRaggedPlanDestroy(tridRaggedPlan_t plan);

EXTERN_C tridStatus_t tridDgtsvRaggedBatch(tridRaggedPlan_t plan,
                                           const double *a, const double *b,
                                           const double *c, double *d,
                                           double *u, int inc, int sync);
EXTERN_C tridStatus_t tridSgtsvRaggedBatch(tridRaggedPlan_t plan,
                                           const float *a, const float *b,
                                           const float *c, float *d, float *u,
                                           int inc, int sync);

/* No reference counterpart: the number of elements every array of a solve
 * with this plan must hold, offsets[nsys] (0 for an empty plan).  Callers
 * that know their arrays' sizes check them against it before a solve. */
EXTERN_C tridStatus_t tridRaggedPlanExtent(tridRaggedPlan_t plan,
                                           long long *elements);

/* Summary of the plan's bins ("R8/cps4: 1234 systems, ..."). */
EXTERN_C const char *tridRaggedPlanDescribe(tridRaggedPlan_t plan);

#endif

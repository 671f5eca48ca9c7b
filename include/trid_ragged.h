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
 * int64 values, non-decreasing) and bins the systems by length into chunked
 * schedules: a power-of-two number of chunks per system (1..512 lanes; up to
 * 64 of them in one wave with an in-wave reduced solve, more across the
 * waves of a work-group) in one of two row families, chunks of 5..8 rows (up
 * to 4096 rows per system) or of 9..16 rows (up to 8192).  Each length class
 * (2^c chunks of the 5..8 family; 4097..8192 rows) takes one family and,
 * past 64 chunks, one of two reduced solves: PCR through LDS, one barrier per
 * level, or PCR within each wave followed by a serial sweep across the
 * waves (three barriers).  Untuned, FP64 takes 5..8 rows up to 4096 rows with
 * PCR through LDS, FP32 (half the registers per row) 9..16 rows throughout
 * with the per-wave reduced solve.  With the autotuner on (tridSetAutotune(1)
 * when the plan is created) the plan times whole solves of a synthetic batch
 * with its own lengths (a uniform sample of the systems past 2^27 unknowns),
 * in buffers of its own, and from the untuned choices changes one class at a
 * time, from the most to the fewest unknowns, keeping a change that makes
 * the whole solve faster by more than 2 %.
 * The plan builds a work list per family ordered from the most to the least
 * expensive item; each solve is one launch of a persistent kernel per
 * family present, whose work-groups pull items from an atomic counter
 * (reset by the kernel itself).  Systems longer than 8192 rows are solved
 * after it: the longest of them one by one, each as a strided batch of one
 * system (tridXmtsvStridedBatch, split-line solve past 8192 rows), the rest
 * by a one-thread-per-system launch; the plan picks the split by a cost
 * model (untuned: a one-system solve ~25 us, the one-thread launch ~0.5 us
 * per row of its longest system; the autotuner measures both on the plan's
 * longest system).
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

/* Summary of the plan's bins ("R8/cps4:1234 ..."), each class's choice
 * ("classes: c3=0w ..." = family 0 rows, per-wave reduced solve; "p" = LDS
 * PCR) and, for a tuned plan, every candidate's time ("tuned: ..."). */
EXTERN_C const char *tridRaggedPlanDescribe(tridRaggedPlan_t plan);

/* Diagnostic (no reference counterpart): plans created afterwards put
 * every length class that the family can hold in row family `family` (0 =
 * chunks of 5..8 rows, 1 = 9..16; -1 = automatic) and use reduced solve
 * `red` past 64 chunks (0 = PCR through LDS, 1 = per-wave PCR and a sweep
 * across waves; -1 = automatic).  Either forced value turns the plan
 * autotuner off.  For tests of every schedule. */
EXTERN_C tridStatus_t tridRaggedForceChoice(int family, int red);

#endif

/*
 * trid_hip_ext.h -- tridsolver_amd extensions with no reference counterpart.
 *
 * Runtime control (stream, autotuner, diagnostics) for the drop-in entry
 * points of trid_hip.h and blktrid_hip.h.  The reference always runs on the
 * null stream and synchronises the device after every call
 * (scalar/src/cuda/trid_cuda.cu:442-443); callers that do nothing with these
 * functions get exactly that behaviour.
 */
#ifndef TRIDSOLVER_AMD_TRID_HIP_EXT_H
#define TRIDSOLVER_AMD_TRID_HIP_EXT_H

#include <stddef.h>
#include <stdint.h>

#include "trid_common.h"

/* Stream on which every entry point of this library enqueues its kernels
 * (a hipStream_t; NULL = the null stream).  Process-wide. */
EXTERN_C tridStatus_t tridSetStream(void *stream);
EXTERN_C void *tridGetStream(void);

/* Autotuner mode: 0 = fixed heuristic table, 1 = time the candidate kernel
 * configurations on the first call for each (precision, layout class, system
 * length, batch size) and keep the fastest.  Default 0; the environment
 * variable TRIDHIP_AUTOTUNE=1 sets 1 at load time. */
EXTERN_C tridStatus_t tridSetAutotune(int mode);

/* Diagnostic: the calling thread's next kernel launch of the library records
 * start / stop (hipEvent_t, either may be NULL) around its own execution
 * (hipExtLaunchKernel timestamps the dispatch packet), so that
 * hipEventElapsedTime(start, stop) is the kernel's duration without the
 * dispatch latency an event recorded before the call adds.  The next
 * strided solve or ADI stage consumes them: a single-kernel strided solve and
 * each tridAdiStage call (fused x stage, y sweep, z sweep: one kernel each)
 * record them; a split-line or one-thread-per-system solve discards them.
 * Other calls (ragged, block, adjoint) leave them pending. */
EXTERN_C tridStatus_t tridSetLaunchEvents(void *start, void *stop);

/* Short text of the last failing call on this thread ("" if none). */
EXTERN_C const char *tridGetLastError(void);
EXTERN_C const char *tridGetStatusString(tridStatus_t status);

/* Describes the kernel configuration the last strided-batch call on this
 * thread used, e.g. "part<double,o0,R8,cps32,spb8>" (for reports and tests). */
EXTERN_C const char *tridGetLastConfig(void);

/* Forces the next strided-batch calls onto one kernel configuration
 * (for tests and the autotuner's own sweeps).  kind: 0 = automatic,
 * 1 = partitioned Thomas with wave/LDS reduced-system PCR, 2 = one thread
 * per system with a global-memory forward sweep (the fallback), 3 = kind 1
 * with the XCD-contiguous work-group order (every XCD takes one contiguous
 * range of systems), 4 = split-line solve (every system cut into segments of
 * rows * cps rows that are reduced, coupled through a reduced system of two
 * rows per segment and solved as one batch; automatic for systems longer
 * than 8192 rows).  rows is the chunk length R of kinds 1, 3 and 4 (4, 8 or
 * 16; also 5, 6, 7 and 12 for kinds 1 and 3 along solvedim 0 with 16-byte
 * aligned arrays, the general staged layout; 0 = automatic), cps the chunks per system (kind 4: per segment; power
 * of two; 0 = automatic), spb the systems (kind 4: segments) per work-group
 * (0 = automatic). */
EXTERN_C tridStatus_t tridForceConfig(int kind, int rows, int cps, int spb);

/* Diagnostic: on = 1 makes every later call use 64-bit element offsets and
 * 64-bit system-index arithmetic, the forms the kernels take for arrays of
 * 4 GiB or more and batches of 2^31 systems or more, whatever the sizes (so
 * that tests reach those forms on small arrays); on = 0 (the default)
 * restores the automatic choice. */
EXTERN_C tridStatus_t tridForceWideIndex(int on);

/* Cache policy of the solvers' streaming loads and stores (the coefficient,
 * right-hand-side and solution arrays): policy 0 = non-temporal (they do not
 * allocate in L2 or the 256 MB last-level cache; best when the arrays exceed
 * that cache), 1 = ordinary (arrays that fit stay in the last-level cache,
 * so the next kernel or call reads them from there), 2 = non-temporal
 * coefficient arrays a, b, c and ordinary d and u (the same as 1 where the
 * coefficients are not arrays: ADI plans, the variable-diffusivity ADI),
 * -1 (the default) = automatic: ordinary when the arrays of the call -- of
 * an ADI plan: u and du -- span at most 384 MiB, else 2 when d and u do,
 * else non-temporal; with autotuning on, strided solves also time the other
 * policies.  Applies to
 * later strided solves and to ADI plans created afterwards; all policies
 * give bitwise equal results.  Solves with 64-bit element offsets are
 * always non-temporal.  No counterpart in the reference (its kernels have no
 * cache hints). */
EXTERN_C tridStatus_t tridForceCachePolicy(int policy);

/* Diagnostic: y[i] = the solvers' reciprocal of x[i] for n device values
 * (double when is_double, else float).  The solvers form every pivot
 * reciprocal with this routine; tests check its error bound. */
EXTERN_C tridStatus_t tridDebugReciprocal(const void *x, void *y, long long n,
                                          int is_double);

/* No-allocation form of tridDmtsvStridedBatch (same arguments, same
 * results).  It has no counterpart in the reference; it is for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  It calls no hipMalloc, hipFree or
 * allocating library routine and never autotunes: the untuned schedule of
 * tridDmtsvStridedBatch (autotuning off) runs, bit for bit the same
 * kernels.  The scratch of split-line solves (systems longer than 8192 rows;
 * solvedim >= 1 from 2048 rows) and of the forced one-thread-per-system
 * kernel comes from a static device array of TRID_NOALLOC_SCRATCH_BYTES that
 * each device holds until the process ends; a solve that needs more returns
 * TRID_STATUS_INVALID_VALUE (the last error names both sizes) before
 * launching anything.  Every call on one device shares that array: calls that
 * overlap on one device (other streams or threads) race. */
#define TRID_NOALLOC_SCRATCH_BYTES (64u << 20)
EXTERN_C tridStatus_t tridDmtsvStridedBatchNoAlloc(const double *a, const double *b,
                                                   const double *c, double *d, double *u,
                                                   int ndim, int solvedim, int *dims,
                                                   int *pads, int *opts, int sync);

#endif

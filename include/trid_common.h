/*
 * trid_common.h -- shared definitions of the tridsolver_amd C ABI.
 *
 * Drop-in counterpart of scalar/include/trid_common.h:39-77 of
 * TobyFlynn/tridsolver: the status enum keeps the reference's names and
 * numeric values (they are part of the ABI), EXTERN_C and MAXDIM keep their
 * meaning.  The reference header's FP/FPPREC/WARP_SIZE macros are build
 * switches of its own sources, not ABI, and are not defined here.
 */
#ifndef TRIDSOLVER_AMD_TRID_COMMON_H
#define TRIDSOLVER_AMD_TRID_COMMON_H

#ifdef __cplusplus
#define EXTERN_C extern "C"
#else
#define EXTERN_C
#endif

/* Largest ndim accepted by the strided-batch entry points.  The reference
 * compiles with MAXDIM 3 (trid_common.h:55); every ndim <= 3 call is
 * accepted unchanged, and up to 8 dimensions are supported. */
#define MAXDIM 8

/* Status codes, numerically identical to scalar/include/trid_common.h:65-76. */
typedef enum {
  TRID_STATUS_SUCCESS = 0,
  TRID_STATUS_NOT_INITIALIZED = 1,
  TRID_STATUS_ALLOC_FAILED = 2,
  TRID_STATUS_INVALID_VALUE = 3,
  TRID_STATUS_ARCH_MISMATCH = 4,
  TRID_STATUS_MAPPING_ERROR = 5,
  TRID_STATUS_EXECUTION_FAILED = 6,
  TRID_STATUS_INTERNAL_ERROR = 7,
  TRID_STATUS_MATRIX_TYPE_NOT_SUPPORTED = 8,
  TRID_STATUS_ZERO_PIVOT = 9
} tridStatus_t;

#endif

/*
 * getrf/double_blocked_getrf_npvt.h -- double-blocked LU without pivoting.
 *
 * Same entry points as the reference (include/getrf/double_blocked_getrf_npvt.h:9-23,
 * src/getrf/double_blocked_getrf_npvt.cpp:94-148), with the rocblas_handle
 * replaced by a mifact_handle.  Outer block nb, inner block b (normalised as
 * n <= 0: return; nb <= 0: n; b <= 0: nb; b = min(b, nb)).  info: first zero
 * pivot as a global 1-based column (0 if none); the factorization continues
 * past it, leaving that column unscaled.
 */
#pragma once

#include "mifact/handle.h"

#ifdef __cplusplus
extern "C" {
#endif

void double_blocked_getrf_npvt(mifact_handle handle, float* A, int lda, int n, int nb, int b,
                               int* info);

void double_blocked_getrf_npvt_double(mifact_handle handle, double* A, int lda, int n, int nb,
                                      int b, int* info);

#ifdef __cplusplus
}
#endif

/*
 * kernels/panel_kernels.h -- the pivot and panel kernels of the path as
 * standalone entry points (device pointers, enqueued on the handle stream).
 *
 * In the reference these are rocSOLVER/rocBLAS internals (no source in the
 * reference tree): the diagonal POTF2 inside rocsolver_spotrf
 * (src/potrf/double_blocked_potrf.cpp:34-56), the panel GETF2 + IAMAX inside
 * rocsolver_sgetrf (src/getrf/getrf_pivot_internal.h:17-37).
 */
#pragma once

#include "mifact/handle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Lower Cholesky of an n x n block in place (recursive, 64-wide POTF2
 * leaves).  *info = 0 or the 1-based column of the first non-positive pivot.
 * Replaces rocsolver_spotrf(lower) at double_blocked_potrf.cpp:38-44. */
void mifact_potrf_block(mifact_handle handle, int n, float* A, int lda, int* info);
void mifact_potrf_block_double(mifact_handle handle, int n, double* A, int lda,
                               int* info);

/* Partial-pivot LU of an m x n panel (m >= n) in place: recursive, with
 * cooperative 32-column GETF2 leaves.  ipiv[n] 1-based (local to the panel),
 * *info = first zero pivot.  Replaces rocsolver_sgetrf at
 * getrf_pivot_internal.h:28-34. */
void mifact_getrf_panel(mifact_handle handle, int m, int n, float* A, int lda,
                        int* ipiv, int* info);
void mifact_getrf_panel_double(mifact_handle handle, int m, int n, double* A,
                               int lda, int* ipiv, int* info);

/* Wave-reduction IAMAX: *result = 1-based index of the first entry of
 * maximum |x| (0 when n <= 0). */
void mifact_isamax(mifact_handle handle, int n, const float* x, int incx, int* result);
void mifact_idamax(mifact_handle handle, int n, const double* x, int incx, int* result);

#ifdef __cplusplus
}
#endif

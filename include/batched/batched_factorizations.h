/*
 * batched/batched_factorizations.h -- uneven batches of small Cholesky and
 * partial-pivot LU factorizations in one persistent launch, and their
 * mixed-precision refinement solves.
 *
 * The reference factors one large matrix per call
 * (include/potrf/double_blocked_potrf.h:9-103,
 *  include/getrf/double_blocked_getrf.h:26-74); a batch of small matrices
 * through those drivers costs one chain of launches per matrix.  Here the
 * whole batch is one launch: each workgroup takes the next matrix from a
 * device work queue (largest first), factors it with 32-column panels in
 * LDS and trailing updates on MFMA (FP32, or BF16 with FP32 accumulation),
 * and moves on until the queue is empty.
 *
 * Layout: matrix i is the n[i] x n[i] column-major FP32 matrix at
 * A + offsets[i] with leading dimension lda[i]; offsets, n and lda are
 * device arrays of `batch` entries.  info[i] (device) receives 0 or the
 * LAPACK info of matrix i.  ipiv for matrix i is at ipiv + ipiv_offsets[i]
 * (1-based row indices local to the matrix, LAPACK convention).
 * Orders up to 4096 are supported for Cholesky and up to 1024 for LU (the
 * LU panel is held in LDS); a matrix beyond the limit, or with lda < n,
 * gets info = -1 and is left untouched.  Every launch is on the handle's
 * stream.
 */
#pragma once

#include "mifact/handle.h"
#include "solve/mixed_refine.h"

#ifdef __cplusplus
extern "C" {
#endif

/* In-place lower Cholesky of every matrix (upper triangles untouched). */
void mifact_potrf_batched(mifact_handle handle, int batch, float* A, const long long* offsets,
                          const int* n, const int* lda, int* info);

/* In-place LU with partial pivoting of every matrix. */
void mifact_getrf_batched(mifact_handle handle, int batch, float* A, const long long* offsets,
                          const int* n, const int* lda, int* ipiv, const long long* ipiv_offsets,
                          int* info);

/* The same with the trailing updates in `precision` (FP32 or BF16/FP16
 * operands with FP32 accumulation). */
void mifact_potrf_batched_ex(mifact_handle handle, mifact_precision precision, int batch,
                             float* A, const long long* offsets, const int* n, const int* lda,
                             int* info);
void mifact_getrf_batched_ex(mifact_handle handle, mifact_precision precision, int batch,
                             float* A, const long long* offsets, const int* n, const int* lda,
                             int* ipiv, const long long* ipiv_offsets, int* info);

/* Batched solves with FP64 refinement, one right-hand side per matrix:
 * factors every A_i (FP32, not modified) into F (same offsets/lda; caller
 * provided) with updates in `precision`, then runs `iterations` refinement
 * steps x_i += M_i^{-1} (b_i - A_i x_i) with the residual and update in FP64.
 * The refinement loop is captured once into a hipGraph and replayed; a
 * matrix whose backward error is already <= tolerance skips its work in the
 * remaining steps.  b and x are FP64 vectors at B + vec_offsets[i] and
 * X + vec_offsets[i], all inside the first vec_elems entries (the library
 * keeps an FP32 staging copy of that size).  Returns after enqueueing (the
 * graph is instantiated, launched and released within the call).
 * berr[i] (device, FP64) receives the final normwise
 * backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf);
 * info[i] the factorization info.  sym = 1: SPD (Cholesky, lower triangle
 * of A read); sym = 0: general (LU; ipiv as above). */
mifact_status mifact_gesv_batched_refine(mifact_handle handle, int sym,
                                         mifact_precision precision, int batch, const float* A,
                                         float* F, const long long* offsets, const int* n,
                                         const int* lda, int* ipiv, const long long* ipiv_offsets,
                                         const double* B, double* X,
                                         const long long* vec_offsets, long long vec_elems,
                                         int iterations, double tolerance, double* berr,
                                         int* info);

#ifdef __cplusplus
}
#endif

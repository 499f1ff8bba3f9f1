/*
 * potrf/double_blocked_potrf.h -- double-blocked lower Cholesky (POTRF) on
 * MI355X, drop-in for the reference's entry points
 * (include/potrf/double_blocked_potrf.h:9-103,
 *  src/potrf/double_blocked_potrf.cpp:294-587).
 *
 * Same symbol names and argument order; the first argument is a
 * mifact_handle (stream + workspace) where the reference takes a
 * rocblas_handle.  A is column-major n x n with leading dimension lda (device
 * pointer, 64-bit addressing inside); on return its lower triangle holds L
 * with A = L L^T.  The strictly upper triangle is not referenced.
 *
 * Argument normalization as in the reference (:309-321): n <= 0 returns;
 * nb <= 0 -> n; b <= 0 -> nb; b = min(b, nb).
 *
 * Schedule (two-level left-looking): for each outer block of nb columns, for
 * each inner panel of b columns: diagonal POTF2 (recursive, 64-wide leaf
 * kernel), panel TRSM (recursive, register-blocked leaves + MFMA GEMM),
 * inner update of the next panel with the outer block's columns (MFMA GEMM,
 * lower trapezoid); then the next outer block is updated with all previous
 * columns by one fused lower-trapezoid MFMA launch (the reference's SYRK +
 * GEMM pair).
 *
 * info (device int): 0 on success, else the 1-based GLOBAL column of the
 * first non-positive pivot (LAPACK semantics).  Deliberate difference: the
 * reference passes one info pointer to every rocSOLVER diagonal call, so it
 * ends up holding only the last block's local status.
 *
 * The syrk_k_tile / gemm_m_tile / thresholds of the reference's variants are
 * accepted for ABI compatibility; they split rocBLAS calls in the reference
 * and have no counterpart in the fused MFMA launches.  _dbtrsm selects the
 * reference's double-blocked TRSM schedule (1024-column micro blocks).
 */
#pragma once

#include "mifact/handle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Allocation-free variant of double_blocked_potrf (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, schedule and results as double_blocked_potrf (FP32), but
 * the diagonal-block inverses live in a fixed 1 MiB device array of the
 * library instead of the handle's workspace, so the call performs no
 * allocation of any kind.  Limit: inner block b <= 4096 after the usual
 * normalization; a larger one prints an error, records
 * mifact_status_invalid_value on the handle, sets *info = -1 and leaves A
 * untouched.  All calls on one GPU share that array: calls that overlap on
 * one GPU (other streams, other threads) race.  The handle's stage profiler
 * must be off. */
void double_blocked_potrf_noalloc(mifact_handle handle, float* A, int lda, int n, int nb, int b,
                                  int* info);

/* reference :9-15 */
void double_blocked_potrf(mifact_handle handle, float* A, int lda, int n,
                          int nb, int b, int* info);
/* reference :17-23 */
void double_blocked_potrf_dbtrsm(mifact_handle handle, float* A, int lda,
                                 int n, int nb, int b, int* info);
/* reference :25-32 */
void double_blocked_potrf_kblocked_syrk(mifact_handle handle, float* A,
                                        int lda, int n, int nb, int b,
                                        int syrk_k_tile, int* info);
/* reference :34-42 */
void double_blocked_potrf_kblocked_syrk_adaptive_gemm(
    mifact_handle handle, float* A, int lda, int n, int nb, int b,
    int syrk_k_tile, int* info);
/* reference :44-55 */
void double_blocked_potrf_kblocked_syrk_adaptive_gemm_configured(
    mifact_handle handle, float* A, int lda, int n, int nb, int b,
    int syrk_k_tile, int gemm_m_tile, int gemm_m_threshold,
    int gemm_k_threshold, int* info);

/* reference :57-63 */
void double_blocked_potrf_double(mifact_handle handle, double* A, int lda,
                                 int n, int nb, int b, int* info);
/* reference :65-71 */
void double_blocked_potrf_double_dbtrsm(mifact_handle handle, double* A,
                                        int lda, int n, int nb, int b,
                                        int* info);
/* reference :73-80 */
void double_blocked_potrf_double_kblocked_syrk(mifact_handle handle,
                                               double* A, int lda, int n,
                                               int nb, int b, int syrk_k_tile,
                                               int* info);
/* reference :82-90 */
void double_blocked_potrf_double_kblocked_syrk_adaptive_gemm(
    mifact_handle handle, double* A, int lda, int n, int nb, int b,
    int syrk_k_tile, int* info);
/* reference :92-103 */
void double_blocked_potrf_double_kblocked_syrk_adaptive_gemm_configured(
    mifact_handle handle, double* A, int lda, int n, int nb, int b,
    int syrk_k_tile, int gemm_m_tile, int gemm_m_threshold,
    int gemm_k_threshold, int* info);

#ifdef __cplusplus
}
#endif

/*
 * compat/rocblas_compat.h -- the reference's exact prototypes, taking a
 * rocblas_handle, for callers that cannot change their call sites
 * (include/potrf/double_blocked_potrf.h, include/potrf/recursive_potrf.h,
 *  include/getrf/double_blocked_getrf.h, include/getrf/
 *  double_blocked_getrf_npvt.h, include/gemm/update_gemm.h, include/syrk/
 *  kblocked_syrk.h, include/trsm/double_blocked_trsm.h, include/trsm/
 *  right_lower_trsm.h of the reference).
 *
 * Implemented by libmifact_compat.so, a thin forwarding library outside
 * the hot path: each call reads the rocblas_handle's stream
 * (rocblas_get_stream, the library's only use of rocBLAS), binds it to a
 * per-thread, per-device mifact_handle and forwards to the same-named
 * entry point of libmifact.so, which it loads privately (RTLD_LOCAL), so
 * the two libraries' identically named symbols never meet.  A caller links
 * libmifact_compat.so (and rocBLAS) in place of the reference's objects
 * and keeps its call sites unchanged.  Semantics are those of the native
 * headers (info is LAPACK-global; the tile knobs are accepted and ignored);
 * the StreamPool of the recursive family is accepted and ignored.
 */
#pragma once

#include <rocblas/rocblas.h>

#include "getrf/getrf_types.h"

#ifdef __cplusplus
class StreamPool;
extern "C" {
#else
typedef struct StreamPool StreamPool;
#endif

/* gemm/update_gemm.h */
void gemm_nt_update(rocblas_handle handle, int rows, int cols, int k, const float* A, int lda,
                    const float* B, int ldb, float* C, int ldc);
void gemm_nt_update_double(rocblas_handle handle, int rows, int cols, int k, const double* A,
                           int lda, const double* B, int ldb, double* C, int ldc);
void gemm_nt_update_m_tiled(rocblas_handle handle, int rows, int cols, int k, const float* A,
                            int lda, const float* B, int ldb, float* C, int ldc, int m_tile);
void gemm_nt_update_m_tiled_double(rocblas_handle handle, int rows, int cols, int k,
                                   const double* A, int lda, const double* B, int ldb, double* C,
                                   int ldc, int m_tile);

/* getrf/double_blocked_getrf.h */
void double_blocked_getrf(rocblas_handle handle, float* A, int lda, int n, int outer_nb,
                          int inner_nb, getrf_outer_trsm_mode outer_trsm_mode, int trsm_block,
                          int* ipiv, int* info);
void double_blocked_getrf_double(rocblas_handle handle, double* A, int lda, int n, int outer_nb,
                                 int inner_nb, getrf_outer_trsm_mode outer_trsm_mode,
                                 int trsm_block, int* ipiv, int* info);
void double_blocked_getrf_batched_pivot(rocblas_handle handle, float* A, int lda, int n,
                                        int outer_nb, int inner_nb,
                                        getrf_outer_trsm_mode outer_trsm_mode, int trsm_block,
                                        int inner_gemm_m_tile, int* ipiv, int* info,
                                        getrf_batched_pivot_workspace* workspace);
void double_blocked_getrf_double_batched_pivot(rocblas_handle handle, double* A, int lda, int n,
                                               int outer_nb, int inner_nb,
                                               getrf_outer_trsm_mode outer_trsm_mode,
                                               int trsm_block, int inner_gemm_m_tile, int* ipiv,
                                               int* info,
                                               getrf_batched_pivot_workspace* workspace);

/* getrf/double_blocked_getrf_npvt.h */
void double_blocked_getrf_npvt(rocblas_handle handle, float* A, int lda, int n, int nb, int b,
                               int* info);
void double_blocked_getrf_npvt_double(rocblas_handle handle, double* A, int lda, int n, int nb,
                                      int b, int* info);

/* potrf/double_blocked_potrf.h */
void double_blocked_potrf(rocblas_handle handle, float* A, int lda, int n, int nb, int b,
                          int* info);
void double_blocked_potrf_dbtrsm(rocblas_handle handle, float* A, int lda, int n, int nb, int b,
                                 int* info);
void double_blocked_potrf_kblocked_syrk(rocblas_handle handle, float* A, int lda, int n, int nb,
                                        int b, int syrk_k_tile, int* info);
void double_blocked_potrf_kblocked_syrk_adaptive_gemm(rocblas_handle handle, float* A, int lda,
                                                      int n, int nb, int b, int syrk_k_tile,
                                                      int* info);
void double_blocked_potrf_kblocked_syrk_adaptive_gemm_configured(rocblas_handle handle, float* A,
                                                                 int lda, int n, int nb, int b,
                                                                 int syrk_k_tile, int gemm_m_tile,
                                                                 int gemm_m_threshold,
                                                                 int gemm_k_threshold, int* info);
void double_blocked_potrf_double(rocblas_handle handle, double* A, int lda, int n, int nb, int b,
                                 int* info);
void double_blocked_potrf_double_dbtrsm(rocblas_handle handle, double* A, int lda, int n, int nb,
                                        int b, int* info);
void double_blocked_potrf_double_kblocked_syrk(rocblas_handle handle, double* A, int lda, int n,
                                               int nb, int b, int syrk_k_tile, int* info);
void double_blocked_potrf_double_kblocked_syrk_adaptive_gemm(rocblas_handle handle, double* A,
                                                             int lda, int n, int nb, int b,
                                                             int syrk_k_tile, int* info);
void double_blocked_potrf_double_kblocked_syrk_adaptive_gemm_configured(rocblas_handle handle,
                                                                        double* A, int lda, int n,
                                                                        int nb, int b,
                                                                        int syrk_k_tile,
                                                                        int gemm_m_tile,
                                                                        int gemm_m_threshold,
                                                                        int gemm_k_threshold,
                                                                        int* info);

/* potrf/recursive_potrf.h */
void recursive_potrf(rocblas_handle handle, StreamPool* pool, float* A, int lda, int n, int nb,
                     int* info);
void recursive_potrf_tiled_trsm(rocblas_handle handle, StreamPool* pool, float* A, int lda, int n,
                                int nb, int* info);
void recursive_potrf_double(rocblas_handle handle, StreamPool* pool, double* A, int lda, int n,
                            int nb, int* info);
void recursive_potrf_mixed_driver(rocblas_handle handle, StreamPool* pool, float* A, int lda,
                                  int n, int nb, int* info);

/* syrk/kblocked_syrk.h */
void syrk_lower_update(rocblas_handle handle, int n, int k, const float* A, int lda, float* C,
                       int ldc);
void syrk_lower_update_double(rocblas_handle handle, int n, int k, const double* A, int lda,
                              double* C, int ldc);
void syrk_lower_update_kblocked(rocblas_handle handle, int n, int k, const float* A, int lda,
                                float* C, int ldc, int k_tile);
void syrk_lower_update_kblocked_double(rocblas_handle handle, int n, int k, const double* A,
                                       int lda, double* C, int ldc, int k_tile);

/* trsm/double_blocked_trsm.h */
void double_blocked_trsm(rocblas_handle handle, float* A, int lda, float* B, int ldb, int n, int m,
                         int nb, int b);
void double_blocked_trsm_double(rocblas_handle handle, double* A, int lda, double* B, int ldb,
                                int n, int m, int nb, int b);

/* trsm/right_lower_trsm.h */
void trsm_right_lower_transpose(rocblas_handle handle, int rows, int cols, const float* A, int lda,
                                float* B, int ldb);
void trsm_right_lower_transpose_double(rocblas_handle handle, int rows, int cols, const double* A,
                                       int lda, double* B, int ldb);

#ifdef __cplusplus
}
#endif

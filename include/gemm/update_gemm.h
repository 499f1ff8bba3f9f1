/*
 * gemm/update_gemm.h -- trailing-update GEMM, C -= A * B^T, on the gfx950
 * MFMA kernels (FP32: v_mfma_f32_32x32x2_f32, FP64: v_mfma_f64_16x16x4_f64).
 *
 * Replaces the reference's rocBLAS-backed wrappers with the same names and
 * argument order (reference: include/gemm/update_gemm.h:9-53,
 * src/gemm/update_gemm.cpp:10-101).  C(rows x cols) -= A(rows x k) *
 * B(cols x k)^T, column-major, device pointers, enqueued on the handle's
 * stream.  The _m_tiled variants keep the reference's signature; the M
 * split was a rocBLAS workaround and is accepted but not needed here (one
 * launch covers all rows; results are identical up to rounding).
 */
#pragma once

#include "mifact/handle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference: include/gemm/update_gemm.h:9-18 */
void gemm_nt_update(mifact_handle handle, int rows, int cols, int k,
                    const float* A, int lda, const float* B, int ldb,
                    float* C, int ldc);
/* reference: include/gemm/update_gemm.h:20-29 */
void gemm_nt_update_double(mifact_handle handle, int rows, int cols, int k,
                           const double* A, int lda, const double* B, int ldb,
                           double* C, int ldc);
/* reference: include/gemm/update_gemm.h:31-41 */
void gemm_nt_update_m_tiled(mifact_handle handle, int rows, int cols, int k,
                            const float* A, int lda, const float* B, int ldb,
                            float* C, int ldc, int m_tile);
/* reference: include/gemm/update_gemm.h:43-53 */
void gemm_nt_update_m_tiled_double(mifact_handle handle, int rows, int cols,
                                   int k, const double* A, int lda,
                                   const double* B, int ldb, double* C,
                                   int ldc, int m_tile);

/* C(rows x cols) -= L(rows x k) * U(k x cols); the reference keeps this as
 * an internal helper (src/getrf/getrf_npvt_internal.h:90-144). */
void gemm_nn_update(mifact_handle handle, int rows, int cols, int k,
                    const float* L, int ldl, const float* U, int ldu,
                    float* C, int ldc);
void gemm_nn_update_double(mifact_handle handle, int rows, int cols, int k,
                           const double* L, int ldl, const double* U, int ldu,
                           double* C, int ldc);

#ifdef __cplusplus
}
#endif

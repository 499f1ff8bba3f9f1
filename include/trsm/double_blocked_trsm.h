/*
 * trsm/double_blocked_trsm.h -- two-level blocked B := B * L^{-T}
 * (reference: include/trsm/double_blocked_trsm.h:9-27,
 * src/trsm/double_blocked_trsm.cpp:99-171): n rows of B, m columns, outer
 * block nb and inner block b.  The diagonal solves use the recursive
 * right-lower-transpose TRSM, the updates the MFMA GEMM.
 */
#pragma once

#include "mifact/handle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference: include/trsm/double_blocked_trsm.h:9-17 */
void double_blocked_trsm(mifact_handle handle, float* A, int lda, float* B,
                         int ldb, int n, int m, int nb, int b);
/* reference: include/trsm/double_blocked_trsm.h:19-27 */
void double_blocked_trsm_double(mifact_handle handle, double* A, int lda,
                                double* B, int ldb, int n, int m, int nb,
                                int b);

#ifdef __cplusplus
}
#endif

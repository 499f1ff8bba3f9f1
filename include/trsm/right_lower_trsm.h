/*
 * trsm/right_lower_trsm.h -- B := B * L^{-T} (Right, Lower, Trans, NonUnit),
 * the POTRF panel solve.  Recursive halving of the columns: 64-wide leaves
 * run a register-blocked substitution kernel (one thread per row of B, L in
 * LDS), the off-diagonal blocks are MFMA GEMM updates.
 *
 * Same names and argument order as the reference
 * (include/trsm/right_lower_trsm.h:9-23, src/trsm/right_lower_trsm.cpp:9-55).
 */
#pragma once

#include "mifact/handle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference: include/trsm/right_lower_trsm.h:9-15 */
void trsm_right_lower_transpose(mifact_handle handle, int rows, int cols,
                                const float* A, int lda, float* B, int ldb);
/* reference: include/trsm/right_lower_trsm.h:17-23 */
void trsm_right_lower_transpose_double(mifact_handle handle, int rows,
                                       int cols, const double* A, int lda,
                                       double* B, int ldb);

#ifdef __cplusplus
}
#endif

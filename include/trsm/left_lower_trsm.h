/*
 * trsm/left_lower_trsm.h -- B := L^{-1} * B (Left, Lower, NoTrans, Unit),
 * the GETRF U-block solve.  Recursive halving of the rows: 64-row leaves run
 * a substitution kernel (one thread per column, B tile staged through LDS),
 * the off-diagonal blocks are MFMA GEMM updates.
 *
 * The reference keeps these as internal helpers:
 *   trsm_left_lower_unit            src/getrf/getrf_npvt_internal.h:42-88
 *   trsm_left_lower_unit_iterative  src/getrf/getrf_pivot_internal.h:64-110
 * (block_size rows per diagonal solve, each followed by a GEMM update).
 */
#pragma once

#include "mifact/handle.h"

#ifdef __cplusplus
extern "C" {
#endif

void trsm_left_lower_unit(mifact_handle handle, int rows, int cols,
                          const float* L, int ldl, float* B, int ldb);
void trsm_left_lower_unit_double(mifact_handle handle, int rows, int cols,
                                 const double* L, int ldl, double* B, int ldb);
void trsm_left_lower_unit_iterative(mifact_handle handle, int rows, int cols,
                                    const float* L, int ldl, float* B, int ldb,
                                    int block_size);
void trsm_left_lower_unit_iterative_double(mifact_handle handle, int rows,
                                           int cols, const double* L, int ldl,
                                           double* B, int ldb, int block_size);

#ifdef __cplusplus
}
#endif

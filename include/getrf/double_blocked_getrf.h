/*
 * getrf/double_blocked_getrf.h -- double-blocked LU with partial pivoting
 * (GETRF) on MI355X, drop-in for the reference's entry points
 * (include/getrf/double_blocked_getrf.h:9-74,
 *  src/getrf/double_blocked_getrf.cpp:125-333,
 *  src/getrf/getrf_batched_pivot_internal.h:445-624).
 *
 * Same symbol names, argument order, workspace struct and enum; the first
 * argument is a mifact_handle where the reference takes a rocblas_handle.
 * A is column-major n x n (lda); on return it holds the unit-lower L
 * (strictly lower part) and U (upper part) with P A = L U.  ipiv[n] is
 * 1-based and global (LAPACK convention).  info (device int): 0, or the
 * 1-based global column of the first exactly-zero pivot; -1 for an invalid
 * batched-pivot workspace.
 *
 * Pivot selection is LAPACK partial pivoting (first row of maximum
 * magnitude), computed by a cooperative multi-workgroup GETF2 leaf; the
 * panel of each inner block is factored recursively (leaves of 32 columns,
 * MFMA GEMM in between).  All pivot bookkeeping stays on the device: the
 * row exchanges of a panel are applied to other columns as a gather from a
 * device-resident row-origin map, with no host round trip (the reference's
 * batched path copies pivots to the host and synchronizes once per panel).
 *
 * double_blocked_getrf applies each panel's exchanges to every other column
 * immediately; double_blocked_getrf_batched_pivot follows the reference's
 * batched schedule (current outer block while factoring, left history and
 * the next outer block once per outer block).  Both give bit-identical
 * factors and pivots.  The batched variant validates the workspace as the
 * reference does (pointers non-null, capacity >= n, column_tile > 0); its
 * device_swap_buffer (capacity x column_tile elements) stages long move
 * lists; the host arrays are accepted but not used.
 */
#pragma once

#include "mifact/handle.h"
#include "getrf/getrf_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference :26-35 */
void double_blocked_getrf(mifact_handle handle, float* A, int lda, int n,
                          int outer_nb, int inner_nb,
                          getrf_outer_trsm_mode outer_trsm_mode,
                          int trsm_block, int* ipiv, int* info);
/* reference :37-46 */
void double_blocked_getrf_double(mifact_handle handle, double* A, int lda,
                                 int n, int outer_nb, int inner_nb,
                                 getrf_outer_trsm_mode outer_trsm_mode,
                                 int trsm_block, int* ipiv, int* info);
/* reference :48-60 */
void double_blocked_getrf_batched_pivot(
    mifact_handle handle, float* A, int lda, int n, int outer_nb,
    int inner_nb, getrf_outer_trsm_mode outer_trsm_mode, int trsm_block,
    int inner_gemm_m_tile, int* ipiv, int* info,
    getrf_batched_pivot_workspace* workspace);
/* reference :62-74 */
void double_blocked_getrf_double_batched_pivot(
    mifact_handle handle, double* A, int lda, int n, int outer_nb,
    int inner_nb, getrf_outer_trsm_mode outer_trsm_mode, int trsm_block,
    int inner_gemm_m_tile, int* ipiv, int* info,
    getrf_batched_pivot_workspace* workspace);

#ifdef __cplusplus
}
#endif

/*
 * getrf/getrf_pivot_kernels.h -- stand-alone row-exchange launchers, same
 * names and arguments as the reference's (include/getrf/
 * getrf_pivot_kernels.h:9-43).  Stream-ordered, no handle, no allocation;
 * launch errors print and exit, as the reference's HIP_CHECK.
 *
 * launch_apply_getrf_panel_pivots: LAPACK laswp of one panel's pivots
 *   (1-based, local to rows [panel, n)) over every column outside
 *   [panel, panel + panel_size); then pivots[k] += panel (global, 1-based).
 *   Implemented as composed row moves (chunks of 256 pivots), not one
 *   sequential swap loop per column.
 * launch_apply_getrf_permutation: for every column of A's first
 *   column_count columns, row targets[m] receives the pre-call entry of row
 *   sources[m] (m < move_count).  temporary (column_tile x move_count
 *   elements) is used only when one column of the move list exceeds 64 KiB;
 *   below that the moves are staged in LDS.
 */
#pragma once

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

void launch_apply_getrf_panel_pivots(float* A, int lda, int n, int panel, int panel_size,
                                     int* pivots, hipStream_t stream);
void launch_apply_getrf_panel_pivots_double(double* A, int lda, int n, int panel, int panel_size,
                                            int* pivots, hipStream_t stream);
void launch_apply_getrf_permutation(float* A, int lda, int column_count, const int* sources,
                                    const int* targets, int move_count, float* temporary,
                                    int column_tile, hipStream_t stream);
void launch_apply_getrf_permutation_double(double* A, int lda, int column_count,
                                           const int* sources, const int* targets,
                                           int move_count, double* temporary, int column_tile,
                                           hipStream_t stream);

#ifdef __cplusplus
}
#endif

/*
 * getrf/getrf_types.h -- the batched-pivot workspace and the outer-TRSM
 * mode of the reference's GETRF interface (include/getrf/
 * double_blocked_getrf.h:9-24), shared by the native header and the
 * rocblas_handle compatibility header.
 */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

/* reference :9-19 (layout unchanged) */
typedef struct getrf_batched_pivot_workspace {
    int* device_sources;
    int* device_targets;
    void* device_swap_buffer;
    int* host_pivots;
    int* host_permutation;
    int* host_sources;
    int* host_targets;
    int capacity;
    int column_tile;
} getrf_batched_pivot_workspace;

/* reference :21-24 */
typedef enum getrf_outer_trsm_mode {
    getrf_outer_trsm_direct = 0,
    getrf_outer_trsm_iterative = 1
} getrf_outer_trsm_mode;

#ifdef __cplusplus
}
#endif

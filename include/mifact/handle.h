/*
 * mifact/handle.h -- library handle for the MI355X double-blocked
 * factorizations.
 *
 * The reference's entry points take a rocblas_handle as their first argument
 * and enqueue every call on that handle's stream
 * (reference: include/potrf/double_blocked_potrf.h:9-103,
 *  include/getrf/double_blocked_getrf.h:26-74, rocblas_get_stream at
 *  src/getrf/getrf_batched_pivot_internal.h:474-475).
 * Here the same position holds a mifact_handle: an opaque object that carries
 * the HIP stream every kernel is launched on, the device it was created on and
 * a lazily grown device workspace arena (pivot permutations, exchange buffers
 * of the cooperative panel kernel).  The handle is not thread-safe; use one
 * handle per host thread, as with rocBLAS.
 */
#pragma once

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mifact_handle_s* mifact_handle;

/* Status codes (0 = success). */
typedef enum mifact_status {
    mifact_status_success = 0,
    mifact_status_invalid_handle = 1,
    mifact_status_invalid_value = 2,
    mifact_status_hip_error = 3,
    mifact_status_allocation_failed = 4,
    mifact_status_kernel_timeout = 5
} mifact_status;

/* Error behaviour on HIP API failures.  The reference's HIP_CHECK prints the
 * error and exits the process (utils/common/error_check.h:9-28); that is the
 * default here too.  mifact_error_return instead records the status, prints
 * it, and returns to the caller (used by test harnesses). */
typedef enum mifact_error_mode {
    mifact_error_exit = 0,
    mifact_error_return = 1
} mifact_error_mode;

/* Creates a handle on the current device, bound to the null stream (like
 * rocblas_create_handle). */
mifact_status mifact_create_handle(mifact_handle* handle);
mifact_status mifact_destroy_handle(mifact_handle handle);
mifact_status mifact_set_stream(mifact_handle handle, hipStream_t stream);
mifact_status mifact_get_stream(mifact_handle handle, hipStream_t* stream);
mifact_status mifact_set_error_mode(mifact_handle handle, mifact_error_mode mode);

/* Returns and clears the first error recorded on the handle since the last
 * call (sticky status for mifact_error_return mode).  It also reports, and
 * clears, the device status word that a cooperative panel kernel or a
 * blocked triangular solve raises when one of its bounded waits expires
 * (mifact_status_kernel_timeout; the GETRF drivers then also leave
 * *info = -2).  Every pivoting driver call clears that word on entry, so the
 * report covers the calls since the last driver entry.  Synchronizes the
 * handle's stream. */
mifact_status mifact_get_last_error(mifact_handle handle);

/* Number of unsuccessful polls after which a bounded device-side wait gives
 * up (the GETF2 leaf's pivot exchange, the blocked TRSV's dependency flags);
 * 0 restores each kernel's default.  A diagnostic knob: a small limit makes
 * the failure path observable without a hang. */
mifact_status mifact_set_wait_limit(mifact_handle handle, long long polls);

/* Pre-sizes the workspace for order n (otherwise grown on first use; growing
 * synchronizes the device). */
mifact_status mifact_reserve_workspace(mifact_handle handle, int n);

/* FP32 update-GEMM tile configuration used by every routine on this handle
 * (a tuning knob for the autotuner): 0 = automatic (default), 1 = 128x128
 * tiles with 16-deep K slices (four waves), 2 = 128x128 / 32-deep, 3 = 64x64
 * / 16-deep, 4 = 256x256 / 16-deep on eight waves. */
mifact_status mifact_set_gemm_config(mifact_handle handle, int config);

/* Rule by which the automatic configuration (config 0) picks FP32 update
 * tiles, the autotuner's choice per routine and order: 0 (default) = 256 x
 * 256 for full-shape updates with >= 4 tiles per CU, 128 x 128 with >= 2,
 * 64 x 64 below; 1 = never 256 x 256; 2 = 256 x 256 also for lower-trapezoid
 * updates; 3 = as 0 but 128 x 128 from 1 tile per CU. */
mifact_status mifact_set_gemm_policy(mifact_handle handle, int policy);

/* Size classes of the batched factorizations' trailing updates (an
 * autotuner knob): a trailing matrix of order up to `potrf_small_max`
 * (Cholesky) or `getrf_small_max` (LU) is updated in 32 x 32 tiles, one per
 * wave; larger ones in 64 x 64 tiles on the workgroup's four waves.
 * Defaults 96 / 96; 0 keeps every update on the 64 x 64 tiles. */
mifact_status mifact_set_batched_tiles(mifact_handle handle, int potrf_small_max,
                                       int getrf_small_max);

/* Library version string. */
const char* mifact_version(void);

/* ------------------------------------------------------------------------
 * Stage profiler (the reference's StageBreakdownProfiler /
 * OperationShapeProfiler, include/utils/benchmark/stage_breakdown_profiler.h:
 * 19-137, re-done without host synchronization): while enabled, every kernel
 * launch of the library is bracketed by a pair of hipEvents on the handle's
 * stream and tagged with its stage and algorithmic work; mifact_profile_end
 * synchronizes once and resolves the times. */
typedef struct mifact_stage_stats {
    char name[32];   /* stage: gemm, potf2, trsm_leaf, getf2_leaf, row_moves, perm,
                        gemm_lowp, solve, batched, gemm_update (128 x 128
                        tile trailing-update GEMM, full shape),
                        gemm_update_256 (256 x 256 tiles, full shape),
                        gemm_update_lower, gemm_update_256_lower (the same
                        tiles, lower-trapezoid SYRK + GEMM updates) */
    int calls;       /* launches */
    double flops;    /* algorithmic flops of those launches */
    double bytes;    /* algorithmic HBM bytes of those launches */
    double ms;       /* summed kernel time (event pairs) */
} mifact_stage_stats;

mifact_status mifact_profile_begin(mifact_handle handle);
/* Records only the stages whose bit is set in stage_mask (bit order: gemm,
 * potf2, trsm_leaf, getf2_leaf, row_moves, perm, gemm_lowp, solve, batched,
 * gemm_update, gemm_update_256, gemm_update_lower, gemm_update_256_lower),
 * so a timed run can bracket just the dominant kernel. */
mifact_status mifact_profile_begin_stages(mifact_handle handle, unsigned stage_mask);
/* Writes up to max_stages entries; *count receives the number of stages. */
mifact_status mifact_profile_end(mifact_handle handle, mifact_stage_stats* stats,
                                 int max_stages, int* count);

#ifdef __cplusplus
}
#endif

/*
 * potrf/recursive_potrf.h -- recursive lower Cholesky, the reference's
 * baseline family (include/potrf/recursive_potrf.h:1-46,
 * src/potrf/recursive_potrf.cpp:17-388).
 *
 * Same names and argument order.  The first argument is a mifact_handle
 * (stream + workspace) where the reference takes a rocblas_handle; the
 * reference's StreamPool* (independent rocBLAS handles that run the batched
 * SYRK's diagonal and off-diagonal blocks concurrently) becomes an opaque
 * pointer that is accepted and ignored: the trailing SYRK is one
 * lower-trapezoid MFMA launch on the handle's stream here.
 *
 * Recursion: n <= nb -> diagonal factorization of the leaf; else n1 = n/2,
 * factor A11, B := A21 L11^-T (recursive TRSM with nb-column leaves), A22 -=
 * A21 A21^T (lower), factor A22.  nb <= 0 means one leaf (nb = n).
 * info (device int): 0, or the 1-based global column of the first
 * non-positive pivot (LAPACK semantics; the reference leaves rocSOLVER's
 * per-leaf local value).
 *
 * recursive_potrf_tiled_trsm differs from recursive_potrf only in how the
 * reference tiles its TRSM update GEMM (src/trsm/recursive_trsm.cpp:82-130);
 * the MFMA GEMM tiles internally, so both compute the same thing here.
 *
 * recursive_potrf_mixed_driver: lower triangle scaled by 65504 / max|diag|,
 * panel-update and trailing-update GEMMs in FP16 with FP32 accumulation
 * (the accumulated matrix stays FP32; the reference rounds it to FP16 too),
 * a leaf that breaks down is restored and retried with diagonal damping
 * 6.5504, doubling, up to 5 times (host-synchronizing per leaf, as in the
 * reference), and the factor is scaled back.  It factors A + D for the
 * damping D it needed: a preconditioner, not an exact Cholesky factor.
 */
#pragma once

#include "mifact/handle.h"

#ifdef __cplusplus
extern "C" {
#endif

void recursive_potrf(mifact_handle handle, void* pool, float* A, int lda, int n, int nb,
                     int* info);
void recursive_potrf_tiled_trsm(mifact_handle handle, void* pool, float* A, int lda, int n,
                                int nb, int* info);
void recursive_potrf_double(mifact_handle handle, void* pool, double* A, int lda, int n, int nb,
                            int* info);
void recursive_potrf_mixed_driver(mifact_handle handle, void* pool, float* A, int lda, int n,
                                  int nb, int* info);

#ifdef __cplusplus
}
#endif

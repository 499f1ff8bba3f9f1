/*
 * solve/mixed_refine.h -- triangular solves from the double-blocked factors,
 * the mixed-precision factorizations and the FP64 iterative-refinement
 * solvers.
 *
 * The reference has no solve path: its drivers stop at the factors
 * (include/potrf/double_blocked_potrf.h:9-103,
 *  include/getrf/double_blocked_getrf.h:26-74) and accuracy is judged by the
 * CPU FP64 residual (include/utils/benchmark/cpu_fp64_residual.h:28-201).
 * These entry points add the solve and the precision pattern on top of the
 * same schedule:
 *   - the trailing SYRK/GEMM updates of POTRF/GETRF run on BF16, FP16 or FP8
 *     e4m3 (per-tile power-of-two scales) MFMA with FP32 accumulation; the diagonal POTF2/GETF2, IAMAX, row moves and
 *     panel TRSM stay FP32;
 *   - the matrix is equilibrated by powers of two first (exact in FP32), so
 *     the low-precision operands stay in range;
 *   - accuracy is recovered by iterative refinement in which only the
 *     residual r = b - A x and the update x += d run in FP64.
 * Column-major device pointers throughout; every call is enqueued on the
 * handle's stream.
 */
#pragma once

#include "mifact/handle.h"
#include "getrf/double_blocked_getrf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic of the trailing updates of a factorization. */
typedef enum mifact_precision {
    mifact_precision_fp32 = 0,  /* FP32 MFMA (the plain path)              */
    mifact_precision_bf16 = 1,  /* BF16 operands, FP32 accumulation        */
    mifact_precision_fp16 = 2,  /* FP16 operands, FP32 accumulation        */
    mifact_precision_fp8 = 3    /* FP8 e4m3 operands with one power-of-two
                                   scale per 32 x 32 operand tile (E8M0 on
                                   v_mfma_scale_f32_32x32x64_f8f6f4), FP32
                                   accumulation                            */
} mifact_precision;

/* C(m x n) += alpha * A(m x k) * op(B) with the products on BF16/FP16/FP8
 * MFMA and FP32 accumulation (operands packed once per call into the
 * handle's pack arena, grown on demand).  trans_b = 1: op(B) = B^T, B stored n x k;
 * trans_b = 0: op(B) = B, B stored k x n.  lower = 1 updates only C(i, j)
 * with i >= j (requires m >= n).  FP32 precision uses the FP32 MFMA GEMM. */
void mifact_gemm_update_lowp(mifact_handle handle, mifact_precision precision, int trans_b,
                             int lower, int m, int n, int k, float alpha, const float* A,
                             int lda, const float* B, int ldb, float* C, int ldc);

/* Double-blocked POTRF / GETRF with the trailing updates in `precision`
 * (same schedule, arguments and info/ipiv semantics as
 * double_blocked_potrf and double_blocked_getrf). */
void double_blocked_potrf_mixed(mifact_handle handle, mifact_precision precision, float* A,
                                int lda, int n, int nb, int b, int* info);
void double_blocked_getrf_mixed(mifact_handle handle, mifact_precision precision, float* A,
                                int lda, int n, int outer_nb, int inner_nb,
                                getrf_outer_trsm_mode outer_trsm_mode, int trsm_block, int* ipiv,
                                int* info);

/* B(n x nrhs) := A^{-1} B from the factors of double_blocked_potrf (L in the
 * lower triangle; LAPACK spotrs with uplo = 'L'). */
void mifact_potrs(mifact_handle handle, int n, int nrhs, const float* L, int ldl, float* B,
                  int ldb);

/* B(n x nrhs) := A^{-1} B from the factors and 1-based pivots of
 * double_blocked_getrf (LAPACK sgetrs with trans = 'N'). */
void mifact_getrs(mifact_handle handle, int n, int nrhs, const float* LU, int ldlu,
                  const int* ipiv, float* B, int ldb);

/* R(n x nrhs) := B - A X in FP64 (A FP32, X and B FP64).  symmetric_lower = 1
 * reads only the lower triangle of A and mirrors it.  Deterministic (fixed
 * reduction order). */
void mifact_residual_fp64(mifact_handle handle, int symmetric_lower, int n, int nrhs,
                          const float* A, int lda, const double* X, int ldx, const double* B,
                          int ldb, double* R, int ldr);

/* Refinement method.  Classical: x += M^{-1} r with the low-precision
 * factors M (converges when kappa(A) times the factorization's backward
 * error is well below one).  GMRES-IR (Carson & Higham, SIAM J. Sci. Comput.
 * 40(2), 2018): each correction solves A d = r by restarted FGMRES in FP64,
 * right-preconditioned by M applied in FP32; converges for kappa(A) up to
 * about the inverse of the FP32 unit roundoff.  Auto: classical steps while
 * they contract, then GMRES-IR from the current iterate. */
typedef enum mifact_refine_method {
    mifact_refine_classical = 0,
    mifact_refine_gmres = 1,
    mifact_refine_auto = 2
} mifact_refine_method;

/* Parameters of the refinement solvers. */
typedef struct mifact_refine_params {
    mifact_precision precision;  /* trailing-update arithmetic                       */
    int nb;                      /* outer block (<= 0: n)                            */
    int b;                       /* inner block (<= 0: nb)                           */
    getrf_outer_trsm_mode outer_trsm_mode;  /* GETRF only                            */
    int trsm_block;              /* GETRF iterative outer TRSM block                 */
    int max_iterations;          /* refinement steps after the first solve (<0: 30)  */
    double tolerance;            /* stop when backward error <= tolerance;
                                    <= 0: sqrt(n) * DBL_EPSILON (LAPACK dsgesv)       */
    int fallback_fp32;           /* 1: when the refinement stagnates with BF16/FP16
                                    updates, refactor with FP32 updates and refine
                                    again (as dsgesv falls back to FP64)              */
    mifact_refine_method method; /* classical, GMRES-IR or automatic                 */
} mifact_refine_params;

/* Outcome of a refinement solve (host struct, filled before return). */
typedef struct mifact_refine_result {
    int info;                    /* factorization info (0: success)                  */
    int iterations;              /* refinement steps taken after the first solve     */
    int converged;               /* 1 when backward_error <= tolerance               */
    double backward_error;       /* max_j ||b_j - A x_j||_inf /
                                    (||A||_inf ||x_j||_inf + ||b_j||_inf)            */
    double first_backward_error; /* the same after the first (unrefined) solve       */
    int fell_back;               /* 1 when the FP32 refactorization was used          */
    int factorizations;          /* factorizations performed (1 or 2)                */
    int gmres_iterations;        /* inner FGMRES iterations (preconditioner applies) */
    double diagonal_shift;       /* POSV: shift added to the equilibrated diagonal when
                                    the BF16/FP16 Cholesky broke down (Higham &
                                    Pranesh 2021); 0 when none was needed            */
} mifact_refine_result;

/* Solves A X = B for SPD A (lower triangle read; FP32, not modified), B and
 * X FP64 (n x nrhs).  `factor` (FP32, n x n, ldf) receives the equilibrated
 * low-precision Cholesky factor.  Synchronizes the handle's stream once per
 * refinement step to test convergence.  Returns mifact_status_success unless
 * an argument or a HIP call failed; numerical outcome in *result. */
mifact_status mifact_posv_refine(mifact_handle handle, int n, int nrhs, const float* A, int lda,
                                 float* factor, int ldf, const double* B, int ldb, double* X,
                                 int ldx, const mifact_refine_params* params,
                                 mifact_refine_result* result);

/* Solves A X = B for general A (FP32, not modified) with partial pivoting;
 * ipiv (device, n ints) receives the 1-based pivots of the equilibrated
 * factorization held in `factor`. */
mifact_status mifact_gesv_refine(mifact_handle handle, int n, int nrhs, const float* A, int lda,
                                 float* factor, int ldf, int* ipiv, const double* B, int ldb,
                                 double* X, int ldx, const mifact_refine_params* params,
                                 mifact_refine_result* result);

#ifdef __cplusplus
}
#endif

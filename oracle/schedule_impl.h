/*
 * oracle/schedule_impl.h -- CPU restatement of the reference's double-blocked
 * POTRF / GETRF schedules (test infrastructure only; included twice by
 * oracle/schedule.c with T = float and T = double).
 *
 * The reference composes rocBLAS 4.3.0 / rocSOLVER 3.20.0 calls (third-party,
 * not vendored).  Their arithmetic is restated from the published LAPACK /
 * BLAS algorithms they implement:
 *   xGEMM  C := C - A op(B)                 (alpha=-1, beta=1)
 *   xSYRK  C_lower := C_lower - A A^T
 *   xTRSM  Right/Lower/Trans/NonUnit and Left/Lower/NoTrans/Unit
 *   xPOTF2 unblocked lower Cholesky (LAPACK spotf2)
 *   xGETF2 unblocked partial-pivot LU with IxAMAX (first index of the max
 *          magnitude; LAPACK sgetf2, reciprocal scaling when |pivot| >= sfmin)
 * Summation order inside a product is plain ascending k in precision T; the
 * rocBLAS kernels' internal order is not published (parity for the floating
 * values is by residual tolerance, not bit equality).
 *
 * Drivers follow the reference loop structure line by line:
 *   double_blocked_potrf_impl   src/potrf/double_blocked_potrf.cpp:294-381
 *   update_next_inner_panel      src/potrf/double_blocked_potrf.cpp:210-249
 *   prepare_next_outer_panel     src/potrf/double_blocked_potrf.cpp:251-292
 *   double_blocked_trsm_impl     src/trsm/double_blocked_trsm.cpp:99-171
 *   double_blocked_getrf_impl    src/getrf/double_blocked_getrf.cpp:125-191
 *   batched-pivot double_blocked_getrf
 *                                src/getrf/getrf_batched_pivot_internal.h:445-624
 *   apply_panel_pivots_kernel    src/getrf/getrf_pivot_kernels.hip:14-58
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUF)

static T* FN(bp)(T* A, int lda, int row, int col)
{
    return A + (size_t)col * (size_t)lda + (size_t)row;
}

/* C(rows x cols) -= A(rows x k) * B(cols x k)^T      (update_gemm.cpp:10-64) */
static void FN(gemm_nt)(int rows, int cols, int k, const T* A, int lda,
                        const T* B, int ldb, T* C, int ldc)
{
    if (rows <= 0 || cols <= 0 || k <= 0) return;
    for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i) {
            T s = 0;
            for (int p = 0; p < k; ++p)
                s += A[(size_t)p * lda + i] * B[(size_t)p * ldb + j];
            C[(size_t)j * ldc + i] -= s;
        }
}

/* C(rows x cols) -= L(rows x k) * U(k x cols)   (getrf_npvt_internal.h:90-144) */
static void FN(gemm_nn)(int rows, int cols, int k, const T* L, int ldl,
                        const T* U, int ldu, T* C, int ldc)
{
    if (rows <= 0 || cols <= 0 || k <= 0) return;
    for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i) {
            T s = 0;
            for (int p = 0; p < k; ++p)
                s += L[(size_t)p * ldl + i] * U[(size_t)j * ldu + p];
            C[(size_t)j * ldc + i] -= s;
        }
}

/* C_lower(n x n) -= A(n x k) A^T, K split in k_tile chunks
 * (kblocked_syrk.cpp:17-96) */
static void FN(syrk_lower)(int n, int k, const T* A, int lda, T* C, int ldc,
                           int k_tile)
{
    if (n <= 0 || k <= 0) return;
    const int step = (k_tile <= 0 || k_tile >= k) ? k : k_tile;
    for (int k0 = 0; k0 < k; k0 += step) {
        const int kc = (step < k - k0) ? step : k - k0;
        for (int j = 0; j < n; ++j)
            for (int i = j; i < n; ++i) {
                T s = 0;
                for (int p = k0; p < k0 + kc; ++p)
                    s += A[(size_t)p * lda + i] * A[(size_t)p * lda + j];
                C[(size_t)j * ldc + i] -= s;
            }
    }
}

/* B(rows x cols) := B * L^{-T}, L lower non-unit (right_lower_trsm.cpp:10-55)
 * LAPACK strsm Right/Lower/Trans: reciprocal of the diagonal. */
static void FN(trsm_rlt)(int rows, int cols, const T* L, int ldl, T* B, int ldb)
{
    if (rows <= 0 || cols <= 0) return;
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) {
            T s = B[(size_t)j * ldb + i];
            for (int p = 0; p < j; ++p)
                s -= B[(size_t)p * ldb + i] * L[(size_t)p * ldl + j];
            B[(size_t)j * ldb + i] = s * ((T)1 / L[(size_t)j * ldl + j]);
        }
}

/* B(rows x cols) := L^{-1} B, L lower unit     (getrf_npvt_internal.h:42-88) */
static void FN(trsm_llu)(int rows, int cols, const T* L, int ldl, T* B, int ldb)
{
    if (rows <= 0 || cols <= 0) return;
    for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i) {
            T s = B[(size_t)j * ldb + i];
            for (int p = 0; p < i; ++p)
                s -= L[(size_t)p * ldl + i] * B[(size_t)j * ldb + p];
            B[(size_t)j * ldb + i] = s;
        }
}

/* Unblocked lower Cholesky (LAPACK xPOTF2 lower).  Returns local info. */
static int FN(potf2)(int n, T* A, int lda)
{
    for (int j = 0; j < n; ++j) {
        T ajj = A[(size_t)j * lda + j];
        for (int p = 0; p < j; ++p)
            ajj -= A[(size_t)p * lda + j] * A[(size_t)p * lda + j];
        if (!(ajj > (T)0)) {
            A[(size_t)j * lda + j] = ajj;
            return j + 1;
        }
        ajj = (T)sqrt((double)ajj);
        A[(size_t)j * lda + j] = ajj;
        const T r = (T)1 / ajj;
        for (int i = j + 1; i < n; ++i) {
            T s = A[(size_t)j * lda + i];
            for (int p = 0; p < j; ++p)
                s -= A[(size_t)p * lda + i] * A[(size_t)p * lda + j];
            A[(size_t)j * lda + i] = s * r;
        }
    }
    return 0;
}

/* IxAMAX: 0-based index of the first entry of maximum magnitude. */
static int FN(iamax)(int n, const T* x)
{
    if (n <= 0) return -1;
    int best = 0;
    T bmax = x[0] < 0 ? -x[0] : x[0];
    for (int i = 1; i < n; ++i) {
        const T v = x[i] < 0 ? -x[i] : x[i];
        if (v > bmax) { bmax = v; best = i; }
    }
    return best;
}

/* Unblocked partial-pivot LU of an m x n panel (LAPACK xGETF2).
 * ipiv is 1-based, local to the panel.  Returns local info. */
static int FN(getf2)(int m, int n, T* A, int lda, int* ipiv)
{
    const T sfmin = SFMIN;
    int info = 0;
    const int mn = m < n ? m : n;
    for (int j = 0; j < mn; ++j) {
        const int jp = j + FN(iamax)(m - j, A + (size_t)j * lda + j);
        ipiv[j] = jp + 1;
        const T piv = A[(size_t)j * lda + jp];
        if (piv != (T)0) {
            if (jp != j)
                for (int c = 0; c < n; ++c) {
                    T t = A[(size_t)c * lda + j];
                    A[(size_t)c * lda + j] = A[(size_t)c * lda + jp];
                    A[(size_t)c * lda + jp] = t;
                }
            const T d = A[(size_t)j * lda + j];
            if ((d < 0 ? -d : d) >= sfmin) {
                const T r = (T)1 / d;
                for (int i = j + 1; i < m; ++i) A[(size_t)j * lda + i] *= r;
            } else {
                for (int i = j + 1; i < m; ++i) A[(size_t)j * lda + i] /= d;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        for (int c = j + 1; c < n; ++c) {
            const T u = A[(size_t)c * lda + j];
            for (int i = j + 1; i < m; ++i)
                A[(size_t)c * lda + i] -= A[(size_t)j * lda + i] * u;
        }
    }
    return info;
}

/* ---------------------------------------------------------------- POTRF */

/* double_blocked_potrf.cpp:210-249 (M-tiling is a rocBLAS call split only) */
static void FN(potrf_inner_update)(T* A, int lda, int n, int outer,
                                   int panel_start, int outer_end, int b)
{
    if (panel_start >= outer_end) return;
    const int panel_cols = (b < outer_end - panel_start) ? b : outer_end - panel_start;
    const int rows = n - panel_start;
    const int k = panel_start - outer;
    FN(gemm_nt)(rows, panel_cols, k, FN(bp)(A, lda, panel_start, outer), lda,
                FN(bp)(A, lda, panel_start, outer), lda,
                FN(bp)(A, lda, panel_start, panel_start), lda);
}

/* double_blocked_potrf.cpp:251-292 */
static void FN(potrf_outer_update)(T* A, int lda, int n, int next_outer,
                                   int nb, int syrk_k_tile)
{
    if (next_outer >= n) return;
    const int next_size = (nb < n - next_outer) ? nb : n - next_outer;
    const int below_start = next_outer + next_size;
    const int below_rows = n - below_start;
    FN(syrk_lower)(next_size, next_outer, FN(bp)(A, lda, next_outer, 0), lda,
                   FN(bp)(A, lda, next_outer, next_outer), lda, syrk_k_tile);
    FN(gemm_nt)(below_rows, next_size, next_outer,
                FN(bp)(A, lda, below_start, 0), lda,
                FN(bp)(A, lda, next_outer, 0), lda,
                FN(bp)(A, lda, below_start, next_outer), lda);
}

/* src/trsm/double_blocked_trsm.cpp:99-171, called with nb = cols and the
 * 1024 micro-block of double_blocked_potrf.cpp:17,92. */
static void FN(dbtrsm)(T* L, int ldl, T* B, int ldb, int n, int m, int nb, int b)
{
    if (n <= 0 || m <= 0) return;
    if (nb <= 0) nb = m;
    if (b <= 0) b = nb;
    if (b > nb) b = nb;
    for (int outer = 0; outer < m; outer += nb) {
        const int outer_size = (nb < m - outer) ? nb : m - outer;
        const int outer_end = outer + outer_size;
        for (int inner = outer; inner < outer_end; inner += b) {
            const int inner_size = (b < outer_end - inner) ? b : outer_end - inner;
            const int panel_start = inner + inner_size;
            FN(trsm_rlt)(n, inner_size, FN(bp)(L, ldl, inner, inner), ldl,
                         FN(bp)(B, ldb, 0, inner), ldb);
            if (panel_start < outer_end) {
                const int pc = (b < outer_end - panel_start) ? b : outer_end - panel_start;
                FN(gemm_nt)(n, pc, panel_start - outer, FN(bp)(B, ldb, 0, outer), ldb,
                            FN(bp)(L, ldl, panel_start, outer), ldl,
                            FN(bp)(B, ldb, 0, panel_start), ldb);
            }
        }
        const int trailing = m - outer_end;
        if (trailing > 0)
            FN(gemm_nt)(n, trailing, outer_size, FN(bp)(B, ldb, 0, outer), ldb,
                        FN(bp)(L, ldl, outer_end, outer), ldl,
                        FN(bp)(B, ldb, 0, outer_end), ldb);
    }
}

/* double_blocked_potrf.cpp:294-381.
 * info_ref receives the reference's semantics (the last diagonal call's
 * local info, overwritten per call); info_global the first failing global
 * column (1-based) or 0. */
void FN(oracle_double_blocked_potrf)(T* A, int lda, int n, int nb, int b,
                                     int use_dbtrsm, int syrk_k_tile,
                                     int* info_ref, int* info_global)
{
    if (info_ref) *info_ref = 0;
    if (info_global) *info_global = 0;
    if (n <= 0) return;
    if (nb <= 0) nb = n;
    if (b <= 0) b = nb;
    if (b > nb) b = nb;
    for (int outer = 0; outer < n; outer += nb) {
        const int outer_size = (nb < n - outer) ? nb : n - outer;
        const int outer_end = outer + outer_size;
        for (int inner = outer; inner < outer_end; inner += b) {
            const int inner_size = (b < outer_end - inner) ? b : outer_end - inner;
            const int panel_start = inner + inner_size;
            const int trailing_rows = n - panel_start;
            T* Aii = FN(bp)(A, lda, inner, inner);
            const int local = FN(potf2)(inner_size, Aii, lda);
            if (info_ref) *info_ref = local;
            if (local != 0 && info_global && *info_global == 0)
                *info_global = inner + local;
            if (use_dbtrsm) {
                const int tb = inner_size < 1024 ? inner_size : 1024;
                if (trailing_rows > 0)
                    FN(dbtrsm)(Aii, lda, FN(bp)(A, lda, panel_start, inner), lda,
                               trailing_rows, inner_size, inner_size, tb);
            } else {
                FN(trsm_rlt)(trailing_rows, inner_size, Aii, lda,
                             FN(bp)(A, lda, panel_start, inner), lda);
            }
            FN(potrf_inner_update)(A, lda, n, outer, panel_start, outer_end, b);
        }
        FN(potrf_outer_update)(A, lda, n, outer_end, nb, syrk_k_tile);
    }
}

/* ---------------------------------------------------------------- GETRF */

/* getrf_pivot_kernels.hip:14-58: swap rows panel+k <-> panel+ipiv[k]-1 in
 * LAPACK order over every column outside [panel, panel+b), then make the
 * panel pivots global. */
static void FN(apply_panel_pivots)(T* A, int lda, int n, int panel, int b, int* ipiv)
{
    for (int c = 0; c < n; ++c) {
        if (c >= panel && c < panel + b) continue;
        T* col = A + (size_t)c * lda;
        for (int k = 0; k < b; ++k) {
            const int r = panel + k, p = panel + ipiv[k] - 1;
            if (p != r) { T t = col[r]; col[r] = col[p]; col[p] = t; }
        }
    }
    for (int k = 0; k < b; ++k) ipiv[k] += panel;
}

/* getrf_pivot_internal.h:64-110 / getrf_batched_pivot_internal.h:225-283 */
static void FN(trsm_llu_iterative)(int rows, int cols, T* L, int ldl, T* B,
                                   int ldb, int block)
{
    if (rows <= 0 || cols <= 0) return;
    const int step = block > 0 ? block : rows;
    for (int row = 0; row < rows; row += step) {
        const int dr = (step < rows - row) ? step : rows - row;
        FN(trsm_llu)(dr, cols, FN(bp)(L, ldl, row, row), ldl, B + row, ldb);
        const int tr = rows - row - dr;
        if (tr > 0)
            FN(gemm_nn)(tr, cols, dr, FN(bp)(L, ldl, row + dr, row), ldl,
                        B + row, ldb, B + row + dr, ldb);
    }
}

/* double_blocked_getrf.cpp:32-71 */
static void FN(getrf_inner_update)(T* A, int lda, int n, int outer,
                                   int next_inner, int outer_end, int inner_nb)
{
    if (next_inner >= outer_end) return;
    const int next_size = (inner_nb < outer_end - next_inner) ? inner_nb : outer_end - next_inner;
    const int history = next_inner - outer;
    FN(trsm_llu)(history, next_size, FN(bp)(A, lda, outer, outer), lda,
                 FN(bp)(A, lda, outer, next_inner), lda);
    FN(gemm_nn)(n - next_inner, next_size, history,
                FN(bp)(A, lda, next_inner, outer), lda,
                FN(bp)(A, lda, outer, next_inner), lda,
                FN(bp)(A, lda, next_inner, next_inner), lda);
}

/* double_blocked_getrf.cpp:73-123 */
static void FN(getrf_outer_update)(T* A, int lda, int n, int next_outer,
                                   int outer_nb, int mode, int trsm_block)
{
    if (next_outer >= n) return;
    const int next_size = (outer_nb < n - next_outer) ? outer_nb : n - next_outer;
    if (mode == 0)
        FN(trsm_llu)(next_outer, next_size, A, lda, FN(bp)(A, lda, 0, next_outer), lda);
    else
        FN(trsm_llu_iterative)(next_outer, next_size, A, lda,
                               FN(bp)(A, lda, 0, next_outer), lda, trsm_block);
    FN(gemm_nn)(n - next_outer, next_size, next_outer,
                FN(bp)(A, lda, next_outer, 0), lda,
                FN(bp)(A, lda, 0, next_outer), lda,
                FN(bp)(A, lda, next_outer, next_outer), lda);
}

/* Row permutation move (gather then scatter, getrf_pivot_kernels.hip:107-206)
 * applied to `columns` columns starting at A: A[dst[m]] = old A[src[m]]. */
static void FN(apply_moves)(T* A, int lda, int columns, const int* src,
                            const int* dst, int moves, T* tmp)
{
    for (int c = 0; c < columns; ++c) {
        T* col = A + (size_t)c * lda;
        for (int m = 0; m < moves; ++m) tmp[m] = col[src[m]];
        for (int m = 0; m < moves; ++m) col[dst[m]] = tmp[m];
    }
}

/* getrf_batched_pivot_internal.h:59-75: rows whose source changed. */
static int FN(collect_moves)(const int* perm, int count, int* src, int* dst)
{
    int moves = 0;
    for (int r = 0; r < count; ++r)
        if (perm[r] != r) { src[moves] = perm[r]; dst[moves] = r; ++moves; }
    return moves;
}

/* double_blocked_getrf_impl (immediate pivots, batched == 0) and the
 * batched-pivot schedule (batched == 1).  ipiv is 1-based global.
 * Returns 0, or -1 when a pivot is invalid (reference sets info = -1). */
int FN(oracle_double_blocked_getrf)(T* A, int lda, int n, int outer_nb,
                                    int inner_nb, int mode, int trsm_block,
                                    int batched, int* ipiv, int* info_ref,
                                    int* info_global)
{
    if (info_ref) *info_ref = 0;
    if (info_global) *info_global = 0;
    if (n <= 0) return 0;
    if (outer_nb <= 0) outer_nb = n;
    if (inner_nb <= 0) inner_nb = outer_nb;
    if (inner_nb > outer_nb) inner_nb = outer_nb;

    int* perm = (int*)malloc((size_t)n * sizeof(int));
    int* src = (int*)malloc((size_t)n * sizeof(int));
    int* dst = (int*)malloc((size_t)n * sizeof(int));
    T* tmp = (T*)malloc((size_t)n * sizeof(T));
    int status = 0;

    for (int outer = 0; outer < n && status == 0; outer += outer_nb) {
        const int outer_size = (outer_nb < n - outer) ? outer_nb : n - outer;
        const int outer_end = outer + outer_size;
        for (int inner = outer; inner < outer_end; inner += inner_nb) {
            const int isz = (inner_nb < outer_end - inner) ? inner_nb : outer_end - inner;
            const int next_inner = inner + isz;
            const int local = FN(getf2)(n - inner, isz, FN(bp)(A, lda, inner, inner),
                                        lda, ipiv + inner);
            if (info_ref) *info_ref = local;
            if (local != 0 && info_global && *info_global == 0)
                *info_global = inner + local;
            if (!batched) {
                FN(apply_panel_pivots)(A, lda, n, inner, isz, ipiv + inner);
            } else {
                /* prepare_panel_permutation (:102-148) */
                const int active = n - inner;
                for (int r = 0; r < active; ++r) perm[r] = r;
                for (int k = 0; k < isz; ++k) {
                    const int lp = ipiv[inner + k] - 1;
                    if (lp < k || lp >= active) { status = -1; break; }
                    int t = perm[k]; perm[k] = perm[lp]; perm[lp] = t;
                    ipiv[inner + k] += inner;
                }
                if (status) break;
                const int moves = FN(collect_moves)(perm, active, src, dst);
                if (inner > outer)
                    FN(apply_moves)(FN(bp)(A, lda, inner, outer), lda, inner - outer,
                                    src, dst, moves, tmp);
                if (next_inner < outer_end)
                    FN(apply_moves)(FN(bp)(A, lda, inner, next_inner), lda,
                                    outer_end - next_inner, src, dst, moves, tmp);
            }
            FN(getrf_inner_update)(A, lda, n, outer, next_inner, outer_end, inner_nb);
        }
        if (status) break;
        if (batched) {
            /* Outer pivot batch (:556-609) */
            if (outer > 0) {
                const int count = n - outer;
                for (int r = 0; r < count; ++r) perm[r] = r;
                for (int row = outer; row < outer_end; ++row) {
                    const int t = row - outer, p = ipiv[row] - 1 - outer;
                    if (t < 0 || t >= count || p < t || p >= count) { status = -1; break; }
                    int s = perm[t]; perm[t] = perm[p]; perm[p] = s;
                }
                if (status) break;
                const int moves = FN(collect_moves)(perm, count, src, dst);
                FN(apply_moves)(FN(bp)(A, lda, outer, 0), lda, outer, src, dst, moves, tmp);
            }
            if (outer_end < n) {
                for (int r = 0; r < n; ++r) perm[r] = r;
                for (int row = 0; row < outer_end; ++row) {
                    const int p = ipiv[row] - 1;
                    if (p < row || p >= n) { status = -1; break; }
                    int s = perm[row]; perm[row] = perm[p]; perm[p] = s;
                }
                if (status) break;
                const int moves = FN(collect_moves)(perm, n, src, dst);
                const int cols = (outer_nb < n - outer_end) ? outer_nb : n - outer_end;
                FN(apply_moves)(FN(bp)(A, lda, 0, outer_end), lda, cols, src, dst, moves, tmp);
            }
        }
        FN(getrf_outer_update)(A, lda, n, outer_end, outer_nb, mode, trsm_block);
    }
    if (status && info_ref) *info_ref = -1;
    if (status && info_global) *info_global = -1;
    free(perm); free(src); free(dst); free(tmp);
    return status;
}

/* Exposed building blocks (used by the tests to check the GPU kernels). */
void FN(oracle_gemm_nt)(int rows, int cols, int k, const T* A, int lda,
                        const T* B, int ldb, T* C, int ldc)
{ FN(gemm_nt)(rows, cols, k, A, lda, B, ldb, C, ldc); }
void FN(oracle_gemm_nn)(int rows, int cols, int k, const T* A, int lda,
                        const T* B, int ldb, T* C, int ldc)
{ FN(gemm_nn)(rows, cols, k, A, lda, B, ldb, C, ldc); }
void FN(oracle_syrk_lower)(int n, int k, const T* A, int lda, T* C, int ldc, int k_tile)
{ FN(syrk_lower)(n, k, A, lda, C, ldc, k_tile); }
void FN(oracle_trsm_rlt)(int rows, int cols, const T* L, int ldl, T* B, int ldb)
{ FN(trsm_rlt)(rows, cols, L, ldl, B, ldb); }
void FN(oracle_trsm_llu)(int rows, int cols, const T* L, int ldl, T* B, int ldb)
{ FN(trsm_llu)(rows, cols, L, ldl, B, ldb); }
int FN(oracle_potf2)(int n, T* A, int lda) { return FN(potf2)(n, A, lda); }
int FN(oracle_getf2)(int m, int n, T* A, int lda, int* ipiv)
{ return FN(getf2)(m, n, A, lda, ipiv); }
int FN(oracle_iamax)(int n, const T* x) { return FN(iamax)(n, x); }
void FN(oracle_apply_panel_pivots)(T* A, int lda, int n, int panel, int b, int* ipiv)
{ FN(apply_panel_pivots)(A, lda, n, panel, b, ipiv); }

#undef FN
#undef CAT
#undef CAT_

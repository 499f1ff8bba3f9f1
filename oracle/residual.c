/*
 * oracle/residual.c -- CPU FP64 residual oracle (test infrastructure only).
 *
 * This file is part of the checker, never of the product: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Restates the reference's header-only CPU FP64 residual
 *   include/utils/benchmark/cpu_fp64_residual.h:28-201
 * in plain C:
 *   - SumOfSquares (LAPACK-style scaled sum of squares)      :29-50
 *   - multiply_accumulate (portable loop or CBLAS dgemm)      :52-71
 *   - full_residual (tiled FP64 product, zero-product skip)   :73-139
 *   - getrf_cpu_fp64_full (ipiv -> permutation, inf on bad)  :145-179
 *   - potrf_cpu_fp64_full                                     :181-201
 * and the ipiv -> row permutation of
 *   src/getrf/benchmark/getrf_benchmark_common.cpp:54-72.
 *
 * Result: ||P A - L U||_F / ||A||_F   or   ||A - L L^T||_F / ||A||_F,
 * with every stored value converted to double BEFORE multiplication.
 * Return codes: the residual (>= 0 or +inf); a negative value -1.0 means
 * "invalid argument" (the reference throws std::invalid_argument there).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

/* Optional BLAS backend: column-major C(m x n) += A(m x k) * B(k x n).
 * Installed by oracle_set_dgemm(); NULL selects the portable loop. */
typedef void (*oracle_dgemm_fn)(int m, int n, int k, const double* a,
                                const double* b, double* c);
static oracle_dgemm_fn g_dgemm = NULL;

void oracle_set_dgemm(oracle_dgemm_fn fn) { g_dgemm = fn; }

typedef struct {
    double scale;
    double sum;
    int finite;
} ssq_t;

/* cpu_fp64_residual.h:29-50 */
static void ssq_add(ssq_t* s, double value)
{
    const double magnitude = fabs(value);
    if (!isfinite(magnitude)) {
        s->finite = 0;
    } else if (magnitude != 0.0) {
        if (s->scale < magnitude) {
            const double ratio = s->scale / magnitude;
            s->sum = 1.0 + s->sum * ratio * ratio;
            s->scale = magnitude;
        } else {
            const double ratio = magnitude / s->scale;
            s->sum += ratio * ratio;
        }
    }
}

/* cpu_fp64_residual.h:52-71 (portable branch) */
static void mult_acc(int m, int n, int k, const double* left,
                     const double* right, double* product)
{
    if (g_dgemm != NULL) {
        g_dgemm(m, n, k, left, right, product);
        return;
    }
    for (int column = 0; column < n; ++column) {
        for (int inner = 0; inner < k; ++inner) {
            const double value = right[(size_t)column * k + inner];
            double* out = product + (size_t)column * m;
            const double* l = left + (size_t)inner * m;
            for (int row = 0; row < m; ++row) {
                out[row] += l[row] * value;
            }
        }
    }
}

/* Accessors for the three operands.  kind selects the factor layout. */
enum { KIND_LU = 0, KIND_LLT = 1 };

typedef struct {
    int kind;
    int is_double;
    const void* factor;
    const void* original;
    const int* permutation; /* NULL: identity */
    size_t lda;
} operands_t;

static double load(const void* base, int is_double, size_t index)
{
    return is_double ? ((const double*)base)[index]
                     : (double)((const float*)base)[index];
}

/* left(row, k): unit-lower L for LU, lower L for LL^T. */
static double left_value(const operands_t* o, int row, int k)
{
    if (row < k) return 0.0;
    if (o->kind == KIND_LU && row == k) return 1.0;
    return load(o->factor, o->is_double, (size_t)k * o->lda + row);
}

/* right(k, column): U for LU, L^T for LL^T. */
static double right_value(const operands_t* o, int k, int column)
{
    if (o->kind == KIND_LU) {
        return k > column ? 0.0
            : load(o->factor, o->is_double, (size_t)column * o->lda + k);
    }
    return column < k ? 0.0
        : load(o->factor, o->is_double, (size_t)k * o->lda + column);
}

static double original_value(const operands_t* o, int row, int column)
{
    const int source = o->permutation ? o->permutation[row] : row;
    return load(o->original, o->is_double,
                (size_t)column * o->lda + (size_t)source);
}

/* cpu_fp64_residual.h:73-139 */
static double full_residual(int n, int tile, const operands_t* o)
{
    if (n <= 0 || tile <= 0) return -1.0;
    if (tile > n) tile = n;
    const size_t capacity = (size_t)tile * tile;
    double* left = (double*)malloc(capacity * sizeof(double));
    double* right = (double*)malloc(capacity * sizeof(double));
    double* product = (double*)malloc(capacity * sizeof(double));
    if (!left || !right || !product) {
        free(left); free(right); free(product);
        return -1.0;
    }
    ssq_t residual = {0.0, 1.0, 1};
    ssq_t norm = {0.0, 1.0, 1};

    for (int column = 0; column < n;) {
        const int columns = (tile < n - column) ? tile : n - column;
        for (int row = 0; row < n;) {
            const int rows = (tile < n - row) ? tile : n - row;
            memset(product, 0, (size_t)rows * columns * sizeof(double));
            /* Both LU and LL^T have zero products for k > min(row, column). */
            const int a = row + rows, b = column + columns;
            const int reduction_end = a < b ? a : b;
            for (int inner = 0; inner < reduction_end;) {
                const int count = (tile < reduction_end - inner)
                    ? tile : reduction_end - inner;
                for (int k = 0; k < count; ++k)
                    for (int i = 0; i < rows; ++i)
                        left[(size_t)k * rows + i] = left_value(o, row + i, inner + k);
                for (int j = 0; j < columns; ++j)
                    for (int k = 0; k < count; ++k)
                        right[(size_t)j * count + k] = right_value(o, inner + k, column + j);
                mult_acc(rows, columns, count, left, right, product);
                inner += count;
            }
            for (int j = 0; j < columns; ++j) {
                for (int i = 0; i < rows; ++i) {
                    const double orig = original_value(o, row + i, column + j);
                    ssq_add(&norm, orig);
                    ssq_add(&residual, product[(size_t)j * rows + i] - orig);
                }
            }
            row += rows;
        }
        column += columns;
    }
    free(left); free(right); free(product);

    if (!residual.finite || !norm.finite) return INFINITY;
    if (norm.scale == 0.0) return residual.scale == 0.0 ? 0.0 : INFINITY;
    return (residual.scale / norm.scale) * sqrt(residual.sum / norm.sum);
}

/* getrf_benchmark_common.cpp:54-72 and cpu_fp64_residual.h:154-164:
 * ipiv (1-based, global) -> permutation; returns 0 when a pivot is invalid. */
int oracle_ipiv_to_permutation(const int* ipiv, int n, int* permutation)
{
    for (int i = 0; i < n; ++i) permutation[i] = i;
    for (int row = 0; row < n; ++row) {
        if (ipiv[row] < row + 1 || ipiv[row] > n) return 0;
        const int p = ipiv[row] - 1;
        const int t = permutation[row];
        permutation[row] = permutation[p];
        permutation[p] = t;
    }
    return 1;
}

static double getrf_residual(const void* lu, const void* a, const int* ipiv,
                             int n, int lda, int pivoting, int tile,
                             int is_double)
{
    if (n <= 0 || lda < n || tile <= 0 || lu == NULL || a == NULL
        || (pivoting && ipiv == NULL))
        return -1.0;
    int* permutation = NULL;
    if (pivoting) {
        permutation = (int*)malloc((size_t)n * sizeof(int));
        if (!permutation) return -1.0;
        if (!oracle_ipiv_to_permutation(ipiv, n, permutation)) {
            free(permutation);
            return INFINITY;
        }
    }
    operands_t o = {KIND_LU, is_double, lu, a, permutation, (size_t)lda};
    const double r = full_residual(n, tile, &o);
    free(permutation);
    return r;
}

static double potrf_residual(const void* l, const void* a, int n, int lda,
                             int tile, int is_double)
{
    if (n <= 0 || lda < n || tile <= 0 || l == NULL || a == NULL)
        return -1.0;
    operands_t o = {KIND_LLT, is_double, l, a, NULL, (size_t)lda};
    return full_residual(n, tile, &o);
}

/* cpu_fp64_residual.h:145-179 */
double oracle_getrf_residual_f32(const float* lu, const float* a,
                                 const int* ipiv, int n, int lda,
                                 int pivoting, int tile)
{
    return getrf_residual(lu, a, ipiv, n, lda, pivoting, tile, 0);
}

double oracle_getrf_residual_f64(const double* lu, const double* a,
                                 const int* ipiv, int n, int lda,
                                 int pivoting, int tile)
{
    return getrf_residual(lu, a, ipiv, n, lda, pivoting, tile, 1);
}

/* cpu_fp64_residual.h:181-201 */
double oracle_potrf_residual_f32(const float* l, const float* a, int n,
                                 int lda, int tile)
{
    return potrf_residual(l, a, n, lda, tile, 0);
}

double oracle_potrf_residual_f64(const double* l, const double* a, int n,
                                 int lda, int tile)
{
    return potrf_residual(l, a, n, lda, tile, 1);
}

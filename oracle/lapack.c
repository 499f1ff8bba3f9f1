/*
 * oracle/lapack.c -- run-time LAPACK/CBLAS provider for the CPU baseline and
 * the FP64 cross-checks (test infrastructure only).
 *
 * The provider is dlopen()ed, in this order:
 *   1. $MIFACT_LAPACK_LIB
 *   2. a system libopenblas / liblapack
 *   3. SciPy's bundled libscipy_openblas (symbols carry a "scipy_" prefix)
 *   4. MKL libmkl_rt
 * Exposed routines: spotrf, sgetrf, dpotrf, dpotrs, dgetrf, dgetrs, plus the
 * CBLAS dgemm that can back the FP64 residual (oracle_set_dgemm).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef void (*potrf_s_t)(const char*, const int*, float*, const int*, int*, size_t);
typedef void (*potrf_d_t)(const char*, const int*, double*, const int*, int*, size_t);
typedef void (*getrf_s_t)(const int*, const int*, float*, const int*, int*, int*);
typedef void (*getrf_d_t)(const int*, const int*, double*, const int*, int*, int*);
typedef void (*potrs_d_t)(const char*, const int*, const int*, const double*,
                          const int*, double*, const int*, int*, size_t);
typedef void (*getrs_d_t)(const char*, const int*, const int*, const double*,
                          const int*, const int*, double*, const int*, int*, size_t);
typedef void (*cblas_dgemm_t)(int, int, int, int, int, int, double, const double*,
                              int, const double*, int, double, double*, int);
typedef void (*set_threads_t)(int);
typedef int (*get_threads_t)(void);
typedef char* (*get_config_t)(void);

static void* g_handle = NULL;
static char g_path[1024];
static potrf_s_t p_spotrf;
static getrf_s_t p_sgetrf;
static potrf_d_t p_dpotrf;
static getrf_d_t p_dgetrf;
static potrs_d_t p_dpotrs;
static getrs_d_t p_dgetrs;
static cblas_dgemm_t p_dgemm;
static set_threads_t p_set_threads;
static get_threads_t p_get_threads;
static get_config_t p_get_config;

static void* sym(const char* prefix, const char* name)
{
    char buf[128];
    snprintf(buf, sizeof buf, "%s%s", prefix, name);
    return dlsym(g_handle, buf);
}

static int bind(const char* prefix)
{
    p_spotrf = (potrf_s_t)sym(prefix, "spotrf_");
    p_sgetrf = (getrf_s_t)sym(prefix, "sgetrf_");
    p_dpotrf = (potrf_d_t)sym(prefix, "dpotrf_");
    p_dgetrf = (getrf_d_t)sym(prefix, "dgetrf_");
    p_dpotrs = (potrs_d_t)sym(prefix, "dpotrs_");
    p_dgetrs = (getrs_d_t)sym(prefix, "dgetrs_");
    p_dgemm = (cblas_dgemm_t)sym(prefix, "cblas_dgemm");
    p_set_threads = (set_threads_t)sym(prefix, "openblas_set_num_threads");
    p_get_threads = (get_threads_t)sym(prefix, "openblas_get_num_threads");
    p_get_config = (get_config_t)sym(prefix, "openblas_get_config");
    if (!p_set_threads) p_set_threads = (set_threads_t)sym("", "MKL_Set_Num_Threads");
    if (!p_get_threads) p_get_threads = (get_threads_t)sym("", "MKL_Get_Max_Threads");
    return p_spotrf && p_sgetrf && p_dpotrf && p_dgetrf && p_dpotrs && p_dgetrs;
}

static int try_open(const char* path)
{
    if (!path || !*path) return 0;
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return 0;
    g_handle = h;
    if (bind("") || bind("scipy_")) {
        snprintf(g_path, sizeof g_path, "%s", path);
        return 1;
    }
    dlclose(h);
    g_handle = NULL;
    return 0;
}

/* scipy_hint: path of SciPy's bundled OpenBLAS (found by the Python side). */
int oracle_lapack_open(const char* scipy_hint)
{
    if (g_handle) return 0;
    if (try_open(getenv("MIFACT_LAPACK_LIB"))) return 0;
    const char* system_libs[] = {"libopenblas.so.0", "libopenblas.so",
                                 "liblapack.so.3", "liblapack.so", NULL};
    for (int i = 0; system_libs[i]; ++i)
        if (try_open(system_libs[i])) return 0;
    if (try_open(scipy_hint)) return 0;
    if (try_open("libmkl_rt.so") || try_open("libmkl_rt.so.2")) return 0;
    return -1;
}

const char* oracle_lapack_path(void) { return g_handle ? g_path : ""; }

const char* oracle_lapack_config(void)
{
    return (g_handle && p_get_config) ? p_get_config() : "";
}

int oracle_lapack_set_threads(int threads)
{
    if (!p_set_threads) return -1;
    p_set_threads(threads);
    return 0;
}

int oracle_lapack_get_threads(void) { return p_get_threads ? p_get_threads() : -1; }

int oracle_spotrf(int n, float* a, int lda)
{
    int info = 0;
    p_spotrf("L", &n, a, &lda, &info, 1);
    return info;
}

int oracle_sgetrf(int m, int n, float* a, int lda, int* ipiv)
{
    int info = 0;
    p_sgetrf(&m, &n, a, &lda, ipiv, &info);
    return info;
}

int oracle_dpotrf(int n, double* a, int lda)
{
    int info = 0;
    p_dpotrf("L", &n, a, &lda, &info, 1);
    return info;
}

int oracle_dgetrf(int m, int n, double* a, int lda, int* ipiv)
{
    int info = 0;
    p_dgetrf(&m, &n, a, &lda, ipiv, &info);
    return info;
}

int oracle_dpotrs(int n, int nrhs, const double* a, int lda, double* b, int ldb)
{
    int info = 0;
    p_dpotrs("L", &n, &nrhs, a, &lda, b, &ldb, &info, 1);
    return info;
}

int oracle_dgetrs(int n, int nrhs, const double* a, int lda, const int* ipiv,
                  double* b, int ldb)
{
    int info = 0;
    p_dgetrs("N", &n, &nrhs, a, &lda, ipiv, b, &ldb, &info, 1);
    return info;
}

/* Column-major C(m x n) += A(m x k) B(k x n) through CBLAS. */
static void blas_dgemm(int m, int n, int k, const double* a, const double* b, double* c)
{
    /* CblasColMajor = 102, CblasNoTrans = 111 */
    p_dgemm(102, 111, 111, m, n, k, 1.0, a, m, b, k, 1.0, c, m);
}

typedef void (*oracle_dgemm_fn)(int, int, int, const double*, const double*, double*);
void oracle_set_dgemm(oracle_dgemm_fn fn);

int oracle_residual_use_blas(int enable)
{
    if (!enable) { oracle_set_dgemm(NULL); return 0; }
    if (!p_dgemm) return -1;
    oracle_set_dgemm(blas_dgemm);
    return 0;
}

// oracle/ref_residual_shim.cpp -- C entry points over the reference's own
// header-only CPU FP64 residual (include/utils/benchmark/cpu_fp64_residual.h
// in the reference tree), compiled in place from the reference sources by
// oracle/Makefile into oracle/_ref/.  Used only by tests to pin the oracle's
// restatement (oracle/residual.c).  std::invalid_argument maps to -1.
#include <stdexcept>

#include "utils/benchmark/cpu_fp64_residual.h"

extern "C" {

double ref_potrf_residual_f32(const float* l, const float* a, int n, int lda, int tile)
{
    try { return benchmark_residual::potrf_cpu_fp64_full(l, a, n, lda, tile); }
    catch (const std::invalid_argument&) { return -1.0; }
}

double ref_potrf_residual_f64(const double* l, const double* a, int n, int lda, int tile)
{
    try { return benchmark_residual::potrf_cpu_fp64_full(l, a, n, lda, tile); }
    catch (const std::invalid_argument&) { return -1.0; }
}

double ref_getrf_residual_f32(const float* lu, const float* a, const int* ipiv,
                              int n, int lda, int pivoting, int tile)
{
    try { return benchmark_residual::getrf_cpu_fp64_full(lu, a, ipiv, n, lda, pivoting != 0, tile); }
    catch (const std::invalid_argument&) { return -1.0; }
}

double ref_getrf_residual_f64(const double* lu, const double* a, const int* ipiv,
                              int n, int lda, int pivoting, int tile)
{
    try { return benchmark_residual::getrf_cpu_fp64_full(lu, a, ipiv, n, lda, pivoting != 0, tile); }
    catch (const std::invalid_argument&) { return -1.0; }
}

}  // extern "C"

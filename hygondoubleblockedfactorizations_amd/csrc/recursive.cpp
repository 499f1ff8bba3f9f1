// Recursive POTRF: the reference's baseline Cholesky family
// (src/potrf/recursive_potrf.cpp:17-388, include/potrf/recursive_potrf.h).
//
// Restated schedule (the algorithm being compared against): split the order
// in halves, factor the leading half, solve the panel below it
// (B := B L11^-T, recursive in the column count with a GEMM update between
// halves, src/trsm/recursive_trsm.cpp:132-161), update the trailing block
// with the panel (SYRK; the reference's bottom-up batched SYRK,
// src/syrk/batched_syrk.cpp:27-144, is one lower-trapezoid MFMA launch here),
// factor the trailing half; orders <= nb are leaves (rocsolver_spotrf in the
// reference, the recursive 64-wide diagonal factorization here).
//
// The mixed driver (recursive_potrf.cpp:241-388) scales the lower triangle
// by SCALE_TARGET / max|diag(A)| (SCALE_TARGET = 65504, the FP16 maximum,
// include/kernels/matrix_scale.h:4; src/kernels/matrix_scale.hip:20-90),
// runs the recursion with the panel-update and trailing-update GEMMs in
// FP16 (FP32 accumulation), retries a leaf that breaks down with a diagonal
// damping lambda starting at SCALE_TARGET * 1e-4 and doubling, up to 5
// times (:250-313), and scales the factor back by sqrt(max / SCALE_TARGET).
// Deliberate difference: the reference's mixed GEMM rounds the accumulated
// matrix C to FP16 as well (rocblas_gemm_ex with FP16 C/D,
// src/gemm/mixed_gemm.cpp:61-90); here C stays FP32.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "blas3.hpp"
#include "potrf/recursive_potrf.h"

namespace mifact {
namespace {

constexpr float kScaleTarget = 65504.0f;  // include/kernels/matrix_scale.h:4
constexpr int kDampingAttempts = 5;       // recursive_potrf.cpp:283

// B(n x m) := B L^-T with L lower non-unit m x m, recursive in m: leaves of
// m <= nb columns are solved with the inverse-based TRSM, the update between
// halves runs in `prec` (src/trsm/recursive_trsm.cpp:132-161, 285-364).
template <typename T>
void recursive_trsm(mifact_handle h, const T* L, long long ldl, T* B, long long ldb, int n, int m,
                    int nb, LowPrecision prec)
{
    if (n <= 0 || m <= 0) return;
    if (m <= nb) {
        trsm_rlt<T>(h, n, m, L, ldl, B, ldb);
        return;
    }
    const int m1 = m / 2;
    const int m2 = m - m1;
    recursive_trsm<T>(h, L, ldl, B, ldb, n, m1, nb, prec);
    // B2 -= B1 L21^T
    if constexpr (std::is_same_v<T, float>) {
        gemm_lowp(h, prec, kOpNT, kShapeFull, n, m2, m1, -1.0f, B, ldb, L + m1, ldl,
                  B + static_cast<size_t>(m1) * ldb, ldb);
    } else {
        gemm<T>(h, kOpNT, kShapeFull, n, m2, m1, T(-1), B, ldb, L + m1, ldl,
                B + static_cast<size_t>(m1) * ldb, ldb);
    }
    recursive_trsm<T>(h, L + static_cast<size_t>(m1) * ldl + m1, ldl,
                      B + static_cast<size_t>(m1) * ldb, ldb, n, m2, nb, prec);
}

// C_lower(n x n) -= A A^T (A n x k) in `prec`.
template <typename T>
void syrk_update(mifact_handle h, const T* A, long long lda, T* C, long long ldc, int n, int k,
                 LowPrecision prec)
{
    if constexpr (std::is_same_v<T, float>) {
        gemm_lowp(h, prec, kOpNT, kShapeLower, n, n, k, -1.0f, A, lda, A, lda, C, ldc);
    } else {
        gemm<T>(h, kOpNT, kShapeLower, n, n, k, T(-1), A, lda, A, lda, C, ldc);
    }
}

__global__ void copy_lower_kernel(const float* __restrict__ src, long long lds, float* dst,
                                  long long ldd, int n)
{
    const int col = blockIdx.x;
    for (int r = col + threadIdx.x; r < n; r += blockDim.x)
        dst[off(r, col, ldd)] = src[off(r, col, lds)];
}

__global__ void add_diagonal_kernel(float* A, long long lda, int n, float lambda)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[off(i, i, lda)] += lambda;
}

__global__ void clear_uint_kernel(unsigned int* p) { *p = 0u; }

// *info = *leaf when *info is still 0 (first failing column wins)
__global__ void keep_first_kernel(int* info, const int* leaf)
{
    if (*info == 0) *info = *leaf;
}

// max |A(i, i)| as float bits (non-negative floats order like their bits)
__global__ void diag_absmax_kernel(const float* __restrict__ A, long long lda, int n,
                                   unsigned int* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicMax(out, __float_as_uint(fabsf(A[off(i, i, lda)])));
}

// Lower triangle *= s, s = SCALE_TARGET / max (forward) or sqrt(max /
// SCALE_TARGET) (back); max <= 0 leaves A unchanged (matrix_scale.hip:40-70).
__global__ void scale_lower_kernel(float* A, long long lda, int n,
                                   const unsigned int* __restrict__ maxbits, int inverse)
{
    const float mx = __uint_as_float(*maxbits);
    float s = 1.0f;
    if (mx > 0.0f) s = inverse ? sqrtf(mx / kScaleTarget) : kScaleTarget / mx;
    const int col = blockIdx.x;
    for (int r = col + threadIdx.x; r < n; r += blockDim.x) A[off(r, col, lda)] *= s;
}

struct MixedLeaf {
    float* backup;      // nb x nb leaf copy
    int* leaf_info;     // device int of one leaf attempt
    int* host_info;     // pinned read-back
    float lambda;       // first damping value (scaled units)
};

// Leaf with the mixed driver's damping retries: factor; when it breaks
// down, restore the leaf, add lambda to its diagonal and retry, doubling
// lambda (recursive_potrf.cpp:250-313; host-synchronizing, as there).
void mixed_leaf(mifact_handle h, float* A, long long lda, int n, int* info, int col0,
                const MixedLeaf& ml)
{
    float* inv = invdiag<float>(h);
    hipLaunchKernelGGL(copy_lower_kernel, dim3(n), dim3(256), 0, h->stream, A, lda, ml.backup,
                       static_cast<long long>(n), n);
    MIFACT_LAUNCHED(h);
    float lambda = ml.lambda;
    for (int attempt = 0; attempt <= kDampingAttempts; ++attempt) {
        if (attempt > 0) {
            hipLaunchKernelGGL(copy_lower_kernel, dim3(n), dim3(256), 0, h->stream, ml.backup,
                               static_cast<long long>(n), A, lda, n);
            MIFACT_LAUNCHED(h);
            hipLaunchKernelGGL(add_diagonal_kernel, dim3(ceil_div(n, 256)), dim3(256), 0,
                               h->stream, A, lda, n, lambda);
            MIFACT_LAUNCHED(h);
            lambda *= 2.0f;
        }
        clear_int(h, ml.leaf_info);
        potrf_diag<float>(h, n, A, lda, ml.leaf_info, col0, inv);
        MIFACT_HIP(h, hipMemcpyAsync(ml.host_info, ml.leaf_info, sizeof(int),
                                     hipMemcpyDeviceToHost, h->stream));
        MIFACT_HIP(h, hipStreamSynchronize(h->stream));
        if (*ml.host_info == 0) return;
    }
    // still not positive definite after the last damping step: report it
    // (the first failing column of the whole factorization wins)
    if (info != nullptr) {
        hipLaunchKernelGGL(keep_first_kernel, dim3(1), dim3(1), 0, h->stream, info, ml.leaf_info);
        MIFACT_LAUNCHED(h);
    }
}

template <typename T>
void recursive_potrf_rec(mifact_handle h, T* A, long long lda, int n, int nb, int* info, int col0,
                         LowPrecision prec, const MixedLeaf* ml)
{
    if (n <= 0) return;
    if (n <= nb) {
        if constexpr (std::is_same_v<T, float>) {
            if (ml != nullptr) {
                mixed_leaf(h, A, lda, n, info, col0, *ml);
                return;
            }
        }
        potrf_diag<T>(h, n, A, lda, info, col0, invdiag<T>(h));
        return;
    }
    const int n1 = n / 2;
    const int n2 = n - n1;
    T* A21 = A + n1;
    T* A22 = A + static_cast<size_t>(n1) * lda + n1;
    recursive_potrf_rec<T>(h, A, lda, n1, nb, info, col0, prec, ml);
    recursive_trsm<T>(h, A, lda, A21, lda, n2, n1, nb, prec);
    syrk_update<T>(h, A21, lda, A22, lda, n2, n1, prec);
    recursive_potrf_rec<T>(h, A22, lda, n2, nb, info, col0 + n1, prec, ml);
}

template <typename T>
void recursive_potrf_entry(mifact_handle h, T* A, int lda, int n, int nb, int* info,
                           const char* name)
{
    if (h == nullptr) {
        report_status(nullptr, mifact_status_invalid_handle, name);
        return;
    }
    if (n <= 0) return;
    if (nb <= 0) nb = n;
    if (!ensure_workspace(h, n)) return;
    if (info != nullptr) clear_int(h, info);
    recursive_potrf_rec<T>(h, A, lda, n, nb, info, 0, kPrecFP32, nullptr);
}

void recursive_potrf_mixed_entry(mifact_handle h, float* A, int lda_i, int n, int nb, int* info)
{
    if (h == nullptr) {
        report_status(nullptr, mifact_status_invalid_handle, "recursive_potrf_mixed_driver");
        return;
    }
    if (n <= 0) return;
    if (nb <= 0) nb = n;
    const long long lda = lda_i;
    // leaf backup (nb x nb) and the scale word come from the workspace's
    // staging area; the workspace must hold an order of at least max(n, nb)
    if (!ensure_workspace(h, n)) return;
    Workspace ws = workspace(h);
    const size_t leaf_bytes = static_cast<size_t>(nb) * nb * sizeof(float);
    if (leaf_bytes + 256 > ws.stage_bytes) {
        // the staging area holds 2n x 16 doubles: enough for leaves up to
        // nb = sqrt(64 n); larger leaves get a workspace of higher order
        const double need = static_cast<double>(leaf_bytes + 256) / (2.0 * 16 * sizeof(double));
        if (!ensure_workspace(h, static_cast<int>(need) + 1)) return;
        ws = workspace(h);
    }
    if (h->host_norms == nullptr || h->host_norms_bytes < 64) {
        if (h->host_norms != nullptr) (void)hipHostFree(h->host_norms);
        h->host_norms = nullptr;
        if (hipHostMalloc(&h->host_norms, 4096) != hipSuccess) {
            report_status(h, mifact_status_allocation_failed, "pinned host buffer failed");
            return;
        }
        h->host_norms_bytes = 4096;
    }
    MixedLeaf ml;
    ml.backup = static_cast<float*>(ws.stage);
    unsigned int* maxbits = reinterpret_cast<unsigned int*>(static_cast<char*>(ws.stage)
                                                            + ws.stage_bytes - 256);
    ml.leaf_info = reinterpret_cast<int*>(maxbits) + 16;
    ml.host_info = static_cast<int*>(h->host_norms);
    ml.lambda = kScaleTarget * 1e-4f;   // recursive_potrf.cpp:375
    if (info != nullptr) clear_int(h, info);
    hipLaunchKernelGGL(clear_uint_kernel, dim3(1), dim3(1), 0, h->stream, maxbits);
    MIFACT_LAUNCHED(h);
    hipLaunchKernelGGL(diag_absmax_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, h->stream, A,
                       lda, n, maxbits);
    MIFACT_LAUNCHED(h);
    hipLaunchKernelGGL(scale_lower_kernel, dim3(n), dim3(256), 0, h->stream, A, lda, n, maxbits,
                       0);
    MIFACT_LAUNCHED(h);
    recursive_potrf_rec<float>(h, A, lda, n, nb, info, 0, kPrecFP16, &ml);
    hipLaunchKernelGGL(scale_lower_kernel, dim3(n), dim3(256), 0, h->stream, A, lda, n, maxbits,
                       1);
    MIFACT_LAUNCHED(h);
}

}  // namespace
}  // namespace mifact

extern "C" {

void recursive_potrf(mifact_handle handle, void* pool, float* A, int lda, int n, int nb,
                     int* info)
{
    (void)pool;
    mifact::recursive_potrf_entry<float>(handle, A, lda, n, nb, info, "recursive_potrf");
}

void recursive_potrf_tiled_trsm(mifact_handle handle, void* pool, float* A, int lda, int n,
                                int nb, int* info)
{
    (void)pool;
    mifact::recursive_potrf_entry<float>(handle, A, lda, n, nb, info,
                                         "recursive_potrf_tiled_trsm");
}

void recursive_potrf_double(mifact_handle handle, void* pool, double* A, int lda, int n, int nb,
                            int* info)
{
    (void)pool;
    mifact::recursive_potrf_entry<double>(handle, A, lda, n, nb, info, "recursive_potrf_double");
}

void recursive_potrf_mixed_driver(mifact_handle handle, void* pool, float* A, int lda, int n,
                                  int nb, int* info)
{
    (void)pool;
    mifact::recursive_potrf_mixed_entry(handle, A, lda, n, nb, info);
}

}  // extern "C"

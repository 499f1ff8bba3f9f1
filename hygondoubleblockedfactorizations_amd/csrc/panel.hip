// Panel kernels for gfx950: diagonal POTF2, register-blocked triangular
// solves, wave-reduction IAMAX.
//
// These are the leaves of the recursive diagonal factorization and of the
// recursive triangular solves; everything above the leaves is MFMA GEMM.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mifact {

constexpr int kLeaf = 64;  // leaf order of POTF2 / TRSM kernels

// ------------------------------------------------------------------ TRSM RLT
// X L^T = B, one thread per row of B: x[] lives in registers, L in LDS (all
// lanes read the same L entry: broadcast).  Coalesced: a column of B is
// contiguous and consecutive threads own consecutive rows.
template <typename T>
__global__ __launch_bounds__(64) void trsm_rlt_kernel(int rows, int s, const T* __restrict__ L,
                                                      long long ldl, T* __restrict__ B,
                                                      long long ldb)
{
    __shared__ T sL[kLeaf][kLeaf + 1];
    __shared__ T sD[kLeaf];
    const int tid = threadIdx.x;
    for (int idx = tid; idx < s * s; idx += 64) {
        const int i = idx % s, j = idx / s;
        if (i >= j) sL[i][j] = L[off(i, j, ldl)];
    }
    __syncthreads();
    if (tid < s) sD[tid] = T(1) / sL[tid][tid];
    __syncthreads();
    const int row = blockIdx.x * 64 + tid;
    if (row >= rows) return;
    T x[kLeaf];
#pragma unroll
    for (int j = 0; j < kLeaf; ++j) x[j] = (j < s) ? B[off(row, j, ldb)] : T(0);
#pragma unroll
    for (int j = 0; j < kLeaf; ++j) {
        if (j < s) {
            T acc = x[j];
#pragma unroll
            for (int k = 0; k < j; ++k) acc -= x[k] * sL[j][k];
            x[j] = acc * sD[j];
        }
    }
#pragma unroll
    for (int j = 0; j < kLeaf; ++j)
        if (j < s) B[off(row, j, ldb)] = x[j];
}

template <typename T>
void trsm_rlt_base(mifact_handle h, int rows, int s, const T* L, long long ldl,
                   T* B, long long ldb)
{
    if (rows <= 0 || s <= 0) return;
    if (s > kLeaf) {
        report_status(h, mifact_status_invalid_value, "trsm_rlt_base: s > 64");
        return;
    }
    StageScope scope(h, kStageTrsmLeaf, static_cast<double>(rows) * s * s,
                     sizeof(T) * (2.0 * rows * s + 0.5 * s * s));
    hipLaunchKernelGGL(trsm_rlt_kernel<T>, dim3(ceil_div(rows, 64)), dim3(64), 0,
                       h->stream, rows, s, L, ldl, B, ldb);
    MIFACT_LAUNCHED(h);
}

// ------------------------------------------------------------------ TRSM LLU
// L X = B (L unit lower), one thread per column of B; the B tile is staged
// through LDS so the global reads and writes are coalesced along columns.
template <typename T>
__global__ __launch_bounds__(64) void trsm_llu_kernel(int s, int cols, const T* __restrict__ L,
                                                      long long ldl, T* __restrict__ B,
                                                      long long ldb)
{
    __shared__ T sL[kLeaf][kLeaf + 1];
    __shared__ T sB[kLeaf][64 + 1];
    const int tid = threadIdx.x;
    const int c0 = blockIdx.x * 64;
    for (int idx = tid; idx < s * s; idx += 64) {
        const int i = idx % s, j = idx / s;
        if (i > j) sL[i][j] = L[off(i, j, ldl)];
    }
    const int nc = (cols - c0) < 64 ? (cols - c0) : 64;
    for (int c = 0; c < nc; ++c)
        for (int r = tid; r < s; r += 64) sB[r][c] = B[off(r, c0 + c, ldb)];
    __syncthreads();
    if (tid < nc) {
        T x[kLeaf];
#pragma unroll
        for (int i = 0; i < kLeaf; ++i) x[i] = (i < s) ? sB[i][tid] : T(0);
#pragma unroll
        for (int i = 0; i < kLeaf; ++i) {
            if (i < s) {
                T acc = x[i];
#pragma unroll
                for (int k = 0; k < i; ++k) acc -= sL[i][k] * x[k];
                x[i] = acc;
            }
        }
#pragma unroll
        for (int i = 0; i < kLeaf; ++i)
            if (i < s) sB[i][tid] = x[i];
    }
    __syncthreads();
    for (int c = 0; c < nc; ++c)
        for (int r = tid; r < s; r += 64) B[off(r, c0 + c, ldb)] = sB[r][c];
}

template <typename T>
void trsm_llu_base(mifact_handle h, int s, int cols, const T* L, long long ldl,
                   T* B, long long ldb)
{
    if (s <= 0 || cols <= 0) return;
    if (s > kLeaf) {
        report_status(h, mifact_status_invalid_value, "trsm_llu_base: s > 64");
        return;
    }
    StageScope scope(h, kStageTrsmLeaf, static_cast<double>(cols) * s * s,
                     sizeof(T) * (2.0 * cols * s + 0.5 * s * s));
    hipLaunchKernelGGL(trsm_llu_kernel<T>, dim3(ceil_div(cols, 64)), dim3(64), 0,
                       h->stream, s, cols, L, ldl, B, ldb);
    MIFACT_LAUNCHED(h);
}

// ------------------------------------------------------------------ POTF2
// Right-looking unblocked Cholesky of an s x s (s <= 64) lower block held in
// LDS by one 256-thread workgroup.  Reciprocal scaling of each column, as
// LAPACK xPOTF2 does.
template <typename T>
__global__ __launch_bounds__(256) void potf2_kernel(int s, T* __restrict__ A, long long lda,
                                                    int* __restrict__ info, int col_offset)
{
    __shared__ T sA[kLeaf][kLeaf + 1];
    __shared__ int failed;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < s * s; idx += 256) {
        const int i = idx % s, j = idx / s;
        if (i >= j) sA[i][j] = A[off(i, j, lda)];
    }
    if (tid == 0) failed = 0;
    __syncthreads();
    int jend = s;
    for (int j = 0; j < s; ++j) {
        if (tid == 0) {
            const T d = sA[j][j];
            if (!(d > T(0))) {
                failed = 1;
                if (info != nullptr) atomicCAS(info, 0, col_offset + j + 1);
            } else {
                sA[j][j] = sqrt(d);
            }
        }
        __syncthreads();
        if (failed) {
            jend = j;
            break;
        }
        const T r = T(1) / sA[j][j];
        for (int i = j + 1 + tid; i < s; i += 256) sA[i][j] *= r;
        __syncthreads();
        // trailing lower update: pairs (i, k) with j < k <= i < s
        const int t = s - j - 1;
        for (int idx = tid; idx < t * t; idx += 256) {
            const int i = j + 1 + idx % t;
            const int k = j + 1 + idx / t;
            if (i >= k) sA[i][k] -= sA[i][j] * sA[k][j];
        }
        __syncthreads();
    }
    (void)jend;
    __syncthreads();
    for (int idx = tid; idx < s * s; idx += 256) {
        const int i = idx % s, j = idx / s;
        if (i >= j) A[off(i, j, lda)] = sA[i][j];
    }
}

template <typename T>
void potf2_base(mifact_handle h, int s, T* A, long long lda, int* info, int col_offset)
{
    if (s <= 0) return;
    if (s > kLeaf) {
        report_status(h, mifact_status_invalid_value, "potf2_base: s > 64");
        return;
    }
    StageScope scope(h, kStagePotf2, static_cast<double>(s) * s * s / 3.0,
                     sizeof(T) * static_cast<double>(s) * s);
    hipLaunchKernelGGL(potf2_kernel<T>, dim3(1), dim3(256), 0, h->stream, s, A, lda,
                       info, col_offset);
    MIFACT_LAUNCHED(h);
}

// ------------------------------------------------------------------ IAMAX
// Single-workgroup IAMAX with wave64 shuffles: max |x|, smallest index on
// exact ties (BLAS IxAMAX returns the first occurrence).  NaN entries never
// win a comparison.
template <typename T>
__device__ inline void amax_combine(T& v, int& i, T ov, int oi)
{
    if (ov > v || (ov == v && oi < i)) {
        v = ov;
        i = oi;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void iamax_kernel(int n, const T* __restrict__ x,
                                                    int incx, int* __restrict__ result)
{
    __shared__ T sv[4];
    __shared__ int si[4];
    const int tid = threadIdx.x;
    T best = T(-1);
    int bi = 0x7fffffff;
    for (int i = tid; i < n; i += 256) {
        const T v = fabs(x[static_cast<size_t>(i) * incx]);
        amax_combine(best, bi, v, i);
    }
    for (int o = 32; o > 0; o >>= 1) {
        const T ov = __shfl_xor(best, o);
        const int oi = __shfl_xor(bi, o);
        amax_combine(best, bi, ov, oi);
    }
    if ((tid & 63) == 0) {
        sv[tid >> 6] = best;
        si[tid >> 6] = bi;
    }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < 4; ++w) amax_combine(best, bi, sv[w], si[w]);
        *result = (n > 0) ? (bi == 0x7fffffff ? 1 : bi + 1) : 0;
    }
}

template <typename T>
void iamax(mifact_handle h, int n, const T* x, int incx, int* result)
{
    hipLaunchKernelGGL(iamax_kernel<T>, dim3(1), dim3(256), 0, h->stream, n, x,
                       incx, result);
    MIFACT_LAUNCHED(h);
}

template void trsm_rlt_base<float>(mifact_handle, int, int, const float*, long long, float*, long long);
template void trsm_rlt_base<double>(mifact_handle, int, int, const double*, long long, double*, long long);
template void trsm_llu_base<float>(mifact_handle, int, int, const float*, long long, float*, long long);
template void trsm_llu_base<double>(mifact_handle, int, int, const double*, long long, double*, long long);
template void potf2_base<float>(mifact_handle, int, float*, long long, int*, int);
template void potf2_base<double>(mifact_handle, int, double*, long long, int*, int);
template void iamax<float>(mifact_handle, int, const float*, int, int*);
template void iamax<double>(mifact_handle, int, const double*, int, int*);

}  // namespace mifact

// Panel kernels for gfx950: diagonal POTF2 (with the inverse of its block),
// batched inverses of 64 x 64 diagonal blocks, wave-reduction IAMAX.
//
// These are the leaves of the recursive diagonal factorization and of the
// recursive triangular solves.  A triangular-solve leaf multiplies by the
// inverse of its 64 x 64 diagonal block on MFMA (as rocBLAS' TRSM does with
// its inverted diagonal blocks); everything above the leaves is MFMA GEMM.
#include <hip/hip_runtime.h>

#include <limits>

#include "kernels.hpp"

namespace mifact {

constexpr int kLeaf = 64;  // leaf order of POTF2 / triangular inverse kernels

// Row i of M = inv(L) for a lower-triangular s x s block held in LDS
// (sL[k][j] = L(k, j)), computed by thread i from the last column to the
// first: M(i,i) = 1/L(i,i) (1 when unit), M(i,j) = -sum_{k>j} M(i,k) L(k,j)
// / L(j,j).  Every LDS read is a broadcast (same address on all lanes).
template <typename T, bool UNIT>
__device__ __attribute__((always_inline)) inline void inverse_row(int s, const T (*sL)[kLeaf + 1],
                                                                 T (&m)[kLeaf])
{
    const int i = threadIdx.x;
#pragma unroll
    for (int j = kLeaf - 1; j >= 0; --j) {
        if (j < s) {
            // four partial sums break the dependent FMA chain
            T acc[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int k = j + 1; k < kLeaf; ++k) acc[(k - j - 1) & 3] += m[k] * sL[k][j];
            const T sum = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            const T dj = UNIT ? T(1) : T(1) / sL[j][j];
            m[j] = (i == j) ? dj : (i > j ? -sum * dj : T(0));
        } else {
            m[j] = T(0);
        }
    }
}

// ------------------------------------------------------------------ POTF2
// Unblocked lower Cholesky of an s x s block (s <= 64) by one wave64: lane i
// holds row i in registers; per column j the pivot is broadcast with a lane
// shuffle, the scaled column goes through LDS and every lane updates its row
// (right-looking; reciprocal scaling as LAPACK xPOTF2).  When inv is not
// null the wave then writes inv(L) (column-major, ld 64): the panel TRSM
// applies it with an MFMA multiply instead of a substitution.
template <typename T>
__global__ __launch_bounds__(64) void potf2_kernel(int s, T* __restrict__ A, long long lda,
                                                   T* __restrict__ inv, int* __restrict__ info,
                                                   int col_offset)
{
    __shared__ T col[kLeaf];
    __shared__ T sL[kLeaf][kLeaf + 1];
    const int i = threadIdx.x;
    T x[kLeaf];
#pragma unroll
    for (int k = 0; k < kLeaf; ++k) x[k] = (i < s && k <= i) ? A[off(i, k, lda)] : T(0);
    bool failed = false;
#pragma unroll
    for (int j = 0; j < kLeaf; ++j) {
        if (!failed && j < s) {
            const T d = __shfl(x[j], j);
            if (!(d > T(0))) {
                failed = true;
                if (i == 0 && info != nullptr) atomicCAS(info, 0, col_offset + j + 1);
            } else {
                const T ljj = sqrt(d);
                const T r = T(1) / ljj;
                if (i == j) {
                    x[j] = ljj;
                } else if (i > j) {
                    x[j] *= r;
                }
                col[i] = x[j];
                __syncthreads();
#pragma unroll
                for (int k = j + 1; k < kLeaf; ++k) x[k] -= x[j] * col[k];
                __syncthreads();
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kLeaf; ++k)
        if (i < s && k <= i) A[off(i, k, lda)] = x[k];
    if (inv == nullptr || failed) return;
#pragma unroll
    for (int k = 0; k < kLeaf; ++k) sL[i][k] = x[k];
    __syncthreads();
    T m[kLeaf];
    inverse_row<T, false>(s, sL, m);
#pragma unroll
    for (int j = 0; j < kLeaf; ++j)
        if (i < s && j < s) inv[j * kLeaf + i] = m[j];
}

// Unpivoted LU of an s x s block (s <= 64) in one wave: lane i owns row i
// in LDS, right-looking column by column (LAPACK sgetf2 arithmetic without
// the row search: multipliers by the reciprocal when |pivot| >= sfmin).  A
// zero pivot records col_offset + j + 1 in *info (first in launch order)
// and leaves its column unscaled.
template <typename T>
__global__ __launch_bounds__(64) void getf2_npvt_kernel(int s, T* A, long long lda, int* info,
                                                        int col_offset)
{
    __shared__ T a[64][65];
    const int lane = threadIdx.x;
    for (int c = 0; c < s; ++c) a[lane][c] = lane < s ? A[off(lane, c, lda)] : T(0);
    __syncthreads();
    const T sfmin = std::numeric_limits<T>::min();
    for (int j = 0; j < s; ++j) {
        const T d = a[j][j];
        if (d == T(0)) {
            if (lane == 0) atomicCAS(info, 0, col_offset + j + 1);
        } else if (lane > j && lane < s) {
            const T l = (fabs(d) >= sfmin) ? a[lane][j] * (T(1) / d) : a[lane][j] / d;
            a[lane][j] = l;
            for (int c = j + 1; c < s; ++c) a[lane][c] -= l * a[j][c];
        }
        __syncthreads();
    }
    if (lane < s)
        for (int c = 0; c < s; ++c) A[off(lane, c, lda)] = a[lane][c];
}

template <typename T>
void getf2_npvt_block(mifact_handle h, int s, T* A, long long lda, int* info, int col_offset)
{
    if (s <= 0) return;
    StageScope scope(h, kStagePotf2, 2.0 * s * s * s / 3.0, 2.0 * sizeof(T) * s * s);
    hipLaunchKernelGGL(getf2_npvt_kernel<T>, dim3(1), dim3(64), 0, h->stream, s, A, lda, info,
                       col_offset);
    MIFACT_LAUNCHED(h);
}

template void getf2_npvt_block<float>(mifact_handle, int, float*, long long, int*, int);
template void getf2_npvt_block<double>(mifact_handle, int, double*, long long, int*, int);

template <typename T>
void potf2_base(mifact_handle h, int s, T* A, long long lda, T* inv, int* info, int col_offset)
{
    if (s <= 0) return;
    if (s > kLeaf) {
        report_status(h, mifact_status_invalid_value, "potf2_base: s > 64");
        return;
    }
    StageScope scope(h, kStagePotf2, static_cast<double>(s) * s * s / 3.0,
                     sizeof(T) * static_cast<double>(s) * s * 1.5);
    hipLaunchKernelGGL(potf2_kernel<T>, dim3(1), dim3(64), 0, h->stream, s, A, lda, inv,
                       info, col_offset);
    MIFACT_LAUNCHED(h);
}

// ------------------------------------------------------------------ TRTRI
// Inverses of the 64 x 64 diagonal blocks of an n x n lower-triangular L
// (unit or non-unit), one wave per block; block b is written at
// inv + 64*64*b (column-major, ld 64).
// UPPER: the block is read transposed (an upper-triangular U as the lower
// U^T), so the result is inv(U^T) = inv(U)^T.
template <typename T, bool UNIT, bool UPPER>
__global__ __launch_bounds__(64) void trtri_kernel(int n, const T* __restrict__ L, long long ldl,
                                                   T* __restrict__ inv)
{
    __shared__ T sL[kLeaf][kLeaf + 1];
    const int b = blockIdx.x;
    const int r0 = b * kLeaf;
    const int s = (n - r0) < kLeaf ? (n - r0) : kLeaf;
    const int i = threadIdx.x;
    const T* Lb = L + off(r0, r0, ldl);
#pragma unroll 8
    for (int k = 0; k < kLeaf; ++k) {
        const bool in = i < s && k < s && k <= i;
        if (UPPER) {
            sL[i][k] = in ? Lb[off(k, i, ldl)] : T(0);  // U^T(i, k) = U(k, i)
        } else {
            sL[i][k] = in ? Lb[off(i, k, ldl)] : T(0);
        }
    }
    __syncthreads();
    T m[kLeaf];
    inverse_row<T, UNIT>(s, sL, m);
    T* ib = inv + static_cast<size_t>(b) * kLeaf * kLeaf;
#pragma unroll
    for (int j = 0; j < kLeaf; ++j)
        if (i < s && j < s) ib[j * kLeaf + i] = m[j];
}

template <typename T>
void trtri_blocks(mifact_handle h, int n, const T* L, long long ldl, T* inv, bool unit, bool upper)
{
    if (n <= 0) return;
    const int nb = ceil_div(n, kLeaf);
    StageScope scope(h, kStageTrsmLeaf, static_cast<double>(nb) * kLeaf * kLeaf * kLeaf / 3.0,
                     sizeof(T) * static_cast<double>(nb) * kLeaf * kLeaf * 1.5);
    const dim3 grid(nb), block(64);
    if (unit && upper) {
        hipLaunchKernelGGL((trtri_kernel<T, true, true>), grid, block, 0, h->stream, n, L, ldl, inv);
    } else if (unit) {
        hipLaunchKernelGGL((trtri_kernel<T, true, false>), grid, block, 0, h->stream, n, L, ldl, inv);
    } else if (upper) {
        hipLaunchKernelGGL((trtri_kernel<T, false, true>), grid, block, 0, h->stream, n, L, ldl, inv);
    } else {
        hipLaunchKernelGGL((trtri_kernel<T, false, false>), grid, block, 0, h->stream, n, L, ldl, inv);
    }
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

template void potf2_base<float>(mifact_handle, int, float*, long long, float*, int*, int);
template void potf2_base<double>(mifact_handle, int, double*, long long, double*, int*, int);
template void trtri_blocks<float>(mifact_handle, int, const float*, long long, float*, bool, bool);
template void trtri_blocks<double>(mifact_handle, int, const double*, long long, double*, bool, bool);
template void iamax<float>(mifact_handle, int, const float*, int, int*);
template void iamax<double>(mifact_handle, int, const double*, int, int*);

}  // namespace mifact

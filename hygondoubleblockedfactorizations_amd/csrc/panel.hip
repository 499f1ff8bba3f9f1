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

// One-wave leaf kernels keep their code small: a wave that runs through
// straight-line code once pays an instruction-cache fill for every line of
// it, so the column loops are runtime loops over groups of kGroup columns
// with fixed register indices, and the registers rotate by kGroup between
// groups (a "window" of W registers holds the live columns).
constexpr int kGroup = 4;
constexpr int kPadL = 64;                    // zero columns in front of each LDS row
constexpr int kLdsLd = kPadL + 2 * kLeaf + 4;  // LDS row stride: padding, 64 columns, overhang

template <typename T>
using Vec4 = T __attribute__((ext_vector_type(4)));

// value of v in lane `lane` (uniform)
template <typename T>
__device__ inline T lane_value(T v, int lane)
{
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
    } else {
        const long long b = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffLL), lane);
        const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
        return __builtin_bit_cast(
            T, (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
    }
}

// Row `ri` of an s x s block, columns 0..63, as 64 unconditional loads
// (column index clamped into the block).  The loads are pinned to registers
// before any select so that none of them becomes a branch around a load
// (which would make them complete one at a time).
template <typename T>
__device__ __attribute__((always_inline)) inline void load_row(const T* __restrict__ A,
                                                               long long lda, int s, int ri,
                                                               T (&x)[kLeaf])
{
    const T* row = A + ri;
#pragma unroll
    for (int t = 0; t < kLeaf; ++t) x[t] = row[static_cast<size_t>(t < s ? t : s - 1) * lda];
#pragma unroll
    for (int t = 0; t < kLeaf; ++t) asm volatile("" : "+v"(x[t]));
}

// One phase of the right-looking Cholesky of a block padded with the
// identity beyond s (so no column needs a size test): columns j0 + c of the
// groups j0 = jbeg, jbeg + 4, ... < jend.  x[t] holds column j0 + t of this
// lane's row.  The pivot of the next column is formed by its own lane
// (x[c + 1] - l^2) while the other lanes' updates run, and one wave's LDS
// operations complete in order, so no barrier is needed.  Finished columns
// go to sL (row i at kPadL + column); r/ljj carry the current pivot; jf
// becomes the first column whose pivot is not positive (ends the phase).
template <typename T, int W>
__device__ __attribute__((always_inline)) inline void chol_phase(int jbeg, int jend,
                                                                 T (&x)[kLeaf], T* col,
                                                                 T (*sL)[kLdsLd], T* rd, int& jf,
                                                                 T& ljj, T& r)
{
    const int i = threadIdx.x;
    for (int j0 = jbeg; j0 < jend && jf == kLeaf; j0 += kGroup) {
#pragma unroll
        for (int c = 0; c < kGroup; ++c) {
            const int j = j0 + c;
            const T lij = (i > j) ? x[c] * r : (i == j ? ljj : T(0));
            x[c] = lij;
            rd[j] = r;
            col[i] = lij;
            const T dn = lane_value(x[c + 1] - lij * lij, (j + 1) & (kLeaf - 1));
            jf = (!(dn > T(0)) && jf == kLeaf && j + 1 < kLeaf) ? j + 1 : jf;
            const T ljj_n = sqrt(dn);
            const T r_n = T(1) / ljj_n;
            T cv[W];
#pragma unroll
            for (int q = 0; q < W; q += 4) {
                const Vec4<T> v = *reinterpret_cast<const Vec4<T>*>(&col[j0 + q]);
                cv[q] = v[0];
                cv[q + 1] = v[1];
                cv[q + 2] = v[2];
                cv[q + 3] = v[3];
            }
#pragma unroll
            for (int t = c + 1; t < W; ++t) x[t] -= lij * cv[t];
            ljj = ljj_n;
            r = r_n;
        }
        *reinterpret_cast<Vec4<T>*>(&sL[i][kPadL + j0]) = Vec4<T>{x[0], x[1], x[2], x[3]};
#pragma unroll
        for (int t = 0; t < W - kGroup; ++t) x[t] = x[t + kGroup];
#pragma unroll
        for (int t = W - kGroup; t < W; ++t) x[t] = T(0);
    }
}

// Inverse M = inv(L) of the 64 x 64 lower-triangular block in sL (identity
// beyond s), row i per lane, from the last column to the first:
// M(i, k) = (delta_ik - sum_{k' > k} M(i, k') L(k', k)) * rd[k]; after step k
// the accumulators of every column j < k gain M(i, k) L(k, j) (row k of L,
// read as a broadcast; the next row is fetched during the current step).
// a[t] <-> column kt - t within a group; the front padding of sL covers the
// window's columns below 0 (FP32 only: FP64 has no registers for the
// prefetch).  Writes M(i, k) for i, k < s at inv + 64 k + i.  A unit
// diagonal is given as rd = 1.
template <typename T>
__device__ __attribute__((always_inline)) inline void inverse_from_lds(int s,
                                                                       const T (*sL)[kLdsLd],
                                                                       const T* rd,
                                                                       T* __restrict__ inv)
{
    constexpr int W = kLeaf;
    constexpr bool kPrefetch = sizeof(T) == 4;
    const int i = threadIdx.x;
    T a[kLeaf];
#pragma unroll
    for (int t = 0; t < kLeaf; ++t) a[t] = T(0);
    for (int kt = kLeaf - 1; kt >= 0; kt -= kGroup) {
        T lv[W];
#pragma unroll
        for (int q = 0; q < W; q += 4) {
            const Vec4<T> v = *reinterpret_cast<const Vec4<T>*>(&sL[kt][kPadL + kt - W + 1 + q]);
            lv[q] = v[0];
            lv[q + 1] = v[1];
            lv[q + 2] = v[2];
            lv[q + 3] = v[3];
        }
#pragma unroll
        for (int c = 0; c < kGroup; ++c) {
            const int k = kt - c;
            T ln[W];
            if (kPrefetch && c + 1 < kGroup) {
#pragma unroll
                for (int q = 0; q < W; q += 4) {
                    const Vec4<T> v =
                        *reinterpret_cast<const Vec4<T>*>(&sL[k - 1][kPadL + kt - W + 1 + q]);
                    ln[q] = v[0];
                    ln[q + 1] = v[1];
                    ln[q + 2] = v[2];
                    ln[q + 3] = v[3];
                }
            }
            const T mk = (((i == k) ? T(1) : T(0)) - a[c]) * rd[k];
            if (i < s && k < s) inv[k * kLeaf + i] = mk;
#pragma unroll
            for (int t = c + 1; t < W; ++t) a[t] += mk * lv[W - 1 - t];
            if (c + 1 < kGroup) {
                if (kPrefetch) {
#pragma unroll
                    for (int q = 0; q < W; ++q) lv[q] = ln[q];
                } else {
#pragma unroll
                    for (int q = 0; q < W; q += 4) {
                        const Vec4<T> v = *reinterpret_cast<const Vec4<T>*>(
                            &sL[k - 1][kPadL + kt - W + 1 + q]);
                        lv[q] = v[0];
                        lv[q + 1] = v[1];
                        lv[q + 2] = v[2];
                        lv[q + 3] = v[3];
                    }
                }
            }
        }
#pragma unroll
        for (int t = 0; t < W - kGroup; ++t) a[t] = a[t + kGroup];
#pragma unroll
        for (int t = W - kGroup; t < W; ++t) a[t] = T(0);
    }
}

// ------------------------------------------------------------------ POTF2
// Rows o .. o + N - 1 of inv(L_b), L_b = L(o:o+N, o:o+N) the lower block in
// sL (1 / diagonal in rd): lane i < N holds row i, m[k] = M(i, k), from the
// last column to the first: M(i, k) = (delta_ik - sum_{k' > k} M(i, k')
// L(k', k)) rd[k] (M L = I; the L entries are LDS broadcasts).
template <typename T, int N>
__device__ __attribute__((always_inline)) inline void lower_inverse_rows(const T (*sL)[kLdsLd],
                                                                          const T* rd, int o,
                                                                          int i, T (&m)[N])
{
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
        T acc = (i == k) ? T(1) : T(0);
#pragma unroll
        for (int kp = k + 1; kp < N; ++kp) acc -= m[kp] * sL[o + kp][kPadL + o + k];
        m[k] = acc * rd[o + k];
        // one row entry at a time: the broadcasts are not all hoisted
        asm volatile("" : "+v"(m[k]));
    }
}

// Unblocked lower Cholesky of an s x s block (s <= 64).  The block is
// loaded into LDS by four waves; wave 0 factors it with lane i holding row i
// in registers (rows beyond s padded with the identity): per column the
// pivot is a square root and the column is scaled by its reciprocal (LAPACK
// xPOTF2 arithmetic), then every lane updates its row (right-looking).  A
// non-positive or NaN pivot records col_offset + j + 1 in *info (first
// failure in launch order), leaves the columns before it factored and
// everything from it on unchanged.  When inv is not null the four waves then
// write inv(L) (column-major, ld 64) as 2 x 2 blocks of 32: inv(L11) and
// inv(L22) by waves 0 and 1 side by side, then -inv(L22) L21 inv(L11) as two
// 32 x 32 products over all threads.  The panel TRSM applies inv(L) with an
// MFMA multiply instead of a substitution.
template <typename T>
__global__ __launch_bounds__(256) void potf2_kernel(int s, T* __restrict__ A, long long lda,
                                                    T* __restrict__ inv, int* __restrict__ info,
                                                    int col_offset)
{
    __shared__ __attribute__((aligned(16))) T col[2 * kLeaf];
    __shared__ __attribute__((aligned(16))) T sL[kLeaf][kLdsLd];
    __shared__ T rd[kLeaf];
    __shared__ T mb[3][32][33];   // inv(L11), inv(L22), L21 inv(L11)
    __shared__ int jf_s;
    const int tid = threadIdx.x;
    const int i = tid & 63;
    const int wv = tid >> 6;
    // the block (lower part, identity beyond s) -> LDS; wave wv: columns
    // 16 wv .. 16 wv + 15 (coalesced down each column) and its share of the
    // zero padding in front of each row
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const int c = 16 * wv + t;
        T v;
        if (i < s) v = (c <= i) ? A[off(i, c, lda)] : T(0);
        else v = (c == i) ? T(1) : T(0);
        sL[i][kPadL + c] = v;
    }
#pragma unroll
    for (int t = 0; t < 16; t += 4)
        *reinterpret_cast<Vec4<T>*>(&sL[i][16 * wv + t]) = Vec4<T>{T(0), T(0), T(0), T(0)};
    if (tid < kLeaf) col[kLeaf + tid] = T(0);
    __syncthreads();
    if (wv == 0) {
        T x[kLeaf];
#pragma unroll
        for (int t = 0; t < kLeaf; t += 4) {
            const Vec4<T> v = *reinterpret_cast<const Vec4<T>*>(&sL[i][kPadL + t]);
            x[t] = v[0];
            x[t + 1] = v[1];
            x[t + 2] = v[2];
            x[t + 3] = v[3];
        }
        int jf = kLeaf;
        const T d0 = lane_value(x[0], 0);
        if (!(d0 > T(0))) jf = 0;
        T ljj = sqrt(d0);
        T r = T(1) / ljj;
        chol_phase<T, 64>(0, 16, x, col, sL, rd, jf, ljj, r);
        chol_phase<T, 48>(16, 32, x, col, sL, rd, jf, ljj, r);
        chol_phase<T, 32>(32, 48, x, col, sL, rd, jf, ljj, r);
        chol_phase<T, 16>(48, kLeaf, x, col, sL, rd, jf, ljj, r);
        if (i == 0) jf_s = jf;
    }
    __syncthreads();
    const int jf = jf_s;
    // the finished columns (all, or those before a failing pivot)
    const int done = jf < s ? jf : s;
    if (i < s) {
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const int c = 16 * wv + t;
            if (c <= i && c < done) A[off(i, c, lda)] = sL[i][kPadL + c];
        }
    }
    if (jf < s) {
        if (tid == 0 && info != nullptr) atomicCAS(info, 0, col_offset + jf + 1);
        return;
    }
    if (inv == nullptr) return;
    if (wv < 2) {
        T m[32];
        lower_inverse_rows<T, 32>(sL, rd, 32 * wv, i, m);
        if (i < 32) {
#pragma unroll
            for (int k = 0; k < 32; ++k) mb[wv][i][k] = m[k];
        }
    }
    __syncthreads();
    // thread: row r, columns c0 .. c0 + 3 of the 32 x 32 off-diagonal products
    const int r = tid & 31;
    const int c0 = (tid >> 5) * 4;
    T acc[4] = {T(0), T(0), T(0), T(0)};
    for (int k = 0; k < 32; ++k) {
        const T l = sL[32 + r][kPadL + k];
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[e] += l * mb[0][k][c0 + e];
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) mb[2][r][c0 + e] = acc[e];
    __syncthreads();
    T m21[4] = {T(0), T(0), T(0), T(0)};
    for (int k = 0; k < 32; ++k) {
        const T g = mb[1][r][k];
#pragma unroll
        for (int e = 0; e < 4; ++e) m21[e] -= g * mb[2][k][c0 + e];
    }
    // inv (ld 64), entries (row, column) with both below s
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int c = c0 + e;
        if (32 + r < s && c < s) inv[c * kLeaf + 32 + r] = m21[e];
        if (r < s && c < s) inv[c * kLeaf + r] = mb[0][r][c];
        if (r < s && 32 + c < s) inv[(32 + c) * kLeaf + r] = T(0);
        if (32 + r < s && 32 + c < s) inv[(32 + c) * kLeaf + 32 + r] = mb[1][r][c];
    }
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
    hipLaunchKernelGGL(potf2_kernel<T>, dim3(1), dim3(256), 0, h->stream, s, A, lda, inv,
                       info, col_offset);
    MIFACT_LAUNCHED(h);
}

// ------------------------------------------------------------------ TRTRI
// Inverses of the 64 x 64 diagonal blocks of an n x n lower-triangular L
// (unit or non-unit), one wave per block; block b is written at
// inv + 64*64*b (column-major, ld 64).
// UPPER: the block is read transposed (an upper-triangular U as the lower
// U^T), so the result is inv(U^T) = inv(U)^T.
// Lane k loads row k of the block (coalesced across lanes), the lower
// triangle of L goes to LDS as rows (identity beyond s), then the inverse
// is formed as in the POTF2 kernel.
template <typename T, bool UNIT, bool UPPER>
__global__ __launch_bounds__(64) void trtri_kernel(int n, const T* __restrict__ L, long long ldl,
                                                   T* __restrict__ inv)
{
    __shared__ __attribute__((aligned(16))) T sL[kLeaf][kLdsLd];
    __shared__ T rd[kLeaf];
    const int b = blockIdx.x;
    const int r0 = b * kLeaf;
    const int s = (n - r0) < kLeaf ? (n - r0) : kLeaf;
    const int k = threadIdx.x;
    const T* Lb = L + off(r0, r0, ldl);
    T x[kLeaf];
    load_row(Lb, ldl, s, k < s ? k : s - 1, x);   // x[t] = block(k, t)
#pragma unroll
    for (int t = 0; t < kPadL; t += 4)
        *reinterpret_cast<Vec4<T>*>(&sL[k][t]) = Vec4<T>{T(0), T(0), T(0), T(0)};
    if (UPPER) {
        // U^T(t, k) = U(k, t): lane k's row of U is column k of U^T
#pragma unroll
        for (int t = 0; t < kLeaf; ++t)
            sL[t][kPadL + k] = (k < s && t < s) ? (k <= t ? x[t] : T(0)) : (t == k ? T(1) : T(0));
    } else {
#pragma unroll
        for (int t = 0; t < kLeaf; t += 4) {
            Vec4<T> v;
#pragma unroll
            for (int e = 0; e < 4; ++e)
                v[e] = (k < s) ? (t + e <= k ? x[t + e] : T(0)) : (t + e == k ? T(1) : T(0));
            *reinterpret_cast<Vec4<T>*>(&sL[k][kPadL + t]) = v;
        }
    }
    __syncthreads();
    rd[k] = UNIT ? T(1) : T(1) / sL[k][kPadL + k];
    __syncthreads();
    inverse_from_lds<T>(s, sL, rd, inv + static_cast<size_t>(b) * kLeaf * kLeaf);
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

// Uneven batches of small Cholesky / LU factorizations in one persistent
// launch, and batched FP64 refinement solves captured in a hipGraph.
//
// Each workgroup (4 waves) loops: take the next matrix id from a device
// work queue (matrices ordered largest first), factor it right-looking with
// 32-column panels, repeat until the queue is drained.  Per panel:
//   POTRF: POTF2 of the 32 x 32 diagonal block in LDS, row-parallel TRSM of
//          the rows below, lower-trapezoid SYRK of the trailing matrix;
//   GETRF: the panel in LDS, column-by-column partial pivoting (smallest
//          index on ties, LAPACK sgetf2 semantics), the panel's row
//          interchanges applied to the other columns in order,
//          column-parallel unit-lower TRSM of the block row, GEMM update
//          (orders up to 1024).
// The trailing updates run as 64 x 64 MFMA tiles (v_mfma_f32_32x32x2_f32,
// or v_mfma_f32_32x32x16_bf16/_f16 with FP32 accumulation for the
// mixed-precision variants).  Same arithmetic contract as the large drivers
// (reference: src/potrf/double_blocked_potrf.cpp:294-381,
// src/getrf/double_blocked_getrf.cpp:125-191), one matrix per workgroup.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>

#include "batched/batched_factorizations.h"
#include "kernels.hpp"

namespace mifact {
namespace {

constexpr int kNB = 32;             // panel width
constexpr int kT = 64;              // trailing-update tile
constexpr int kMaxOrder = 4096;
constexpr int kBuckets = 256;       // work-queue size buckets (16 orders each)

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

struct BatchArgs {
    int small_max;       // trailing orders up to this take the 32 x 32 per-wave tiles
    float* A;
    const long long* offsets;
    const int* n;
    const int* lda;
    int* ipiv;
    const long long* ipiv_offsets;
    int* info;
    const int* order;
    unsigned* counter;
    int batch;
};

// The 64 x 64 tiles of a trailing update C -= L R (C m x nc, K = kb), one
// after another on the workgroup's four waves (2 x 2 blocks of
// v_mfma_f32_32x32x*), each tile in ceil(kb / 32) steps of 32-wide k slices
// that accumulate before its block of C is written once.  Software
// pipelined: while step s's products run, the operands of step s + 1 are
// read from global memory into registers, and a tile's block of C is read
// at its first step; two operand buffers (k-major, 2 x kNB x 65 floats
// each) alternate in `scratch`, one workgroup barrier per step.
// la(i, k): left operand element (row i); rb(k, j): right operand element
// (column j).  LOWER: only the tiles and elements with row >= column.
// MAXSL: the largest slice count the caller passes (kb <= MAXSL * 32); 1
// compiles the single-slice sweep.
constexpr int kTileFloats = 2 * kNB * (kT + 1);

template <int PREC, bool LOWER, int MAXSL, typename LA, typename RB>
__device__ void tile_sweep(float* C, long long ldc, int m, int nc, int kb, LA la, RB rb, float* scratch)
{
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int tiles_r = (m + kT - 1) / kT, tiles_c = (nc + kT - 1) / kT;
    const int slices = MAXSL == 1 ? 1 : (kb + kNB - 1) / kNB;
    constexpr int PER = kNB * kT / 256;   // operand elements per thread and operand
    float na[PER], nbv[PER];
    auto fetch = [&](int r0, int c0, int k0) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = tid + 256 * q;
            const int k = k0 + e / kT, i = e % kT;
            float va = 0.0f, vb = 0.0f;
            if (k < kb) {
                if (r0 + i < m) va = la(r0 + i, k);
                if (c0 + i < nc) vb = rb(k, c0 + i);
            }
            na[q] = va;
            nbv[q] = vb;
        }
    };
    int ti = 0, tj = 0, sl = 0;
    fetch(0, 0, 0);
    const int row = wm * 32 + (lane & 31);
    const int col = wn * 32 + (lane & 31);
    float cv[16];
    floatx16 acc;
    for (int t = 0;; ++t) {
        float (*sa)[kT + 1] = reinterpret_cast<float (*)[kT + 1]>(scratch + (t & 1) * kTileFloats);
        float (*sb)[kT + 1] = sa + kNB;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = tid + 256 * q;
            sa[e / kT][e % kT] = na[q];
            sb[e / kT][e % kT] = nbv[q];
        }
        __syncthreads();
        const int r0 = ti * kT, c0 = tj * kT;
        // next step: the next k slice, else down the tile column, then the next one
        int nsl = sl + 1, ni = ti, nj = tj;
        if (nsl == slices) {
            nsl = 0;
            ++ni;
            if (ni >= tiles_r) {
                ++nj;
                ni = LOWER ? nj : 0;
            }
        }
        const bool more = nj < tiles_c;
        // the products run as C^T -= R^T L^T, so accumulator r holds C row
        // `gr` (the lane's, consecutive down a column: two 128-byte lines
        // per load) and column 8 (r / 4) + 4 (lane / 32) + r % 4
        const int gr = r0 + row;
        if (sl == 0) {
            // this tile's C block, read before the next step's operands
            // (loads complete in issue order, so the store waits for these only)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int gc = c0 + wn * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const bool ok = gc < nc && gr < m && (!LOWER || gr >= gc);
                cv[r] = ok ? C[off(gr, gc, ldc)] : 0.0f;
                acc[r] = 0.0f;
            }
        }
        if (more) fetch(ni * kT, nj * kT, nsl * kNB);
        if (PREC == kPrecFP32) {
#pragma unroll
            for (int kk = 0; kk < kNB; kk += 2) {
                const float a = sa[kk + (lane >> 5)][row];
                const float b = sb[kk + (lane >> 5)][col];
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(b, a, acc, 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < kNB; kk += 16) {
                const int k0 = kk + 8 * (lane >> 5);
                if (PREC == kPrecBF16) {
                    bf16x8 a, b;
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        a[e] = static_cast<__bf16>(sa[k0 + e][row]);
                        b[e] = static_cast<__bf16>(sb[k0 + e][col]);
                    }
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, a, acc, 0, 0, 0);
                } else {
                    halfx8 a, b;
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        a[e] = static_cast<_Float16>(sa[k0 + e][row]);
                        b[e] = static_cast<_Float16>(sb[k0 + e][col]);
                    }
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(b, a, acc, 0, 0, 0);
                }
            }
        }
        if (sl == slices - 1) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int gc = c0 + wn * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (gc < nc && gr < m && (!LOWER || gr >= gc)) C[off(gr, gc, ldc)] = cv[r] - acc[r];
            }
        }
        if (!more) break;
        ti = ni;
        tj = nj;
        sl = nsl;
    }
    // every wave is done with the buffers and with C before the caller goes on
    __syncthreads();
}

// Size class of the small trailing updates: one 32 x 32 tile per wave, four
// tiles per pass, no workgroup barrier inside a pass.  Each wave stages its
// operands one 32-wide k slice at a time (32 x 32 of each, k-major) in its
// own 2 x 32 x 33 slice of `scratch` (4 x 2112 floats).  Tiles t = 0 ..
// count - 1 in row-major order over the lower trapezoid (LOWER) or the
// full rectangle of 32-blocks of the m x nc block C.
constexpr int kSmallT = 32;
constexpr int kSmallScratch = 4 * 2 * kNB * (kSmallT + 1);

template <int PREC, bool LOWER, typename LA, typename RB>
__device__ void tile_update_small(float* C, long long ldc, int m, int nc, int kb, LA la, RB rb,
                                  float* scratch)
{
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    float (*sa)[kSmallT + 1] = reinterpret_cast<float (*)[kSmallT + 1]>(scratch + wave * 2 * kNB * (kSmallT + 1));
    float (*sb)[kSmallT + 1] = sa + kNB;
    const int tr = (m + kSmallT - 1) / kSmallT, tc = (nc + kSmallT - 1) / kSmallT;
    // LOWER: row-major over the trapezoid ti >= tj (tj < tc): row ti holds
    // min(ti + 1, tc) tiles
    const int full = tc < tr ? tc : tr;   // rows 0 .. full - 1 hold ti + 1 tiles
    const int count = LOWER ? full * (full + 1) / 2 + (tr - full) * tc : tr * tc;
    for (int t = wave; t < count; t += 4) {
        int ti, tj;
        if (LOWER) {
            if (t < full * (full + 1) / 2) {   // t = ti (ti + 1) / 2 + tj, tj <= ti
                ti = 0;
                while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
                tj = t - ti * (ti + 1) / 2;
            } else {
                const int u = t - full * (full + 1) / 2;
                ti = full + u / tc;
                tj = u % tc;
            }
        } else {
            ti = t % tr;
            tj = t / tr;
        }
        const int r0 = ti * kSmallT, c0 = tj * kSmallT;
        floatx16 acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
        const int rc = lane & 31;
        for (int k0 = 0; k0 < kb; k0 += kNB) {
            for (int e = lane; e < kNB * kSmallT; e += 64) {
                const int i = e % kSmallT, k = e / kSmallT;
                float va = 0.0f, vb = 0.0f;
                if (k0 + k < kb) {
                    if (r0 + i < m) va = la(r0 + i, k0 + k);
                    if (c0 + i < nc) vb = rb(k0 + k, c0 + i);
                }
                sa[k][i] = va;
                sb[k][i] = vb;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (PREC == kPrecFP32) {
#pragma unroll
                for (int kk = 0; kk < kNB; kk += 2)
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(sb[kk + (lane >> 5)][rc],
                                                                sa[kk + (lane >> 5)][rc], acc, 0, 0, 0);
            } else {
#pragma unroll
                for (int kk = 0; kk < kNB; kk += 16) {
                    const int k1 = kk + 8 * (lane >> 5);
                    if (PREC == kPrecBF16) {
                        bf16x8 a, b;
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            a[e] = static_cast<__bf16>(sa[k1 + e][rc]);
                            b[e] = static_cast<__bf16>(sb[k1 + e][rc]);
                        }
                        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, a, acc, 0, 0, 0);
                    } else {
                        halfx8 a, b;
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            a[e] = static_cast<_Float16>(sa[k1 + e][rc]);
                            b[e] = static_cast<_Float16>(sb[k1 + e][rc]);
                        }
                        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(b, a, acc, 0, 0, 0);
                    }
                }
            }
            // the wave's slice is restaged by its next k slice or tile
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        // C^T -= R^T L^T: accumulator r holds C row rc (the lane's),
        // column 8 (r / 4) + 4 (lane / 32) + r % 4
        const int gr = r0 + rc;
        if (gr < m) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int gc = c0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (gc < nc && (!LOWER || gr >= gc)) C[off(gr, gc, ldc)] -= acc[r];
            }
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------- Cholesky

__device__ inline float lane_value(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Right-looking lower Cholesky of one matrix by one workgroup (LAPACK
// spotrf semantics: info = the first column whose pivot is not positive and
// finite; the block being factored is written back as far as it got), in
// outer blocks of 64 columns made of two 32-column panels.  Per panel:
//   1. POTF2 of the diagonal block on wave 0 alone, lane i holding row i in
//      registers (pivots and column entries broadcast with readlane, no
//      barrier); L11 and 1 / diag(L11) go to LDS, L11 back to global memory;
//   2. TRSM A21 := A21 L11^{-T}, one row per thread in registers, L11 read
//      as workgroup-uniform LDS broadcasts.
// Between the two panels the second one's 32 columns are updated by the
// first (a 32-wide strip); after both, the trailing matrix (lower) -=
// P P^T with the 64-column P on the MFMA tiles, so each element of the
// trailing matrix is read and written once per 64 columns.
// A last panel narrower than 32 (kb < 32) reads its column kb - 1 (and, in
// the POTF2, its row kb - 1) in place of the missing ones: every access
// stays inside the matrix and nothing computed from the copies is stored.
constexpr int kCholPanels = 4;   // 32-column panels per outer block

template <int PREC>
__device__ void chol_matrix(float* A, long long lda, int n, int* info, int small_max)
{
    __shared__ __align__(16) float lt[kNB][kNB];   // L11 (row-major; entries above the diagonal unused)
    __shared__ float rdiag[kNB];                   // 1 / L(j, j)
    constexpr int kScratch = kSmallScratch > 2 * kTileFloats ? kSmallScratch : 2 * kTileFloats;
    __shared__ float scratch[kScratch];            // the trailing update's operand buffers
    __shared__ int failed;
    const int tid = threadIdx.x;
    // the panel A(k:n, k:k+kb); false when a pivot fails (info set)
    auto panel = [&](int k, int kb) -> bool {
        // 1. POTF2
        if (tid < 64) {
            const int lane = tid;
            // the row walked column by column with one pointer (precomputed
            // per-column addresses would hold 64 registers)
            const float* p = A + off(k + imin(lane, kb - 1), k, lda);
            float x[kNB];
#pragma unroll
            for (int c = 0; c < kNB; ++c) {
                x[c] = *p;
                if (c + 1 < kb) p += lda;
            }
            int bad = -1;
            // fully unrolled (x[] stays in registers): the column loop has
            // no early exit, later columns are skipped by uniform branches
#pragma unroll
            for (int j = 0; j < kNB; ++j) {
                if (j < kb && bad < 0) {
                    const float djj = lane_value(x[j], j);
                    if (!(djj > 0.0f) || !isfinite(djj)) {
                        bad = j;
                    } else {
                        const float s = sqrtf(djj);
                        x[j] = lane == j ? s : (lane > j ? x[j] / s : x[j]);
#pragma unroll
                        for (int c = j + 1; c < kNB; ++c) x[c] -= x[j] * lane_value(x[j], c);
                    }
                }
            }
            if (lane < kNB) {
#pragma unroll
                for (int c = 0; c < kNB; c += 4)
                    *reinterpret_cast<float4*>(&lt[lane][c]) = make_float4(x[c], x[c + 1], x[c + 2], x[c + 3]);
                rdiag[lane] = 1.0f / x[lane];
            }
            if (tid == 0) {
                failed = bad >= 0;
                if (bad >= 0) *info = k + bad + 1;
            }
        }
        __syncthreads();
        // the factored block (lower) back to global memory
        for (int e = tid; e < kNB * kNB; e += 256) {
            const int i = e % kNB, c = e / kNB;
            if (i < kb && c <= i) A[off(k + i, k + c, lda)] = lt[i][c];
        }
        if (failed) return false;
        const int r1 = k + kb;
        // 2. TRSM
        for (int r = r1 + tid; r < n; r += 256) {
            const float* p = A + off(r, k, lda);
            float x[kNB];
#pragma unroll
            for (int j = 0; j < kNB; ++j) {
                x[j] = *p;
                if (j + 1 < kb) p += lda;
            }
#pragma unroll
            for (int j = 0; j < kNB; ++j) {
                float v = x[j];
#pragma unroll
                for (int c = 0; c < j; ++c) v -= x[c] * lt[j][c];
                x[j] = v * rdiag[j];
                // row j + 1 of L11 is read after x[j] is done, not all 528
                // entries ahead of the arithmetic (they would spill)
                asm volatile("" : "+v"(x[j]) : : "memory");
            }
            float* q = A + off(r, k, lda);
#pragma unroll
            for (int j = 0; j < kNB; ++j) {
                if (j < kb) *q = x[j];
                if (j + 1 < kb) q += lda;
            }
        }
        __syncthreads();
        return true;
    };
    // A(r1:n, r1:r1+nc) (lower) -= P P^T with P = A(r1:n, k:k+kw)
    auto update = [&](int r1, int nc, int k, int kw) {
        const int m = n - r1;
        float* C = A + off(r1, r1, lda);
        const float* Pp = A + off(r1, k, lda);
        auto la = [&](int i, int kk) { return Pp[off(i, kk, lda)]; };
        auto rb = [&](int kk, int j) { return Pp[off(j, kk, lda)]; };
        if (m <= small_max) tile_update_small<PREC, true>(C, lda, m, nc, kw, la, rb, scratch);
        else tile_sweep<PREC, true, kCholPanels>(C, lda, m, nc, kw, la, rb, scratch);
    };
    // per outer block of kCholPanels panels: before panel q its 32 columns
    // by panels 0 .. q - 1 of the block (a strip, K = 32 q), then the panel;
    // after the last one the trailing matrix by the whole block (a runtime
    // loop: one copy of each step's code)
    for (int k = 0; k < n; k += kCholPanels * kNB) {
        const int kbb = imin(kCholPanels * kNB, n - k);
        const int nq = (kbb + kNB - 1) / kNB;
#pragma nounroll
        for (int q = 0; q <= nq; ++q) {
            const int kp = imin(k + q * kNB, k + kbb);
            if (q > 0 && kp < n) {
                const bool strip = q < nq;
                update(kp, strip ? imin(kNB, k + kbb - kp) : n - kp, k, kp - k);
            }
            if (q < nq && !panel(kp, imin(kNB, k + kbb - kp))) return;
        }
    }
}

// ---------------------------------------------------------------- LU

// Wavefront max reduction of the device library (declared by the HIP
// headers only under HIP_ENABLE_EXTRA_WARP_SYNC_TYPES).
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_max_u64(
    unsigned long long);

constexpr int kLuMaxOrder = 1024;   // the whole panel (up to 1024 x 32) sits in LDS
constexpr int kPanelLd = kNB + 1;   // LDS row stride: conflict-free column walks

// Right-looking LU with partial pivoting of one matrix by one workgroup
// (LAPACK sgetrf semantics: IxAMAX first of ties, reciprocal scaling when
// |pivot| >= FLT_MIN, the row interchanges of each panel applied in order to
// every other column, info = first zero pivot), in outer blocks of 64
// columns made of two 32-column panels.  Per panel:
//   1. the m x kb panel (m = n - k <= 1024 rows) is copied into LDS and
//      factored there column by column, one workgroup barrier per column:
//      each wave publishes its arg-max candidate for column j with a copy
//      of the candidate row; after the barrier every thread picks the pivot
//      row from the four candidates; the wave owning the pivot row
//      exchanges it with row j, and every thread applies the scaling and
//      the rank-1 update to its own rows and tracks its candidate for
//      column j + 1;
//   2. the factored panel is written back; its interchanges are composed
//      into one move list (at most 2 x 32 rows); the second panel of a
//      block applies its moves to the first one's columns too;
//   3. each column right of the panel (one per thread, in registers): the
//      moves, (second panel) minus the first panel's part, L_ba U_a, then
//      U := L11^{-1} (.) (unit lower, L11 from LDS);
//   4. first panel: the second panel's columns -= L21 U12 (a 32-wide strip
//      on the MFMA tiles); second panel: the trailing matrix -= the two
//      panels' L21 times their U rows (K = 64), so each trailing element is
//      read and written once per 64 columns.
// The columns left of a block are not read again, so the interchanges of
// the later blocks reach them once, at the end: per 32 columns, the rows
// below the block staged in LDS and written to their final rows.
// Every global access is to (row, column) of the n x n matrix with both
// below n (the caller checks lda >= n), every LDS access below its bounds.
template <int PREC>
__device__ void lu_matrix(float* A, long long lda, int n, int* ipiv, int* info, int small_max)
{
    static_assert(kLuMaxOrder * kPanelLd >= kSmallScratch, "small-tile staging inside the panel");
    static_assert(kLuMaxOrder * kPanelLd >= 2 * kTileFloats, "tile buffers inside the panel");
    __shared__ float pnl[kLuMaxOrder][kPanelLd];
    __shared__ float cand[2][4][kPanelLd];      // per column parity and wave: the candidate row
    __shared__ float cand_v[2][4];              // its |value| (-1: no candidate)
    __shared__ int cand_i[2][4];                // its row
    __shared__ int piv_s[kNB];
    __shared__ int rowmap[kLuMaxOrder];         // panel rows: whose content ends where
    __shared__ int piv_all[kLuMaxOrder];        // every column's pivot row (absolute)
    __shared__ int mv_src[2 * kNB], mv_dst[2 * kNB];
    __shared__ float lba[kNB][kNB + 1];         // the first panel's rows of the second panel (L_ba)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr float sfmin = FLT_MIN;
    for (int k = 0; k < n; k += kNB) {
        const int kb = imin(kNB, n - k);
        const int m = n - k;
        const bool second = (k / kNB) & 1;     // the second panel of its block
        const int k0 = second ? k - kNB : k;   // the block's first column
        // 1. panel -> LDS: thread tid takes rows tid, tid + 256, ..., each
        //    row's 32 loads in flight together (coalesced across the wave:
        //    consecutive rows of one column); columns past kb are zero
        for (int r = tid; r < m; r += 256) {
            const float* p = A + off(k + r, k, lda);
            float v[kNB];
#pragma unroll
            for (int c = 0; c < kNB; ++c) {
                v[c] = c < kb ? *p : 0.0f;
                if (c + 1 < kb) p += lda;
            }
#pragma unroll
            for (int c = 0; c < kNB; ++c) pnl[r][c] = v[c];
        }
        __syncthreads();
        // wave arg max of (bv, bi) (first index among equal magnitudes; NaN
        // never wins: bv only takes values that compared greater), published
        // with a copy of the winning row; rows written by this wave's lanes
        // in the same pass are seen (a wave's LDS accesses complete in
        // order).  One wavefront max reduction of the key (|v| bits, ~index):
        // non-negative floats order like their bit patterns, and on equal
        // magnitudes the smaller index has the larger complement
        auto publish = [&](int par, float bv, int bi) {
            const unsigned long long key =
                (bi < m && bv >= 0.0f)
                    ? (static_cast<unsigned long long>(__float_as_uint(bv)) << 32)
                          | static_cast<unsigned int>(~bi)
                    : 0ULL;
            const unsigned long long kmax = __ockl_wfred_max_u64(key);
            if (kmax == 0ULL) {
                bv = -1.0f;
                bi = m;
            } else {
                bv = __uint_as_float(static_cast<unsigned int>(kmax >> 32));
                bi = ~static_cast<int>(static_cast<unsigned int>(kmax));
            }
            if (lane == 0) {
                cand_v[par][wave] = bv;
                cand_i[par][wave] = bi;
            }
            if (lane < kNB) cand[par][wave][lane] = bi < m ? pnl[bi][lane] : 0.0f;
        };
        {
            float bv = -1.0f;
            int bi = m;
            for (int r = tid; r < m; r += 256) {
                const float v = fabsf(pnl[r][0]);
                if (v > bv) {
                    bv = v;
                    bi = r;
                }
            }
            publish(0, bv, bi);
        }
        __syncthreads();
        for (int j = 0; j < kb; ++j) {
            const int par = j & 1;
            float pv = cand_v[par][0];
            int p = cand_i[par][0], w = 0;
#pragma unroll
            for (int q = 1; q < 4; ++q)
                if (cand_v[par][q] > pv || (cand_v[par][q] == pv && cand_i[par][q] < p)) {
                    pv = cand_v[par][q];
                    p = cand_i[par][q];
                    w = q;
                }
            // no candidate (all NaN): keep the diagonal
            const bool none = p < j || p >= m;
            if (none) p = j;
            const float* prow = none ? &pnl[j][0] : &cand[par][w][0];
            const float piv = prow[j];
            if (tid == 0) {
                ipiv[k + j] = k + p + 1;
                piv_s[j] = p;
                piv_all[k + j] = k + p;
                if (piv == 0.0f && *info == 0) *info = k + j + 1;
            }
            const bool scale = piv != 0.0f;
            const bool recip = fabsf(piv) >= sfmin;
            const float rinv = 1.0f / piv;
            // the exchange, by the wave that owns row p (its lane p % 64
            // updates row p below, after these stores: a wave's LDS
            // accesses complete in order); no other wave touches rows j, p
            if (p != j && wave == ((p & 255) >> 6) && lane < kNB) {
                const float t = pnl[j][lane];
                pnl[j][lane] = prow[lane];
                pnl[p][lane] = t;
            }
            float bv = -1.0f;
            int bi = m;
            for (int r = tid; r < m; r += 256) {
                if (r <= j) continue;
                float* dst = &pnl[r][0];
                if (scale) {
                    const float x = dst[j];
                    const float l = recip ? x * rinv : x / piv;
                    dst[j] = l;
                    for (int c = j + 1; c < kb; ++c) dst[c] -= l * prow[c];
                }
                if (j + 1 < kb) {
                    const float v = fabsf(dst[j + 1]);
                    if (v > bv) {
                        bv = v;
                        bi = r;
                    }
                }
            }
            if (j + 1 < kb) publish(par ^ 1, bv, bi);
            __syncthreads();
        }
        // 2. panel -> global (the same row split); the panel's interchanges
        //    to the other columns
        for (int r = tid; r < m; r += 256) {
            float* p = A + off(k + r, k, lda);
            float v[kNB];
#pragma unroll
            for (int c = 0; c < kNB; ++c) v[c] = pnl[r][c];
#pragma unroll
            for (int c = 0; c < kNB; ++c) {
                if (c < kb) *p = v[c];
                if (c + 1 < kb) p += lda;
            }
        }
        const int c1 = k + kb;           // the first column right of the panel
        const bool right = c1 < n;       // (then kb = 32)
        if (!right && !second) continue;
        // the panel's interchanges, composed: rowmap[r] = the row whose
        // content ends at row r, for r in {0..kb-1} and the pivot rows
        if (tid < kb) {
            rowmap[tid] = tid;
            rowmap[piv_s[tid]] = piv_s[tid];
        }
        __syncthreads();
        if (tid == 0) {
            for (int jj = 0; jj < kb; ++jj) {
                const int p = piv_s[jj];
                if (p != jj) {
                    const int t = rowmap[jj];
                    rowmap[jj] = rowmap[p];
                    rowmap[p] = t;
                }
            }
        }
        __syncthreads();
        // move list: entry t < kb for row t, entry kb + u for pivot row
        // piv_s[u] below the panel's first kb rows (its first occurrence;
        // its new content is one of the first kb rows' old contents);
        // src -1 where nothing moves
        if (tid < 2 * kNB) {
            int dst = -1;
            if (tid < kb) {
                dst = tid;
            } else if (tid >= kNB && tid - kNB < kb) {
                const int u = tid - kNB;
                const int p = piv_s[u];
                bool first = p >= kb;
                for (int v = 0; v < u; ++v) first = first && piv_s[v] != p;
                if (first) dst = p;
            }
            const int src = dst >= 0 ? rowmap[dst] : dst;
            mv_dst[tid] = dst;
            mv_src[tid] = (dst >= 0 && src != dst) ? src : -1;
        }
        __syncthreads();
        if (second) {
            // 2b. the moves on the first panel's 32 columns (one per
            //     thread; every read before any write); their rows of this
            //     panel (L_ba) to LDS for step 3
            if (tid < kNB) {
                float* p = A + off(k, k0 + tid, lda);
                float x[kNB], y[kNB];
#pragma unroll
                for (int i = 0; i < kNB; ++i) x[i] = i < kb ? p[mv_src[i] >= 0 ? mv_src[i] : i] : 0.0f;
#pragma unroll
                for (int u = 0; u < kNB; ++u) y[u] = mv_src[kNB + u] >= 0 ? p[mv_src[kNB + u]] : 0.0f;
#pragma unroll
                for (int u = 0; u < kNB; ++u)
                    if (mv_src[kNB + u] >= 0) p[mv_dst[kNB + u]] = y[u];
#pragma unroll
                for (int i = 0; i < kNB; ++i) {
                    if (mv_src[i] >= 0) p[i] = x[i];
                    lba[i][tid] = x[i];
                }
            }
            __syncthreads();
        }
        if (right) {
            // 3. the columns right of the panel, one per thread in
            //    registers: rows 0..31 read from their sources, the pivot
            //    rows below written from theirs (every read before any
            //    write); second panel: minus L_ba times the first panel's U
            //    rows of the column; then U := L11^{-1} (.) (unit lower L11
            //    = pnl rows 0..31, read as workgroup-uniform LDS broadcasts)
            for (int c = c1 + tid; c < n; c += 256) {
                float* p = A + off(k, c, lda);
                float x[kNB], y[kNB];
#pragma unroll
                for (int i = 0; i < kNB; ++i) x[i] = p[mv_src[i] >= 0 ? mv_src[i] : i];
#pragma unroll
                for (int u = 0; u < kNB; ++u) y[u] = mv_src[kNB + u] >= 0 ? p[mv_src[kNB + u]] : 0.0f;
#pragma unroll
                for (int u = 0; u < kNB; ++u)
                    if (mv_src[kNB + u] >= 0) p[mv_dst[kNB + u]] = y[u];
                if (second) {
                    const float* pa = A + off(k0, c, lda);
                    float xa[kNB];
#pragma unroll
                    for (int q = 0; q < kNB; ++q) xa[q] = pa[q];
#pragma unroll
                    for (int i = 0; i < kNB; ++i) {
                        float v = x[i];
#pragma unroll
                        for (int q = 0; q < kNB; ++q) v -= lba[i][q] * xa[q];
                        x[i] = v;
                        // row i + 1 of L_ba is read after x[i] is done
                        asm volatile("" : "+v"(x[i]) : : "memory");
                    }
                }
#pragma unroll
                for (int i = 1; i < kNB; ++i) {
                    float v = x[i];
#pragma unroll
                    for (int q = 0; q < i; ++q) v -= pnl[i][q] * x[q];
                    x[i] = v;
                    // row i + 1 of L11 is read after x[i] is done (all 496
                    // entries ahead of the arithmetic would spill)
                    asm volatile("" : "+v"(x[i]) : : "memory");
                }
#pragma unroll
                for (int i = 0; i < kNB; ++i) p[i] = x[i];
            }
            __syncthreads();
        }
        // 4. the updates (the panel array is free once step 3 is done)
        if (!second) {
            // the second panel's columns by this panel (a 32-wide strip);
            // the columns after them wait for the second panel
            const int r1 = k + kNB;
            const int ms = n - r1;
            float* C = A + off(r1, r1, lda);
            const float* L21 = A + off(r1, k, lda);
            const float* U12 = A + off(k, r1, lda);
            auto la = [&](int i, int kk) { return L21[off(i, kk, lda)]; };
            auto rb = [&](int kk, int jc) { return U12[off(kk, jc, lda)]; };
            if (ms <= small_max) tile_update_small<PREC, false>(C, lda, ms, imin(kNB, ms), kNB, la, rb, &pnl[0][0]);
            else tile_sweep<PREC, false, 1>(C, lda, ms, imin(kNB, ms), kNB, la, rb, &pnl[0][0]);
            continue;
        }
        if (!right) continue;
        // the trailing matrix by both panels of the block (K = 64)
        const int mr = n - c1;
        float* C = A + off(c1, c1, lda);
        const float* L21 = A + off(c1, k0, lda);
        const float* U12 = A + off(k0, c1, lda);
        auto la = [&](int i, int kk) { return L21[off(i, kk, lda)]; };
        auto rb = [&](int kk, int jc) { return U12[off(kk, jc, lda)]; };
        if (mr <= small_max) tile_update_small<PREC, false>(C, lda, mr, mr, 2 * kNB, la, rb, &pnl[0][0]);
        else tile_sweep<PREC, false, 2>(C, lda, mr, mr, 2 * kNB, la, rb, &pnl[0][0]);
    }
    // 5. the interchanges of the later blocks, applied to the columns of
    //    each earlier block once at the end (no step reads those columns
    //    after their block): inv = the composition of the later blocks'
    //    exchanges (rowmap), built backwards; per 32 columns, rows [kn, n)
    //    staged in LDS, then each moved row written to inv(row)
    const int blocks = (n + 2 * kNB - 1) / (2 * kNB);
    if (blocks > 1) {
        for (int r = tid; r < n; r += 256) rowmap[r] = r;
        __syncthreads();
        for (int b = blocks - 2; b >= 0; --b) {
            const int kn = (b + 1) * 2 * kNB;
            if (tid == 0) {
                for (int j = imin(kn + 2 * kNB, n) - 1; j >= kn; --j) {
                    const int p = piv_all[j];
                    const int v = rowmap[j];
                    rowmap[j] = rowmap[p];
                    rowmap[p] = v;
                }
            }
            __syncthreads();
            for (int h = 0; h < 2; ++h) {
                const int cb = b * 2 * kNB + h * kNB;
                for (int r = kn + tid; r < n; r += 256) {
                    const float* q = A + off(r, cb, lda);
                    float v[kNB];
#pragma unroll
                    for (int c = 0; c < kNB; ++c) {
                        v[c] = *q;
                        if (c + 1 < kNB) q += lda;
                    }
#pragma unroll
                    for (int c = 0; c < kNB; ++c) pnl[r - kn][c] = v[c];
                }
                __syncthreads();
                for (int r = kn + tid; r < n; r += 256) {
                    const int d = rowmap[r];
                    if (d == r) continue;
                    float* q = A + off(d, cb, lda);
                    float v[kNB];
#pragma unroll
                    for (int c = 0; c < kNB; ++c) v[c] = pnl[r - kn][c];
#pragma unroll
                    for (int c = 0; c < kNB; ++c) {
                        *q = v[c];
                        if (c + 1 < kNB) q += lda;
                    }
                }
                __syncthreads();
            }
        }
    }
}

// ---------------------------------------------------------------- queue

__device__ inline int next_item(unsigned* counter)
{
    __shared__ int item;
    __syncthreads();
    if (threadIdx.x == 0) item = static_cast<int>(atomicAdd(counter, 1u));
    __syncthreads();
    return item;
}

template <int PREC, bool LU>
__global__ __launch_bounds__(256) void batched_kernel(BatchArgs a)
{
    for (;;) {
        const int q = next_item(a.counter);
        if (q >= a.batch) return;            // queue drained: every workgroup exits
        const int id = a.order[q];
        const int n = a.n[id];
        const long long lda = a.lda[id];
        int* info = a.info + id;
        // orders beyond the path's limit or lda < n: LAPACK's illegal-argument
        // info, nothing touched
        if (n <= 0 || n > (LU ? kLuMaxOrder : kMaxOrder) || lda < n) {
            if (threadIdx.x == 0) *info = n == 0 ? 0 : -1;
            continue;
        }
        float* A = a.A + a.offsets[id];
        if (threadIdx.x == 0) *info = 0;
        __syncthreads();
        if constexpr (LU) lu_matrix<PREC>(A, lda, n, a.ipiv + a.ipiv_offsets[id], info, a.small_max);
        else chol_matrix<PREC>(A, lda, n, info, a.small_max);
    }
}

// Work order: ids sorted by decreasing order n (counting sort on n / 16),
// and the queue counter reset (one workgroup).
__global__ __launch_bounds__(1024) void order_kernel(const int* n, int batch, int* order,
                                                     unsigned* counter)
{
    __shared__ int hist[kBuckets];
    __shared__ int start[kBuckets];
    for (int b = threadIdx.x; b < kBuckets; b += 1024) hist[b] = 0;
    __syncthreads();
    auto bucket = [&](int i) {
        const int v = n[i] / 16;
        return kBuckets - 1 - (v < 0 ? 0 : (v >= kBuckets ? kBuckets - 1 : v));
    };
    for (int i = threadIdx.x; i < batch; i += 1024) atomicAdd(&hist[bucket(i)], 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int b = 0; b < kBuckets; ++b) {
            start[b] = s;
            s += hist[b];
        }
        *counter = 0u;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < batch; i += 1024) order[atomicAdd(&start[bucket(i)], 1)] = i;
}

// ---------------------------------------------------------------- solves

// t := M^{-1} t for one matrix (t FP32 in global memory, staged in LDS for
// the substitutions): LU: P, unit L, U; Cholesky: L, L^T.  Column-oriented
// substitutions, two barriers per column.
template <bool LU>
__device__ void tri_solve(const float* F, long long ldf, int n, const int* ipiv, float* t)
{
    __shared__ float ts[kMaxOrder];
    const int tid = threadIdx.x;
    for (int i = tid; i < n; i += 256) ts[i] = t[i];
    __syncthreads();
    if (LU) {
        if (tid == 0) {
            for (int j = 0; j < n; ++j) {
                const int p = ipiv[j] - 1;
                if (p != j && p >= 0 && p < n) {
                    const float s = ts[j];
                    ts[j] = ts[p];
                    ts[p] = s;
                }
            }
        }
        __syncthreads();
    }
    for (int j = 0; j < n; ++j) {            // forward
        if (!LU) {
            if (tid == 0) ts[j] = ts[j] / F[off(j, j, ldf)];
            __syncthreads();
        }
        const float tj = ts[j];
        for (int i = j + 1 + tid; i < n; i += 256) ts[i] -= F[off(i, j, ldf)] * tj;
        __syncthreads();
    }
    for (int j = n - 1; j >= 0; --j) {       // backward
        if (tid == 0) ts[j] = ts[j] / F[off(j, j, ldf)];
        __syncthreads();
        const float tj = ts[j];
        for (int i = tid; i < j; i += 256)
            ts[i] -= (LU ? F[off(i, j, ldf)] : F[off(j, i, ldf)]) * tj;
        __syncthreads();
    }
    for (int i = tid; i < n; i += 256) t[i] = ts[i];
    __syncthreads();
}

struct SolveArgs {
    const float* A;
    const float* F;
    const long long* offsets;
    const int* n;
    const int* lda;
    const int* ipiv;
    const long long* ipiv_offsets;
    const double* B;
    double* X;
    const long long* vec_offsets;
    float* t;                  // FP32 staging, vec_offsets layout
    double* berr;
    const int* info;
    double tol;
    int batch;
    int sym;
};

// One refinement step per matrix (one workgroup per matrix):
// r = b - A x (FP64, A FP32), berr, and unless converged x += M^{-1} r.
// first: x = 0 on entry (x := M^{-1} b).
template <bool LU>
__global__ __launch_bounds__(256) void refine_step_kernel(SolveArgs a, int first)
{
    const int id = blockIdx.x;
    if (id >= a.batch) return;
    const int n = a.n[id];
    if (n <= 0 || n > kMaxOrder || a.info[id] != 0) return;
    if (!first && a.berr[id] <= a.tol) return;       // converged: skip
    const float* A = a.A + a.offsets[id];
    const float* F = a.F + a.offsets[id];
    const long long lda = a.lda[id];
    const double* b = a.B + a.vec_offsets[id];
    double* x = a.X + a.vec_offsets[id];
    float* t = a.t + a.vec_offsets[id];
    const int tid = threadIdx.x;
    __shared__ double red[256];
    __shared__ double red2[256];
    __shared__ double red3[256];
    // r = b - A x; norms
    double rmax = 0.0, xmax = 0.0, bmax = 0.0, amax = 0.0;
    for (int i = tid; i < n; i += 256) {
        double s = 0.0, as = 0.0;
        if (!first) {
            for (int c = 0; c < n; ++c) {
                const double av = (!LU && c > i) ? A[off(c, i, lda)] : A[off(i, c, lda)];
                s = fma(av, x[c], s);
                as += fabs(av);
            }
        }
        const double r = b[i] - s;
        t[i] = static_cast<float>(r);
        rmax = fmax(rmax, fabs(r));
        xmax = fmax(xmax, first ? 0.0 : fabs(x[i]));
        bmax = fmax(bmax, fabs(b[i]));
        amax = fmax(amax, as);
    }
    red[tid] = rmax;
    red2[tid] = xmax;
    red3[tid] = amax;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) {
            red[tid] = fmax(red[tid], red[tid + o]);
            red2[tid] = fmax(red2[tid], red2[tid + o]);
            red3[tid] = fmax(red3[tid], red3[tid + o]);
        }
        __syncthreads();
    }
    const double rn = red[0], xn = red2[0], an = red3[0];
    __syncthreads();
    red[tid] = bmax;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) red[tid] = fmax(red[tid], red[tid + o]);
        __syncthreads();
    }
    const double bn = red[0];
    if (!first) {
        const double den = an * xn + bn;
        const double e = den > 0.0 ? rn / den : (rn > 0.0 ? INFINITY : 0.0);
        if (tid == 0) a.berr[id] = e;
        if (e <= a.tol) return;
    } else if (tid == 0) {
        a.berr[id] = INFINITY;
    }
    __syncthreads();
    tri_solve<LU>(F, lda, n, LU ? a.ipiv + a.ipiv_offsets[id] : nullptr, t);
    for (int i = tid; i < n; i += 256) x[i] = (first ? 0.0 : x[i]) + static_cast<double>(t[i]);
}

// Final backward error of every matrix (after the last update).
template <bool LU>
__global__ __launch_bounds__(256) void berr_kernel(SolveArgs a)
{
    const int id = blockIdx.x;
    if (id >= a.batch) return;
    const int n = a.n[id];
    if (n <= 0 || n > kMaxOrder || a.info[id] != 0) return;
    if (a.berr[id] <= a.tol) return;
    const float* A = a.A + a.offsets[id];
    const long long lda = a.lda[id];
    const double* b = a.B + a.vec_offsets[id];
    const double* x = a.X + a.vec_offsets[id];
    const int tid = threadIdx.x;
    __shared__ double r1[256], r2[256], r3[256], r4[256];
    double rmax = 0.0, xmax = 0.0, bmax = 0.0, amax = 0.0;
    for (int i = tid; i < n; i += 256) {
        double s = 0.0, as = 0.0;
        for (int c = 0; c < n; ++c) {
            const double av = (!LU && c > i) ? A[off(c, i, lda)] : A[off(i, c, lda)];
            s = fma(av, x[c], s);
            as += fabs(av);
        }
        rmax = fmax(rmax, fabs(b[i] - s));
        xmax = fmax(xmax, fabs(x[i]));
        bmax = fmax(bmax, fabs(b[i]));
        amax = fmax(amax, as);
    }
    r1[tid] = rmax;
    r2[tid] = xmax;
    r3[tid] = bmax;
    r4[tid] = amax;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) {
            r1[tid] = fmax(r1[tid], r1[tid + o]);
            r2[tid] = fmax(r2[tid], r2[tid + o]);
            r3[tid] = fmax(r3[tid], r3[tid + o]);
            r4[tid] = fmax(r4[tid], r4[tid + o]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        const double den = r4[0] * r2[0] + r3[0];
        a.berr[id] = den > 0.0 ? r1[0] / den : (r1[0] > 0.0 ? INFINITY : 0.0);
    }
}

// F = A (copy of every matrix; lower triangle for SPD)
__global__ __launch_bounds__(256) void copy_batch_kernel(const float* A, float* F,
                                                         const long long* offsets, const int* n,
                                                         const int* lda, int batch, int sym)
{
    const int id = blockIdx.x;
    if (id >= batch) return;
    const int nn = n[id];
    if (nn <= 0 || nn > kMaxOrder) return;
    const long long ld = lda[id];
    const float* a = A + offsets[id];
    float* f = F + offsets[id];
    for (int c = 0; c < nn; ++c)
        for (int i = threadIdx.x; i < nn; i += 256)
            if (!sym || i >= c) f[off(i, c, ld)] = a[off(i, c, ld)];
}

// ---------------------------------------------------------------- host

struct BatchScratch {
    int* order = nullptr;
    unsigned* counter = nullptr;
    float* t = nullptr;
};

bool batch_scratch(mifact_handle h, int batch, size_t t_elems, BatchScratch* s)
{
    const size_t need = 256 + sizeof(int) * static_cast<size_t>(batch) + 256
                      + sizeof(float) * t_elems + 256;
    if (h->batch_ws == nullptr || h->batch_ws_bytes < need) {
        if (h->batch_ws != nullptr) {
            if (hipStreamSynchronize(h->stream) != hipSuccess) return false;
            (void)hipFree(h->batch_ws);
            h->batch_ws = nullptr;
            h->batch_ws_bytes = 0;
        }
        if (hipMalloc(&h->batch_ws, need) != hipSuccess) {
            h->batch_ws = nullptr;
            report_status(h, mifact_status_allocation_failed, "batch scratch allocation failed");
            return false;
        }
        h->batch_ws_bytes = need;
    }
    char* base = static_cast<char*>(h->batch_ws);
    s->counter = reinterpret_cast<unsigned*>(base);
    s->order = reinterpret_cast<int*>(base + 256);
    const size_t o = 256 + ((sizeof(int) * static_cast<size_t>(batch) + 255) & ~size_t(255));
    s->t = reinterpret_cast<float*>(base + o);
    return true;
}

void launch_batched(mifact_handle h, bool lu, LowPrecision prec, int batch, float* A,
                    const long long* offsets, const int* n, const int* lda, int* ipiv,
                    const long long* ipiv_offsets, int* info, const BatchScratch& s)
{
    hipLaunchKernelGGL(order_kernel, dim3(1), dim3(1024), 0, h->stream, n, batch, s.order,
                       s.counter);
    MIFACT_LAUNCHED(h);
    BatchArgs a{lu ? h->batched_small_max_lu : h->batched_small_max_chol,
                A, offsets, n, lda, ipiv, ipiv_offsets, info, s.order, s.counter, batch};
    // the LU workgroup holds a 1024 x 32 panel in LDS: one per CU
    const int grid = imax(1, imin(batch, h->num_cu * (lu ? 1 : 4)));
    StageScope scope(h, kStageBatched, 0.0, 0.0);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, h->stream, a);
    };
    if (lu) {
        if (prec == kPrecBF16) go(batched_kernel<kPrecBF16, true>);
        else if (prec == kPrecFP16) go(batched_kernel<kPrecFP16, true>);
        else go(batched_kernel<kPrecFP32, true>);
    } else {
        if (prec == kPrecBF16) go(batched_kernel<kPrecBF16, false>);
        else if (prec == kPrecFP16) go(batched_kernel<kPrecFP16, false>);
        else go(batched_kernel<kPrecFP32, false>);
    }
    MIFACT_LAUNCHED(h);
}

bool valid_prec(int p) { return p == kPrecFP32 || p == kPrecBF16 || p == kPrecFP16; }

void batched_entry(mifact_handle h, bool lu, int prec, int batch, float* A,
                   const long long* offsets, const int* n, const int* lda, int* ipiv,
                   const long long* ipiv_offsets, int* info)
{
    if (h == nullptr) {
        report_status(nullptr, mifact_status_invalid_handle, "batched factorization");
        return;
    }
    if (batch <= 0) return;
    if (!valid_prec(prec) || A == nullptr || offsets == nullptr || n == nullptr
        || lda == nullptr || info == nullptr || (lu && (ipiv == nullptr || ipiv_offsets == nullptr))) {
        report_status(h, mifact_status_invalid_value, "batched factorization: arguments");
        return;
    }
    BatchScratch s;
    if (!batch_scratch(h, batch, 0, &s)) return;
    launch_batched(h, lu, static_cast<LowPrecision>(prec), batch, A, offsets, n, lda, ipiv,
                   ipiv_offsets, info, s);
}

}  // namespace
}  // namespace mifact

using namespace mifact;

extern "C" {

void mifact_potrf_batched(mifact_handle handle, int batch, float* A, const long long* offsets,
                          const int* n, const int* lda, int* info)
{
    batched_entry(handle, false, kPrecFP32, batch, A, offsets, n, lda, nullptr, nullptr, info);
}

void mifact_getrf_batched(mifact_handle handle, int batch, float* A, const long long* offsets,
                          const int* n, const int* lda, int* ipiv, const long long* ipiv_offsets,
                          int* info)
{
    batched_entry(handle, true, kPrecFP32, batch, A, offsets, n, lda, ipiv, ipiv_offsets, info);
}

void mifact_potrf_batched_ex(mifact_handle handle, mifact_precision precision, int batch,
                             float* A, const long long* offsets, const int* n, const int* lda,
                             int* info)
{
    batched_entry(handle, false, precision, batch, A, offsets, n, lda, nullptr, nullptr, info);
}

void mifact_getrf_batched_ex(mifact_handle handle, mifact_precision precision, int batch,
                             float* A, const long long* offsets, const int* n, const int* lda,
                             int* ipiv, const long long* ipiv_offsets, int* info)
{
    batched_entry(handle, true, precision, batch, A, offsets, n, lda, ipiv, ipiv_offsets, info);
}

mifact_status mifact_gesv_batched_refine(mifact_handle handle, int sym,
                                         mifact_precision precision, int batch, const float* A,
                                         float* F, const long long* offsets, const int* n,
                                         const int* lda, int* ipiv, const long long* ipiv_offsets,
                                         const double* B, double* X,
                                         const long long* vec_offsets, long long vec_elems,
                                         int iterations, double tolerance, double* berr,
                                         int* info)
{
    mifact_handle h = handle;
    if (h == nullptr) {
        report_status(nullptr, mifact_status_invalid_handle, "mifact_gesv_batched_refine");
        return mifact_status_invalid_handle;
    }
    if (batch <= 0) return mifact_status_success;
    const bool lu = sym == 0;
    if (!valid_prec(precision) || iterations < 0 || vec_elems <= 0 || A == nullptr
        || F == nullptr || offsets == nullptr || n == nullptr || lda == nullptr || B == nullptr
        || X == nullptr || vec_offsets == nullptr || berr == nullptr || info == nullptr
        || (lu && (ipiv == nullptr || ipiv_offsets == nullptr))) {
        report_status(h, mifact_status_invalid_value, "mifact_gesv_batched_refine: arguments");
        return mifact_status_invalid_value;
    }
    const mifact_status before = h->last_error;
    h->last_error = mifact_status_success;
    auto finish = [&]() {
        mifact_status s = h->last_error;
        if (before != mifact_status_success) h->last_error = before;
        return s;
    };
    BatchScratch s;
    if (!batch_scratch(h, batch, static_cast<size_t>(vec_elems), &s)) return finish();
    hipLaunchKernelGGL(copy_batch_kernel, dim3(batch), dim3(256), 0, h->stream, A, F, offsets, n,
                       lda, batch, sym ? 1 : 0);
    MIFACT_HIP_RET(h, hipGetLastError(), finish());
    launch_batched(h, lu, static_cast<LowPrecision>(precision), batch, F, offsets, n, lda, ipiv,
                   ipiv_offsets, info, s);
    if (h->last_error != mifact_status_success) return finish();

    // The refinement loop (first solve, `iterations` steps, final backward
    // errors) is captured once on a private stream and replayed on the
    // handle's stream as one graph launch.
    if (h->capture_stream == nullptr) {
        MIFACT_HIP_RET(h, hipStreamCreateWithFlags(&h->capture_stream, hipStreamNonBlocking),
                       finish());
    }
    SolveArgs sa{A, F, offsets, n, lda, ipiv, ipiv_offsets, B, X, vec_offsets, s.t, berr, info,
                 tolerance, batch, sym};
    hipStream_t cs = h->capture_stream;
    MIFACT_HIP_RET(h, hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), finish());
    for (int it = 0; it <= iterations; ++it) {
        if (lu) hipLaunchKernelGGL(refine_step_kernel<true>, dim3(batch), dim3(256), 0, cs, sa, it == 0);
        else hipLaunchKernelGGL(refine_step_kernel<false>, dim3(batch), dim3(256), 0, cs, sa, it == 0);
    }
    if (lu) hipLaunchKernelGGL(berr_kernel<true>, dim3(batch), dim3(256), 0, cs, sa);
    else hipLaunchKernelGGL(berr_kernel<false>, dim3(batch), dim3(256), 0, cs, sa);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &graph);
    MIFACT_HIP_RET(h, ec, finish());
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    MIFACT_HIP_RET(h, ei, finish());
    {
        StageScope scope(h, kStageBatched, 0.0, 0.0);
        const hipError_t el = hipGraphLaunch(exec, h->stream);
        if (el != hipSuccess) {
            (void)hipGraphExecDestroy(exec);
            MIFACT_HIP_RET(h, el, finish());
        }
    }
    // the executable graph may be destroyed once its launch is enqueued
    (void)hipGraphExecDestroy(exec);
    return finish();
}

}  // extern "C"

// MFMA update GEMM / SYRK tiles for gfx950.
//
//   C(M x N) += alpha * A(M x K) * op(B),   column-major, 64-bit offsets.
//
// FP32 runs on v_mfma_f32_32x32x2_f32, FP64 on v_mfma_f64_16x16x4_f64.  A
// 256-thread workgroup (four wave64s) owns a BM x BN tile of C; each wave
// owns a (BM/WAVES_M) x (BN/WAVES_N) sub-tile made of MFMA blocks whose
// accumulators stay in registers for the whole K loop.  K advances in BK
// slices staged global -> registers -> LDS with double buffering (one
// barrier per slice; the next slice's global loads are in flight while the
// current slice is multiplied).  Tile order is XCD-aware: the hardware
// deals workgroup i to XCD i % 8, so consecutive tile indices are remapped
// to give each XCD a contiguous run of tiles (shared A/B panels stay in that
// XCD's L2).  kShapeLower enumerates only tiles on or below the diagonal and
// masks the strictly-upper part of diagonal tiles (SYRK / trapezoid).
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mifact {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;

// v_mfma_f32_32x32x2_f32: A lane l holds A[l%32][l/32]; B lane l holds
// B[l/32][l%32]; acc[r] holds D[8*(r/4) + 4*(l/32) + r%4][l%32].
template <> struct Mfma<float> {
    using acc_t = floatx16;
    static constexpr int BLK = 32;
    static constexpr int KSTEP = 2;
    static constexpr int GROUPS = 4;      // groups of 4 output rows per lane
    static constexpr int ROW_STRIDE = 1;  // rows of a group are consecutive
    __device__ static int op_row(int lane) { return lane & 31; }
    __device__ static int op_k(int lane) { return lane >> 5; }
    __device__ static int out_row(int lane, int g) { return g * 8 + (lane >> 5) * 4; }
    __device__ static int out_col(int lane) { return lane & 31; }
    __device__ static acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
};

// v_mfma_f64_16x16x4_f64: A lane l holds A[l%16][l/16]; B lane l holds
// B[l/16][l%16]; acc[r] holds D[l/16 + 4*r][l%16] (rows interleaved by 4).
template <> struct Mfma<double> {
    using acc_t = doublex4;
    static constexpr int BLK = 16;
    static constexpr int KSTEP = 4;
    static constexpr int GROUPS = 1;
    static constexpr int ROW_STRIDE = 4;
    __device__ static int op_row(int lane) { return lane & 15; }
    __device__ static int op_k(int lane) { return lane >> 4; }
    __device__ static int out_row(int lane, int) { return lane >> 4; }
    __device__ static int out_col(int lane) { return lane & 15; }
    __device__ static acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

template <typename T> struct Vec;
template <> struct Vec<float> { using type = float4; static constexpr int N = 4; };
template <> struct Vec<double> { using type = double2; static constexpr int N = 2; };

template <typename T>
struct GemmArgs {
    int M, N, K;
    T alpha;
    const T* A;
    long long lda;
    const T* B;
    long long ldb;
    T* C;
    long long ldc;
    int tiles_m, tiles_n;
};

// Full tile grid: XCD remap, then GROUP_M row-tiles per group.
__device__ inline void tile_full(int bid, int tiles_m, int tiles_n, int& tm, int& tn)
{
    const int total = tiles_m * tiles_n;
    const int per = total / 8;
    int L = bid;
    if (bid < per * 8) {
        L = (bid % 8) * per + bid / 8;
    }
    constexpr int GROUP_M = 8;
    const int in_group = GROUP_M * tiles_n;
    const int g = L / in_group;
    const int first_m = g * GROUP_M;
    const int rem = tiles_m - first_m;
    const int gsz = rem < GROUP_M ? rem : GROUP_M;
    const int r = L - g * in_group;
    tm = first_m + r % gsz;
    tn = r / gsz;
}

// Lower trapezoid (square tiles, tm >= tn), row-major over the kept tiles.
__device__ inline void tile_lower(int bid, int tiles_m, int tiles_n, int& tm, int& tn)
{
    const long long tn_l = tiles_n;
    const long long tri = tn_l * (tn_l + 1) / 2;
    const long long total = tri + static_cast<long long>(tiles_m - tiles_n) * tn_l;
    const long long per = total / 8;
    long long L = bid;
    if (bid < per * 8) {
        L = (bid % 8) * per + bid / 8;
    }
    if (L < tri) {
        long long r = static_cast<long long>((sqrtf(8.0f * static_cast<float>(L) + 1.0f) - 1.0f) * 0.5f);
        while (r * (r + 1) / 2 > L) --r;
        while ((r + 1) * (r + 2) / 2 <= L) ++r;
        tm = static_cast<int>(r);
        tn = static_cast<int>(L - r * (r + 1) / 2);
    } else {
        const long long q = L - tri;
        tm = tiles_n + static_cast<int>(q / tn_l);
        tn = static_cast<int>(q % tn_l);
    }
}

template <typename T, int BM, int BN, int BK, int WAVES_M, int WAVES_N, int OP,
          int SHAPE, bool VEC, bool BETA0>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs<T> p)
{
    using MF = Mfma<T>;
    using acc_t = typename MF::acc_t;
    using vec_t = typename Vec<T>::type;
    constexpr int VW = Vec<T>::N;
    constexpr int BLK = MF::BLK;
    constexpr int WM = BM / WAVES_M;
    constexpr int WN = BN / WAVES_N;
    constexpr int FM = WM / BLK;
    constexpr int FN = WN / BLK;
    constexpr int PAD = (sizeof(T) == 4) ? 32 : 8;
    constexpr int SA = BM + PAD;
    constexpr int SB = BN + PAD;
    constexpr int A_VECS = BM * BK / VW / 256;  // vectors per thread, A tile
    constexpr int B_VECS = BN * BK / VW / 256;
    static_assert(A_VECS >= 1 && B_VECS >= 1, "tile too small for 256 threads");
    static_assert(WAVES_M * WAVES_N == 4, "four waves per workgroup");

    __shared__ T sA[2][BK][SA];
    __shared__ T sB[2][BK][SB];

    int tm, tn;
    if (SHAPE == kShapeLower) {
        tile_lower(blockIdx.x, p.tiles_m, p.tiles_n, tm, tn);
    } else {
        tile_full(blockIdx.x, p.tiles_m, p.tiles_n, tm, tn);
    }
    const int m0 = tm * BM;
    const int n0 = tn * BN;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WAVES_M;
    const int wn = wave / WAVES_M;

    const int M = p.M, N = p.N, K = p.K;
    // No __restrict__: in overwrite mode C may alias A or B (each workgroup
    // reads every input element of its tile before the epilogue writes).
    const T* A = p.A;
    const T* B = p.B;

    vec_t ra[A_VECS];
    vec_t rb[B_VECS];

    // Global -> registers for the K slice starting at k0 (zero padded).
    auto load_slice = [&](int k0) {
#pragma unroll
        for (int v = 0; v < A_VECS; ++v) {
            const int f = tid + v * 256;
            const int kk = f / (BM / VW);
            const int ii = (f % (BM / VW)) * VW;
            const int row = m0 + ii;
            const int k = k0 + kk;
            T tmp[VW];
            if (VEC && k < K && row + VW <= M) {
                const vec_t x = *reinterpret_cast<const vec_t*>(A + off(row, k, p.lda));
#pragma unroll
                for (int e = 0; e < VW; ++e) tmp[e] = x[e];
            } else {
#pragma unroll
                for (int e = 0; e < VW; ++e)
                    tmp[e] = (k < K && row + e < M) ? A[off(row + e, k, p.lda)] : T(0);
            }
#pragma unroll
            for (int e = 0; e < VW; ++e) ra[v][e] = tmp[e];
        }
#pragma unroll
        for (int v = 0; v < B_VECS; ++v) {
            const int f = tid + v * 256;
            T tmp[VW];
            if (OP == kOpNT) {
                const int kk = f / (BN / VW);
                const int jj = (f % (BN / VW)) * VW;
                const int col = n0 + jj;
                const int k = k0 + kk;
                if (VEC && k < K && col + VW <= N) {
                    const vec_t x = *reinterpret_cast<const vec_t*>(B + off(col, k, p.ldb));
#pragma unroll
                    for (int e = 0; e < VW; ++e) tmp[e] = x[e];
                } else {
#pragma unroll
                    for (int e = 0; e < VW; ++e)
                        tmp[e] = (k < K && col + e < N) ? B[off(col + e, k, p.ldb)] : T(0);
                }
            } else {
                const int jj = f / (BK / VW);
                const int kk = (f % (BK / VW)) * VW;
                const int col = n0 + jj;
                const int k = k0 + kk;
                if (VEC && col < N && k + VW <= K) {
                    const vec_t x = *reinterpret_cast<const vec_t*>(B + off(k, col, p.ldb));
#pragma unroll
                    for (int e = 0; e < VW; ++e) tmp[e] = x[e];
                } else {
#pragma unroll
                    for (int e = 0; e < VW; ++e)
                        tmp[e] = (col < N && k + e < K) ? B[off(k + e, col, p.ldb)] : T(0);
                }
            }
#pragma unroll
            for (int e = 0; e < VW; ++e) rb[v][e] = tmp[e];
        }
    };

    // Registers -> LDS buffer `buf`.
    auto store_slice = [&](int buf) {
#pragma unroll
        for (int v = 0; v < A_VECS; ++v) {
            const int f = tid + v * 256;
            const int kk = f / (BM / VW);
            const int ii = (f % (BM / VW)) * VW;
            *reinterpret_cast<vec_t*>(&sA[buf][kk][ii]) = ra[v];
        }
#pragma unroll
        for (int v = 0; v < B_VECS; ++v) {
            const int f = tid + v * 256;
            if (OP == kOpNT) {
                const int kk = f / (BN / VW);
                const int jj = (f % (BN / VW)) * VW;
                *reinterpret_cast<vec_t*>(&sB[buf][kk][jj]) = rb[v];
            } else {
                const int jj = f / (BK / VW);
                const int kk = (f % (BK / VW)) * VW;
#pragma unroll
                for (int e = 0; e < VW; ++e) sB[buf][kk + e][jj] = rb[v][e];
            }
        }
    };

    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < static_cast<int>(sizeof(acc_t) / sizeof(T)); ++r)
                acc[i][j][r] = T(0);

    const int nk = (K + BK - 1) / BK;
    if (nk > 0) {
        load_slice(0);
        store_slice(0);
    }
    __syncthreads();

    const int orow = MF::op_row(lane);
    const int ok = MF::op_k(lane);
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) {
            load_slice((kt + 1) * BK);
        }
#pragma unroll
        for (int s = 0; s < BK / MF::KSTEP; ++s) {
            const int k = s * MF::KSTEP + ok;
            T a[FM], b[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) a[i] = sA[cur][k][wm * WM + i * BLK + orow];
#pragma unroll
            for (int j = 0; j < FN; ++j) b[j] = sB[cur][k][wn * WN + j * BLK + orow];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(a[i], b[j], acc[i][j]);
        }
        if (kt + 1 < nk) {
            store_slice(cur ^ 1);
        }
        __syncthreads();
    }

    // Epilogue: C += alpha * acc, in groups of 4 consecutive rows.
    const T alpha = p.alpha;
    T* C = p.C;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int col = n0 + wn * WN + j * BLK + MF::out_col(lane);
            if (col >= N) continue;
#pragma unroll
            for (int g = 0; g < MF::GROUPS; ++g) {
                constexpr int RS = MF::ROW_STRIDE;
                const int row = m0 + wm * WM + i * BLK + MF::out_row(lane, g);
                T* c = C + off(row, col, p.ldc);
                const bool full = (row + 3 * RS < M) && (SHAPE != kShapeLower || row >= col);
                if (full) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        if (BETA0) c[e * RS] = alpha * acc[i][j][g * 4 + e];
                        else c[e * RS] += alpha * acc[i][j][g * 4 + e];
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int r = row + e * RS;
                        if (r < M && (SHAPE != kShapeLower || r >= col)) {
                            if (BETA0) c[e * RS] = alpha * acc[i][j][g * 4 + e];
                            else c[e * RS] += alpha * acc[i][j][g * 4 + e];
                        }
                    }
                }
            }
        }
    }
}

template <typename T, int BM, int BN, int BK, int WAVES_M, int WAVES_N>
static void launch_cfg(mifact_handle h, GemmOp op, GemmShape shape, bool vec, bool beta0,
                       const GemmArgs<T>& a0)
{
    GemmArgs<T> a = a0;
    a.tiles_m = ceil_div(a.M, BM);
    a.tiles_n = ceil_div(a.N, BN);
    long long blocks;
    if (shape == kShapeLower) {
        const long long tn = a.tiles_n;
        blocks = tn * (tn + 1) / 2 + static_cast<long long>(a.tiles_m - a.tiles_n) * tn;
    } else {
        blocks = static_cast<long long>(a.tiles_m) * a.tiles_n;
    }
    if (blocks <= 0) return;
    const double mnk_c = (shape == kShapeLower)
        ? (static_cast<double>(a.M) * a.N - 0.5 * static_cast<double>(a.N) * (a.N - 1))
        : static_cast<double>(a.M) * a.N;
    StageScope scope(h, kStageGemm, 2.0 * mnk_c * a.K,
                     sizeof(T) * ((static_cast<double>(a.M) + a.N) * a.K + 2.0 * mnk_c));
    const dim3 grid(static_cast<unsigned>(blocks));
    const dim3 block(256);
#define MIFACT_GEMM_LAUNCH(OPV, SHV, VV)                                                    \
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WAVES_M, WAVES_N, OPV, SHV, VV, false>),    \
                       grid, block, 0, h->stream, a)
#define MIFACT_GEMM_LAUNCH0(OPV, VV)                                                         \
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WAVES_M, WAVES_N, OPV, kShapeFull, VV, true>), \
                       grid, block, 0, h->stream, a)
    if (beta0) {
        if (op == kOpNT) {
            if (vec) MIFACT_GEMM_LAUNCH0(kOpNT, true);
            else MIFACT_GEMM_LAUNCH0(kOpNT, false);
        } else {
            if (vec) MIFACT_GEMM_LAUNCH0(kOpNN, true);
            else MIFACT_GEMM_LAUNCH0(kOpNN, false);
        }
    } else if (op == kOpNT) {
        if (shape == kShapeFull) {
            if (vec) MIFACT_GEMM_LAUNCH(kOpNT, kShapeFull, true);
            else MIFACT_GEMM_LAUNCH(kOpNT, kShapeFull, false);
        } else {
            if (vec) MIFACT_GEMM_LAUNCH(kOpNT, kShapeLower, true);
            else MIFACT_GEMM_LAUNCH(kOpNT, kShapeLower, false);
        }
    } else {
        if (shape == kShapeFull) {
            if (vec) MIFACT_GEMM_LAUNCH(kOpNN, kShapeFull, true);
            else MIFACT_GEMM_LAUNCH(kOpNN, kShapeFull, false);
        } else {
            if (vec) MIFACT_GEMM_LAUNCH(kOpNN, kShapeLower, true);
            else MIFACT_GEMM_LAUNCH(kOpNN, kShapeLower, false);
        }
    }
#undef MIFACT_GEMM_LAUNCH
#undef MIFACT_GEMM_LAUNCH0
    MIFACT_LAUNCHED(h);
}

template <typename T>
static bool aligned_for_vectors(const void* p, long long ld)
{
    constexpr int VW = Vec<T>::N;
    return (reinterpret_cast<uintptr_t>(p) % (VW * sizeof(T)) == 0) && (ld % VW == 0);
}

template <typename T>
static void dispatch(mifact_handle h, GemmOp op, GemmShape shape, bool beta0, int M, int N, int K,
                     T alpha, const T* A, long long lda, const T* B, long long ldb, T* C,
                     long long ldc)
{
    if (M <= 0 || N <= 0 || K <= 0) return;
    if (shape == kShapeLower && M < N) {
        // Lower trapezoid assumes the diagonal runs from (0,0) within M >= N.
        report_status(h, mifact_status_invalid_value, "lower-shape gemm needs M >= N");
        return;
    }
    GemmArgs<T> a{M, N, K, alpha, A, lda, B, ldb, C, ldc, 0, 0};
    const bool vec = aligned_for_vectors<T>(A, lda) && aligned_for_vectors<T>(B, ldb);
    if constexpr (sizeof(T) == 4) {
        const long long big = static_cast<long long>(ceil_div(M, 128)) * ceil_div(N, 128);
        if (!beta0 && N > 64 && M > 64 && big >= 2LL * h->num_cu) {
            launch_cfg<T, 128, 128, 16, 2, 2>(h, op, shape, vec, beta0, a);
        } else {
            launch_cfg<T, 64, 64, 16, 2, 2>(h, op, shape, vec, beta0, a);
        }
    } else {
        launch_cfg<T, 64, 64, 16, 2, 2>(h, op, shape, vec, beta0, a);
    }
}

template <typename T>
void gemm(mifact_handle h, GemmOp op, GemmShape shape, int M, int N, int K, T alpha, const T* A,
          long long lda, const T* B, long long ldb, T* C, long long ldc)
{
    dispatch<T>(h, op, shape, false, M, N, K, alpha, A, lda, B, ldb, C, ldc);
}

template <typename T>
void gemm_overwrite(mifact_handle h, GemmOp op, int M, int N, int K, T alpha, const T* A,
                    long long lda, const T* B, long long ldb, T* C, long long ldc)
{
    dispatch<T>(h, op, kShapeFull, true, M, N, K, alpha, A, lda, B, ldb, C, ldc);
}

template void gemm<float>(mifact_handle, GemmOp, GemmShape, int, int, int, float,
                          const float*, long long, const float*, long long, float*,
                          long long);
template void gemm<double>(mifact_handle, GemmOp, GemmShape, int, int, int, double,
                           const double*, long long, const double*, long long,
                           double*, long long);
template void gemm_overwrite<float>(mifact_handle, GemmOp, int, int, int, float, const float*,
                                    long long, const float*, long long, float*, long long);
template void gemm_overwrite<double>(mifact_handle, GemmOp, int, int, int, double, const double*,
                                     long long, const double*, long long, double*, long long);

}  // namespace mifact

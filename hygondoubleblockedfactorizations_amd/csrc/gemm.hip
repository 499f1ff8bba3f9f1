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

#include <type_traits>

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
    int cvec;  // C base and ldc allow vector access
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
    // CHECK = false: interior tile and a full K slice, no bounds tests.
    auto load_slice = [&](int k0, auto check_tag) {
        constexpr bool CHECK = decltype(check_tag)::value;
#pragma unroll
        for (int v = 0; v < A_VECS; ++v) {
            const int f = tid + v * 256;
            T tmp[VW];
            if (OP != kOpTN) {
                const int kk = f / (BM / VW);
                const int ii = (f % (BM / VW)) * VW;
                const int row = m0 + ii;
                const int k = k0 + kk;
                if (VEC && (!CHECK || (k < K && row + VW <= M))) {
                    const vec_t x = *reinterpret_cast<const vec_t*>(A + off(row, k, p.lda));
#pragma unroll
                    for (int e = 0; e < VW; ++e) tmp[e] = x[e];
                } else {
#pragma unroll
                    for (int e = 0; e < VW; ++e)
                        tmp[e] = (k < K && row + e < M) ? A[off(row + e, k, p.lda)] : T(0);
                }
            } else {
                // A stored K x M (op(A) = A^T): K-contiguous vectors
                const int ii = f / (BK / VW);
                const int kk = (f % (BK / VW)) * VW;
                const int row = m0 + ii;
                const int k = k0 + kk;
                if (VEC && (!CHECK || (row < M && k + VW <= K))) {
                    const vec_t x = *reinterpret_cast<const vec_t*>(A + off(k, row, p.lda));
#pragma unroll
                    for (int e = 0; e < VW; ++e) tmp[e] = x[e];
                } else {
#pragma unroll
                    for (int e = 0; e < VW; ++e)
                        tmp[e] = (row < M && k + e < K) ? A[off(k + e, row, p.lda)] : T(0);
                }
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
                if (VEC && (!CHECK || (k < K && col + VW <= N))) {
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
                if (VEC && (!CHECK || (col < N && k + VW <= K))) {
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
            if (OP != kOpTN) {
                const int kk = f / (BM / VW);
                const int ii = (f % (BM / VW)) * VW;
                *reinterpret_cast<vec_t*>(&sA[buf][kk][ii]) = ra[v];
            } else {
                const int ii = f / (BK / VW);
                const int kk = (f % (BK / VW)) * VW;
#pragma unroll
                for (int e = 0; e < VW; ++e) sA[buf][kk + e][ii] = ra[v][e];
            }
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
    const int nk_full = K / BK;  // slices with no K tail
    const bool interior = VEC && (m0 + BM <= M) && (n0 + BN <= N);
    const int orow = MF::op_row(lane);
    const int ok = MF::op_k(lane);

    auto multiply = [&](int cur) {
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
    };

    // Full slices of interior tiles load without bounds tests (uniform branch).
    auto fetch = [&](int kt) {
        if (interior && kt < nk_full) load_slice(kt * BK, std::false_type{});
        else load_slice(kt * BK, std::true_type{});
    };

    if (nk > 0) {
        fetch(0);
        store_slice(0);
    }
    __syncthreads();

    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) fetch(kt + 1);
        multiply(cur);
        if (kt + 1 < nk) store_slice(cur ^ 1);
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
                if constexpr (RS == 1 && sizeof(T) == 4) {
                    // four consecutive rows of one column: one 16-byte access
                    if (full && p.cvec) {
                        float4* c4 = reinterpret_cast<float4*>(c);
                        float4 v = BETA0 ? float4{0.f, 0.f, 0.f, 0.f} : *c4;
                        v.x += alpha * acc[i][j][g * 4 + 0];
                        v.y += alpha * acc[i][j][g * 4 + 1];
                        v.z += alpha * acc[i][j][g * 4 + 2];
                        v.w += alpha * acc[i][j][g * 4 + 3];
                        *c4 = v;
                        continue;
                    }
                }
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
    // trailing updates on the large tiles are their own stage (the dominant
    // kernel of the factorizations); small and in-place launches stay "gemm"
    const Stage stage = (BM >= 128 && !beta0) ? kStageGemmUpdate : kStageGemm;
    StageScope scope(h, stage, 2.0 * mnk_c * a.K,
                     sizeof(T) * ((static_cast<double>(a.M) + a.N) * a.K + 2.0 * mnk_c));
    const dim3 grid(static_cast<unsigned>(blocks));
    const dim3 block(256);
    auto go = [&](auto op_tag, auto shape_tag, auto vec_tag, auto beta_tag) {
        constexpr int OPV = decltype(op_tag)::value;
        constexpr int SHV = decltype(shape_tag)::value;
        constexpr bool VV = decltype(vec_tag)::value;
        constexpr bool B0 = decltype(beta_tag)::value;
        hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WAVES_M, WAVES_N, OPV, SHV, VV, B0>), grid,
                           block, 0, h->stream, a);
    };
    using NT = std::integral_constant<int, kOpNT>;
    using NN = std::integral_constant<int, kOpNN>;
    using TN = std::integral_constant<int, kOpTN>;
    using FULL = std::integral_constant<int, kShapeFull>;
    using LOWER = std::integral_constant<int, kShapeLower>;
    auto by_vec = [&](auto op_tag, auto shape_tag, auto beta_tag) {
        if (vec) go(op_tag, shape_tag, std::true_type{}, beta_tag);
        else go(op_tag, shape_tag, std::false_type{}, beta_tag);
    };
    if (beta0) {
        if (op == kOpNT) by_vec(NT{}, FULL{}, std::true_type{});
        else if (op == kOpNN) by_vec(NN{}, FULL{}, std::true_type{});
        else by_vec(TN{}, FULL{}, std::true_type{});
    } else if (op == kOpTN) {
        by_vec(TN{}, FULL{}, std::false_type{});
    } else if (op == kOpNT) {
        if (shape == kShapeFull) by_vec(NT{}, FULL{}, std::false_type{});
        else by_vec(NT{}, LOWER{}, std::false_type{});
    } else {
        if (shape == kShapeFull) by_vec(NN{}, FULL{}, std::false_type{});
        else by_vec(NN{}, LOWER{}, std::false_type{});
    }
    MIFACT_LAUNCHED(h);
}

template <typename T>
static bool aligned_for_vectors(const void* p, long long ld)
{
    constexpr int VW = Vec<T>::N;
    return (reinterpret_cast<uintptr_t>(p) % (VW * sizeof(T)) == 0) && (ld % VW == 0);
}

// ------------------------------------------------------------------ low precision
// C(fp32) += alpha * A * op(B) with A, B read as FP32 and multiplied on
// v_mfma_f32_32x32x16_{bf16,f16} (FP32 accumulation).  The FP32 tiles are
// converted while they are staged into LDS, where rows are stored
// k-contiguous: each lane's MFMA operand (8 consecutive k of one row) is a
// single 16-byte LDS read.  A lane l of the 32x32x16 operands holds
// row/column l%32 and k = 8*(l/32) .. +7; the accumulator layout is that of
// the FP32 32x32 instruction.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

template <typename TC> struct LowMfma;
template <> struct LowMfma<__bf16> {
    using v8 = bf16x8;
    __device__ static floatx16 mma(v8 a, v8 b, floatx16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};
template <> struct LowMfma<_Float16> {
    using v8 = halfx8;
    __device__ static floatx16 mma(v8 a, v8 b, floatx16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};

template <typename TC>
__device__ inline unsigned int pack2(float lo, float hi)
{
    const TC a = static_cast<TC>(lo);
    const TC b = static_cast<TC>(hi);
    const unsigned short ua = __builtin_bit_cast(unsigned short, a);
    const unsigned short ub = __builtin_bit_cast(unsigned short, b);
    return static_cast<unsigned int>(ua) | (static_cast<unsigned int>(ub) << 16);
}

template <typename TC, int OP, int SHAPE>
__global__ __launch_bounds__(256) void gemm_lowp_kernel(GemmArgs<float> p)
{
    using LM = LowMfma<TC>;
    using v8 = typename LM::v8;
    constexpr int BM = 128, BN = 128, BK = 32;
    constexpr int LDK = BK + 8;       // 80-byte rows: 16-byte aligned, conflict-free b128 reads
    constexpr int FM = 2, FN = 2;     // wave tile 64 x 64 = 2 x 2 blocks of 32 x 32
    __shared__ __attribute__((aligned(16))) TC sA[2][BM][LDK];
    __shared__ __attribute__((aligned(16))) TC sB[2][BN][LDK];

    int tm, tn;
    if (SHAPE == kShapeLower) tile_lower(blockIdx.x, p.tiles_m, p.tiles_n, tm, tn);
    else tile_full(blockIdx.x, p.tiles_m, p.tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % 2, wn = wave / 2;
    const int M = p.M, N = p.N, K = p.K;
    const float* A = p.A;
    const float* B = p.B;
    const bool vec = p.cvec >> 1;     // operands vector-aligned (bit 1)

    // staging registers: A and NT-B as (4 rows) x (2 k) pairs, NN-B as (1 col) x (4 k)
    float4 ra[2][2];
    float4 rb[4];

    // M-contiguous operand: item f -> rows r0 + 4 (f % 32) .. +3, k = k0 + 2 (f / 32), +1
    auto load_mk = [&](const float* X, long long ldx, int R, int r0, int k0, float4 (&dst)[2][2]) {
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const int f = tid + v * 256;
            const int r = r0 + 4 * (f & 31);
            const int k = k0 + 2 * (f >> 5);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int kk = k + h;
                float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
                if (kk < K) {
                    if (vec && r + 4 <= R) {
                        x = *reinterpret_cast<const float4*>(X + off(r, kk, ldx));
                    } else {
                        if (r + 0 < R) x.x = X[off(r + 0, kk, ldx)];
                        if (r + 1 < R) x.y = X[off(r + 1, kk, ldx)];
                        if (r + 2 < R) x.z = X[off(r + 2, kk, ldx)];
                        if (r + 3 < R) x.w = X[off(r + 3, kk, ldx)];
                    }
                }
                dst[v][h] = x;
            }
        }
    };
    auto store_mk = [&](TC (*S)[LDK], const float4 (&src)[2][2]) {
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const int f = tid + v * 256;
            const int r = 4 * (f & 31);
            const int k = 2 * (f >> 5);
            const float4 a = src[v][0], b = src[v][1];
            *reinterpret_cast<unsigned int*>(&S[r + 0][k]) = pack2<TC>(a.x, b.x);
            *reinterpret_cast<unsigned int*>(&S[r + 1][k]) = pack2<TC>(a.y, b.y);
            *reinterpret_cast<unsigned int*>(&S[r + 2][k]) = pack2<TC>(a.z, b.z);
            *reinterpret_cast<unsigned int*>(&S[r + 3][k]) = pack2<TC>(a.w, b.w);
        }
    };
    // K-contiguous operand (NN B): item f -> column c0 + f / 8, k = k0 + 4 (f % 8) .. +3
    auto load_kn = [&](int k0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int f = tid + v * 256;
            const int c = n0 + (f >> 3);
            const int k = k0 + 4 * (f & 7);
            float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
            if (c < N) {
                if (vec && k + 4 <= K) {
                    x = *reinterpret_cast<const float4*>(B + off(k, c, p.ldb));
                } else {
                    if (k + 0 < K) x.x = B[off(k + 0, c, p.ldb)];
                    if (k + 1 < K) x.y = B[off(k + 1, c, p.ldb)];
                    if (k + 2 < K) x.z = B[off(k + 2, c, p.ldb)];
                    if (k + 3 < K) x.w = B[off(k + 3, c, p.ldb)];
                }
            }
            rb[v] = x;
        }
    };
    auto store_kn = [&](TC (*S)[LDK]) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int f = tid + v * 256;
            const int c = f >> 3;
            const int k = 4 * (f & 7);
            uint2 w;
            w.x = pack2<TC>(rb[v].x, rb[v].y);
            w.y = pack2<TC>(rb[v].z, rb[v].w);
            *reinterpret_cast<uint2*>(&S[c][k]) = w;
        }
    };
    float4 rbt[2][2];
    auto fetch = [&](int k0) {
        load_mk(A, p.lda, M, m0, k0, ra);
        if (OP == kOpNT) load_mk(B, p.ldb, N, n0, k0, rbt);
        else load_kn(k0);
    };
    auto stage = [&](int buf) {
        store_mk(sA[buf], ra);
        if (OP == kOpNT) store_mk(sB[buf], rbt);
        else store_kn(sB[buf]);
    };

    floatx16 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    const int nk = (K + BK - 1) / BK;
    if (nk > 0) {
        fetch(0);
        stage(0);
    }
    __syncthreads();
    const int orow = lane & 31;
    const int okh = (lane >> 5) * 8;
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) fetch((kt + 1) * BK);
#pragma unroll
        for (int ks = 0; ks < BK / 16; ++ks) {
            v8 a[FM], b[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i)
                a[i] = *reinterpret_cast<const v8*>(&sA[cur][wm * 64 + i * 32 + orow][ks * 16 + okh]);
#pragma unroll
            for (int j = 0; j < FN; ++j)
                b[j] = *reinterpret_cast<const v8*>(&sB[cur][wn * 64 + j * 32 + orow][ks * 16 + okh]);
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = LM::mma(a[i], b[j], acc[i][j]);
        }
        if (kt + 1 < nk) stage(cur ^ 1);
        __syncthreads();
    }

    const float alpha = p.alpha;
    float* C = p.C;
    const bool cvec = p.cvec & 1;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int col = n0 + wn * 64 + j * 32 + (lane & 31);
            if (col >= N) continue;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int row = m0 + wm * 64 + i * 32 + g * 8 + (lane >> 5) * 4;
                float* c = C + off(row, col, p.ldc);
                const bool full = (row + 3 < M) && (SHAPE != kShapeLower || row >= col);
                if (full && cvec) {
                    float4* c4 = reinterpret_cast<float4*>(c);
                    float4 v = *c4;
                    v.x += alpha * acc[i][j][g * 4 + 0];
                    v.y += alpha * acc[i][j][g * 4 + 1];
                    v.z += alpha * acc[i][j][g * 4 + 2];
                    v.w += alpha * acc[i][j][g * 4 + 3];
                    *c4 = v;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int r = row + e;
                        if (r < M && (SHAPE != kShapeLower || r >= col)) c[e] += alpha * acc[i][j][g * 4 + e];
                    }
                }
            }
        }
    }
}

void gemm_lowp(mifact_handle h, LowPrecision prec, GemmOp op, GemmShape shape, int M, int N, int K,
               float alpha, const float* A, long long lda, const float* B, long long ldb, float* C,
               long long ldc)
{
    if (M <= 0 || N <= 0 || K <= 0) return;
    if (prec == kPrecFP32) {
        gemm<float>(h, op, shape, M, N, K, alpha, A, lda, B, ldb, C, ldc);
        return;
    }
    if (op == kOpTN || (shape == kShapeLower && M < N)) {
        report_status(h, mifact_status_invalid_value, "gemm_lowp: unsupported operand form");
        return;
    }
    GemmArgs<float> a{M, N, K, alpha, A, lda, B, ldb, C, ldc, 0, 0, 0};
    const bool vec = aligned_for_vectors<float>(A, lda) && aligned_for_vectors<float>(B, ldb);
    a.cvec = (aligned_for_vectors<float>(C, ldc) ? 1 : 0) | (vec ? 2 : 0);
    a.tiles_m = ceil_div(M, 128);
    a.tiles_n = ceil_div(N, 128);
    long long blocks;
    double mnk_c;
    if (shape == kShapeLower) {
        const long long tn = a.tiles_n;
        blocks = tn * (tn + 1) / 2 + static_cast<long long>(a.tiles_m - a.tiles_n) * tn;
        mnk_c = static_cast<double>(M) * N - 0.5 * static_cast<double>(N) * (N - 1);
    } else {
        blocks = static_cast<long long>(a.tiles_m) * a.tiles_n;
        mnk_c = static_cast<double>(M) * N;
    }
    StageScope scope(h, kStageGemmLowp, 2.0 * mnk_c * K,
                     4.0 * ((static_cast<double>(M) + N) * K + 2.0 * mnk_c));
    const dim3 grid(static_cast<unsigned>(blocks)), block(256);
    auto go = [&](auto tc_tag, auto op_tag, auto shape_tag) {
        using TC = typename decltype(tc_tag)::type;
        constexpr int OPV = decltype(op_tag)::value;
        constexpr int SHV = decltype(shape_tag)::value;
        hipLaunchKernelGGL((gemm_lowp_kernel<TC, OPV, SHV>), grid, block, 0, h->stream, a);
    };
    auto by_shape = [&](auto tc_tag, auto op_tag) {
        if (shape == kShapeLower) go(tc_tag, op_tag, std::integral_constant<int, kShapeLower>{});
        else go(tc_tag, op_tag, std::integral_constant<int, kShapeFull>{});
    };
    auto by_op = [&](auto tc_tag) {
        if (op == kOpNT) by_shape(tc_tag, std::integral_constant<int, kOpNT>{});
        else by_shape(tc_tag, std::integral_constant<int, kOpNN>{});
    };
    if (prec == kPrecBF16) by_op(std::type_identity<__bf16>{});
    else by_op(std::type_identity<_Float16>{});
    MIFACT_LAUNCHED(h);
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
    GemmArgs<T> a{M, N, K, alpha, A, lda, B, ldb, C, ldc, 0, 0, 0};
    const bool vec = aligned_for_vectors<T>(A, lda) && aligned_for_vectors<T>(B, ldb);
    a.cvec = aligned_for_vectors<T>(C, ldc) ? 1 : 0;
    if constexpr (sizeof(T) == 4) {
        // Large tiles only where they fill the chip; the overwrite (triangular
        // solve leaf) form always uses 64 x 64 so that one tile spans s <= 64.
        const long long big = static_cast<long long>(ceil_div(M, 128)) * ceil_div(N, 128);
        const bool large = !beta0 && N > 64 && M > 64 && big >= 2LL * h->num_cu;
        int cfg = h->gemm_config;
        if (beta0 || (cfg == 0 && !large)) cfg = 3;
        if (cfg == 0) cfg = 1;
        switch (cfg) {
        case 2: launch_cfg<T, 128, 128, 32, 2, 2>(h, op, shape, vec, beta0, a); break;
        case 3: launch_cfg<T, 64, 64, 16, 2, 2>(h, op, shape, vec, beta0, a); break;
        case 4: launch_cfg<T, 256, 256, 16, 2, 2>(h, op, shape, vec, beta0, a); break;
        default: launch_cfg<T, 128, 128, 16, 2, 2>(h, op, shape, vec, beta0, a); break;
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

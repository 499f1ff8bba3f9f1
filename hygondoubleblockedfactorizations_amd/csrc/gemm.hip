// MFMA update GEMM / SYRK tiles for gfx950.
//
//   C(M x N) += alpha * A(M x K) * op(B),   column-major, 64-bit offsets.
//
// FP32 runs on v_mfma_f32_32x32x2_f32, FP64 on v_mfma_f64_16x16x4_f64.  A
// workgroup of WAVES_M x WAVES_N wave64s (four, or eight for the 256 x 256
// tile, whose 64 x 128 wave tiles keep their 128 accumulators in VGPRs at
// two waves per SIMD) owns a BM x BN tile of C; each wave
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
__global__ __launch_bounds__(64 * WAVES_M * WAVES_N) void gemm_kernel(GemmArgs<T> p)
{
    constexpr int NTHR = 64 * WAVES_M * WAVES_N;
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
    constexpr int A_VECS = BM * BK / VW / NTHR;  // vectors per thread, A tile
    constexpr int B_VECS = BN * BK / VW / NTHR;
    static_assert(A_VECS >= 1 && B_VECS >= 1, "tile too small for the workgroup");
    static_assert(A_VECS * VW * NTHR == BM * BK && B_VECS * VW * NTHR == BN * BK,
                  "tile not a whole number of vectors per thread");

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
            const int f = tid + v * NTHR;
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
            const int f = tid + v * NTHR;
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
            const int f = tid + v * NTHR;
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
            const int f = tid + v * NTHR;
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
    // trailing updates on the large tiles are their own stages, one per tile
    // size and shape (each one kernel: the dominant kernels of the
    // factorizations); small and in-place launches stay "gemm"
    const bool lower = shape == kShapeLower;
    const Stage stage = beta0 ? kStageGemm
                      : BM >= 256 ? (lower ? kStageGemmUpdate256Lower : kStageGemmUpdate256)
                      : BM >= 128 ? (lower ? kStageGemmUpdateLower : kStageGemmUpdate)
                                  : kStageGemm;
    StageScope scope(h, stage, 2.0 * mnk_c * a.K,
                     sizeof(T) * ((static_cast<double>(a.M) + a.N) * a.K + 2.0 * mnk_c));
    const dim3 grid(static_cast<unsigned>(blocks));
    const dim3 block(64 * WAVES_M * WAVES_N);
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
// C(fp32) += alpha * A * op(B) with the products on MFMA in BF16, FP16 or
// FP8 e4m3, FP32 accumulation.  Both operands are first packed, once per
// update, into low-precision copies in the handle's pack arena (rows padded
// with zeros to a multiple of 256, k to a multiple of 128) laid out in
// slabs: slab s holds 128 bytes of k (64 16-bit or 128 FP8 values) of every
// row, row after row, so one GEMM stage of a tile is one contiguous block
// and every lane's MFMA operand is one 16- or 32-byte vector in global
// memory, LDS and registers; the GEMM streams 2 (or 1) bytes per element
// instead of 4.  The FP8 copy carries one power-of-two scale per 32 x 32
// tile (rows x k), stored as an E8M0 byte: the exponent e that brings the
// tile's largest magnitude into [128, 256) (e4m3 holds up to 448), so that
// x = q * 2^e with q the e4m3 value; v_mfma_scale_f32_32x32x64_f8f6f4
// applies 2^(ea + eb) to each lane's 32-element k-block product.

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

struct Fp8Tag {};

template <typename TC>
struct LowTraits {
    static constexpr int kBytes = 2;
};
template <>
struct LowTraits<Fp8Tag> {
    static constexpr int kBytes = 1;
};

constexpr int kPackRowAlign = 256;   // packed rows: multiple of the largest tile
constexpr int kPackKAlign = 128;     // packed k: multiple of one 128-byte FP8 stage
constexpr int kPackTile = 64;        // pack kernel tile (rows x k)

template <typename TC>
__device__ inline unsigned int pack2(float lo, float hi)
{
    const TC a = static_cast<TC>(lo);
    const TC b = static_cast<TC>(hi);
    const unsigned short ua = __builtin_bit_cast(unsigned short, a);
    const unsigned short ub = __builtin_bit_cast(unsigned short, b);
    return static_cast<unsigned int>(ua) | (static_cast<unsigned int>(ub) << 16);
}

// Packs the R x K operand X into slabs (zero padded): element (r, k) at byte
// (k EB / 128) rows_p 128 + r 128 + (k EB) % 128.  TRANS: X is stored
// column-major with rows contiguous, X(r, k) = X[r + k ld]; otherwise k is
// contiguous, X(r, k) = X[k + r ld].  One workgroup per 64 x 64 tile,
// through LDS.  FP8: the scale of each 32 x 32 sub-tile (one per wave) goes
// to scale[r / 32][k / 32] (row stride k_p / 32).
template <typename TC, bool TRANS>
__global__ __launch_bounds__(256) void pack_kernel(const float* __restrict__ X, long long ld,
                                                   int R, int K, unsigned char* __restrict__ out,
                                                   unsigned char* __restrict__ scale, int rows_p,
                                                   int k_p)
{
    constexpr int EB = LowTraits<TC>::kBytes;
    __shared__ float t[kPackTile][kPackTile + 1];
    const int r0 = blockIdx.x * kPackTile, k0 = blockIdx.y * kPackTile;
    const int tid = threadIdx.x;
    for (int e = tid; e < kPackTile * kPackTile; e += 256) {
        const int rr = TRANS ? (e & 63) : (e >> 6);
        const int kk = TRANS ? (e >> 6) : (e & 63);
        const int r = r0 + rr, k = k0 + kk;
        float v = 0.0f;
        if (r < R && k < K) v = TRANS ? X[off(r, k, ld)] : X[off(k, r, ld)];
        t[rr][kk] = v;
    }
    __syncthreads();
    // this tile's 64 k values: one slab (16-bit) or half of one (FP8)
    const size_t kb = static_cast<size_t>(k0) * EB;
    unsigned char* row_base = out + (kb / 128) * rows_p * 128 + static_cast<size_t>(r0) * 128 + kb % 128;
    if constexpr (EB == 2) {
        for (int c = tid; c < kPackTile * 8; c += 256) {
            const int rr = c >> 3, kc = (c & 7) * 8;
            uint4 w;
            w.x = pack2<TC>(t[rr][kc + 0], t[rr][kc + 1]);
            w.y = pack2<TC>(t[rr][kc + 2], t[rr][kc + 3]);
            w.z = pack2<TC>(t[rr][kc + 4], t[rr][kc + 5]);
            w.w = pack2<TC>(t[rr][kc + 6], t[rr][kc + 7]);
            *reinterpret_cast<uint4*>(row_base + static_cast<size_t>(rr) * 128 + kc * 2) = w;
        }
    } else {
        __shared__ int sexp[4];
        const int w = tid >> 6, l = tid & 63;
        const int sr = (w & 1) * 32, sk = (w >> 1) * 32;
        float m = 0.0f;
        for (int e = l; e < 32 * 32; e += 64) m = fmaxf(m, fabsf(t[sr + (e >> 5)][sk + (e & 31)]));
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if (l == 0) {
            int ex = 0;
            if (m > 0.0f && isfinite(m)) ex = ilogbf(m) - 7;   // m * 2^-ex in [128, 256)
            ex = ex < -127 ? -127 : (ex > 127 ? 127 : ex);
            sexp[w] = ex;
            scale[static_cast<size_t>((r0 + sr) >> 5) * (k_p >> 5) + ((k0 + sk) >> 5)] =
                static_cast<unsigned char>(ex + 127);
        }
        __syncthreads();
        const int rr = tid >> 2, kc = (tid & 3) * 16;
        const float sc = ldexpf(1.0f, -sexp[(rr >> 5) | ((kc >> 5) << 1)]);
        int q[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const float* v = &t[rr][kc + 4 * g];
            int word = __builtin_amdgcn_cvt_pk_fp8_f32(v[0] * sc, v[1] * sc, 0, false);
            word = __builtin_amdgcn_cvt_pk_fp8_f32(v[2] * sc, v[3] * sc, word, true);
            q[g] = word;
        }
        *reinterpret_cast<uint4*>(row_base + static_cast<size_t>(rr) * 128 + kc) =
            make_uint4(q[0], q[1], q[2], q[3]);
    }
}

struct PackedGemmArgs {
    int M, N;
    int k_p;                       // packed k (multiple of 128)
    float alpha;
    const unsigned char* A;        // slabs of rows_a x 128 bytes, rows_a >= tiles_m * BM
    const unsigned char* B;        // slabs of rows_b x 128 bytes, rows_b >= tiles_n * BN
    long long slab_a, slab_b;      // bytes per slab (rows x 128)
    const unsigned char* sa;       // FP8: [rows / 32][k_p / 32] E8M0 scales
    const unsigned char* sb;
    float* C;
    long long ldc;
    int tiles_m, tiles_n;
    int cvec;                      // C base and ldc allow float4 access
};

typedef int intx8 __attribute__((ext_vector_type(8)));

// One stage of a tile: BM (BN) rows x 128 bytes of a slab, contiguous; thread
// tid copies the 16-byte chunks tid + 256 v.
template <int ROWS>
__device__ __attribute__((always_inline)) inline void stage_fetch(const unsigned char* __restrict__ src,
                                                                  uint4 (&r)[ROWS / 32])
{
    const unsigned char* q = src + threadIdx.x * 16;
#pragma unroll
    for (int v = 0; v < ROWS / 32; ++v) r[v] = *reinterpret_cast<const uint4*>(q + v * 4096);
}

// Chunk tid + 256 v is row tid / 8 + 32 v, bytes (tid % 8) * 16 .. +15.
template <int ROWS, int LDB>
__device__ __attribute__((always_inline)) inline void stage_store(unsigned char* dst,
                                                                  const uint4 (&r)[ROWS / 32])
{
    unsigned char* q = dst + (threadIdx.x >> 3) * LDB + (threadIdx.x & 7) * 16;
#pragma unroll
    for (int v = 0; v < ROWS / 32; ++v) *reinterpret_cast<uint4*>(q + v * 32 * LDB) = r[v];
}

// FP8: the scale words (4 k-blocks of one stage) of COUNT consecutive
// 32-row blocks starting at block rb.
template <int COUNT>
__device__ __attribute__((always_inline)) inline void scale_fetch(const unsigned char* __restrict__ sc,
                                                                  int kbs, int rb, int st,
                                                                  unsigned int (&w)[COUNT])
{
#pragma unroll
    for (int i = 0; i < COUNT; ++i)
        w[i] = *reinterpret_cast<const unsigned int*>(sc + static_cast<size_t>(rb + i) * kbs + st * 4);
}

// C(M x N) += alpha * A_p B_p^T on MFMA.  BM x BN tile per workgroup of four
// waves (2 x 2, each (BM/2) x (BN/2) as 32 x 32 MFMA blocks), 128 bytes of
// k per stage, LDS double-buffered (row stride 144 bytes).  16-bit: four
// v_mfma_f32_32x32x16 steps per stage, lane l holding row l % 32 and k =
// 8 (l / 32) .. +7 of the step; FP8: two v_mfma_scale_f32_32x32x64 steps,
// lane l (h = l / 32) holding k = 16 h .. +15 and 32 + 16 h .. +15 of the
// step: the instruction's 32-k scale block b is made of bytes 16 b .. +15
// of lanes x and x + 32 and takes its scale from lane x + 32 b, so lane
// half h supplies the scale of k-block h (checked on the GPU with one-hot
// scales).
template <typename TC, int BM, int BN, int SHAPE>
__global__ __launch_bounds__(256) void gemm_packed_kernel(PackedGemmArgs p)
{
    constexpr int EB = LowTraits<TC>::kBytes;
    constexpr int SB = 128;                 // bytes of k per row per stage
    constexpr int LDB = SB + 16;            // LDS row stride (bytes)
    constexpr int FM = BM / 64, FN = BN / 64;
    constexpr int VA = BM / 32, VB = BN / 32;   // 16-byte chunks per thread per stage
    __shared__ __attribute__((aligned(16))) unsigned char sA[2][BM * LDB];
    __shared__ __attribute__((aligned(16))) unsigned char sB[2][BN * LDB];

    int tm, tn;
    if (SHAPE == kShapeLower) tile_lower(blockIdx.x, p.tiles_m, p.tiles_n, tm, tn);
    else tile_full(blockIdx.x, p.tiles_m, p.tiles_n, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int nst = p.k_p * EB / SB;                          // stages = slabs
    const unsigned char* ga = p.A + static_cast<size_t>(m0) * SB;
    const unsigned char* gb = p.B + static_cast<size_t>(n0) * SB;
    const int kbs = p.k_p >> 5;
    const int rba = (m0 + wm * (BM / 2)) >> 5, rbb = (n0 + wn * (BN / 2)) >> 5;

    uint4 ra[VA], rb[VB];
    unsigned int rsa[FM], rsb[FN];   // FP8: 4 k-block scales (one stage) per 32-row block
    floatx16 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    stage_fetch<BM>(ga, ra);
    stage_fetch<BN>(gb, rb);
    if constexpr (EB == 1) {
        scale_fetch<FM>(p.sa, kbs, rba, 0, rsa);
        scale_fetch<FN>(p.sb, kbs, rbb, 0, rsb);
    }
    stage_store<BM, LDB>(sA[0], ra);
    stage_store<BN, LDB>(sB[0], rb);
    unsigned int csa[FM], csb[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i) csa[i] = EB == 1 ? rsa[i] : 0u;
#pragma unroll
    for (int j = 0; j < FN; ++j) csb[j] = EB == 1 ? rsb[j] : 0u;
    __syncthreads();
    const int arow = wm * (BM / 2) + (lane & 31);
    const int brow = wn * (BN / 2) + (lane & 31);
    const int khalf = lane >> 5;
    // The prefetch and the staging run unconditionally (the last stage
    // re-reads its own slab into the idle buffer): conditional copies would
    // keep the prefetch registers in memory.
    for (int st = 0; st < nst; ++st) {
        const int cur = st & 1;
        const int nx = st + 1 < nst ? st + 1 : st;
        stage_fetch<BM>(ga + static_cast<size_t>(nx) * p.slab_a, ra);
        stage_fetch<BN>(gb + static_cast<size_t>(nx) * p.slab_b, rb);
        if constexpr (EB == 1) {
            scale_fetch<FM>(p.sa, kbs, rba, nx, rsa);
            scale_fetch<FN>(p.sb, kbs, rbb, nx, rsb);
        }
        const unsigned char* la = &sA[cur][arow * LDB];
        const unsigned char* lb = &sB[cur][brow * LDB];
        if constexpr (EB == 2) {
            using LM = LowMfma<TC>;
            using v8 = typename LM::v8;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                v8 a[FM], b[FN];
#pragma unroll
                for (int i = 0; i < FM; ++i)
                    a[i] = *reinterpret_cast<const v8*>(la + i * 32 * LDB + ks * 32 + khalf * 16);
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    b[j] = *reinterpret_cast<const v8*>(lb + j * 32 * LDB + ks * 32 + khalf * 16);
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j) acc[i][j] = LM::mma(a[i], b[j], acc[i][j]);
            }
        } else {
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                intx8 a[FM], b[FN];
                int ea[FM], eb[FN];
                const int sh = 8 * (ks * 2 + khalf);
#pragma unroll
                for (int i = 0; i < FM; ++i) {
                    const unsigned char* q = la + i * 32 * LDB + ks * 64 + khalf * 16;
                    const uint4 lo = *reinterpret_cast<const uint4*>(q);
                    const uint4 hi = *reinterpret_cast<const uint4*>(q + 32);
                    a[i] = intx8{static_cast<int>(lo.x), static_cast<int>(lo.y),
                                 static_cast<int>(lo.z), static_cast<int>(lo.w),
                                 static_cast<int>(hi.x), static_cast<int>(hi.y),
                                 static_cast<int>(hi.z), static_cast<int>(hi.w)};
                    ea[i] = static_cast<int>((csa[i] >> sh) & 0xffu);
                }
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    const unsigned char* q = lb + j * 32 * LDB + ks * 64 + khalf * 16;
                    const uint4 lo = *reinterpret_cast<const uint4*>(q);
                    const uint4 hi = *reinterpret_cast<const uint4*>(q + 32);
                    b[j] = intx8{static_cast<int>(lo.x), static_cast<int>(lo.y),
                                 static_cast<int>(lo.z), static_cast<int>(lo.w),
                                 static_cast<int>(hi.x), static_cast<int>(hi.y),
                                 static_cast<int>(hi.z), static_cast<int>(hi.w)};
                    eb[j] = static_cast<int>((csb[j] >> sh) & 0xffu);
                }
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                            a[i], b[j], acc[i][j], 0, 0, 0, ea[i], 0, eb[j]);
            }
        }
        stage_store<BM, LDB>(sA[cur ^ 1], ra);
        stage_store<BN, LDB>(sB[cur ^ 1], rb);
        if constexpr (EB == 1) {
#pragma unroll
            for (int i = 0; i < FM; ++i) csa[i] = rsa[i];
#pragma unroll
            for (int j = 0; j < FN; ++j) csb[j] = rsb[j];
        }
        __syncthreads();
    }

    const float alpha = p.alpha;
    float* C = p.C;
    const int M = p.M, N = p.N;
    const bool cvec = p.cvec & 1;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int col = n0 + wn * (BN / 2) + j * 32 + (lane & 31);
            if (col >= N) continue;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int row = m0 + wm * (BM / 2) + i * 32 + g * 8 + khalf * 4;
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

namespace {

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Grows the handle's pack arena to at least `bytes` (synchronizing the
// stream before freeing the old one).
bool ensure_pack(mifact_handle h, size_t bytes)
{
    if (h->pack_ws != nullptr && h->pack_ws_bytes >= bytes) return true;
    if (h->pack_ws != nullptr) {
        if (hipStreamSynchronize(h->stream) != hipSuccess) {
            report_status(h, mifact_status_hip_error, "gemm_lowp: synchronize before growing");
            return false;
        }
        (void)hipFree(h->pack_ws);
        h->pack_ws = nullptr;
        h->pack_ws_bytes = 0;
    }
    if (hipMalloc(&h->pack_ws, bytes) != hipSuccess) {
        h->pack_ws = nullptr;
        report_status(h, mifact_status_allocation_failed, "gemm_lowp: pack arena");
        return false;
    }
    h->pack_ws_bytes = bytes;
    return true;
}

template <typename TC>
void launch_pack(mifact_handle h, const float* X, long long ld, int R, int K, bool trans,
                 unsigned char* out, unsigned char* scale, int rows_p, int k_p)
{
    const dim3 grid(rows_p / kPackTile, k_p / kPackTile), block(256);
    if (trans) {
        hipLaunchKernelGGL((pack_kernel<TC, true>), grid, block, 0, h->stream, X, ld, R, K, out,
                           scale, rows_p, k_p);
    } else {
        hipLaunchKernelGGL((pack_kernel<TC, false>), grid, block, 0, h->stream, X, ld, R, K, out,
                           scale, rows_p, k_p);
    }
}

template <typename TC>
void gemm_lowp_typed(mifact_handle h, GemmOp op, GemmShape shape, int M, int N, int K,
                     float alpha, const float* A, long long lda, const float* B, long long ldb,
                     float* C, long long ldc)
{
    constexpr int EB = LowTraits<TC>::kBytes;
    const int k_p = static_cast<int>(round_up(K, kPackKAlign));
    const int rows_a = static_cast<int>(round_up(M, kPackRowAlign));
    // A lower-trapezoid C -= A A(0:N, :)^T reads B as the first N rows of A.
    const bool share = shape == kShapeLower && op == kOpNT && B == A && ldb == lda;
    const int rows_b = share ? 0 : static_cast<int>(round_up(N, kPackRowAlign));
    const size_t bytes_a = static_cast<size_t>(rows_a) * k_p * EB;
    const size_t bytes_b = static_cast<size_t>(rows_b) * k_p * EB;
    const size_t sc_a = EB == 1 ? round_up(static_cast<size_t>(rows_a / 32) * (k_p / 32), 256) : 0;
    const size_t sc_b = EB == 1 ? round_up(static_cast<size_t>(rows_b / 32) * (k_p / 32), 256) : 0;
    if (!ensure_pack(h, round_up(bytes_a, 256) + round_up(bytes_b, 256) + sc_a + sc_b)) return;
    unsigned char* base = static_cast<unsigned char*>(h->pack_ws);
    unsigned char* pa = base;
    unsigned char* pb = share ? pa : base + round_up(bytes_a, 256);
    unsigned char* sa = EB == 1 ? base + round_up(bytes_a, 256) + round_up(bytes_b, 256) : nullptr;
    unsigned char* sb = EB == 1 ? (share ? sa : sa + sc_a) : nullptr;

    const long long slab_a = static_cast<long long>(rows_a) * 128;
    PackedGemmArgs a{M, N, k_p, alpha, pa, pb, slab_a,
                     share ? slab_a : static_cast<long long>(rows_b) * 128, sa, sb, C, ldc, 0, 0, 0};
    a.cvec = aligned_for_vectors<float>(C, ldc) ? 1 : 0;
    // 256 x 256 tiles where they fill the chip, 128 x 128 otherwise
    const long long big = static_cast<long long>(ceil_div(M, 256)) * ceil_div(N, 256);
    const bool large = big >= h->num_cu;
    const int T = large ? 256 : 128;
    a.tiles_m = ceil_div(M, T);
    a.tiles_n = ceil_div(N, T);
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
                     4.0 * (static_cast<double>(M) + N) * K + EB * (static_cast<double>(rows_a) + rows_b) * k_p * 2.0
                         + 8.0 * mnk_c);
    launch_pack<TC>(h, A, lda, M, K, true, pa, sa, rows_a, k_p);
    MIFACT_LAUNCHED(h);
    if (!share) {
        launch_pack<TC>(h, B, ldb, N, K, op == kOpNT, pb, sb, rows_b, k_p);
        MIFACT_LAUNCHED(h);
    }
    const dim3 grid(static_cast<unsigned>(blocks)), block(256);
    if (large) {
        if (shape == kShapeLower)
            hipLaunchKernelGGL((gemm_packed_kernel<TC, 256, 256, kShapeLower>), grid, block, 0, h->stream, a);
        else
            hipLaunchKernelGGL((gemm_packed_kernel<TC, 256, 256, kShapeFull>), grid, block, 0, h->stream, a);
    } else {
        if (shape == kShapeLower)
            hipLaunchKernelGGL((gemm_packed_kernel<TC, 128, 128, kShapeLower>), grid, block, 0, h->stream, a);
        else
            hipLaunchKernelGGL((gemm_packed_kernel<TC, 128, 128, kShapeFull>), grid, block, 0, h->stream, a);
    }
    MIFACT_LAUNCHED(h);
}

}  // namespace

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
    switch (prec) {
    case kPrecBF16:
        gemm_lowp_typed<__bf16>(h, op, shape, M, N, K, alpha, A, lda, B, ldb, C, ldc);
        break;
    case kPrecFP16:
        gemm_lowp_typed<_Float16>(h, op, shape, M, N, K, alpha, A, lda, B, ldb, C, ldc);
        break;
    case kPrecFP8:
        gemm_lowp_typed<Fp8Tag>(h, op, shape, M, N, K, alpha, A, lda, B, ldb, C, ldc);
        break;
    default:
        report_status(h, mifact_status_invalid_value, "gemm_lowp: precision");
        break;
    }
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
        // Large tiles only where they give every CU several of them: 256 x 256
        // (eight waves) for full updates from four tiles per CU, 128 x 128
        // from two; 64 x 64 otherwise.  The overwrite (triangular-solve leaf)
        // form always uses 64 x 64 so that one tile spans s <= 64.
        auto tiles = [&](int t) {
            const long long tm = ceil_div(M, t), tn = ceil_div(N, t);
            return shape == kShapeLower ? tn * (tn + 1) / 2 + (tm - tn) * tn : tm * tn;
        };
        int cfg = h->gemm_config;
        if (beta0) {
            cfg = 3;
        } else if (cfg == 0) {
            const int pol = h->gemm_policy;
            const bool big_ok = pol != 1 && (shape == kShapeFull || pol == 2);
            const long long mid = pol == 3 ? h->num_cu : 2LL * h->num_cu;
            if (big_ok && tiles(256) >= 4LL * h->num_cu) cfg = 4;
            else if (M > 64 && N > 64 && tiles(128) >= mid) cfg = 1;
            else cfg = 3;
        }
        switch (cfg) {
        case 2: launch_cfg<T, 128, 128, 32, 2, 2>(h, op, shape, vec, beta0, a); break;
        case 3: launch_cfg<T, 64, 64, 16, 2, 2>(h, op, shape, vec, beta0, a); break;
        case 4: launch_cfg<T, 256, 256, 16, 4, 2>(h, op, shape, vec, beta0, a); break;
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

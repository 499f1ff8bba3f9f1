// Partial-pivoting kernels for gfx950: the cooperative GETF2 leaf and the
// device-resident row permutation machinery.
//
// GETF2 leaf.  An m x w panel (w <= 32) is spread over G workgroups of 256
// threads; each thread keeps RPT whole rows of the panel in registers for
// the entire leaf.  Per column j:
//   1. every workgroup finds its local pivot candidate (wave64 shuffle
//      reduction, then across the four waves through LDS): max |a(i,j)|,
//      smallest row on exact ties (IxAMAX convention);
//   2. the owner of the candidate row publishes (value, row, full row, row
//      origin) in an exchange buffer; workgroup 0 also publishes row j;
//   3. one grid barrier (monotonic 64-bit counter, bounded spin);
//   4. every workgroup reduces the G candidates, swaps (owner of the pivot
//      row takes row j, owner of row j takes the pivot row), scales the
//      column by the reciprocal of the pivot (|pivot| >= FLT_MIN, else
//      divides, as LAPACK xGETF2) and applies the rank-1 update to its rows.
// The exchange buffers are double-buffered by column parity so a workgroup
// that runs ahead cannot overwrite records another one still reads.
//
// Row permutations.  orig[p] records which original row's content sits at
// row p.  A snapshot inv[orig[p]] = p taken at the start of a phase lets any
// later moment express "rows exchanged since the snapshot" as a gather:
// content now at p came from row inv[orig[p]] at snapshot time.  Applying
// that gather to columns that missed the phase's exchanges needs no host
// round trip and no sequential replay of the LAPACK swap sequence.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <utility>

#include "kernels.hpp"

namespace mifact {

constexpr int kLeafW = 32;
constexpr int kMaxG = 128;  // workgroups of a leaf (rows <= 128 * 256 * 4)

template <typename T> __device__ inline T sfmin();
template <> __device__ inline float sfmin<float>() { return FLT_MIN; }
template <> __device__ inline double sfmin<double>() { return DBL_MIN; }

template <typename T>
__device__ inline void amax_pick(T& v, int& i, T ov, int oi)
{
    if (ov > v || (ov == v && oi < i)) {
        v = ov;
        i = oi;
    }
}

// DPP row/bank moves (lanes whose source is out of range keep their own
// value, which is neutral for the idempotent arg-max combine).
template <int CTRL>
__device__ inline int dpp_move(int v)
{
    return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ inline float dpp_move(float v)
{
    return __int_as_float(dpp_move<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ inline double dpp_move(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_move<CTRL>(static_cast<int>(b));
    const int hi = dpp_move<CTRL>(static_cast<int>(b >> 32));
    return __longlong_as_double((static_cast<long long>(hi) << 32)
                                | static_cast<unsigned int>(lo));
}

template <int CTRL, typename T>
__device__ inline void dpp_step(T& v, int& i)
{
    const T ov = dpp_move<CTRL>(v);
    const int oi = dpp_move<CTRL>(i);
    amax_pick(v, i, ov, oi);
}

__device__ inline float lane63(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }
__device__ inline double lane63(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), 63);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), 63);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Wave64 arg-max with DPP butterflies inside rows of 16 (quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_ror 4, row_ror 8) and row broadcasts (15, 31);
// the full result lands in lane 63 and is returned wave-uniform.
template <typename T>
__device__ inline void wave_argmax(T& v, int& i)
{
    dpp_step<0xb1>(v, i);
    dpp_step<0x4e>(v, i);
    dpp_step<0x124>(v, i);
    dpp_step<0x128>(v, i);
    dpp_step<0x142>(v, i);
    dpp_step<0x143>(v, i);
    v = lane63(v);
    i = __builtin_amdgcn_readlane(i, 63);
}

// Exchange records.  For column j every workgroup publishes one record
// (its pivot candidate: row index, |value|, the whole candidate row and the
// row's origin tag) and workgroup 0 also publishes row j.  A record is RW
// 64-bit words {payload: 32 bits, epoch: 32 bits}, each written by a single
// coherent (sc1) 64-bit store, so a reader that sees column j's epoch in a
// word also sees that word's payload.  Readers poll the records themselves
// until every word carries the epoch: no counter, no fences, one round trip
// when the data is already there.  Records are double-buffered by column
// parity: a workgroup can only overwrite a parity slot (column j + 2) after
// it has read every record of column j + 1, which every workgroup publishes
// only after it has finished reading column j.
template <typename T>
struct Rec {
    static constexpr int PW = sizeof(T) / 4;           // payload words per T
    static constexpr int IDX = 0;                       // candidate row
    static constexpr int VAL = 1;                       // |value|
    static constexpr int ROW = 1 + PW;                  // kLeafW values
    static constexpr int TAG = ROW + PW * kLeafW;       // row origin
    static constexpr int RW = TAG + 1;                  // words per record
};

template <typename T>
struct LeafArgs {
    T* A;
    long long lda;
    int n;
    int c0;
    int w;
    int m;
    int* ipiv;
    int* orig;
    int* info;
    unsigned long long* xrec;  // [2][kMaxG + 1][RW]
    unsigned int epoch0;       // epoch of column j is epoch0 + j
    int* status;
    int G;
};

__device__ inline void xstore64(unsigned long long* p, unsigned int payload, unsigned int epoch)
{
    const unsigned long long v = (static_cast<unsigned long long>(epoch) << 32) | payload;
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline unsigned long long xload64(const unsigned long long* p)
{
    return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void to_words(float v, unsigned int* w) { w[0] = __float_as_uint(v); }
__device__ inline void to_words(double v, unsigned int* w)
{
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
    w[0] = static_cast<unsigned int>(b);
    w[1] = static_cast<unsigned int>(b >> 32);
}
template <typename T> __device__ inline T from_words(const unsigned int* w);
template <> __device__ inline float from_words<float>(const unsigned int* w)
{
    return __uint_as_float(w[0]);
}
template <> __device__ inline double from_words<double>(const unsigned int* w)
{
    const unsigned long long b = (static_cast<unsigned long long>(w[1]) << 32) | w[0];
    return __longlong_as_double(static_cast<long long>(b));
}

// Writes one record (one thread).
template <typename T>
__device__ inline void publish(unsigned long long* rec, unsigned int epoch, int idx, T val,
                               const T (&row)[kLeafW], int tag)
{
    using R = Rec<T>;
    unsigned int w[R::PW];
    xstore64(rec + R::IDX, static_cast<unsigned int>(idx), epoch);
    to_words(val, w);
#pragma unroll
    for (int q = 0; q < R::PW; ++q) xstore64(rec + R::VAL + q, w[q], epoch);
#pragma unroll
    for (int c = 0; c < kLeafW; ++c) {
        to_words(row[c], w);
#pragma unroll
        for (int q = 0; q < R::PW; ++q) xstore64(rec + R::ROW + R::PW * c + q, w[q], epoch);
    }
    xstore64(rec + R::TAG, static_cast<unsigned int>(tag), epoch);
}

// Shared state of one leaf workgroup.
template <typename T>
struct LeafShared {
    T wv[4];
    int wi[4];
    unsigned int rec[kMaxG + 1][Rec<T>::RW];  // payloads of every record of the column
    int p;
    int abort;
};

// Column J of the leaf (J is a compile-time constant so the register rows
// x[r][J] stay in registers).  Returns false when the bounded wait expired.
template <typename T, int RPT, int J>
__device__ __attribute__((always_inline)) inline bool leaf_column(
    const LeafArgs<T>& a, LeafShared<T>& sh, T (&x)[RPT][kLeafW], int (&tag)[RPT], int base)
{
    using R = Rec<T>;
    const int tid = threadIdx.x;
    const int g = blockIdx.x;
    const int G = a.G;
    const int m = a.m;
    const int w = a.w;
    constexpr int j = J;
    const unsigned int epoch = a.epoch0 + static_cast<unsigned int>(j);
    unsigned long long* slots =
        a.xrec + static_cast<size_t>(j & 1) * (kMaxG + 1) * R::RW;

    // 1. local candidate: max |a(i,j)| over own rows i >= j, smallest row on ties
    T best = T(-1);
    int bi = 0x7fffffff;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int lrow = base + r * 256 + tid;
        if (lrow >= j && lrow < m) amax_pick(best, bi, fabs(x[r][j]), lrow);
    }
    wave_argmax(best, bi);
    if ((tid & 63) == 0) {
        sh.wv[tid >> 6] = best;
        sh.wi[tid >> 6] = bi;
    }
    __syncthreads();
    T wb = sh.wv[0];
    int wi = sh.wi[0];
#pragma unroll
    for (int q = 1; q < 4; ++q) amax_pick(wb, wi, sh.wv[q], sh.wi[q]);

    // 2. publish the candidate (and row j from workgroup 0)
    if (wi == 0x7fffffff) {
        if (tid == 0) {
            T zero[kLeafW];
#pragma unroll
            for (int c = 0; c < kLeafW; ++c) zero[c] = T(0);
            publish<T>(slots + static_cast<size_t>(g) * R::RW, epoch, 0x7fffffff, T(-1), zero, -1);
        }
    } else {
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int lrow = base + r * 256 + tid;
            if (lrow == wi) publish<T>(slots + static_cast<size_t>(g) * R::RW, epoch, wi, wb, x[r], tag[r]);
        }
    }
    if (g == 0 && tid == j) publish<T>(slots + static_cast<size_t>(G) * R::RW, epoch, j, T(0), x[0], tag[0]);

    // 3. gather every record of this column into LDS (poll until all carry the epoch)
    const int total = (G + 1) * R::RW;
    constexpr int kPoll = ((kMaxG + 1) * R::RW + 255) / 256;
    unsigned long long spins = 0;
    while (true) {
        // issue every load of this thread before consuming any
        unsigned long long v[kPoll];
#pragma unroll
        for (int q = 0; q < kPoll; ++q) {
            const int e = tid + q * 256;
            v[q] = (e < total) ? xload64(slots + e) : 0ULL;
        }
        int miss = 0;
#pragma unroll
        for (int q = 0; q < kPoll; ++q) {
            const int e = tid + q * 256;
            if (e < total) {
                if (static_cast<unsigned int>(v[q] >> 32) != epoch) {
                    miss = 1;
                } else {
                    (&sh.rec[0][0])[e] = static_cast<unsigned int>(v[q]);
                }
            }
        }
        if (!__syncthreads_or(miss)) break;
        if (++spins > (1ULL << 18)) {
            if (tid == 0) atomicExch(a.status, 1);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }

    // 4. global pivot: max over the G candidates, smallest row on ties
    if (tid < 64) {
        T gb = T(-1);
        int gi = 0x7fffffff;
        for (int t = tid; t < G; t += 64)
            amax_pick(gb, gi, from_words<T>(&sh.rec[t][R::VAL]), static_cast<int>(sh.rec[t][R::IDX]));
        wave_argmax(gb, gi);
        if (tid == 0) sh.p = (gi == 0x7fffffff) ? j : gi;
    }
    __syncthreads();
    const int p = sh.p;
    const int gp = p / (256 * RPT);
    const unsigned int* prec = sh.rec[gp];
    const unsigned int* jrec = sh.rec[G];
    const T piv = from_words<T>(prec + R::ROW + R::PW * j);
    if (piv != T(0)) {
        if (p != j) {
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int lrow = base + r * 256 + tid;
                if (lrow == p) {
#pragma unroll
                    for (int c = 0; c < kLeafW; ++c) x[r][c] = from_words<T>(jrec + R::ROW + R::PW * c);
                    tag[r] = static_cast<int>(jrec[R::TAG]);
                } else if (lrow == j) {
#pragma unroll
                    for (int c = 0; c < kLeafW; ++c) x[r][c] = from_words<T>(prec + R::ROW + R::PW * c);
                    tag[r] = static_cast<int>(prec[R::TAG]);
                }
            }
        }
        const bool recip = fabs(piv) >= sfmin<T>();
        const T rcp = T(1) / piv;
        T u[kLeafW];
#pragma unroll
        for (int c = j + 1; c < kLeafW; ++c) u[c] = from_words<T>(prec + R::ROW + R::PW * c);
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int lrow = base + r * 256 + tid;
            if (lrow > j && lrow < m) {
                const T l = recip ? x[r][j] * rcp : x[r][j] / piv;
                x[r][j] = l;
#pragma unroll
                for (int c = j + 1; c < kLeafW; ++c) x[r][c] -= l * u[c];
            }
        }
    } else if (g == 0 && tid == 0) {
        atomicCAS(a.info, 0, a.c0 + j + 1);
    }
    if (g == 0 && tid == 0) a.ipiv[a.c0 + j] = a.c0 + p + 1;
    (void)w;
    return true;
}

template <typename T, int RPT, int... Js>
__device__ __attribute__((always_inline)) inline bool leaf_columns(
    const LeafArgs<T>& a, LeafShared<T>& sh, T (&x)[RPT][kLeafW], int (&tag)[RPT], int base,
    std::integer_sequence<int, Js...>)
{
    bool ok = true;
    ((ok = ok && (Js >= a.w || leaf_column<T, RPT, Js>(a, sh, x, tag, base))), ...);
    return ok;
}

template <typename T, int RPT>
__global__ __launch_bounds__(256) void getf2_leaf_kernel(LeafArgs<T> a)
{
    const int tid = threadIdx.x;
    const int base = blockIdx.x * 256 * RPT;
    __shared__ LeafShared<T> sh;
    // A leaf whose predecessor timed out would wait on stale records: every
    // later leaf exits at once (the host reports the timeout).
    if (__hip_atomic_load(a.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;

    T x[RPT][kLeafW];
    int tag[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int lrow = base + r * 256 + tid;
        const bool in = lrow < a.m;
#pragma unroll
        for (int c = 0; c < kLeafW; ++c)
            x[r][c] = (in && c < a.w) ? a.A[off(a.c0 + lrow, a.c0 + c, a.lda)] : T(0);
        tag[r] = in ? a.orig[a.c0 + lrow] : -1;
    }

    if (!leaf_columns<T, RPT>(a, sh, x, tag, base, std::make_integer_sequence<int, kLeafW>{}))
        return;

#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int lrow = base + r * 256 + tid;
        if (lrow < a.m) {
#pragma unroll
            for (int c = 0; c < kLeafW; ++c)
                if (c < a.w) a.A[off(a.c0 + lrow, a.c0 + c, a.lda)] = x[r][c];
            a.orig[a.c0 + lrow] = tag[r];
        }
    }
}

template <typename T, int RPT>
static void launch_leaf(mifact_handle h, LeafArgs<T>& args)
{
    const double m = args.m, w = args.w;
    StageScope scope(h, kStageGetf2Leaf, m * w * w - w * w * w / 3.0,
                     sizeof(T) * 2.0 * m * w + 8.0 * m);
    void* params[] = {&args};
    hipError_t e = hipLaunchCooperativeKernel(
        reinterpret_cast<const void*>(&getf2_leaf_kernel<T, RPT>), dim3(args.G),
        dim3(256), params, 0, h->stream);
    if (e != hipSuccess) {
        report_hip(h, e, "hipLaunchCooperativeKernel(getf2_leaf)", __FILE__, __LINE__);
        return;
    }
    h->leaf_epoch += static_cast<unsigned int>(args.w);
}

template <typename T>
void getf2_leaf(mifact_handle h, T* A, long long lda, int n, int c0, int w,
                int* ipiv, int* orig, int* info)
{
    const int m = n - c0;
    if (w <= 0 || m <= 0) return;
    if (w > kLeafW) {
        report_status(h, mifact_status_invalid_value, "getf2_leaf: w > 32");
        return;
    }
    Workspace ws = workspace(h);
    LeafArgs<T> args;
    args.A = A;
    args.lda = lda;
    args.n = n;
    args.c0 = c0;
    args.w = w;
    args.m = m;
    args.ipiv = ipiv;
    args.orig = orig;
    args.info = info;
    args.xrec = reinterpret_cast<unsigned long long*>(ws.exchange);
    args.epoch0 = h->leaf_epoch;
    args.status = h->kernel_status;
    // Rows per thread: keep G (the workgroups that exchange per column) at 64
    // or below where the registers allow.
    int rpt = 1;
    while (rpt < 4 && ceil_div(m, 256 * rpt) > 64) rpt *= 2;
    args.G = ceil_div(m, 256 * rpt);
    if (args.G > kMaxG) {
        report_status(h, mifact_status_invalid_value, "getf2_leaf: panel taller than 131072 rows");
        return;
    }
    if (rpt == 1) launch_leaf<T, 1>(h, args);
    else if (rpt == 2) launch_leaf<T, 2>(h, args);
    else launch_leaf<T, 4>(h, args);
}

// ------------------------------------------------------------------ perms

__global__ void perm_identity_kernel(int* orig, int n)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) orig[p] = p;
}

__global__ void perm_snapshot_kernel(const int* orig, int* inv, int r0, int n)
{
    const int p = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) inv[orig[p]] = p;
}

__global__ void clear_int_kernel(int* p) { *p = 0; }

void perm_identity(mifact_handle h, int* orig, int n)
{
    if (n <= 0) return;
    StageScope scope(h, kStagePerm, 0.0, 4.0 * n);
    hipLaunchKernelGGL(perm_identity_kernel, dim3(ceil_div(n, 256)), dim3(256), 0,
                       h->stream, orig, n);
    MIFACT_LAUNCHED(h);
}

void perm_snapshot(mifact_handle h, const int* orig, int* inv, int r0, int n)
{
    if (n - r0 <= 0) return;
    StageScope scope(h, kStagePerm, 0.0, 12.0 * (n - r0));
    hipLaunchKernelGGL(perm_snapshot_kernel, dim3(ceil_div(n - r0, 256)), dim3(256), 0,
                       h->stream, orig, inv, r0, n);
    MIFACT_LAUNCHED(h);
}

void clear_int(mifact_handle h, int* p)
{
    hipLaunchKernelGGL(clear_int_kernel, dim3(1), dim3(1), 0, h->stream, p);
    MIFACT_LAUNCHED(h);
}

// Candidate moves: t < K -> row p0 + t; t >= K -> pivot row of step p0+t-K.
__global__ void build_moves_kernel(int p0, int K, const int* __restrict__ ipiv,
                                   const int* __restrict__ orig, const int* __restrict__ inv,
                                   int* __restrict__ src, int* __restrict__ dst)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * K) return;
    const int pos = (t < K) ? p0 + t : ipiv[p0 + t - K] - 1;
    const int s = inv[orig[pos]];
    src[t] = (s != pos) ? s : -1;
    dst[t] = pos;
}

// A tile of CT columns per workgroup: gather every moved entry of those
// columns into LDS, barrier, scatter.  All reads of a column precede all its
// writes, so the in-place permutation is race-free.
template <typename T>
__global__ __launch_bounds__(256) void apply_moves_lds_kernel(T* __restrict__ A, long long lda,
                                                              int c_begin, int c_end, int CT,
                                                              int nm, const int* __restrict__ src,
                                                              const int* __restrict__ dst)
{
    extern __shared__ unsigned char smem_raw[];
    T* vals = reinterpret_cast<T*>(smem_raw);
    const int c0 = c_begin + blockIdx.x * CT;
    const int nc = (c_end - c0) < CT ? (c_end - c0) : CT;
    // Move lists are read once per workgroup; the gathers of all its columns
    // are issued back to back (independent loads in flight) before any use.
    for (int t0 = threadIdx.x; t0 < nm; t0 += 256 * 4) {
        int sv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int t = t0 + u * 256;
            sv[u] = (t < nm) ? src[t] : -1;
        }
        for (int c = 0; c < nc; ++c) {
            const T* col = A + off(0, c0 + c, lda);
            T v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = (sv[u] >= 0) ? col[sv[u]] : T(0);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = t0 + u * 256;
                if (t < nm) vals[c * nm + t] = v[u];
            }
        }
    }
    __syncthreads();
    for (int t0 = threadIdx.x; t0 < nm; t0 += 256 * 4) {
        int sv[4], dv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int t = t0 + u * 256;
            sv[u] = (t < nm) ? src[t] : -1;
            dv[u] = (t < nm) ? dst[t] : 0;
        }
        for (int c = 0; c < nc; ++c) {
            T* col = A + off(0, c0 + c, lda);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (sv[u] >= 0) col[dv[u]] = vals[c * nm + t0 + u * 256];
        }
    }
}

// Global staging variant for very long move lists: gather a column chunk to
// tmp, then scatter (two launches per chunk).
template <typename T>
__global__ void gather_moves_kernel(const T* __restrict__ A, long long lda, int c0, int nc,
                                    int nm, const int* __restrict__ src, T* __restrict__ tmp)
{
    const size_t total = static_cast<size_t>(nc) * nm;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int t = static_cast<int>(i % nm);
        const int c = static_cast<int>(i / nm);
        const int s = src[t];
        if (s >= 0) tmp[i] = A[off(s, c0 + c, lda)];
    }
}

template <typename T>
__global__ void scatter_moves_kernel(T* __restrict__ A, long long lda, int c0, int nc, int nm,
                                     const int* __restrict__ src, const int* __restrict__ dst,
                                     const T* __restrict__ tmp)
{
    const size_t total = static_cast<size_t>(nc) * nm;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int t = static_cast<int>(i % nm);
        const int c = static_cast<int>(i / nm);
        if (src[t] >= 0) A[off(dst[t], c0 + c, lda)] = tmp[i];
    }
}

template <typename T>
void apply_row_moves(mifact_handle h, T* A, long long lda, int c_begin, int c_end, int p0,
                     int K, const int* ipiv, const int* orig, const int* inv, T* tmp,
                     long long tmp_elems)
{
    if (c_end <= c_begin || K <= 0) return;
    Workspace ws = workspace(h);
    const int nm = 2 * K;
    StageScope scope(h, kStageRowMoves, 0.0,
                     2.0 * sizeof(T) * static_cast<double>(nm) * (c_end - c_begin) + 20.0 * nm);
    int* src = ws.moves;
    int* dst = ws.moves + 2 * static_cast<size_t>(h->ws_n);
    hipLaunchKernelGGL(build_moves_kernel, dim3(ceil_div(nm, 256)), dim3(256), 0, h->stream,
                       p0, K, ipiv, orig, inv, src, dst);
    MIFACT_LAUNCHED(h);
    const int ncols = c_end - c_begin;
    constexpr size_t kLdsBudget = 64 * 1024;
    const size_t per_col = static_cast<size_t>(nm) * sizeof(T);
    if (per_col <= kLdsBudget) {
        int CT = static_cast<int>(kLdsBudget / per_col);
        if (CT > 16) CT = 16;
        const size_t lds = static_cast<size_t>(CT) * per_col;
        hipLaunchKernelGGL(apply_moves_lds_kernel<T>, dim3(ceil_div(ncols, CT)), dim3(256), lds,
                           h->stream, A, lda, c_begin, c_end, CT, nm, src, dst);
        MIFACT_LAUNCHED(h);
        return;
    }
    if (tmp == nullptr || tmp_elems < nm) {
        report_status(h, mifact_status_invalid_value, "apply_row_moves: staging buffer too small");
        return;
    }
    const int chunk = static_cast<int>(tmp_elems / nm);
    for (int c0 = c_begin; c0 < c_end; c0 += chunk) {
        const int nc = (c_end - c0) < chunk ? (c_end - c0) : chunk;
        const size_t total = static_cast<size_t>(nc) * nm;
        const int blocks = static_cast<int>(total / 256 + 1 < 65535 ? total / 256 + 1 : 65535);
        hipLaunchKernelGGL(gather_moves_kernel<T>, dim3(blocks), dim3(256), 0, h->stream, A, lda,
                           c0, nc, nm, src, tmp);
        MIFACT_LAUNCHED(h);
        hipLaunchKernelGGL(scatter_moves_kernel<T>, dim3(blocks), dim3(256), 0, h->stream, A, lda,
                           c0, nc, nm, src, dst, tmp);
        MIFACT_LAUNCHED(h);
    }
}

template void getf2_leaf<float>(mifact_handle, float*, long long, int, int, int, int*, int*, int*);
template void getf2_leaf<double>(mifact_handle, double*, long long, int, int, int, int*, int*, int*);
template void apply_row_moves<float>(mifact_handle, float*, long long, int, int, int, int,
                                     const int*, const int*, const int*, float*, long long);
template void apply_row_moves<double>(mifact_handle, double*, long long, int, int, int, int,
                                      const int*, const int*, const int*, double*, long long);

}  // namespace mifact

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
constexpr int kMaxG = 128;           // workgroups of a leaf (rows <= 128 * 256 * 4)
constexpr int kIdxNone = 0x7fffffff; // "no candidate" row index

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

// Exchange records of one leaf column, double-buffered by column parity.
// Every word is 64 bits {payload: 32 bits, epoch: 32 bits} written by one
// coherent 64-bit store, so a reader that sees the column's epoch in a word
// also sees that word's payload; readers poll the words they need until all
// carry the epoch (no counter, no fence).  Per parity:
//   S[g]  small record of workgroup g: its candidate row and |value|;
//   R[g]  the candidate row itself (kLeafW values) and its origin tag;
//   J     row j (written by workgroup 0, read only by the pivot's owner).
// Every workgroup polls the G small records, picks the pivot, then reads one
// row record: G * SW + RW words per column instead of every candidate row.
// A workgroup can only overwrite a parity slot (column j + 2) after it has
// read every small record of column j + 1, which each workgroup publishes
// only after it has finished reading column j.
template <typename T>
struct Rec {
    static constexpr int PW = sizeof(T) / 4;       // payload words per T
    static constexpr int SIDX = 0;
    static constexpr int SVAL = 1;
    static constexpr int SW = 1 + PW;
    static constexpr int RTAG = PW * kLeafW;
    static constexpr int RW = RTAG + 1;
    static constexpr int S_OFF = 0;
    static constexpr int R_OFF = kMaxG * SW;
    static constexpr int J_OFF = R_OFF + kMaxG * RW;
    static constexpr int PARITY = J_OFF + RW;
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
    unsigned long long* xrec;       // [2][Rec<T>::PARITY]
    unsigned int epoch0;            // epoch of column j is epoch0 + j
    int* status;
    unsigned long long spin_limit;  // unsuccessful polls before giving up
    int G;
    T* inv;                         // inverse of the leaf's unit-lower block (ld 64) or null
    T* inv_zero;                    // 32 x 32 block (ld 64) to clear, or null
    int* mv_src;                    // move list of the leaf's pivots (2w entries) or null
    int* mv_dst;
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

// One thread writes a small record.
template <typename T>
__device__ inline void publish_small(unsigned long long* rec, unsigned int epoch, int idx, T val)
{
    using R = Rec<T>;
    unsigned int w[R::PW];
    to_words(val, w);
#pragma unroll
    for (int q = 0; q < R::PW; ++q) xstore64(rec + R::SVAL + q, w[q], epoch);
    xstore64(rec + R::SIDX, static_cast<unsigned int>(idx), epoch);
}

// One thread writes a row record and keeps an LDS copy of its payload.
template <typename T>
__device__ inline void publish_row(unsigned long long* rec, unsigned int* lds, unsigned int epoch,
                                   const T (&row)[kLeafW], int tag)
{
    using R = Rec<T>;
    unsigned int w[R::PW];
#pragma unroll
    for (int c = 0; c < kLeafW; ++c) {
        to_words(row[c], w);
#pragma unroll
        for (int q = 0; q < R::PW; ++q) {
            xstore64(rec + R::PW * c + q, w[q], epoch);
            lds[R::PW * c + q] = w[q];
        }
    }
    xstore64(rec + R::RTAG, static_cast<unsigned int>(tag), epoch);
    lds[R::RTAG] = static_cast<unsigned int>(tag);
}

// Shared state of one leaf workgroup.
template <typename T>
struct LeafShared {
    T wv[4];
    int wi[4];
    unsigned int s[kMaxG][Rec<T>::SW];  // small records of the column
    unsigned int prow[Rec<T>::RW];      // pivot row (polled)
    unsigned int mine[Rec<T>::RW];      // this workgroup's candidate row
    unsigned int jrow[Rec<T>::RW];      // row j (workgroup 0: own copy; pivot owner: polled)
    int p;
    int gp;
    int none;                           // no finite candidate in the column
    int piv[kLeafW];                    // local pivot row of each column (workgroup 0)
    T lb[kLeafW][kLeafW + 1];           // epilogue: the unit-lower diagonal block
};

// Polls `cnt` words at `src` (thread e reads word e, e < cnt <= 256 * NP)
// into `dst` until every word carries `epoch`.  Returns false when the
// bounded wait expired (uniform across the workgroup).
template <int NP>
__device__ inline bool poll_words(const unsigned long long* src, int cnt, unsigned int* dst,
                                  unsigned int epoch, unsigned long long limit)
{
    const int tid = threadIdx.x;
    unsigned long long spins = 0;
    while (true) {
        unsigned long long v[NP];
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const int e = tid + q * 256;
            v[q] = (e < cnt) ? xload64(src + e) : 0ULL;
        }
        int miss = 0;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const int e = tid + q * 256;
            if (e < cnt) {
                if (static_cast<unsigned int>(v[q] >> 32) != epoch) miss = 1;
                else dst[e] = static_cast<unsigned int>(v[q]);
            }
        }
        if (!__syncthreads_or(miss)) return true;
        if (++spins > limit) return false;
        __builtin_amdgcn_s_sleep(1);
    }
}

// Register rotation: column j of the leaf always sits at index 0 of the
// register rows, so the column loop is a runtime loop over fixed indices.
template <typename T, int RPT>
__device__ inline void rotate_rows(T (&x)[RPT][kLeafW])
{
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const T t0 = x[r][0];
#pragma unroll
        for (int c = 0; c < kLeafW - 1; ++c) x[r][c] = x[r][c + 1];
        x[r][kLeafW - 1] = t0;
    }
}

template <typename T, int RPT>
__global__ __launch_bounds__(256) void getf2_leaf_kernel(LeafArgs<T> a)
{
    using R = Rec<T>;
    const int tid = threadIdx.x;
    const int g = blockIdx.x;
    const int G = a.G;
    const int m = a.m;
    const int w = a.w;
    const int base = g * 256 * RPT;
    __shared__ LeafShared<T> sh;
    // A leaf whose predecessor timed out would wait on stale records: every
    // later leaf exits at once (the host reports the timeout).
    if (__hip_atomic_load(a.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;

    // Columns never reached (an expired wait) stay unexchanged: identity local
    // pivots, and ipiv entries that are valid row indices.
    if (tid < kLeafW) sh.piv[tid] = tid;
    if (g == 0 && tid < w) a.ipiv[a.c0 + tid] = a.c0 + tid + 1;
    T x[RPT][kLeafW];
    int tag[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int lrow = base + r * 256 + tid;
        const bool in = lrow < m;
#pragma unroll
        for (int c = 0; c < kLeafW; ++c)
            x[r][c] = (in && c < w) ? a.A[off(a.c0 + lrow, a.c0 + c, a.lda)] : T(0);
        tag[r] = in ? a.orig[a.c0 + lrow] : -1;
    }

    constexpr int kPollS = (kMaxG * R::SW + 255) / 256;
    for (int j = 0; j < w; ++j) {
        const unsigned int epoch = a.epoch0 + static_cast<unsigned int>(j);
        unsigned long long* slot = a.xrec + static_cast<size_t>(j & 1) * R::PARITY;

        // 1. local candidate: max |a(i, j)| over own rows i >= j, smallest row on ties
        T best = T(-1);
        int bi = kIdxNone;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int lrow = base + r * 256 + tid;
            if (lrow >= j && lrow < m) amax_pick(best, bi, fabs(x[r][0]), lrow);
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

        // 2. publish: the candidate row, then its small record; row j from workgroup 0
        if (wi == kIdxNone) {
            if (tid == 0) publish_small<T>(slot + R::S_OFF + g * R::SW, epoch, kIdxNone, T(-1));
        } else {
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int lrow = base + r * 256 + tid;
                if (lrow == wi) {
                    publish_row<T>(slot + R::R_OFF + g * R::RW, sh.mine, epoch, x[r], tag[r]);
                    publish_small<T>(slot + R::S_OFF + g * R::SW, epoch, wi, wb);
                }
            }
        }
        if (g == 0 && tid == j) publish_row<T>(slot + R::J_OFF, sh.jrow, epoch, x[0], tag[0]);

        // 3. every small record of the column
        if (!poll_words<kPollS>(slot + R::S_OFF, G * R::SW, &sh.s[0][0], epoch, a.spin_limit)) {
            if (tid == 0) {
                atomicExch(a.status, 1);
                atomicCAS(a.info, 0, -2);
            }
            return;
        }

        // 4. global pivot: max over the G candidates, smallest row on ties
        if (tid < 64) {
            T gb = T(-1);
            int gi = kIdxNone;
            for (int t = tid; t < G; t += 64)
                amax_pick(gb, gi, from_words<T>(&sh.s[t][R::SVAL]), static_cast<int>(sh.s[t][R::SIDX]));
            wave_argmax(gb, gi);
            if (tid == 0) {
                const int p = (gi == kIdxNone) ? j : gi;
                sh.p = p;
                sh.gp = p / (256 * RPT);
                sh.none = gi == kIdxNone;
            }
        }
        __syncthreads();
        const int p = sh.p;
        const int gp = sh.gp;
        const bool none = sh.none != 0;

        // 5. the pivot row (own copy when this workgroup owns it) and, for the
        //    owner of the pivot row, row j
        const bool need_p = !none && gp != g;
        const bool need_j = !none && gp == g && p != j && g != 0;
        if (need_p || need_j) {
            const unsigned long long* src = slot + (need_p ? R::R_OFF + gp * R::RW : R::J_OFF);
            unsigned int* dst = need_p ? sh.prow : sh.jrow;
            if (!poll_words<1>(src, R::RW, dst, epoch, a.spin_limit)) {
                if (tid == 0) {
                    atomicExch(a.status, 1);
                    atomicCAS(a.info, 0, -2);
                }
                return;
            }
        }
        const unsigned int* prow = need_p ? sh.prow : sh.mine;
        const unsigned int* jrow = sh.jrow;

        // 6. swap, scale by the reciprocal of the pivot (|pivot| >= sfmin, else
        //    divide, as LAPACK xGETF2), rank-1 update of the later columns
        const T piv = none ? T(0) : from_words<T>(prow);
        if (piv != T(0)) {
            if (p != j) {
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    const int lrow = base + r * 256 + tid;
                    if (lrow == p) {
#pragma unroll
                        for (int c = 0; c < kLeafW; ++c) x[r][c] = from_words<T>(jrow + R::PW * c);
                        tag[r] = static_cast<int>(jrow[R::RTAG]);
                    } else if (lrow == j) {
#pragma unroll
                        for (int c = 0; c < kLeafW; ++c) x[r][c] = from_words<T>(prow + R::PW * c);
                        tag[r] = static_cast<int>(prow[R::RTAG]);
                    }
                }
            }
            const bool recip = fabs(piv) >= sfmin<T>();
            const T rcp = T(1) / piv;
            // rotated index c is column j + c; indices past kLeafW - 1 - j hold
            // finished columns and take no update
            T u[kLeafW];
#pragma unroll
            for (int c = 1; c < kLeafW; ++c)
                u[c] = (c < kLeafW - j) ? from_words<T>(prow + R::PW * c) : T(0);
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int lrow = base + r * 256 + tid;
                if (lrow > j && lrow < m) {
                    const T l = recip ? x[r][0] * rcp : x[r][0] / piv;
                    x[r][0] = l;
#pragma unroll
                    for (int c = 1; c < kLeafW; ++c) x[r][c] -= l * u[c];
                }
            }
        } else if (g == 0 && tid == 0) {
            atomicCAS(a.info, 0, a.c0 + j + 1);
        }
        if (g == 0 && tid == 0) {
            a.ipiv[a.c0 + j] = a.c0 + p + 1;
            sh.piv[j] = p;
        }
        rotate_rows<T, RPT>(x);
    }
    for (int j = w; j < kLeafW; ++j) rotate_rows<T, RPT>(x);

#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int lrow = base + r * 256 + tid;
        if (lrow < m) {
#pragma unroll
            for (int c = 0; c < kLeafW; ++c)
                if (c < w) a.A[off(a.c0 + lrow, a.c0 + c, a.lda)] = x[r][c];
            a.orig[a.c0 + lrow] = tag[r];
        }
    }
    if (g != 0) return;

    // Epilogue (workgroup 0).
    // (a) Move list of the leaf's exchanges: candidate t < w is row t, t >= w
    //     the pivot row of step t - w; its content came from the row found
    //     by undoing the exchanges in reverse order.
    __syncthreads();
    if (a.mv_src != nullptr && tid < 2 * w) {
        const int dst = tid < w ? tid : sh.piv[tid - w];
        int pos = dst;
        for (int k = w - 1; k >= 0; --k) {
            const int pk = sh.piv[k];
            if (pos == k) pos = pk;
            else if (pos == pk) pos = k;
        }
        a.mv_src[tid] = (pos != dst) ? a.c0 + pos : -1;
        a.mv_dst[tid] = a.c0 + dst;
    }
    // (b) Inverse of the unit-lower w x w diagonal block (identity rows past w):
    //     wave v, half h solves L x = e_c for columns c = 8 v + 4 h + q by
    //     forward substitution, lane (l % 32) holding row l % 32.
    if (a.inv == nullptr) return;
    if (tid < kLeafW) {
#pragma unroll
        for (int c = 0; c < kLeafW; ++c)
            sh.lb[tid][c] = (tid < w && c < tid) ? x[0][c] : T(0);
    }
    __syncthreads();
    const int lane = tid & 63;
    const int i = lane & 31;
    const int half = lane & 32;
    T xv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = (tid >> 6) * 8 + (half >> 3) + q;
        xv[q] = (i == c) ? T(1) : T(0);
    }
#pragma unroll
    for (int k = 0; k < kLeafW; ++k) {
        const T lik = sh.lb[i][k];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const T xk = __shfl(xv[q], half + k);
            if (i > k) xv[q] -= lik * xk;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = (tid >> 6) * 8 + (half >> 3) + q;
        a.inv[i + 64 * c] = xv[q];
        if (a.inv_zero != nullptr) a.inv_zero[i + 64 * c] = T(0);
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
void getf2_leaf(mifact_handle h, T* A, long long lda, int n, int c0, int w, int* ipiv,
                int* orig, int* info, T* inv, T* inv_zero, int* mv_src, int* mv_dst)
{
    const int m = n - c0;
    if (w <= 0 || m <= 0) return;
    if (w > kLeafW) {
        report_status(h, mifact_status_invalid_value, "getf2_leaf: w > 32");
        return;
    }
    static_assert(2 * Rec<double>::PARITY * 8 <= kExchangeBytes / 2,
                  "leaf exchange records exceed their half of the exchange area");
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
    args.spin_limit = h->wait_limit > 0 ? static_cast<unsigned long long>(h->wait_limit)
                                        : (1ULL << 18);
    args.inv = inv;
    args.inv_zero = inv_zero;
    args.mv_src = mv_src;
    args.mv_dst = mv_dst;
    // Rows per thread: keep G (the workgroups that exchange per column) at 32
    // or below where the registers allow.
    int rpt = 1;
    while (rpt < 4 && ceil_div(m, 256 * rpt) > 32) rpt *= 2;
    args.G = ceil_div(m, 256 * rpt);
    if (args.G > kMaxG) {
        report_status(h, mifact_status_invalid_value, "getf2_leaf: panel taller than 131072 rows");
        return;
    }
    if (rpt == 1) launch_leaf<T, 1>(h, args);
    else if (rpt == 2) launch_leaf<T, 2>(h, args);
    else launch_leaf<T, 4>(h, args);
}

// Completes the inverse of a 64 x 64 unit-lower diagonal block from the
// inverses of its two 32 x 32 halves, which the two leaves of the block
// left on its diagonal:  inv([A 0; B C]) = [inv(A) 0; -inv(C) B inv(A)  inv(C)].
// B is read from the factors after the block's last row exchanges.
template <typename T>
__global__ __launch_bounds__(256) void inv_pair_kernel(const T* __restrict__ L, long long ldl,
                                                       T* __restrict__ inv)
{
    __shared__ T sb[32][33];
    __shared__ T sa[32][33];
    __shared__ T sc[32][33];
    __shared__ T st[32][33];
    const int tid = threadIdx.x;
    for (int e = tid; e < 1024; e += 256) {
        const int i = e & 31, k = e >> 5;
        sb[i][k] = L[off(32 + i, k, ldl)];
        sa[i][k] = inv[i + 64 * k];
        sc[i][k] = inv[32 + i + 64 * (32 + k)];
    }
    __syncthreads();
    for (int e = tid; e < 1024; e += 256) {
        const int i = e & 31, k = e >> 5;
        T s = T(0);
#pragma unroll 8
        for (int q = 0; q < 32; ++q) s += sb[i][q] * sa[q][k];
        st[i][k] = s;
    }
    __syncthreads();
    for (int e = tid; e < 1024; e += 256) {
        const int i = e & 31, k = e >> 5;
        T s = T(0);
#pragma unroll 8
        for (int q = 0; q < 32; ++q) s += sc[i][q] * st[q][k];
        inv[32 + i + 64 * k] = -s;
    }
}

template <typename T>
void inv_pair(mifact_handle h, const T* L, long long ldl, T* inv)
{
    StageScope scope(h, kStageTrsmLeaf, 2.0 * 2 * 32 * 32 * 32, sizeof(T) * 5.0 * 1024);
    hipLaunchKernelGGL(inv_pair_kernel<T>, dim3(1), dim3(256), 0, h->stream, L, ldl, inv);
    MIFACT_LAUNCHED(h);
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
// Kernels that take row indices from ipiv or a move list skip their work
// when a GETF2 leaf's bounded wait expired (status != 0): that leaf left its
// pivots and move list unwritten, and the factorization is already failed
// (info = -2).
__device__ inline bool leaf_failed(const int* status)
{
    return __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

__global__ void build_moves_kernel(int p0, int K, const int* __restrict__ ipiv,
                                   const int* __restrict__ orig, const int* __restrict__ inv,
                                   int* __restrict__ src, int* __restrict__ dst,
                                   const int* status)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * K || leaf_failed(status)) return;
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
                                                              const int* __restrict__ dst,
                                                              const int* status)
{
    if (leaf_failed(status)) return;
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
                                    int nm, const int* __restrict__ src, T* __restrict__ tmp,
                                    const int* status)
{
    if (leaf_failed(status)) return;
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
                                     const T* __restrict__ tmp, const int* status)
{
    if (leaf_failed(status)) return;
    const size_t total = static_cast<size_t>(nc) * nm;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int t = static_cast<int>(i % nm);
        const int c = static_cast<int>(i / nm);
        if (src[t] >= 0) A[off(dst[t], c0 + c, lda)] = tmp[i];
    }
}

template <typename T>
void apply_moves(mifact_handle h, T* A, long long lda, int c_begin, int c_end, int nm,
                 const int* src, const int* dst, T* tmp, long long tmp_elems)
{
    if (c_end <= c_begin || nm <= 0) return;
    Workspace ws = workspace(h);
    StageScope scope(h, kStageRowMoves, 0.0,
                     2.0 * sizeof(T) * static_cast<double>(nm) * (c_end - c_begin) + 8.0 * nm);
    const int ncols = c_end - c_begin;
    constexpr size_t kLdsBudget = 64 * 1024;
    const size_t per_col = static_cast<size_t>(nm) * sizeof(T);
    if (per_col <= kLdsBudget) {
        int CT = static_cast<int>(kLdsBudget / per_col);
        if (CT > 16) CT = 16;
        const size_t lds = static_cast<size_t>(CT) * per_col;
        hipLaunchKernelGGL(apply_moves_lds_kernel<T>, dim3(ceil_div(ncols, CT)), dim3(256), lds,
                           h->stream, A, lda, c_begin, c_end, CT, nm, src, dst,
                           h->kernel_status);
        MIFACT_LAUNCHED(h);
        return;
    }
    // Global staging: the caller's buffer when it holds at least one column
    // of the move list, else the workspace's own (2n x kStageCols doubles, so
    // at least kStageCols columns of any move list of order <= n).
    if (tmp == nullptr || tmp_elems < nm) {
        tmp = static_cast<T*>(ws.stage);
        tmp_elems = static_cast<long long>(ws.stage_bytes / sizeof(T));
    }
    const int chunk = static_cast<int>(tmp_elems / nm);
    for (int c0 = c_begin; c0 < c_end; c0 += chunk) {
        const int nc = (c_end - c0) < chunk ? (c_end - c0) : chunk;
        const size_t total = static_cast<size_t>(nc) * nm;
        const int blocks = static_cast<int>(total / 256 + 1 < 65535 ? total / 256 + 1 : 65535);
        hipLaunchKernelGGL(gather_moves_kernel<T>, dim3(blocks), dim3(256), 0, h->stream, A, lda,
                           c0, nc, nm, src, tmp, h->kernel_status);
        MIFACT_LAUNCHED(h);
        hipLaunchKernelGGL(scatter_moves_kernel<T>, dim3(blocks), dim3(256), 0, h->stream, A, lda,
                           c0, nc, nm, src, dst, tmp, h->kernel_status);
        MIFACT_LAUNCHED(h);
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
    int* src = ws.moves;
    int* dst = ws.moves + 2 * static_cast<size_t>(h->ws_n);
    {
        StageScope scope(h, kStagePerm, 0.0, 20.0 * nm);
        hipLaunchKernelGGL(build_moves_kernel, dim3(ceil_div(nm, 256)), dim3(256), 0, h->stream,
                           p0, K, ipiv, orig, inv, src, dst, h->kernel_status);
        MIFACT_LAUNCHED(h);
    }
    apply_moves<T>(h, A, lda, c_begin, c_end, nm, src, dst, tmp, tmp_elems);
}

template void getf2_leaf<float>(mifact_handle, float*, long long, int, int, int, int*, int*, int*,
                                float*, float*, int*, int*);
template void getf2_leaf<double>(mifact_handle, double*, long long, int, int, int, int*, int*,
                                 int*, double*, double*, int*, int*);
template void inv_pair<float>(mifact_handle, const float*, long long, float*);
template void inv_pair<double>(mifact_handle, const double*, long long, double*);
template void apply_moves<float>(mifact_handle, float*, long long, int, int, int, const int*,
                                 const int*, float*, long long);
template void apply_moves<double>(mifact_handle, double*, long long, int, int, int, const int*,
                                  const int*, double*, long long);
template void apply_row_moves<float>(mifact_handle, float*, long long, int, int, int, int,
                                     const int*, const int*, const int*, float*, long long);
template void apply_row_moves<double>(mifact_handle, double*, long long, int, int, int, int,
                                      const int*, const int*, const int*, double*, long long);

}  // namespace mifact

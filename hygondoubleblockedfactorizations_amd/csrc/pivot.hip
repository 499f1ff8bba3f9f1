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

constexpr int kMaxG = 128;           // workgroups of a leaf (rows <= 128 * 512)
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

// Wavefront reductions of the device library (declared by the HIP headers
// only under HIP_ENABLE_EXTRA_WARP_SYNC_TYPES).
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_max_u64(
    unsigned long long);
extern "C" __device__ __attribute__((const)) double __ockl_wfred_max_f64(double);

// Single-reduction wave arg-max for the leaf (max |v|, smallest label on
// ties; no candidate: v < 0 or label kIdxNone).  FP32 packs (|v| bits,
// complement of the label) into one 64-bit key -- non-negative floats order
// like their bit patterns -- so one wavefront max reduction decides both;
// FP64 takes the max value, then the smallest label among the lanes holding
// it.  Returns the wave-uniform result.
__device__ inline unsigned long long cand_key(float v, int i)
{
    return (i == kIdxNone || !(v >= 0.0f))
               ? 0ULL
               : (static_cast<unsigned long long>(__float_as_uint(v)) << 32)
                     | static_cast<unsigned int>(~i);
}

__device__ inline void key_to_cand(unsigned long long k, float& v, int& i)
{
    if (k == 0ULL) {
        v = -1.0f;
        i = kIdxNone;
    } else {
        v = __uint_as_float(static_cast<unsigned int>(k >> 32));
        i = ~static_cast<int>(static_cast<unsigned int>(k));
    }
}

__device__ inline void wave_argmax_fast(float& v, int& i)
{
    key_to_cand(__ockl_wfred_max_u64(cand_key(v, i)), v, i);
}

__device__ inline void wave_argmax_fast(double& v, int& i)
{
    const bool ok = i != kIdxNone && v >= 0.0;
    const double mx = __ockl_wfred_max_f64(ok ? v : -1.0);
    const int mi = __ockl_wfred_min_i32((ok && v == mx) ? i : kIdxNone);
    if (mx >= 0.0) {
        v = mx;
        i = mi;
    } else {
        v = -1.0;
        i = kIdxNone;
    }
}

// Exchange records of one leaf column, double-buffered by column parity.
// Every word is 64 bits {payload: 32 bits, epoch: 32 bits} written by one
// coherent 64-bit store, so a reader that sees the column's epoch in a word
// also sees that word's payload; readers poll until every word carries the
// epoch (no counter, no fence).  Workgroup g publishes record g: its pivot
// candidate (row, |value|), the candidate's whole row (column order) and
// its origin tag.  Every workgroup gathers all G records in one round trip,
// so the pivot row is at hand as soon as the pivot is known.  A workgroup
// can only overwrite a parity slot (column j + 2) after it has gathered
// every record of column j + 1, which each workgroup publishes only after
// it has finished with column j.
template <typename T>
struct Rec {
    static constexpr int LW = kLeafWidth<T>;
    static constexpr int PW = sizeof(T) / 4;   // payload words per T
    static constexpr int ROW = 0;               // LW values
    static constexpr int TAG = PW * LW;         // the row's origin tag
    static constexpr int VAL = TAG + 1;         // |value| of the candidate
    static constexpr int IDX = VAL + PW;        // candidate row (label)
    static constexpr int XCC = IDX + 1;         // the publisher's XCD
    static constexpr int RW = XCC + 1;          // words per record
    static constexpr int PARITY = kMaxG * RW;
};

constexpr int kLeafThreads = 512;    // 8 waves: rows of the panel in registers
constexpr int kLeafWaves = kLeafThreads / 64;
constexpr int kLeafQ = 16;           // columns per phase: the live window shrinks by kLeafQ
constexpr int kLeafU = 4;            // columns per group: registers rotate by kLeafU

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
    int G;                          // workgroups that hold rows
    int stride;                     // launched blocks per holding workgroup
    int one_xcd;                    // launch shape puts every holding workgroup on one XCD
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

// The same word for readers on the writer's own XCD only: a plain store that
// stays in the XCD's L2, which the readers' device-scope loads hit (about
// half the latency of a store written through to memory).
__device__ inline void xstore64_xcd(unsigned long long* p, unsigned int payload, unsigned int epoch)
{
    const unsigned long long v = (static_cast<unsigned long long>(epoch) << 32) | payload;
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ inline unsigned long long xload64(const unsigned long long* p)
{
    return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline unsigned int word_of(float v, int) { return __float_as_uint(v); }
__device__ inline unsigned int word_of(double v, int half)
{
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
    return static_cast<unsigned int>(half ? (b >> 32) : b);
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

template <typename T>
using LVec4 = T __attribute__((ext_vector_type(4)));

// Shared state of one leaf workgroup.
template <typename T>
struct LeafShared {
    static constexpr int LW = kLeafWidth<T>;
    // each wave's candidate row in register order (window, then the columns
    // finished in earlier phases), its multiplier of the previous column
    // and its origin tag
    alignas(16) T stage[kLeafWaves][LW];
    // the previous column's pivot row in the register order of the current
    // step, zero outside the columns that still take its update
    alignas(16) T urow[2][LW];
    T stage_l[kLeafWaves];
    unsigned int stage_tag[kLeafWaves];
    T wv[kLeafWaves];
    int wi[kLeafWaves];
    unsigned int rec[kMaxG][Rec<T>::RW];  // payloads of every record of the column
    unsigned int rec_sink[kLeafThreads];  // the first poll's out-of-range slots land here
    int abort;                            // a bounded wait expired
    int piv[LW];                          // local pivot row of each column (workgroup 0)
    T lb[LW][LW + 1];                     // workgroup 0: the leaf's pivot rows
};

// Register index, at a phase whose live window is W wide, of column c < LW - W
// (finished in an earlier phase): each phase leaves its kLeafQ columns, in
// order, just above the next window, so the blocks of earlier phases lie
// above in reverse phase order.
template <int LW>
__device__ inline int leaf_index(int W, int c)
{
    const int done = (LW - W) / kLeafQ;   // phases finished
    return W + kLeafQ * (done - 1 - c / kLeafQ) + c % kLeafQ;
}

// Column j = base + U of a leaf (see getf2_leaf_kernel).  Register index i
// of x holds column base + i for i < LW - base, the columns finished in this
// phase (wrapped) above that up to W, and above W the columns of earlier
// phases (leaf_index).
// lp carries the multipliers of column j - 1, whose update of the columns
// after j is deferred to this step so that it runs while the column's
// records are in flight.  Returns false when a bounded wait expired
// (uniform over the workgroup).
template <typename T, int RPT, int NP, int W, int U>
__device__ __attribute__((always_inline)) inline bool leaf_step(
    const LeafArgs<T>& a, LeafShared<T>& sh, T (&x)[RPT][kLeafWidth<T>], int (&tag)[RPT],
    int (&lab)[RPT], T (&lp)[RPT], bool& local, int base, int g)
{
    using R = Rec<T>;
    constexpr int LW = kLeafWidth<T>;
    constexpr int QP = LW - W;     // columns finished in earlier phases
    const int j = base + U;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int G = a.G;
    const unsigned int epoch = a.epoch0 + static_cast<unsigned int>(j);
    unsigned long long* slot = a.xrec + static_cast<size_t>(j & 1) * R::PARITY;
    const int par = j & 1;
    const bool has_prev = j > 0;

    // 1. local candidate of column j (current in x[r][U]); each wave's
    //    winning lane stages its row
    T best = T(-1);
    int bi = kIdxNone;
#pragma unroll
    for (int r = 0; r < RPT; ++r)
        if (lab[r] >= j && lab[r] != kIdxNone) amax_pick(best, bi, fabs(x[r][U]), lab[r]);
    wave_argmax_fast(best, bi);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        if (bi != kIdxNone && lab[r] == bi) {
#pragma unroll
            for (int q = 0; q < LW; q += 4)
                *reinterpret_cast<LVec4<T>*>(&sh.stage[wave][q]) =
                    LVec4<T>{x[r][q], x[r][q + 1], x[r][q + 2], x[r][q + 3]};
            sh.stage_l[wave] = lp[r];
            sh.stage_tag[wave] = static_cast<unsigned int>(tag[r]);
        }
    }
    if (lane == 0) {
        sh.wv[wave] = best;
        sh.wi[wave] = bi;
    }
    __syncthreads();

    // 2. the workgroup's candidate: best over the waves
    T wb = sh.wv[0];
    int wi = sh.wi[0];
    int qw = 0;
#pragma unroll
    for (int q = 1; q < kLeafWaves; ++q) {
        const T ov = sh.wv[q];
        const int oi = sh.wi[q];
        if (ov > wb || (ov == wb && oi < wi)) {
            wb = ov;
            wi = oi;
            qw = q;
        }
    }

    // 3. publish the record, one word per thread; the row's columns after j
    //    take the deferred update of column j - 1 here (the same fused
    //    multiply-add its owner applies below)
    if (tid < R::RW) {
        unsigned int word;
        if (tid == R::XCC) {
            word = xcd_id();
        } else if (wi == kIdxNone) {
            word = (tid == R::TAG) ? 0xffffffffu
                 : (tid >= R::VAL && tid < R::IDX) ? word_of(T(-1), tid - R::VAL)
                 : (tid == R::IDX) ? static_cast<unsigned int>(kIdxNone) : 0u;
        } else if (tid < R::TAG) {
            const int c = tid / R::PW;
            const int i = (c < QP) ? leaf_index<LW>(W, c) : ((c >= base) ? c - base : c - base + W);
            T v = sh.stage[qw][i];
            if (has_prev && i < W) v = fma(-sh.stage_l[qw], sh.urow[par ^ 1][i], v);
            word = word_of(v, tid % R::PW);
        } else if (tid == R::TAG) {
            word = sh.stage_tag[qw];
        } else if (tid < R::IDX) {
            word = word_of(wb, tid - R::VAL);
        } else {
            word = static_cast<unsigned int>(wi);
        }
        if (local) xstore64_xcd(slot + static_cast<size_t>(g) * R::RW + tid, word, epoch);
        else xstore64(slot + static_cast<size_t>(g) * R::RW + tid, word, epoch);
    }

    // 4. deferred update of column j - 1 on own rows, columns after j
    if (has_prev) {
#pragma unroll
        for (int q = 0; q < W; q += 4) {
            if (q + 3 <= U) continue;
            const LVec4<T> u4 = *reinterpret_cast<const LVec4<T>*>(&sh.urow[par ^ 1][q]);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (q + e <= U) continue;
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    x[r][q + e] = fma(-lp[r], u4[e], x[r][q + e]);
                    // pinned here: not sunk past the gather below (the point
                    // is to run while the records travel), and the
                    // pivot-row reads are not all held at once
                    asm volatile("" : "+v"(x[r][q + e]));
                }
            }
        }
    }

    // 5. gather: each thread polls its own words of the G records until all
    //    carry the column's epoch; one barrier joins the workgroup.  The
    //    first poll issues every slot's load without branches: a slot past
    //    the records re-reads word idx - total (its own word, no shared hot
    //    spot) into a per-thread sink, and only words not yet published are
    //    polled again.
    {
        const int total = G * R::RW;
        unsigned int pending = 0;
        {
            unsigned long long v[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const int idx = tid + q * kLeafThreads;
                v[q] = xload64(slot + (idx < total ? idx : idx - total));
            }
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const int idx = tid + q * kLeafThreads;
                const bool valid = idx < total;
                unsigned int* dst = valid ? &sh.rec[0][0] + idx : &sh.rec_sink[tid];
                *dst = static_cast<unsigned int>(v[q]);
                pending |= (valid && static_cast<unsigned int>(v[q] >> 32) != epoch) ? (1u << q)
                                                                                       : 0u;
            }
        }
        unsigned long long spins = 0;
        while (pending != 0) {
            if (++spins > a.spin_limit) {
                sh.abort = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            unsigned long long v[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q)
                v[q] = ((pending >> q) & 1u) ? xload64(slot + tid + q * kLeafThreads) : 0ULL;
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                if (((pending >> q) & 1u) && static_cast<unsigned int>(v[q] >> 32) == epoch) {
                    (&sh.rec[0][0])[tid + q * kLeafThreads] = static_cast<unsigned int>(v[q]);
                    pending &= ~(1u << q);
                }
            }
        }
    }
    __syncthreads();
    if (sh.abort) {
        if (tid == 0) {
            atomicExch(a.status, 1);
            atomicCAS(a.info, 0, -2);
        }
        return false;
    }
    // after the leaf's first column: if every record came from this
    // workgroup's XCD, later columns publish with XCD-local stores
    if (U == 0 && j == 0 && a.one_xcd) {
        const unsigned int x0 = sh.rec[0][R::XCC];
        bool same = true;
        for (int t = lane; t < G; t += 64) same = same && sh.rec[t][R::XCC] == x0;
        local = __all(same);
    }

    // 6. global pivot, computed by every wave from the gathered records (lane
    //    t takes record t, then the wave arg-max): max over the G candidates,
    //    smallest label on ties (the key carries the record index in its low
    //    8 bits)
    T gb = T(-1);
    int gk = kIdxNone;
    // one record per lane (two passes only for the variant sized beyond 64
    // workgroups)
    constexpr int kPasses = NP > (64 * R::RW + kLeafThreads - 1) / kLeafThreads ? 2 : 1;
#pragma unroll
    for (int ps = 0; ps < kPasses; ++ps) {
        const int t = lane + 64 * ps;
        if (t < G) {
            const int idx = static_cast<int>(sh.rec[t][R::IDX]);
            amax_pick(gb, gk, from_words<T>(&sh.rec[t][R::VAL]),
                      idx == kIdxNone ? kIdxNone : (idx << 8) | t);
        }
    }
    wave_argmax_fast(gb, gk);
    const bool none = gk == kIdxNone;
    const int p = none ? j : (gk >> 8);
    const unsigned int* prow = sh.rec[none ? 0 : (gk & 0xff)];
    const T piv = none ? T(0) : from_words<T>(prow + R::ROW + R::PW * j);
    T l[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) l[r] = T(0);
    if (piv != T(0)) {
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            if (lab[r] == p) lab[r] = j;
            else if (lab[r] == j) lab[r] = p;
        }
        // |pivot| >= sfmin: multiply by the reciprocal, else divide (xGETF2);
        // the test is uniform, so only one of the two paths runs
        if (fabs(piv) >= sfmin<T>()) {
            const T rcp = T(1) / piv;
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const bool upd = lab[r] > j && lab[r] != kIdxNone;
                const T lv = x[r][U] * rcp;
                l[r] = upd ? lv : T(0);
                if (upd) x[r][U] = lv;
            }
        } else {
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const bool upd = lab[r] > j && lab[r] != kIdxNone;
                const T lv = x[r][U] / piv;
                l[r] = upd ? lv : T(0);
                if (upd) x[r][U] = lv;
            }
        }
    } else if (g == 0 && tid == 0) {
        atomicCAS(a.info, 0, a.c0 + j + 1);
    }

    // 7. column j + 1 takes this pivot's update now (its candidates come
    //    next); the pivot row goes to urow in the register order of step
    //    j + 1 for the deferred update of the other columns
    {
        const T un = (piv != T(0) && j + 1 < LW) ? from_words<T>(prow + R::ROW + R::PW * (j + 1))
                                                  : T(0);
#pragma unroll
        for (int r = 0; r < RPT; ++r) x[r][U + 1] = fma(-l[r], un, x[r][U + 1]);
    }
    if (tid < LW) {
        constexpr int UN = (U == kLeafU - 1) ? 0 : U + 1;
        const int base_n = (U == kLeafU - 1) ? base + kLeafU : base;
        const int i = tid;
        T v = T(0);
        if (piv != T(0) && i > UN && i < LW - base_n)
            v = from_words<T>(prow + R::ROW + R::PW * (base_n + i));
        sh.urow[par][i] = v;
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) lp[r] = l[r];

    // workgroup 0 keeps the pivot row (its multipliers before column j are
    // final) and the pivot
    if (g == 0 && tid < LW) sh.lb[j][tid] = none ? T(0) : from_words<T>(prow + R::ROW + R::PW * tid);
    if (g == 0 && tid == 0) {
        a.ipiv[a.c0 + j] = a.c0 + p + 1;
        sh.piv[j] = p;
    }
    return true;
}

// Rotates the live window x[r][0, W) left by kLeafU.
template <typename T, int RPT, int W>
__device__ __attribute__((always_inline)) inline void leaf_rotate(T (&x)[RPT][kLeafWidth<T>])
{
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        T t[kLeafU];
#pragma unroll
        for (int u = 0; u < kLeafU; ++u) t[u] = x[r][u];
#pragma unroll
        for (int c = 0; c < W - kLeafU; ++c) x[r][c] = x[r][c + kLeafU];
#pragma unroll
        for (int u = 0; u < kLeafU; ++u) x[r][W - kLeafU + u] = t[u];
    }
}

// One phase: columns [LW - W, LW - W + kLeafQ) in groups of kLeafU (a
// runtime loop over fixed register indices, the window rotating by kLeafU
// after each group); at its end the phase's columns sit, in order, at
// [W - kLeafQ, W), outside the next window.
template <typename T, int RPT, int NP, int W>
__device__ __attribute__((always_inline)) inline bool leaf_phase(
    const LeafArgs<T>& a, LeafShared<T>& sh, T (&x)[RPT][kLeafWidth<T>], int (&tag)[RPT],
    int (&lab)[RPT], T (&lp)[RPT], bool& local, int g)
{
    constexpr int LW = kLeafWidth<T>;
    constexpr int QP = LW - W;
    for (int base = QP; base < QP + kLeafQ; base += kLeafU) {
        if (base < a.w && !leaf_step<T, RPT, NP, W, 0>(a, sh, x, tag, lab, lp, local, base, g))
            return false;
        if (base + 1 < a.w && !leaf_step<T, RPT, NP, W, 1>(a, sh, x, tag, lab, lp, local, base, g))
            return false;
        if (base + 2 < a.w && !leaf_step<T, RPT, NP, W, 2>(a, sh, x, tag, lab, lp, local, base, g))
            return false;
        if (base + 3 < a.w && !leaf_step<T, RPT, NP, W, 3>(a, sh, x, tag, lab, lp, local, base, g))
            return false;
        leaf_rotate<T, RPT, W>(x);
    }
    return true;
}

// Cooperative partial-pivot LU of an m x w leaf panel (w <= LW: 64 columns
// in FP32, 32 in FP64).  Blocks b with b % stride == 0 hold rows: workgroup
// g = b / stride loads rows [g * 512 * RPT, (g + 1) * 512 * RPT) of the
// panel, RPT per thread, into registers for the whole leaf; the other
// blocks exit at once.  The hardware deals consecutive blocks to the XCDs
// in turn, so a stride of 8 keeps the exchanging workgroups on one XCD.
// Rows never move between threads: each register row carries a label, its
// current row index in the LAPACK sense, and a row exchange swaps two
// labels.  At the end every row is stored at its label.  Per column j:
//   1. local candidate: wave64 arg-max, then across the 8 waves in LDS
//      (max |a(i, j)| over labels i >= j, smallest label on ties: IxAMAX);
//   2. publish the record (the candidate row in column order);
//   3. apply column j - 1's update to the columns after j of own rows while
//      the records travel;
//   4. gather every record of the column (one round trip once all are out);
//   5. the global pivot p (every thread reduces the G records itself),
//      relabel j <-> p, scale by the reciprocal of the pivot (|pivot| >=
//      sfmin, else divide, as LAPACK xGETF2) and update column j + 1 only.
// Each element takes its updates in pivot order with the same fused
// multiply-add, so the factors are those of the plain column-by-column
// elimination.  The live register window shrinks by kLeafQ columns per phase
// (finished columns leave it), which bounds the update work.  Workgroup 0
// keeps each column's pivot row: they are the leaf's first rows, used by the
// epilogue (inverse of the unit-lower block).
template <typename T, int RPT, int NP>
__global__ __launch_bounds__(kLeafThreads) void getf2_leaf_kernel(LeafArgs<T> a)
{
    constexpr int LW = kLeafWidth<T>;
    constexpr int ROWS = kLeafThreads * RPT;
    if (static_cast<int>(blockIdx.x) % a.stride != 0) return;
    const int tid = threadIdx.x;
    const int g = static_cast<int>(blockIdx.x) / a.stride;
    const int m = a.m;
    const int w = a.w;
    const int base = g * ROWS;
    __shared__ LeafShared<T> sh;
    // A leaf whose predecessor timed out would wait on stale records: every
    // later leaf exits at once (the host reports the timeout).
    if (__hip_atomic_load(a.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;

    // Columns never reached (an expired wait) stay unexchanged: identity local
    // pivots, and ipiv entries that are valid row indices.
    if (tid < LW) sh.piv[tid] = tid;
    if (tid == 0) sh.abort = 0;
    if (g == 0 && tid < w) a.ipiv[a.c0 + tid] = a.c0 + tid + 1;
    T x[RPT][LW];
    int tag[RPT];
    int lab[RPT];
    T lp[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int lrow = base + r * kLeafThreads + tid;
        const bool in = lrow < m;
#pragma unroll
        for (int c = 0; c < LW; ++c)
            x[r][c] = (in && c < w) ? a.A[off(a.c0 + lrow, a.c0 + c, a.lda)] : T(0);
        tag[r] = in ? a.orig[a.c0 + lrow] : -1;
        lab[r] = in ? lrow : kIdxNone;
        lp[r] = T(0);
    }

    // the first column publishes with device-scope stores (and its XCD);
    // later ones with XCD-local stores if every record came from one XCD
    bool local = false;
    bool ok = leaf_phase<T, RPT, NP, LW>(a, sh, x, tag, lab, lp, local, g);
    if constexpr (LW > kLeafQ) ok = ok && leaf_phase<T, RPT, NP, LW - kLeafQ>(a, sh, x, tag, lab, lp, local, g);
    if constexpr (LW > 2 * kLeafQ) ok = ok && leaf_phase<T, RPT, NP, LW - 2 * kLeafQ>(a, sh, x, tag, lab, lp, local, g);
    if constexpr (LW > 3 * kLeafQ) ok = ok && leaf_phase<T, RPT, NP, LW - 3 * kLeafQ>(a, sh, x, tag, lab, lp, local, g);
    static_assert(LW <= 4 * kLeafQ && LW % kLeafQ == 0, "leaf of one to four phases");
    if (!ok) return;

    // after the last phase register i holds column leaf_index<LW>(0, i)
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        if (lab[r] != kIdxNone) {
            const int row = a.c0 + lab[r];
#pragma unroll
            for (int i = 0; i < LW; ++i) {
                constexpr int NPH = LW / kLeafQ;
                const int c = kLeafQ * (NPH - 1 - i / kLeafQ) + i % kLeafQ;
                if (c < w) a.A[off(row, a.c0 + c, a.lda)] = x[r][i];
            }
            a.orig[row] = tag[r];
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
    // (b) Inverse of the unit-lower w x w diagonal block (identity rows past
    //     w) by forward substitution: wave v solves L x = e_c for its LW / 8
    //     columns c, lane i < LW holding row i.
    if (a.inv == nullptr) return;
    constexpr int CPW = LW / kLeafWaves;
    const int lane = tid & 63;
    const int i = lane;
    const int wv = tid >> 6;
    T xv[CPW];
#pragma unroll
    for (int q = 0; q < CPW; ++q) xv[q] = (i == wv * CPW + q) ? T(1) : T(0);
    for (int k = 0; k < LW; ++k) {
        const T lik = (i < w && i < LW && k < i) ? sh.lb[i][k] : T(0);
#pragma unroll
        for (int q = 0; q < CPW; ++q) {
            const T xk = __shfl(xv[q], k);
            if (i > k) xv[q] -= lik * xk;
        }
    }
    if (i < LW) {
#pragma unroll
        for (int q = 0; q < CPW; ++q) {
            const int c = wv * CPW + q;
            a.inv[i + 64 * c] = xv[q];
            if (a.inv_zero != nullptr) a.inv_zero[i + 64 * c] = T(0);
        }
    }
}

template <typename T, int RPT, int NP>
static void launch_leaf(mifact_handle h, LeafArgs<T>& args)
{
    const double m = args.m, w = args.w;
    StageScope scope(h, kStageGetf2Leaf, m * w * w - w * w * w / 3.0,
                     sizeof(T) * 2.0 * m * w + 8.0 * m);
    void* params[] = {&args};
    hipError_t e = hipLaunchCooperativeKernel(
        reinterpret_cast<const void*>(&getf2_leaf_kernel<T, RPT, NP>), dim3(args.G * args.stride),
        dim3(kLeafThreads), params, 0, h->stream);
    if (e != hipSuccess) {
        report_hip(h, e, "hipLaunchCooperativeKernel(getf2_leaf)", __FILE__, __LINE__);
        return;
    }
    h->leaf_epoch += static_cast<unsigned int>(args.w);
}

template <typename T, int RPT>
static void launch_leaf_np(mifact_handle h, LeafArgs<T>& args)
{
    // words gathered per thread: up to 16, 32, 64 or kMaxG records
    constexpr int kSmall = (16 * Rec<T>::RW + kLeafThreads - 1) / kLeafThreads;
    constexpr int kMid = (32 * Rec<T>::RW + kLeafThreads - 1) / kLeafThreads;
    constexpr int kWide = (64 * Rec<T>::RW + kLeafThreads - 1) / kLeafThreads;
    constexpr int kLarge = (kMaxG * Rec<T>::RW + kLeafThreads - 1) / kLeafThreads;
    static_assert(kLarge <= 32, "poll mask is 32 bits");
    if (args.G <= 16) launch_leaf<T, RPT, kSmall>(h, args);
    else if (args.G <= 32) launch_leaf<T, RPT, kMid>(h, args);
    else if (args.G <= 64) launch_leaf<T, RPT, kWide>(h, args);
    else launch_leaf<T, RPT, kLarge>(h, args);
}

template <typename T>
void getf2_leaf(mifact_handle h, T* A, long long lda, int n, int c0, int w, int* ipiv,
                int* orig, int* info, T* inv, T* inv_zero, int* mv_src, int* mv_dst)
{
    const int m = n - c0;
    if (w <= 0 || m <= 0) return;
    if (w > kLeafWidth<T>) {
        report_status(h, mifact_status_invalid_value, "getf2_leaf: w above the leaf width");
        return;
    }
    static_assert(2 * Rec<float>::PARITY * 8 <= kExchangeBytes / 2
                      && 2 * Rec<double>::PARITY * 8 <= kExchangeBytes / 2,
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
    // One row per thread (LW values in registers; a second row would not
    // fit beside the rotating window without spills).
    args.G = ceil_div(m, kLeafThreads);
    if (args.G > kMaxG) {
        report_status(h, mifact_status_invalid_value, "getf2_leaf: panel too tall");
        return;
    }
    // One CU per launched block: at most num_cu blocks, stride 8 (one XCD)
    // where that fits.
    int stride = 8;
    while (stride > 1 && args.G * stride > h->num_cu) stride /= 2;
    args.stride = stride;
    args.one_xcd = stride == 8;
    launch_leaf_np<T, 1>(h, args);
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
    return status != nullptr
        && __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
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

// ------------------------------------------------------------------ pivot launchers
//
// The reference's stand-alone pivot launchers (include/getrf/
// getrf_pivot_kernels.h:9-43, src/getrf/getrf_pivot_kernels.hip:14-206):
// stream-ordered, no handle, no allocation.

// Launch errors exit, as the reference's HIP_CHECK does (error_check.h:9-28).
inline void report_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) report_hip(nullptr, e, what, __FILE__, __LINE__);
}

constexpr int kPanelChunk = 256;   // pivots composed per step
template <typename T> constexpr int kPanelCols = 64 / sizeof(T);   // columns per workgroup

// LAPACK laswp of one panel's local pivots over the columns outside the
// panel: swaps rows panel + k <-> panel + pivots[k] - 1, k = 0 .. ps - 1 in
// order.  A workgroup owns 16 / 8 columns (FP32 / FP64); pivots are taken kPanelChunk
// at a time: the chunk's swaps are composed into a move list in LDS (each
// candidate row's source found by undoing the chunk's swaps in reverse),
// the moved entries of the workgroup's columns are gathered into LDS, then
// scattered.  Chunks run in order, so the composition equals the sequence.
template <typename T>
__global__ __launch_bounds__(256) void panel_pivots_kernel(T* A, long long lda, int n, int panel,
                                                           int ps, const int* __restrict__ piv)
{
    __shared__ int sp[kPanelChunk];
    __shared__ int src[2 * kPanelChunk];
    __shared__ int dst[2 * kPanelChunk];
    __shared__ T vals[kPanelCols<T>][2 * kPanelChunk];   // 32 KiB
    const int tid = threadIdx.x;
    const int ext = n - ps;                       // columns outside the panel
    const int e0 = blockIdx.x * kPanelCols<T>;
    const int ncols = (ext - e0) < kPanelCols<T> ? (ext - e0) : kPanelCols<T>;
    for (int k0 = 0; k0 < ps; k0 += kPanelChunk) {
        const int kc = (ps - k0) < kPanelChunk ? (ps - k0) : kPanelChunk;
        for (int t = tid; t < kc; t += 256) sp[t] = piv[k0 + t] - 1 - k0;  // chunk-local
        __syncthreads();
        for (int t = tid; t < 2 * kc; t += 256) {
            const int d = t < kc ? t : sp[t - kc];
            int pos = d;
            for (int k = kc - 1; k >= 0; --k) {
                const int pk = sp[k];
                if (pos == k) pos = pk;
                else if (pos == pk) pos = k;
            }
            src[t] = (pos != d) ? panel + k0 + pos : -1;
            dst[t] = panel + k0 + d;
        }
        __syncthreads();
        for (int e = tid; e < ncols * 2 * kc; e += 256) {
            const int c = e / (2 * kc), t = e % (2 * kc);
            const int ec = e0 + c;
            const int col = ec < panel ? ec : ec + ps;
            if (src[t] >= 0) vals[c][t] = A[off(src[t], col, lda)];
        }
        __syncthreads();
        for (int e = tid; e < ncols * 2 * kc; e += 256) {
            const int c = e / (2 * kc), t = e % (2 * kc);
            const int ec = e0 + c;
            const int col = ec < panel ? ec : ec + ps;
            if (src[t] >= 0) A[off(dst[t], col, lda)] = vals[c][t];
        }
        __syncthreads();
    }
}

__global__ void globalize_pivots_kernel(int* piv, int panel, int ps)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ps) piv[i] += panel;
}

template <typename T>
void launch_panel_pivots(T* A, int lda, int n, int panel, int ps, int* pivots, hipStream_t s)
{
    if (ps <= 0) return;
    const int ext = n - ps;
    if (ext > 0) {
        hipLaunchKernelGGL(panel_pivots_kernel<T>, dim3(ceil_div(ext, kPanelCols<T>)), dim3(256), 0,
                           s, A, static_cast<long long>(lda), n, panel, ps, pivots);
        report_launch("launch_apply_getrf_panel_pivots");
    }
    hipLaunchKernelGGL(globalize_pivots_kernel, dim3(ceil_div(ps, 256)), dim3(256), 0, s, pivots,
                       panel, ps);
    report_launch("launch_apply_getrf_panel_pivots");
}

// Row moves over columns [0, column_count): row targets[m] receives the
// pre-call entry of row sources[m].  The LDS kernel when a column of the
// move list fits in LDS, else the caller's temporary in column tiles (the
// reference's arrangement).
template <typename T>
void launch_permutation(T* A, int lda, int column_count, const int* sources, const int* targets,
                        int move_count, T* temporary, int column_tile, hipStream_t s)
{
    if (column_count <= 0 || move_count <= 0) return;
    constexpr size_t kLdsBudget = 64 * 1024;
    const size_t per_col = static_cast<size_t>(move_count) * sizeof(T);
    if (per_col <= kLdsBudget) {
        int CT = static_cast<int>(kLdsBudget / per_col);
        if (CT > 16) CT = 16;
        hipLaunchKernelGGL(apply_moves_lds_kernel<T>, dim3(ceil_div(column_count, CT)), dim3(256),
                           static_cast<size_t>(CT) * per_col, s, A, static_cast<long long>(lda), 0,
                           column_count, CT, move_count, sources, targets,
                           static_cast<const int*>(nullptr));
        report_launch("launch_apply_getrf_permutation");
        return;
    }
    if (temporary == nullptr || column_tile <= 0) return;
    for (int c0 = 0; c0 < column_count; c0 += column_tile) {
        const int nc = (column_count - c0) < column_tile ? (column_count - c0) : column_tile;
        const size_t total = static_cast<size_t>(nc) * move_count;
        const int blocks = static_cast<int>(total / 256 + 1 < 65535 ? total / 256 + 1 : 65535);
        hipLaunchKernelGGL(gather_moves_kernel<T>, dim3(blocks), dim3(256), 0, s, A,
                           static_cast<long long>(lda), c0, nc, move_count, sources, temporary,
                           static_cast<const int*>(nullptr));
        report_launch("launch_apply_getrf_permutation");
        hipLaunchKernelGGL(scatter_moves_kernel<T>, dim3(blocks), dim3(256), 0, s, A,
                           static_cast<long long>(lda), c0, nc, move_count, sources, targets,
                           temporary, static_cast<const int*>(nullptr));
        report_launch("launch_apply_getrf_permutation");
    }
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

extern "C" {

void launch_apply_getrf_panel_pivots(float* A, int lda, int n, int panel, int panel_size,
                                     int* pivots, hipStream_t stream)
{
    mifact::launch_panel_pivots<float>(A, lda, n, panel, panel_size, pivots, stream);
}

void launch_apply_getrf_panel_pivots_double(double* A, int lda, int n, int panel, int panel_size,
                                            int* pivots, hipStream_t stream)
{
    mifact::launch_panel_pivots<double>(A, lda, n, panel, panel_size, pivots, stream);
}

void launch_apply_getrf_permutation(float* A, int lda, int column_count, const int* sources,
                                    const int* targets, int move_count, float* temporary,
                                    int column_tile, hipStream_t stream)
{
    mifact::launch_permutation<float>(A, lda, column_count, sources, targets, move_count,
                                      temporary, column_tile, stream);
}

void launch_apply_getrf_permutation_double(double* A, int lda, int column_count,
                                           const int* sources, const int* targets,
                                           int move_count, double* temporary, int column_tile,
                                           hipStream_t stream)
{
    mifact::launch_permutation<double>(A, lda, column_count, sources, targets, move_count,
                                       temporary, column_tile, stream);
}

}  // extern "C"

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
constexpr int kMaxG = 256;

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
    T* xval;       // [2][kMaxG + 1]
    int* xidx;     // [2][kMaxG + 1]
    T* xrow;       // [2][kMaxG + 1][kLeafW]
    int* xtag;     // [2][kMaxG + 1]
    unsigned long long* bar;
    unsigned long long bar_base;
    int* status;
    int G;
};

// Grid barrier over G workgroups.  Returns false when the bounded wait
// expired (the kernel then exits; the host reports a timeout).
__device__ inline bool grid_barrier(unsigned long long* bar, unsigned long long target,
                                    int* status, int* s_abort)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(bar, 1ULL);
        unsigned long long spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++spins > (1ULL << 24)) {
                atomicExch(status, 1);
                *s_abort = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __threadfence();
    }
    __syncthreads();
    return *s_abort == 0;
}

// Shared state of one leaf workgroup.
template <typename T>
struct LeafShared {
    T wv[4];
    int wi[4];
    T prow[kLeafW];
    T jrow[kLeafW];
    int ptag, jtag, p;
    int abort;
};

// Column J of the leaf (J is a compile-time constant so the register rows
// x[r][J] stay in registers).  Returns false when the grid barrier timed out.
template <typename T, int RPT, int J>
__device__ __attribute__((always_inline)) inline bool leaf_column(
    const LeafArgs<T>& a, LeafShared<T>& sh, T (&x)[RPT][kLeafW], int (&tag)[RPT], int base)
{
    const int tid = threadIdx.x;
    const int g = blockIdx.x;
    const int G = a.G;
    const int m = a.m;
    const int w = a.w;
    const int R = kMaxG + 1;
    constexpr int j = J;
    const int par = j & 1;
    // 1. local candidate
    T best = T(-1);
    int bi = 0x7fffffff;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int lrow = base + r * 256 + tid;
        if (lrow >= j && lrow < m) amax_pick(best, bi, fabs(x[r][j]), lrow);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T ov = __shfl_xor(best, o);
        const int oi = __shfl_xor(bi, o);
        amax_pick(best, bi, ov, oi);
    }
    if ((tid & 63) == 0) {
        sh.wv[tid >> 6] = best;
        sh.wi[tid >> 6] = bi;
    }
    __syncthreads();
    T wb = sh.wv[0];
    int wi = sh.wi[0];
#pragma unroll
    for (int q = 1; q < 4; ++q) amax_pick(wb, wi, sh.wv[q], sh.wi[q]);

    // 2. publish
    const size_t rec = static_cast<size_t>(par) * R + g;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int lrow = base + r * 256 + tid;
        if (lrow == wi) {
            a.xval[rec] = wb;
            a.xidx[rec] = wi;
#pragma unroll
            for (int c = 0; c < kLeafW; ++c) a.xrow[rec * kLeafW + c] = x[r][c];
            a.xtag[rec] = tag[r];
        }
    }
    if (wi == 0x7fffffff && tid == 0) {
        a.xval[rec] = T(-1);
        a.xidx[rec] = 0x7fffffff;
    }
    if (g == 0 && tid == j) {
        const size_t jrec = static_cast<size_t>(par) * R + G;
#pragma unroll
        for (int c = 0; c < kLeafW; ++c) a.xrow[jrec * kLeafW + c] = x[0][c];
        a.xtag[jrec] = tag[0];
    }

    // 3. grid barrier
    if (!grid_barrier(a.bar, a.bar_base + static_cast<unsigned long long>(j + 1) * G, a.status,
                      &sh.abort)) {
        return false;
    }

    // 4. global pivot
    best = T(-1);
    bi = 0x7fffffff;
    for (int t = tid; t < G; t += 256) {
        const size_t rr = static_cast<size_t>(par) * R + t;
        amax_pick(best, bi, a.xval[rr], a.xidx[rr]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const T ov = __shfl_xor(best, o);
        const int oi = __shfl_xor(bi, o);
        amax_pick(best, bi, ov, oi);
    }
    if ((tid & 63) == 0) {
        sh.wv[tid >> 6] = best;
        sh.wi[tid >> 6] = bi;
    }
    __syncthreads();
    if (tid == 0) {
        T gb = sh.wv[0];
        int gi = sh.wi[0];
        for (int q = 1; q < 4; ++q) amax_pick(gb, gi, sh.wv[q], sh.wi[q]);
        sh.p = (gi == 0x7fffffff) ? j : gi;
    }
    __syncthreads();
    const int p = sh.p;
    const int gp = p / (256 * RPT);
    if (tid < kLeafW) {
        const bool c_in = tid < w;
        sh.prow[tid] = c_in ? a.xrow[(static_cast<size_t>(par) * R + gp) * kLeafW + tid] : T(0);
        sh.jrow[tid] = c_in ? a.xrow[(static_cast<size_t>(par) * R + G) * kLeafW + tid] : T(0);
    }
    if (tid == 0) {
        sh.ptag = a.xtag[static_cast<size_t>(par) * R + gp];
        sh.jtag = a.xtag[static_cast<size_t>(par) * R + G];
    }
    __syncthreads();

    const T piv = sh.prow[j];
    if (piv != T(0)) {
        if (p != j) {
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int lrow = base + r * 256 + tid;
                if (lrow == p) {
#pragma unroll
                    for (int c = 0; c < kLeafW; ++c) x[r][c] = sh.jrow[c];
                    tag[r] = sh.jtag;
                } else if (lrow == j) {
#pragma unroll
                    for (int c = 0; c < kLeafW; ++c) x[r][c] = sh.prow[c];
                    tag[r] = sh.ptag;
                }
            }
        }
        const bool recip = fabs(piv) >= sfmin<T>();
        const T rcp = T(1) / piv;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int lrow = base + r * 256 + tid;
            if (lrow > j && lrow < m) {
                const T l = recip ? x[r][j] * rcp : x[r][j] / piv;
                x[r][j] = l;
#pragma unroll
                for (int c = j + 1; c < kLeafW; ++c) x[r][c] -= l * sh.prow[c];
            }
        }
    } else if (g == 0 && tid == 0) {
        atomicCAS(a.info, 0, a.c0 + j + 1);
    }
    if (g == 0 && tid == 0) a.ipiv[a.c0 + j] = a.c0 + p + 1;
    return true;
}

template <typename T, int RPT, int... Js>
__device__ __attribute__((always_inline)) inline void leaf_columns(
    const LeafArgs<T>& a, LeafShared<T>& sh, T (&x)[RPT][kLeafW], int (&tag)[RPT], int base,
    std::integer_sequence<int, Js...>)
{
    bool ok = true;
    ((ok = ok && (Js >= a.w || leaf_column<T, RPT, Js>(a, sh, x, tag, base))), ...);
}

template <typename T, int RPT>
__global__ __launch_bounds__(256) void getf2_leaf_kernel(LeafArgs<T> a)
{
    const int tid = threadIdx.x;
    const int base = blockIdx.x * 256 * RPT;
    __shared__ LeafShared<T> sh;
    // A leaf whose predecessor timed out leaves the grid-barrier counter out
    // of step: every later leaf exits at once (the host reports the timeout).
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
    if (tid == 0) sh.abort = 0;
    __syncthreads();

    leaf_columns<T, RPT>(a, sh, x, tag, base, std::make_integer_sequence<int, kLeafW>{});
    if (sh.abort) return;

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
    h->barrier_base += static_cast<unsigned long long>(args.w) * args.G;
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
    char* ex = reinterpret_cast<char*>(ws.exchange);
    const size_t R = kMaxG + 1;
    args.xval = reinterpret_cast<T*>(ex);
    args.xrow = args.xval + 2 * R;
    args.xidx = reinterpret_cast<int*>(args.xrow + 2 * R * kLeafW);
    args.xtag = args.xidx + 2 * R;
    args.bar = h->barrier;
    args.bar_base = h->barrier_base;
    args.status = h->kernel_status;
    // Rows per thread: keep G (the number of workgroups that exchange per
    // column) at 64 or below where possible.
    int rpt = 1;
    while (rpt < 4 && ceil_div(m, 256 * rpt) > 64) rpt *= 2;
    args.G = ceil_div(m, 256 * rpt);
    if (args.G > kMaxG) {
        report_status(h, mifact_status_invalid_value, "getf2_leaf: panel too tall");
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
    for (int c = 0; c < nc; ++c) {
        const T* col = A + off(0, c0 + c, lda);
        for (int t = threadIdx.x; t < nm; t += 256) {
            const int s = src[t];
            vals[c * nm + t] = (s >= 0) ? col[s] : T(0);
        }
    }
    __syncthreads();
    for (int c = 0; c < nc; ++c) {
        T* col = A + off(0, c0 + c, lda);
        for (int t = threadIdx.x; t < nm; t += 256) {
            if (src[t] >= 0) col[dst[t]] = vals[c * nm + t];
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

// Triangular solves from the double-blocked factors, power-of-two
// equilibration, the FP64 residual and the mixed-precision refinement
// solvers (POSV / GESV with FP64 iterative refinement).
//
// The reference stops at the factors; its accuracy oracle is the CPU FP64
// residual (include/utils/benchmark/cpu_fp64_residual.h:28-201).  The solve
// path here follows LAPACK's mixed-precision dsgesv/dsposv scheme (factor in
// lower precision, residual and update in FP64, normwise backward-error
// stopping test ||r|| <= tol (||A|| ||x|| + ||b||)).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "blas3.hpp"
#include "drivers.hpp"
#include "solve/mixed_refine.h"

namespace mifact {
namespace {

constexpr int kTile = 64;          // residual tile (rows x columns)
constexpr int kRhsPass = 16;       // right-hand sides per residual pass
constexpr int kTargetBlocks = 2048;
constexpr int kPermLds = 36864;    // ipiv -> permutation in LDS up to this order

int round_up(int x, int m) { return (x + m - 1) / m * m; }

// ---------------------------------------------------------------- arena

// Solver scratch for order n and nrhs right-hand sides.
struct SolveArena {
    float* t = nullptr;          // ldt x nrhs FP32 correction / right-hand side
    int ldt = 0;
    float* inv_a = nullptr;      // inverses of the 64 x 64 diagonal blocks (L)
    float* inv_b = nullptr;      // inv(U)^T blocks (GETRF)
    float* rs = nullptr;         // row scale (powers of two)
    float* cs = nullptr;         // column scale
    int* perm = nullptr;         // factor row p came from row perm[p]
    int* pinv = nullptr;         // pinv[perm[p]] = p
    unsigned* rowmax = nullptr;  // |a| row maxima (float bits)
    double* part = nullptr;      // residual partials: chunks x kRhsPass x n
    unsigned long long* norms = nullptr;  // [rnorm | xnorm | bnorm] x nrhs, anorm
    int* info = nullptr;
    double* V = nullptr;         // FGMRES Krylov basis, ldt x (restart + 1)
    double* Z = nullptr;         // preconditioned directions, ldt x restart
    double* dpart = nullptr;     // dot-product partials
    double* dvals = nullptr;     // [h1 | h2 | ||w||^2 | y] on the device
    unsigned long long* ticket = nullptr;  // TRSV ticket counter (monotonic)
    unsigned long long* xtag = nullptr;    // TRSV solved blocks as (epoch, value) words, ldt x kRhsPass
};

int residual_chunks(int n)
{
    const int tiles = ceil_div(n, kTile);
    const int want = ceil_div(kTargetBlocks, imax(tiles, 1));
    return imax(1, imin(tiles, want));
}

size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

constexpr int kDotChunks = 64;
constexpr int kGmresRestart = 30;

// Pinned host staging: [norms (3 nrhs + 1) | info (8 bytes) | FGMRES values].
size_t host_gmres_offset(int nrhs)
{
    return sizeof(unsigned long long) * (3 * static_cast<size_t>(nrhs) + 2);
}

size_t arena_bytes(int n, int nrhs, SolveArena* out, char* base)
{
    const int ldt = round_up(imax(n, 1), 64);
    const size_t blocks = static_cast<size_t>(ceil_div(imax(n, 1), 64));
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at_ = o;
        o += align256(bytes);
        return base == nullptr ? nullptr : base + at_;
    };
    SolveArena a;
    a.ldt = ldt;
    a.t = reinterpret_cast<float*>(take(sizeof(float) * ldt * static_cast<size_t>(nrhs)));
    a.inv_a = reinterpret_cast<float*>(take(sizeof(float) * 4096 * blocks));
    a.inv_b = reinterpret_cast<float*>(take(sizeof(float) * 4096 * blocks));
    a.rs = reinterpret_cast<float*>(take(sizeof(float) * ldt));
    a.cs = reinterpret_cast<float*>(take(sizeof(float) * ldt));
    a.perm = reinterpret_cast<int*>(take(sizeof(int) * ldt));
    a.pinv = reinterpret_cast<int*>(take(sizeof(int) * ldt));
    a.rowmax = reinterpret_cast<unsigned*>(take(sizeof(unsigned) * ldt));
    // chunks(m) * round_up(m, 64) <= (kTargetBlocks + tiles) * 64 for every
    // order m <= n: the partials of any smaller call fit as well
    a.part = reinterpret_cast<double*>(
        take(sizeof(double) * kRhsPass * 64
             * (static_cast<size_t>(kTargetBlocks) + static_cast<size_t>(blocks) + 1)));
    a.norms = reinterpret_cast<unsigned long long*>(
        take(sizeof(unsigned long long) * (3 * static_cast<size_t>(nrhs) + 1)));
    a.info = reinterpret_cast<int*>(take(sizeof(int)));
    a.V = reinterpret_cast<double*>(take(sizeof(double) * ldt * (kGmresRestart + 1)));
    a.Z = reinterpret_cast<double*>(take(sizeof(double) * ldt * kGmresRestart));
    a.dpart = reinterpret_cast<double*>(take(sizeof(double) * kDotChunks * (kGmresRestart + 1)));
    a.dvals = reinterpret_cast<double*>(take(sizeof(double) * (4 * kGmresRestart + 8)));
    a.ticket = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long)));
    a.xtag = reinterpret_cast<unsigned long long*>(
        take(sizeof(unsigned long long) * ldt * static_cast<size_t>(kRhsPass)));
    if (out != nullptr) *out = a;
    return o;
}

// The arena's layout is a function of its capacity (largest order and
// right-hand-side count requested so far), so every offset -- in particular
// the TRSV tagged words and ticket counter -- stays put between calls until the
// arena is reallocated (and zeroed) for a larger capacity.
bool ensure_arena(mifact_handle h, int n, int nrhs, SolveArena* a)
{
    const int cap_n = imax(h->solve_cap_n, imax(n, 1));
    const int cap_r = imax(h->solve_cap_nrhs, imax(nrhs, 1));
    const size_t need = arena_bytes(cap_n, cap_r, nullptr, nullptr);
    if (h->solve_ws == nullptr || cap_n != h->solve_cap_n || cap_r != h->solve_cap_nrhs) {
        if (h->solve_ws != nullptr) {
            if (hipStreamSynchronize(h->stream) != hipSuccess) {
                report_status(h, mifact_status_hip_error, "stream sync before solver arena growth");
                return false;
            }
            (void)hipFree(h->solve_ws);
            h->solve_ws = nullptr;
            h->solve_ws_bytes = 0;
        }
        void* p = nullptr;
        if (hipMalloc(&p, need) != hipSuccess) {
            report_status(h, mifact_status_allocation_failed, "solver arena allocation failed");
            return false;
        }
        // zero once: TRSV tagged words (epochs start at 1) and the ticket counter
        if (hipMemset(p, 0, need) != hipSuccess) {
            (void)hipFree(p);
            report_status(h, mifact_status_hip_error, "solver arena initialization failed");
            return false;
        }
        h->solve_ws = p;
        h->solve_ws_bytes = need;
        h->solve_cap_n = cap_n;
        h->solve_cap_nrhs = cap_r;
        h->trsv_epoch = 0;
        h->trsv_ticket_base = 0;
    }
    const size_t host_need = host_gmres_offset(cap_r) + sizeof(double) * (4 * kGmresRestart + 8);
    if (h->host_norms == nullptr || h->host_norms_bytes < host_need) {
        if (h->host_norms != nullptr) {
            (void)hipHostFree(h->host_norms);
            h->host_norms = nullptr;
        }
        void* p = nullptr;
        if (hipHostMalloc(&p, host_need, hipHostMallocDefault) != hipSuccess) {
            report_status(h, mifact_status_allocation_failed, "pinned host buffer failed");
            return false;
        }
        h->host_norms = p;
        h->host_norms_bytes = host_need;
    }
    arena_bytes(cap_n, cap_r, a, static_cast<char*>(h->solve_ws));
    return true;
}

// ---------------------------------------------------------------- kernels

__device__ inline unsigned long long abs_bits(double v)
{
    return static_cast<unsigned long long>(__double_as_longlong(fabs(v)));
}

__device__ inline float pow2_scale_for(float m)
{
    // 2^-e with m = f 2^e, f in [0.5, 1): the scaled maximum lies in [0.5, 1).
    if (!(m > 0.0f) || !isfinite(m)) return 1.0f;
    int e;
    (void)frexpf(m, &e);
    return ldexpf(1.0f, -e);
}

template <typename T>
__global__ void fill_kernel(T* x, long long ld, int rows, int cols, T v)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long total = static_cast<long long>(rows) * cols;
    if (i >= total) return;
    const int r = static_cast<int>(i % rows);
    const int c = static_cast<int>(i / rows);
    x[off(r, c, ld)] = v;
}

template <typename T>
void fill(mifact_handle h, T* x, long long ld, int rows, int cols, T v)
{
    const long long total = static_cast<long long>(rows) * cols;
    if (total <= 0) return;
    hipLaunchKernelGGL(fill_kernel<T>, dim3(static_cast<unsigned>((total + 255) / 256)),
                       dim3(256), 0, h->stream, x, ld, rows, cols, v);
    MIFACT_LAUNCHED(h);
}

// s_i = 2^-floor(e_i / 2) for a_ii = f 2^e_i: s_i^2 a_ii in [0.5, 2).
__global__ void sym_scale_kernel(const float* A, long long lda, int n, float* s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = A[off(i, i, lda)];
    float v = 1.0f;
    if (d > 0.0f && isfinite(d)) {
        int e;
        (void)frexpf(d, &e);
        v = ldexpf(1.0f, -(e >> 1));
    }
    s[i] = v;
}

// Row maxima of |A| over a chunk of columns (thread per row), combined with
// an atomic max on the float bits (non-negative floats order as integers).
__global__ __launch_bounds__(256) void rowmax_kernel(const float* A, long long lda, int n,
                                                     int cols_per_chunk, unsigned* rowmax)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int c0 = blockIdx.y * cols_per_chunk;
    const int c1 = imin(n, c0 + cols_per_chunk);
    float m = 0.0f;
    for (int c = c0; c < c1; ++c) m = fmaxf(m, fabsf(A[off(i, c, lda)]));
    atomicMax(rowmax + i, __float_as_uint(m));
}

__global__ void row_scale_kernel(const unsigned* rowmax, int n, float* rs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rs[i] = pow2_scale_for(__uint_as_float(rowmax[i]));
}

// c_j from max_i |r_i a_ij| (one wave per column).
__global__ __launch_bounds__(256) void col_scale_kernel(const float* A, long long lda, int n,
                                                        const float* rs, float* cs)
{
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= n) return;
    float m = 0.0f;
    for (int i = lane; i < n; i += 64) m = fmaxf(m, fabsf(rs[i] * A[off(i, j, lda)]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if (lane == 0) cs[j] = pow2_scale_for(m);
}

// F = diag(rs) A diag(cs) (SYM: lower triangle only, cs = rs).
template <bool SYM>
__global__ __launch_bounds__(256) void scale_copy_kernel(const float* A, long long lda, float* F,
                                                         long long ldf, int n, const float* rs,
                                                         const float* cs)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float ri = rs[i];
    const int c0 = blockIdx.y * 8;
    const int c1 = imin(n, c0 + 8);
    for (int c = c0; c < c1; ++c) {
        if (SYM && i < c) continue;
        F[off(i, c, ldf)] = ri * A[off(i, c, lda)] * cs[c];
    }
}

// Partial products of the residual over 64 x 64 tiles.  Block (rb, chunk)
// covers rows [64 rb, 64 rb + 64) and the chunk's column tiles, writing
// part[chunk][k][row] (k: right-hand side within the pass).
//   SYM: A symmetric with only its lower triangle read (upper tiles are the
//        stored lower tiles transposed, diagonal tiles are mirrored);
//   ABS: row sums of |A| (X unused, one output).
template <int NR, bool SYM, bool ABS>
__global__ __launch_bounds__(256) void residual_partial_kernel(
    const float* A, long long lda, int n, const double* X, long long ldx, int nr,
    int tiles_per_chunk, double* part, long long ldp)
{
    __shared__ float tile[kTile][kTile + 1];   // tile[col][row]
    __shared__ double xs[kTile][NR];
    __shared__ double red[4][kTile][NR];
    const int t = threadIdx.x;
    const int r = t & 63;
    const int g = t >> 6;
    const int r0 = blockIdx.x * kTile;
    const int chunk = blockIdx.y;
    const int ntiles = ceil_div(n, kTile);
    double acc[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) acc[k] = 0.0;

    for (int q = 0; q < tiles_per_chunk; ++q) {
        const int ct = chunk * tiles_per_chunk + q;
        if (ct >= ntiles) break;
        const int c0 = ct * kTile;
        const bool mirror = SYM && c0 > r0;
        const int sr0 = mirror ? c0 : r0;
        const int sc0 = mirror ? r0 : c0;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int idx = t + 256 * e;
            const int lr = idx & 63;
            const int lc = idx >> 6;
            float v = 0.0f;
            if (sr0 + lr < n && sc0 + lc < n && (!SYM || sr0 + lr >= sc0 + lc)) {
                v = A[off(sr0 + lr, sc0 + lc, lda)];
            }
            tile[lc][lr] = v;
        }
        if (!ABS) {
            for (int idx = t; idx < kTile * NR; idx += 256) {
                const int j = idx / NR;
                const int k = idx % NR;
                xs[j][k] = (c0 + j < n && k < nr) ? X[off(c0 + j, k, ldx)] : 0.0;
            }
        }
        __syncthreads();
#pragma unroll 4
        for (int cc = 0; cc < 16; ++cc) {
            const int c = g * 16 + cc;
            float a;
            if (!SYM || r0 > c0) a = tile[c][r];
            else if (mirror) a = tile[r][c];
            else a = (r >= c) ? tile[c][r] : tile[r][c];
            if (ABS) {
                acc[0] += fabs(static_cast<double>(a));
            } else {
#pragma unroll
                for (int k = 0; k < NR; ++k) acc[k] = fma(static_cast<double>(a), xs[c][k], acc[k]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) red[g][r][k] = acc[k];
    __syncthreads();
    if (g == 0 && r0 + r < n) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (k >= nr) break;
            const double v = (red[0][r][k] + red[1][r][k]) + (red[2][r][k] + red[3][r][k]);
            part[(static_cast<size_t>(chunk) * kRhsPass + k) * ldp + r0 + r] = v;
        }
    }
}

// ||A||_inf = max_i sum over chunks of the |A| row partials.
__global__ void rowsum_max_kernel(const double* part, long long ldp, int chunks, int n,
                                  unsigned long long* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < chunks; ++c) s += part[static_cast<size_t>(c) * kRhsPass * ldp + i];
    atomicMax(out, abs_bits(s));
}

// r = b - sum of the partials (chunks = 0: r = b).  Writes r (FP64, when R
// is given), t[pinv[i]] = rs_i r_i in FP32 (when t is given) and the column
// maxima |r| (atomic max on the double bits).
__global__ void residual_reduce_kernel(const double* part, long long ldp, int chunks, int n,
                                       int nr, const double* B, long long ldb, double* R,
                                       long long ldr, float* t, long long ldt, const float* rs,
                                       const int* pinv, unsigned long long* rnorm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= n || k >= nr) return;
    double y = 0.0;
    for (int c = 0; c < chunks; ++c) y += part[(static_cast<size_t>(c) * kRhsPass + k) * ldp + i];
    const double r = B != nullptr ? B[off(i, k, ldb)] - y : y;   // B null: r = A x
    if (R != nullptr) R[off(i, k, ldr)] = r;
    if (t != nullptr) {
        const int p = pinv != nullptr ? pinv[i] : i;
        t[off(p, k, ldt)] = static_cast<float>(static_cast<double>(rs[i]) * r);
    }
    if (rnorm != nullptr) atomicMax(rnorm + k, abs_bits(r));
}

// x_i += cs_i t_i (FP64) and the column maxima |x|.
__global__ void update_kernel(double* X, long long ldx, int n, int nrhs, const float* t,
                              long long ldt, const float* cs, unsigned long long* xnorm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= n || k >= nrhs) return;
    double* x = X + off(i, k, ldx);
    const double v = *x + static_cast<double>(cs[i]) * static_cast<double>(t[off(i, k, ldt)]);
    *x = v;
    atomicMax(xnorm + k, abs_bits(v));
}

// pinv[perm[p]] = p
__global__ void invert_perm_kernel(const int* perm, int n, int* pinv)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) pinv[perm[p]] = p;
}

// LAPACK ipiv (1-based, applied in order) -> permutation perm (row p of the
// factors came from row perm[p]).  The exchanges are sequential by
// definition; one thread applies them to an LDS copy of the permutation.
__global__ __launch_bounds__(256) void ipiv_to_perm_lds_kernel(const int* ipiv, int n, int* perm)
{
    __shared__ int s[kPermLds];
    __shared__ int piv[1024];
    for (int i = threadIdx.x; i < n; i += 256) s[i] = i;
    for (int k0 = 0; k0 < n; k0 += 1024) {
        __syncthreads();
        for (int i = threadIdx.x; i < 1024 && k0 + i < n; i += 256) piv[i] = ipiv[k0 + i] - 1;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int k1 = imin(n, k0 + 1024);
            for (int k = k0; k < k1; ++k) {
                const int p = piv[k - k0];
                if (p != k && p >= 0 && p < n) {
                    const int a = s[k];
                    s[k] = s[p];
                    s[p] = a;
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256) perm[i] = s[i];
}

// Same for orders beyond the LDS copy: the permutation lives in perm itself.
__global__ void ipiv_to_perm_global_kernel(const int* ipiv, int n, int* perm)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
        const int p = ipiv[k] - 1;
        if (p != k && p >= 0 && p < n) {
            const int a = perm[k];
            perm[k] = perm[p];
            perm[p] = a;
        }
    }
}

// Y[p] = B[perm[p]] (gather, FP32) or B[i] = Y[i] (copy back).
__global__ void gather_rows_kernel(const float* B, long long ldb, int n, int nrhs, const int* perm,
                                   float* Y, long long ldy)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (p >= n || k >= nrhs) return;
    const int src = perm != nullptr ? perm[p] : p;
    Y[off(p, k, ldy)] = B[off(src, k, ldb)];
}

// F(i, i) += sigma
__global__ void add_diag_kernel(float* F, long long ldf, int n, float sigma)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) F[off(i, i, ldf)] += sigma;
}

// ---------------------------------------------------------------- FGMRES vectors (FP64)

// partial[c * kDotChunks + chunk] = sum over the chunk's rows of V(i, c) w(i)
__global__ __launch_bounds__(256) void dot_partial_kernel(const double* V, long long ldv,
                                                          const double* w, int n, double* partial)
{
    __shared__ double red[256];
    const int c = blockIdx.y;
    const int per = (n + kDotChunks - 1) / kDotChunks;
    const int i0 = blockIdx.x * per;
    const int i1 = imin(n, i0 + per);
    const double* v = V + static_cast<size_t>(c) * ldv;
    double s = 0.0;
    for (int i = i0 + threadIdx.x; i < i1; i += 256) s = fma(v[i], w[i], s);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[c * kDotChunks + blockIdx.x] = red[0];
}

// out[c] = sum of the chunk partials of column c (fixed order)
__global__ void dot_reduce_kernel(const double* partial, int k, double* out)
{
    const int c = threadIdx.x;
    if (c >= k) return;
    double s = 0.0;
    for (int j = 0; j < kDotChunks; ++j) s += partial[c * kDotChunks + j];
    out[c] = s;
}

// w -= V(:, 0:k) h   (h on the device)
__global__ void gemv_sub_kernel(const double* V, long long ldv, int k, const double* h, double* w,
                                int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < k; ++c) s = fma(V[static_cast<size_t>(c) * ldv + i], h[c], s);
    w[i] -= s;
}

// x += Z(:, 0:k) y   (y on the device)
__global__ void gemv_add_kernel(const double* Z, long long ldz, int k, const double* y, double* x,
                                int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < k; ++c) s = fma(Z[static_cast<size_t>(c) * ldz + i], y[c], s);
    x[i] += s;
}

__global__ void scale_f64_kernel(const double* w, int n, double alpha, double* v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = alpha * w[i];
}

// t[pinv[i]] = rs_i v_i in FP32 (preconditioner input)
__global__ void to_fp32_kernel(const double* v, int n, const float* rs, const int* pinv, float* t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = pinv != nullptr ? pinv[i] : i;
    t[p] = static_cast<float>(static_cast<double>(rs[i]) * v[i]);
}

// z_i = cs_i t_i in FP64 (preconditioner output)
__global__ void from_fp32_kernel(const float* t, int n, const float* cs, double* z)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = static_cast<double>(cs[i]) * static_cast<double>(t[i]);
}

__global__ void absmax_kernel(const double* x, int n, unsigned long long* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicMax(out, abs_bits(x[i]));
}

// ---------------------------------------------------------------- blocked TRSV

// Triangular solves with few right-hand sides in one launch.  Row block i
// (64 rows) is owned by the workgroup that draws ticket t (t = i forward,
// nblk - 1 - i backward); it accumulates M_ij x_j over its dependency blocks
// as they appear, then applies the 64 x 64 diagonal inverse and publishes
// x_i as (epoch, value) words in `xtag` (ld `ldx`): a reader polls the words
// it needs until every one carries the call's epoch, so no separate ready
// flag (and no completion wait before one) is needed.  Tickets are handed out in
// dispatch order, so every block waited on belongs to a workgroup that is
// already running: the chain always progresses.  Waits are bounded; on
// expiry the kernel raises the handle's status word and finishes.
enum TrsvMode { kFwdLower = 0, kBwdUpper = 1, kBwdLowerT = 2 };

constexpr int kTrsvSpin = 1 << 24;

template <int MODE, int NR>
__global__ __launch_bounds__(256) void trsv_kernel(const float* F, long long ldf, int n,
                                                   const float* inv, float* T, long long ldt,
                                                   int nr, unsigned long long* xtag, int ldx,
                                                   int epoch,
                                                   unsigned long long* ticket,
                                                   unsigned long long ticket_base, int* status,
                                                   int spin_limit)
{
    __shared__ float tile[64][65];
    __shared__ float dinv[64][65];   // the diagonal block's inverse, loaded up front
    __shared__ float xs[64][NR];
    __shared__ float part[4][64][NR];
    __shared__ int bsh;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int nblk = (n + 63) / 64;
    if (tid == 0) {
        const unsigned long long t = atomicAdd(ticket, 1ull) - ticket_base;
        bsh = MODE == kFwdLower ? static_cast<int>(t) : nblk - 1 - static_cast<int>(t);
    }
    __syncthreads();
    const int bi = bsh;
    if (bi < 0 || bi >= nblk) return;
    const int r0 = bi * 64;
    const int rb = imin(64, n - r0);
    float acc[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) acc[k] = 0.0f;
    // the inverse does not depend on other blocks: fetch it before waiting
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int idx = tid + 256 * e;
        const int lr = idx & 63, lc = idx >> 6;
        dinv[lc][lr] = inv[static_cast<size_t>(bi) * 4096 + lc * 64 + lr];
    }

    const int ndep = MODE == kFwdLower ? bi : nblk - 1 - bi;
    for (int q = 0; q < ndep; ++q) {
        const int bj = MODE == kFwdLower ? q : nblk - 1 - q;
        const int c0 = bj * 64;
        const int cb = imin(64, n - c0);
        // source block of M_ij (Lower-T reads L_ji and transposes)
        const int sr = MODE == kBwdLowerT ? c0 : r0;
        const int sc = MODE == kBwdLowerT ? r0 : c0;
        const int srows = MODE == kBwdLowerT ? cb : rb;
        const int scols = MODE == kBwdLowerT ? rb : cb;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int idx = tid + 256 * e;
            const int lr = idx & 63, lc = idx >> 6;
            tile[lc][lr] = (lr < srows && lc < scols) ? F[off(sr + lr, sc + lc, ldf)] : 0.0f;
        }
        // x_j: each word polled until it carries this call's epoch
        for (int e = tid; e < 64 * NR; e += 256) {
            const int c = e / NR, k = e % NR;
            float v = 0.0f;
            if (c < cb && k < nr) {
                unsigned long long* wp = xtag + static_cast<size_t>(k) * ldx + c0 + c;
                int spins = 0;
                unsigned long long word;
                while (static_cast<int>((word = __hip_atomic_load(wp, __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT)) >> 32)
                       != epoch) {
                    if (++spins > spin_limit) {
                        atomicExch(status, 3);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                v = __uint_as_float(static_cast<unsigned int>(word));
            }
            xs[c][k] = v;
        }
        __syncthreads();
#pragma unroll 4
        for (int cc = 0; cc < 16; ++cc) {
            const int c = w * 16 + cc;
            const float m = MODE == kBwdLowerT ? tile[lane][c] : tile[c][lane];
#pragma unroll
            for (int k = 0; k < NR; ++k) acc[k] = fmaf(m, xs[c][k], acc[k]);
        }
        __syncthreads();
    }
    // s = b_i - sum; x_i = D_i^{-1} s with the stored inverse block
#pragma unroll
    for (int k = 0; k < NR; ++k) part[w][lane][k] = acc[k];
    __syncthreads();
    for (int e = tid; e < 64 * NR; e += 256) {
        const int r = e / NR, k = e % NR;
        float v = 0.0f;
        if (r < rb && k < nr) {
            v = T[off(r0 + r, k, ldt)]
              - ((part[0][r][k] + part[1][r][k]) + (part[2][r][k] + part[3][r][k]));
        }
        xs[r][k] = v;
    }
    __syncthreads();
    // x(r) = sum_c Dinv(r, c) s(c); backward modes hold the transposed inverse
    float y[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) y[k] = 0.0f;
#pragma unroll 4
    for (int cc = 0; cc < 16; ++cc) {
        const int c = w * 16 + cc;
        const float d = (MODE == kFwdLower) ? dinv[c][lane] : dinv[lane][c];
#pragma unroll
        for (int k = 0; k < NR; ++k) y[k] = fmaf(d, xs[c][k], y[k]);
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) part[w][lane][k] = y[k];
    __syncthreads();
    for (int e = tid; e < 64 * NR; e += 256) {
        const int r = e / NR, k = e % NR;
        if (r < rb && k < nr) {
            const float x = (part[0][r][k] + part[1][r][k]) + (part[2][r][k] + part[3][r][k]);
            // the tagged word for the readers in this launch (device scope),
            // the plain value for the caller
            const unsigned long long word =
                (static_cast<unsigned long long>(static_cast<unsigned int>(epoch)) << 32)
                | __float_as_uint(x);
            __hip_atomic_store(xtag + static_cast<size_t>(k) * ldx + r0 + r, word, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            T[off(r0 + r, k, ldt)] = x;
        }
    }
}

// ---------------------------------------------------------------- host pieces

void ipiv_to_perm(mifact_handle h, const int* ipiv, int n, int* perm)
{
    if (n <= 0) return;
    if (n <= kPermLds) {
        hipLaunchKernelGGL(ipiv_to_perm_lds_kernel, dim3(1), dim3(256), 0, h->stream, ipiv, n,
                           perm);
    } else {
        hipLaunchKernelGGL(ipiv_to_perm_global_kernel, dim3(1), dim3(64), 0, h->stream, ipiv, n,
                           perm);
    }
    MIFACT_LAUNCHED(h);
}

void invert_perm(mifact_handle h, const int* perm, int n, int* pinv)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(invert_perm_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, h->stream, perm,
                       n, pinv);
    MIFACT_LAUNCHED(h);
}

void gather_rows(mifact_handle h, const float* B, long long ldb, int n, int nrhs, const int* perm,
                 float* Y, long long ldy)
{
    if (n <= 0 || nrhs <= 0) return;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(ceil_div(n, 256), nrhs), dim3(256), 0, h->stream,
                       B, ldb, n, nrhs, perm, Y, ldy);
    MIFACT_LAUNCHED(h);
}

// Partials of A X (or of |A| row sums when abs) for one pass of <= 16
// right-hand sides.
void residual_partials(mifact_handle h, bool sym, bool abs_rows, const float* A, long long lda,
                       int n, const double* X, long long ldx, int nr, double* part, long long ldp,
                       int chunks)
{
    const int tiles = ceil_div(n, kTile);
    const int per = ceil_div(tiles, chunks);
    const dim3 grid(tiles, chunks), block(256);
    StageScope scope(h, kStageSolve, abs_rows ? 1.0 * n * n : 2.0 * n * n * nr,
                     (sym ? 2.0 : 4.0) * n * n);
    auto go = [&](auto nr_tag, auto sym_tag, auto abs_tag) {
        constexpr int NR = decltype(nr_tag)::value;
        constexpr bool SY = decltype(sym_tag)::value;
        constexpr bool AB = decltype(abs_tag)::value;
        hipLaunchKernelGGL((residual_partial_kernel<NR, SY, AB>), grid, block, 0, h->stream, A, lda,
                           n, X, ldx, nr, per, part, ldp);
    };
    using F = std::false_type;
    using T = std::true_type;
    if (abs_rows) {
        if (sym) go(std::integral_constant<int, 1>{}, T{}, T{});
        else go(std::integral_constant<int, 1>{}, F{}, T{});
    } else if (nr == 1) {
        if (sym) go(std::integral_constant<int, 1>{}, T{}, F{});
        else go(std::integral_constant<int, 1>{}, F{}, F{});
    } else if (nr <= 4) {
        if (sym) go(std::integral_constant<int, 4>{}, T{}, F{});
        else go(std::integral_constant<int, 4>{}, F{}, F{});
    } else {
        if (sym) go(std::integral_constant<int, kRhsPass>{}, T{}, F{});
        else go(std::integral_constant<int, kRhsPass>{}, F{}, F{});
    }
    MIFACT_LAUNCHED(h);
}

// R/t = B - A X for all right-hand sides, in passes of 16.
void residual(mifact_handle h, bool sym, const float* A, long long lda, int n, int nrhs,
              const double* X, long long ldx, const double* B, long long ldb, double* R,
              long long ldr, float* t, long long ldt, const float* rs, const int* pinv,
              unsigned long long* rnorm, double* part, bool x_is_zero)
{
    const int chunks = residual_chunks(n);
    const long long ldp = round_up(n, 64);
    for (int k0 = 0; k0 < nrhs; k0 += kRhsPass) {
        const int nr = imin(kRhsPass, nrhs - k0);
        if (!x_is_zero) {
            residual_partials(h, sym, false, A, lda, n, X + off(0, k0, ldx), ldx, nr, part, ldp,
                              chunks);
        }
        hipLaunchKernelGGL(residual_reduce_kernel, dim3(ceil_div(n, 256), nr), dim3(256), 0,
                           h->stream, part, ldp, x_is_zero ? 0 : chunks, n, nr,
                           B != nullptr ? B + off(0, k0, ldb) : nullptr,
                           ldb, R != nullptr ? R + off(0, k0, ldr) : nullptr, ldr,
                           t != nullptr ? t + off(0, k0, ldt) : nullptr, ldt, rs, pinv,
                           rnorm != nullptr ? rnorm + k0 : nullptr);
        MIFACT_LAUNCHED(h);
    }
}

// ||A||_inf into *out (FP64 bits).
void norm_inf(mifact_handle h, bool sym, const float* A, long long lda, int n, double* part,
              unsigned long long* out)
{
    const int chunks = residual_chunks(n);
    const long long ldp = round_up(n, 64);
    residual_partials(h, sym, true, A, lda, n, nullptr, 0, 1, part, ldp, chunks);
    hipLaunchKernelGGL(rowsum_max_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, h->stream, part,
                       ldp, chunks, n, out);
    MIFACT_LAUNCHED(h);
}

// Power-of-two equilibration and the scaled copy F = R A C.
void equilibrate(mifact_handle h, bool sym, const float* A, long long lda, int n, float* F,
                 long long ldf, const SolveArena& w)
{
    const dim3 rows(ceil_div(n, 256));
    if (sym) {
        hipLaunchKernelGGL(sym_scale_kernel, rows, dim3(256), 0, h->stream, A, lda, n, w.rs);
        MIFACT_LAUNCHED(h);
        hipLaunchKernelGGL(scale_copy_kernel<true>, dim3(ceil_div(n, 256), ceil_div(n, 8)),
                           dim3(256), 0, h->stream, A, lda, F, ldf, n, w.rs, w.rs);
        MIFACT_LAUNCHED(h);
        return;
    }
    fill<unsigned>(h, w.rowmax, n, n, 1, 0u);
    const int cols_per_chunk = 256;
    hipLaunchKernelGGL(rowmax_kernel, dim3(ceil_div(n, 256), ceil_div(n, cols_per_chunk)),
                       dim3(256), 0, h->stream, A, lda, n, cols_per_chunk, w.rowmax);
    MIFACT_LAUNCHED(h);
    hipLaunchKernelGGL(row_scale_kernel, rows, dim3(256), 0, h->stream, w.rowmax, n, w.rs);
    MIFACT_LAUNCHED(h);
    hipLaunchKernelGGL(col_scale_kernel, dim3(ceil_div(n, 4)), dim3(256), 0, h->stream, A, lda, n,
                       w.rs, w.cs);
    MIFACT_LAUNCHED(h);
    hipLaunchKernelGGL(scale_copy_kernel<false>, dim3(ceil_div(n, 256), ceil_div(n, 8)),
                       dim3(256), 0, h->stream, A, lda, F, ldf, n, w.rs, w.cs);
    MIFACT_LAUNCHED(h);
}

// One blocked TRSV launch per pass of <= 16 right-hand sides.
void trsv(mifact_handle h, TrsvMode mode, int n, int nrhs, const float* F, long long ldf,
          float* T, long long ldt, const float* inv, const SolveArena& w)
{
    const int nblk = ceil_div(n, 64);
    for (int k0 = 0; k0 < nrhs; k0 += kRhsPass) {
        const int nr = imin(kRhsPass, nrhs - k0);
        h->trsv_epoch += 1;
        if (h->trsv_epoch <= 0) h->trsv_epoch = 1;
        const int epoch = h->trsv_epoch;
        const unsigned long long base = h->trsv_ticket_base;
        h->trsv_ticket_base += static_cast<unsigned long long>(nblk);
        float* Tk = T + off(0, k0, ldt);
        auto go = [&](auto mode_tag, auto nr_tag) {
            constexpr int M = decltype(mode_tag)::value;
            constexpr int R = decltype(nr_tag)::value;
            hipLaunchKernelGGL((trsv_kernel<M, R>), dim3(nblk), dim3(256), 0, h->stream, F, ldf,
                               n, inv, Tk, ldt, nr, w.xtag, w.ldt, epoch, w.ticket, base,
                               h->kernel_status,
                               h->wait_limit > 0 && h->wait_limit < kTrsvSpin
                                   ? static_cast<int>(h->wait_limit) : kTrsvSpin);
        };
        auto by_nr = [&](auto mode_tag) {
            if (nr == 1) go(mode_tag, std::integral_constant<int, 1>{});
            else if (nr <= 4) go(mode_tag, std::integral_constant<int, 4>{});
            else go(mode_tag, std::integral_constant<int, kRhsPass>{});
        };
        if (mode == kFwdLower) by_nr(std::integral_constant<int, kFwdLower>{});
        else if (mode == kBwdUpper) by_nr(std::integral_constant<int, kBwdUpper>{});
        else by_nr(std::integral_constant<int, kBwdLowerT>{});
        MIFACT_LAUNCHED(h);
    }
}

// FP32 solves with the factors and the inverses of their diagonal blocks:
// blocked single-launch TRSV per triangle.
void solve_chol(mifact_handle h, int n, int nrhs, const float* L, long long ldl, float* B,
                long long ldb, const float* inv, const SolveArena& w)
{
    trsv(h, kFwdLower, n, nrhs, L, ldl, B, ldb, inv, w);
    trsv(h, kBwdLowerT, n, nrhs, L, ldl, B, ldb, inv, w);
}

void solve_lu(mifact_handle h, int n, int nrhs, const float* LU, long long ldlu, float* B,
              long long ldb, const float* inv_l, const float* inv_ut, const SolveArena& w)
{
    trsv(h, kFwdLower, n, nrhs, LU, ldlu, B, ldb, inv_l, w);
    trsv(h, kBwdUpper, n, nrhs, LU, ldlu, B, ldb, inv_ut, w);
}

double bits_to_double(unsigned long long b)
{
    double d;
    std::memcpy(&d, &b, sizeof(d));
    return d;
}

bool valid_ld(int ld, int n) { return ld >= (n > 1 ? n : 1); }

bool launched(mifact_handle h, const char* what)
{
    if (hipGetLastError() != hipSuccess) {
        report_status(h, mifact_status_hip_error, what);
        return false;
    }
    return true;
}

bool to_host(mifact_handle h, void* dst, const void* src, size_t bytes)
{
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream) != hipSuccess
        || hipStreamSynchronize(h->stream) != hipSuccess) {
        report_status(h, mifact_status_hip_error, "refine: read-back");
        return false;
    }
    return h->last_error == mifact_status_success;
}

// out[0:k] = V(:, 0:k)^T w (deterministic two-stage reduction).
void dots(mifact_handle h, const double* V, long long ldv, int k, const double* w, int n,
          double* part, double* out)
{
    hipLaunchKernelGGL(dot_partial_kernel, dim3(kDotChunks, k), dim3(256), 0, h->stream, V, ldv, w,
                       n, part);
    hipLaunchKernelGGL(dot_reduce_kernel, dim3(1), dim3(64), 0, h->stream, part, k, out);
}

// Everything a refinement step needs to apply the factors.
struct Precond {
    bool sym;
    int n;
    const float* F;
    long long ldf;
    const SolveArena* w;
    const float* rs;
    const float* cs;
    const int* pinv;
};

// z = M^{-1} v: FP64 -> scaled, permuted FP32 -> two triangular solves -> FP64.
void apply_precond(mifact_handle h, const Precond& P, const double* v, double* z)
{
    const int n = P.n;
    const dim3 g(ceil_div(n, 256)), b(256);
    hipLaunchKernelGGL(to_fp32_kernel, g, b, 0, h->stream, v, n, P.rs, P.pinv, P.w->t);
    {
        StageScope scope(h, kStageSolve, 2.0 * n * n, 4.0 * n * n);
        if (P.sym) solve_chol(h, n, 1, P.F, P.ldf, P.w->t, P.w->ldt, P.w->inv_a, *P.w);
        else solve_lu(h, n, 1, P.F, P.ldf, P.w->t, P.w->ldt, P.w->inv_a, P.w->inv_b, *P.w);
    }
    hipLaunchKernelGGL(from_fp32_kernel, g, b, 0, h->stream, P.w->t, n, P.cs, z);
}

// FGMRES-IR (Carson & Higham 2018, with the flexible right-preconditioned
// GMRES of Saad 1993 so that an FP32-applied preconditioner is admissible):
// for each right-hand side, outer steps x += d where A d = r is solved by
// restarted FGMRES(30) in FP64 with classical Gram-Schmidt applied twice.
// Returns false on a HIP error; the backward error of each column ends in
// *berr_max.
bool gmres_refine(mifact_handle h, const Precond& P, const float* A, long long lda, int nrhs,
                  const double* B, long long ldb, double* X, long long ldx, double tol,
                  int max_outer, mifact_refine_result* res, double* berr_max)
{
    const int n = P.n;
    const SolveArena& w = *P.w;
    const int m = imin(kGmresRestart, n);
    const long long ldv = w.ldt;
    const double inner_tol = 1e-6;
    double* hv = reinterpret_cast<double*>(static_cast<char*>(h->host_norms)
                                           + host_gmres_offset(nrhs));
    unsigned long long* scratch_bits = w.norms;   // [rinf, xinf, binf, ainf] per call
    const unsigned long long* hb = static_cast<const unsigned long long*>(h->host_norms);
    double* d_h1 = w.dvals;
    double* d_h2 = w.dvals + (kGmresRestart + 1);
    double* d_nrm = w.dvals + 2 * (kGmresRestart + 1);
    double* d_y = w.dvals + 2 * (kGmresRestart + 1) + 2;
    const dim3 g1(ceil_div(n, 256)), b1(256);
    std::vector<double> H(static_cast<size_t>(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m);
    *berr_max = 0.0;
    bool all = true;

    // ||A||_inf (bits at scratch_bits[3]) is computed once per column pass.
    for (int k = 0; k < nrhs; ++k) {
        double* xk = X + static_cast<size_t>(k) * ldx;
        const double* bk = B + static_cast<size_t>(k) * ldb;
        double berr = INFINITY, prev_outer = INFINITY;
        for (int outer = 0;; ++outer) {
            // r = b - A x into V(:, 0); |r|, |x|, |b|, ||A|| maxima; ||r||_2^2
            fill<unsigned long long>(h, scratch_bits, 3, 3, 1, 0ull);
            residual(h, P.sym, A, lda, n, 1, xk, ldx, bk, ldb, w.V, ldv, nullptr, 0, nullptr,
                     nullptr, scratch_bits, w.part, false);
            hipLaunchKernelGGL(absmax_kernel, g1, b1, 0, h->stream, xk, n, scratch_bits + 1);
            hipLaunchKernelGGL(absmax_kernel, g1, b1, 0, h->stream, bk, n, scratch_bits + 2);
            dots(h, w.V, ldv, 1, w.V, n, w.dpart, d_nrm);
            if (!launched(h, "fgmres residual")) return false;
            if (!to_host(h, hv, d_nrm, sizeof(double))) return false;
            if (!to_host(h, h->host_norms, scratch_bits, 4 * sizeof(unsigned long long)))
                return false;
            const double rn = bits_to_double(hb[0]);
            const double xn = bits_to_double(hb[1]);
            const double bn = bits_to_double(hb[2]);
            const double an = bits_to_double(hb[3]);
            const double den = an * xn + bn;
            berr = den > 0.0 ? rn / den : (rn > 0.0 ? INFINITY : 0.0);
            // stop on convergence, at the step limit, or when a full FGMRES
            // cycle reduced the backward error by less than 10% (stagnation)
            if (berr <= tol || outer >= max_outer || !std::isfinite(berr)
                || berr > 0.9 * prev_outer)
                break;
            prev_outer = berr;
            res->iterations += 1;
            const double beta = std::sqrt(hv[0]);
            if (!(beta > 0.0)) break;
            hipLaunchKernelGGL(scale_f64_kernel, g1, b1, 0, h->stream, w.V, n, 1.0 / beta, w.V);
            std::fill(gv.begin(), gv.end(), 0.0);
            gv[0] = beta;
            int used = 0;
            for (int j = 0; j < m; ++j) {
                double* vj = w.V + static_cast<size_t>(j) * ldv;
                double* vn = w.V + static_cast<size_t>(j + 1) * ldv;
                double* zj = w.Z + static_cast<size_t>(j) * ldv;
                apply_precond(h, P, vj, zj);
                // vn = A z_j (FP64)
                residual(h, P.sym, A, lda, n, 1, zj, ldv, nullptr, 0, vn, ldv, nullptr, 0, nullptr,
                         nullptr, nullptr, w.part, false);
                // two passes of classical Gram-Schmidt against V(:, 0:j+1)
                dots(h, w.V, ldv, j + 1, vn, n, w.dpart, d_h1);
                hipLaunchKernelGGL(gemv_sub_kernel, g1, b1, 0, h->stream, w.V, ldv, j + 1, d_h1, vn,
                                   n);
                dots(h, w.V, ldv, j + 1, vn, n, w.dpart, d_h2);
                hipLaunchKernelGGL(gemv_sub_kernel, g1, b1, 0, h->stream, w.V, ldv, j + 1, d_h2, vn,
                                   n);
                dots(h, vn, ldv, 1, vn, n, w.dpart, d_nrm);
                if (!launched(h, "fgmres step")) return false;
                if (!to_host(h, hv, w.dvals, sizeof(double) * (2 * (kGmresRestart + 1) + 1)))
                    return false;
                res->gmres_iterations += 1;
                double* col = &H[static_cast<size_t>(j) * (m + 1)];
                for (int i = 0; i <= j; ++i) col[i] = hv[i] + hv[kGmresRestart + 1 + i];
                const double hn = std::sqrt(std::max(hv[2 * (kGmresRestart + 1)], 0.0));
                // previous rotations, then the one that annihilates hn
                for (int i = 0; i < j; ++i) {
                    const double a = cs[i] * col[i] + sn[i] * col[i + 1];
                    col[i + 1] = -sn[i] * col[i] + cs[i] * col[i + 1];
                    col[i] = a;
                }
                const double r = std::hypot(col[j], hn);
                cs[j] = r > 0.0 ? col[j] / r : 1.0;
                sn[j] = r > 0.0 ? hn / r : 0.0;
                col[j] = r;
                gv[j + 1] = -sn[j] * gv[j];
                gv[j] = cs[j] * gv[j];
                used = j + 1;
                if (!(hn > 0.0) || std::fabs(gv[j + 1]) <= inner_tol * beta) break;
                hipLaunchKernelGGL(scale_f64_kernel, g1, b1, 0, h->stream, vn, n, 1.0 / hn, vn);
            }
            // y = R^{-1} g (upper triangular), x += Z y
            for (int i = used - 1; i >= 0; --i) {
                double acc = gv[i];
                for (int c = i + 1; c < used; ++c) acc -= H[static_cast<size_t>(c) * (m + 1) + i] * y[c];
                const double d = H[static_cast<size_t>(i) * (m + 1) + i];
                y[i] = d != 0.0 ? acc / d : 0.0;
            }
            if (hipMemcpyAsync(d_y, y.data(), sizeof(double) * used, hipMemcpyHostToDevice,
                               h->stream) != hipSuccess) {
                report_status(h, mifact_status_hip_error, "fgmres: coefficients upload");
                return false;
            }
            hipLaunchKernelGGL(gemv_add_kernel, g1, b1, 0, h->stream, w.Z, ldv, used, d_y, xk, n);
            if (!launched(h, "fgmres update")) return false;
        }
        if (!(berr <= tol)) all = false;
        *berr_max = (berr > *berr_max || std::isnan(berr)) ? berr : *berr_max;
    }
    // hv/y are host staging: the upload above completes before the next
    // read-back synchronizes, and nothing is left in flight on return.
    if (hipStreamSynchronize(h->stream) != hipSuccess) return false;
    res->converged = all ? 1 : 0;
    return true;
}

// The refinement solver shared by POSV (sym) and GESV.
mifact_status refine_solve(mifact_handle h, bool sym, int n, int nrhs, const float* A, int lda_i,
                           float* F, int ldf_i, int* ipiv, const double* B, int ldb_i, double* X,
                           int ldx_i, const mifact_refine_params* prm, mifact_refine_result* res)
{
    const char* name = sym ? "mifact_posv_refine" : "mifact_gesv_refine";
    if (h == nullptr) {
        report_status(nullptr, mifact_status_invalid_handle, name);
        return mifact_status_invalid_handle;
    }
    if (res == nullptr || prm == nullptr || n < 0 || nrhs < 0 || !valid_ld(lda_i, n)
        || !valid_ld(ldf_i, n) || !valid_ld(ldb_i, n) || !valid_ld(ldx_i, n)
        || (n > 0 && nrhs > 0
            && (A == nullptr || F == nullptr || B == nullptr || X == nullptr
                || (!sym && ipiv == nullptr)))) {
        report_status(h, mifact_status_invalid_value, name);
        return mifact_status_invalid_value;
    }
    std::memset(res, 0, sizeof(*res));
    if (n == 0 || nrhs == 0) {
        res->converged = 1;
        return mifact_status_success;
    }
    const long long lda = lda_i, ldf = ldf_i, ldb = ldb_i, ldx = ldx_i;
    const LowPrecision prec = static_cast<LowPrecision>(prm->precision);
    if (prec != kPrecFP32 && prec != kPrecBF16 && prec != kPrecFP16 && prec != kPrecFP8) {
        report_status(h, mifact_status_invalid_value, "refine: unknown precision");
        return mifact_status_invalid_value;
    }
    const int max_it = prm->max_iterations < 0 ? 30 : prm->max_iterations;
    const mifact_refine_method method = prm->method;
    if (method != mifact_refine_classical && method != mifact_refine_gmres
        && method != mifact_refine_auto) {
        report_status(h, mifact_status_invalid_value, "refine: unknown method");
        return mifact_status_invalid_value;
    }
    const double tol = prm->tolerance > 0.0 ? prm->tolerance : std::sqrt(static_cast<double>(n)) * DBL_EPSILON;
    const mifact_status before = h->last_error;
    h->last_error = mifact_status_success;
    auto failed = [&]() { return h->last_error != mifact_status_success; };
    auto finish = [&](mifact_status s) {
        if (s == mifact_status_success && failed()) s = h->last_error;
        if (before != mifact_status_success) h->last_error = before;
        return s;
    };

    if (!ensure_workspace(h, n)) return finish(mifact_status_allocation_failed);
    SolveArena w;
    if (!ensure_arena(h, n, nrhs, &w)) return finish(mifact_status_allocation_failed);
    unsigned long long* rnorm = w.norms;
    unsigned long long* xnorm = w.norms + nrhs;
    unsigned long long* bnorm = w.norms + 2 * nrhs;
    unsigned long long* anorm = w.norms + 3 * nrhs;
    const long long ldp = round_up(n, 64);

    // ||A||_inf and the power-of-two scales (once)
    fill<unsigned long long>(h, w.norms, 3 * nrhs + 1, 3 * nrhs + 1, 1, 0ull);
    norm_inf(h, sym, A, lda, n, w.part, anorm);
    const float* cs = sym ? w.rs : w.cs;
    const int* pinv = sym ? nullptr : w.pinv;
    int* hinfo = reinterpret_cast<int*>(static_cast<unsigned long long*>(h->host_norms)
                                        + 3 * nrhs + 1);
    const unsigned long long* hn = static_cast<const unsigned long long*>(h->host_norms);

    // One factorization in `p` followed by refinement; false on a HIP error.
    auto attempt = [&](LowPrecision p, bool first, float shift) -> bool {
        equilibrate(h, sym, A, lda, n, F, ldf, w);
        if (shift > 0.0f) {
            hipLaunchKernelGGL(add_diag_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, h->stream, F,
                               ldf, n, shift);
            if (!launched(h, "refine: diagonal shift")) return false;
        }
        if (sym) {
            double_blocked_potrf_impl<float>(h, F, ldf_i, n, prm->nb, prm->b, w.info, false, p);
        } else {
            double_blocked_getrf_impl<float>(h, F, ldf_i, n, prm->nb, prm->b,
                                             prm->outer_trsm_mode, prm->trsm_block, ipiv, w.info,
                                             false, nullptr, 0, p);
            if (hipMemcpyAsync(w.perm, workspace(h).orig, sizeof(int) * n,
                               hipMemcpyDeviceToDevice, h->stream) != hipSuccess) {
                report_status(h, mifact_status_hip_error, "refine: permutation copy");
                return false;
            }
            invert_perm(h, w.perm, n, w.pinv);
        }
        res->factorizations += 1;
        if (hipMemcpyAsync(hinfo, w.info, sizeof(int), hipMemcpyDeviceToHost, h->stream)
                != hipSuccess
            || hipStreamSynchronize(h->stream) != hipSuccess) {
            report_status(h, mifact_status_hip_error, "refine: info read-back");
            return false;
        }
        if (failed()) return false;
        res->info = *hinfo;
        res->converged = 0;
        if (res->info != 0) return true;
        if (sym) {
            trtri_blocks<float>(h, n, F, ldf, w.inv_a, false, false);
        } else {
            trtri_blocks<float>(h, n, F, ldf, w.inv_a, true, false);
            trtri_blocks<float>(h, n, F, ldf, w.inv_b, false, true);
        }
        // x = 0; t = R b (permuted); ||b||
        fill<double>(h, X, ldx, n, nrhs, 0.0);
        fill<unsigned long long>(h, bnorm, nrhs, nrhs, 1, 0ull);
        residual(h, sym, A, lda, n, nrhs, X, ldx, B, ldb, nullptr, 0, w.t, w.ldt, w.rs, pinv,
                 bnorm, w.part, true);
        double prev = INFINITY;
        for (int it = 0;; ++it) {
            {
                StageScope scope(h, kStageSolve, 2.0 * n * n * nrhs, 4.0 * n * n);
                if (sym) solve_chol(h, n, nrhs, F, ldf, w.t, w.ldt, w.inv_a, w);
                else solve_lu(h, n, nrhs, F, ldf, w.t, w.ldt, w.inv_a, w.inv_b, w);
            }
            fill<unsigned long long>(h, w.norms, 2 * nrhs, 2 * nrhs, 1, 0ull);
            hipLaunchKernelGGL(update_kernel, dim3(ceil_div(n, 256), nrhs), dim3(256), 0,
                               h->stream, X, ldx, n, nrhs, w.t, w.ldt, cs, xnorm);
            if (hipGetLastError() != hipSuccess) {
                report_status(h, mifact_status_hip_error, "refine: update launch");
                return false;
            }
            residual(h, sym, A, lda, n, nrhs, X, ldx, B, ldb, nullptr, 0, w.t, w.ldt, w.rs, pinv,
                     rnorm, w.part, false);
            if (hipMemcpyAsync(h->host_norms, w.norms,
                               sizeof(unsigned long long) * (3 * nrhs + 1),
                               hipMemcpyDeviceToHost, h->stream) != hipSuccess
                || hipStreamSynchronize(h->stream) != hipSuccess) {
                report_status(h, mifact_status_hip_error, "refine: norm read-back");
                return false;
            }
            if (failed()) return false;
            const double an = bits_to_double(hn[3 * nrhs]);
            double berr = 0.0;
            for (int k = 0; k < nrhs; ++k) {
                const double rn = bits_to_double(hn[k]);
                const double xn = bits_to_double(hn[nrhs + k]);
                const double bn = bits_to_double(hn[2 * nrhs + k]);
                const double den = an * xn + bn;
                const double e = den > 0.0 ? rn / den : (rn > 0.0 ? INFINITY : 0.0);
                berr = (e > berr || std::isnan(e)) ? e : berr;
            }
            if (first && it == 0) res->first_backward_error = berr;
            res->backward_error = berr;
            if (berr <= tol) {
                res->converged = 1;
                res->iterations += it;
                return true;
            }
            // stop at the step limit, on divergence or on stagnation; GMRES-IR
            // takes over after the first solve (or, in auto mode, here)
            if (it >= max_it || !(berr < prev) || (it >= 2 && berr > 0.9 * prev)
                || (method == mifact_refine_gmres && first)) {
                res->iterations += it;
                return true;
            }
            prev = berr;
        }
    };

    if (!attempt(prec, true, 0.0f)) return finish(h->last_error);
    if (sym && res->info > 0 && prec != kPrecFP32) {
        // A BF16/FP16/FP8 Cholesky of an ill-conditioned matrix can lose
        // definiteness; factor the equilibrated matrix plus c u_low I instead
        // (its diagonal lies in [0.5, 2)) and let the refinement remove the
        // shift (Higham & Pranesh, SIAM J. Sci. Comput. 43(1), 2021).
        const float shift = prec == kPrecBF16 ? 0x1p-6f : (prec == kPrecFP8 ? 0x1p-2f : 0x1p-9f);
        res->diagonal_shift = shift;
        if (!attempt(prec, true, shift)) return finish(h->last_error);
    }
    // GMRES-IR from the current iterate with the current factors as the
    // preconditioner, when the classical steps stopped short
    auto krylov = [&]() -> bool {
        if (res->info != 0 || res->converged || method == mifact_refine_classical) return true;
        const Precond pc{sym, n, F, ldf, &w, w.rs, cs, pinv};
        double berr = 0.0;
        // ||A||_inf for the FGMRES backward errors (bits in norms[3])
        fill<unsigned long long>(h, w.norms + 3, 1, 1, 1, 0ull);
        norm_inf(h, sym, A, lda, n, w.part, w.norms + 3);
        if (!gmres_refine(h, pc, A, lda, nrhs, B, ldb, X, ldx, tol, max_it, res, &berr))
            return false;
        res->backward_error = berr;
        return true;
    };
    if (!krylov()) return finish(h->last_error);
    if (!res->converged && prec != kPrecFP32 && prm->fallback_fp32 != 0) {
        // FP32 factors; at large n and kappa (n u_32 kappa > 1) their
        // classical steps can stall too, and FGMRES finishes
        res->fell_back = 1;
        res->diagonal_shift = 0.0;
        if (!attempt(kPrecFP32, false, 0.0f)) return finish(h->last_error);
        if (!krylov()) return finish(h->last_error);
    }
    return finish(mifact_status_success);
}

}  // namespace
}  // namespace mifact

using namespace mifact;

extern "C" {

void mifact_gemm_update_lowp(mifact_handle handle, mifact_precision precision, int trans_b,
                             int lower, int m, int n, int k, float alpha, const float* A,
                             int lda, const float* B, int ldb, float* C, int ldc)
{
    if (handle == nullptr) {
        report_status(nullptr, mifact_status_invalid_handle, "mifact_gemm_update_lowp");
        return;
    }
    if (m <= 0 || n <= 0 || k <= 0) return;
    gemm_lowp(handle, static_cast<LowPrecision>(precision), trans_b ? kOpNT : kOpNN,
              lower ? kShapeLower : kShapeFull, m, n, k, alpha, A, lda, B, ldb, C, ldc);
}

void double_blocked_potrf_mixed(mifact_handle handle, mifact_precision precision, float* A,
                                int lda, int n, int nb, int b, int* info)
{
    double_blocked_potrf_impl<float>(handle, A, lda, n, nb, b, info, false,
                                     static_cast<LowPrecision>(precision));
}

void double_blocked_getrf_mixed(mifact_handle handle, mifact_precision precision, float* A,
                                int lda, int n, int outer_nb, int inner_nb,
                                getrf_outer_trsm_mode outer_trsm_mode, int trsm_block, int* ipiv,
                                int* info)
{
    if (handle == nullptr) {
        report_status(nullptr, mifact_status_invalid_handle, "double_blocked_getrf_mixed");
        return;
    }
    double_blocked_getrf_impl<float>(handle, A, lda, n, outer_nb, inner_nb, outer_trsm_mode,
                                     trsm_block, ipiv, info, false, nullptr, 0,
                                     static_cast<LowPrecision>(precision));
}

void mifact_potrs(mifact_handle handle, int n, int nrhs, const float* L, int ldl, float* B,
                  int ldb)
{
    if (handle == nullptr) {
        report_status(nullptr, mifact_status_invalid_handle, "mifact_potrs");
        return;
    }
    if (n <= 0 || nrhs <= 0) return;
    SolveArena w;
    if (!ensure_arena(handle, n, 1, &w)) return;
    trtri_blocks<float>(handle, n, L, ldl, w.inv_a, false, false);
    StageScope scope(handle, kStageSolve, 2.0 * n * n * nrhs, 4.0 * n * n);
    solve_chol(handle, n, nrhs, L, ldl, B, ldb, w.inv_a, w);
}

void mifact_getrs(mifact_handle handle, int n, int nrhs, const float* LU, int ldlu,
                  const int* ipiv, float* B, int ldb)
{
    if (handle == nullptr) {
        report_status(nullptr, mifact_status_invalid_handle, "mifact_getrs");
        return;
    }
    if (n <= 0 || nrhs <= 0) return;
    SolveArena w;
    if (!ensure_arena(handle, n, nrhs, &w)) return;
    trtri_blocks<float>(handle, n, LU, ldlu, w.inv_a, true, false);
    trtri_blocks<float>(handle, n, LU, ldlu, w.inv_b, false, true);
    ipiv_to_perm(handle, ipiv, n, w.perm);
    gather_rows(handle, B, ldb, n, nrhs, w.perm, w.t, w.ldt);
    {
        StageScope scope(handle, kStageSolve, 2.0 * n * n * nrhs, 4.0 * n * n);
        solve_lu(handle, n, nrhs, LU, ldlu, w.t, w.ldt, w.inv_a, w.inv_b, w);
    }
    gather_rows(handle, w.t, w.ldt, n, nrhs, nullptr, B, ldb);
}

void mifact_residual_fp64(mifact_handle handle, int symmetric_lower, int n, int nrhs,
                          const float* A, int lda, const double* X, int ldx, const double* B,
                          int ldb, double* R, int ldr)
{
    if (handle == nullptr) {
        report_status(nullptr, mifact_status_invalid_handle, "mifact_residual_fp64");
        return;
    }
    if (n <= 0 || nrhs <= 0) return;
    SolveArena w;
    if (!ensure_arena(handle, n, 1, &w)) return;
    residual(handle, symmetric_lower != 0, A, lda, n, nrhs, X, ldx, B, ldb, R, ldr, nullptr, 0,
             nullptr, nullptr, nullptr, w.part, false);
}

mifact_status mifact_posv_refine(mifact_handle handle, int n, int nrhs, const float* A, int lda,
                                 float* factor, int ldf, const double* B, int ldb, double* X,
                                 int ldx, const mifact_refine_params* params,
                                 mifact_refine_result* result)
{
    return refine_solve(handle, true, n, nrhs, A, lda, factor, ldf, nullptr, B, ldb, X, ldx,
                        params, result);
}

mifact_status mifact_gesv_refine(mifact_handle handle, int n, int nrhs, const float* A, int lda,
                                 float* factor, int ldf, int* ipiv, const double* B, int ldb,
                                 double* X, int ldx, const mifact_refine_params* params,
                                 mifact_refine_result* result)
{
    return refine_solve(handle, false, n, nrhs, A, lda, factor, ldf, ipiv, B, ldb, X, ldx,
                        params, result);
}

}  // extern "C"

// Double-blocked POTRF driver and the BLAS-3 C-ABI entry points.
//
// Schedule restated from the reference (src/potrf/double_blocked_potrf.cpp:
// double_blocked_potrf_impl :294-381, update_next_inner_panel :210-249,
// prepare_next_outer_panel :251-292); the rocBLAS/rocSOLVER calls are
// replaced by the gfx950 kernels through the recursive building blocks.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <type_traits>

#include "blas3.hpp"
#include "drivers.hpp"
#include "gemm/update_gemm.h"
#include "kernels/panel_kernels.h"
#include "potrf/double_blocked_potrf.h"
#include "syrk/kblocked_syrk.h"
#include "trsm/double_blocked_trsm.h"
#include "trsm/left_lower_trsm.h"
#include "trsm/right_lower_trsm.h"

namespace mifact {
namespace {

constexpr int kDbTrsmMicroBlock = 1024;  // double_blocked_potrf.cpp:17

// src/trsm/double_blocked_trsm.cpp:99-171: n rows of B, m columns of L.
template <typename T>
void double_blocked_trsm_impl(mifact_handle h, const T* L, long long ldl, T* B,
                              long long ldb, int n, int m, int nb, int b)
{
    if (n <= 0 || m <= 0) return;
    if (nb <= 0) nb = m;
    if (b <= 0) b = nb;
    b = imin(b, nb);
    for (int outer = 0; outer < m; outer += nb) {
        const int outer_size = imin(nb, m - outer);
        const int outer_end = outer + outer_size;
        for (int inner = outer; inner < outer_end; inner += b) {
            const int inner_size = imin(b, outer_end - inner);
            const int panel_start = inner + inner_size;
            trsm_rlt<T>(h, n, inner_size, at(L, ldl, inner, inner), ldl, at(B, ldb, 0, inner),
                        ldb);
            if (panel_start < outer_end) {
                const int pc = imin(b, outer_end - panel_start);
                gemm<T>(h, kOpNT, kShapeFull, n, pc, panel_start - outer, T(-1),
                        at(B, ldb, 0, outer), ldb, at(L, ldl, panel_start, outer), ldl,
                        at(B, ldb, 0, panel_start), ldb);
            }
        }
        const int trailing = m - outer_end;
        if (trailing > 0) {
            gemm<T>(h, kOpNT, kShapeFull, n, trailing, outer_size, T(-1), at(B, ldb, 0, outer),
                    ldb, at(L, ldl, outer_end, outer), ldl, at(B, ldb, 0, outer_end), ldb);
        }
    }
}

// Trailing update C -= A B^T: FP32/FP64 MFMA, or BF16/FP16 MFMA with FP32
// accumulation for the mixed-precision factorization (FP32 data only).
template <typename T>
void update_nt(mifact_handle h, LowPrecision prec, GemmShape shape, int M, int N, int K, const T* A,
               long long lda, const T* B, long long ldb, T* C, long long ldc)
{
    if constexpr (std::is_same_v<T, float>) {
        gemm_lowp(h, prec, kOpNT, shape, M, N, K, -1.0f, A, lda, B, ldb, C, ldc);
    } else {
        gemm<T>(h, kOpNT, shape, M, N, K, T(-1), A, lda, B, ldb, C, ldc);
    }
}

}  // namespace

template <typename T>
void double_blocked_potrf_impl(mifact_handle h, T* A, int lda_i, int n, int nb, int b,
                               int* info, bool use_dbtrsm, LowPrecision prec, T* inv_scratch)
{
    if (h == nullptr) {
        report_status(nullptr, mifact_status_invalid_handle, "double_blocked_potrf: null handle");
        return;
    }
    if (n <= 0) return;
    if (nb <= 0) nb = n;
    if (b <= 0) b = nb;
    b = imin(b, nb);
    const long long lda = lda_i;
    // The diagonal blocks' inverses (the only scratch of this schedule):
    // (b + 64) x 64 elements, from the workspace or from the caller.
    T* inv = inv_scratch;
    if (inv == nullptr) {
        if (!ensure_workspace(h, n)) return;
        inv = invdiag<T>(h);
    }
    if (info != nullptr) clear_int(h, info);

    // Outer updates are applied right-looking, block by block, as soon as an
    // outer block is factored: block k + 1 -= L_k L_k^T on the main stream
    // (its panel work needs it next), blocks k + 2 .. -= L_k L_k^T on the side
    // stream, where they overlap block k + 1's panel work.  Every block
    // receives the contributions of the blocks before it in block order, so
    // the result does not depend on the overlap.
    bool side_pending = false;   // a side update (it wrote block k + 1) may be running
    for (int outer = 0; outer < n; outer += nb) {
        const int outer_size = imin(nb, n - outer);
        const int outer_end = outer + outer_size;
        for (int inner = outer; inner < outer_end; inner += b) {
            const int inner_size = imin(b, outer_end - inner);
            const int panel_start = inner + inner_size;
            const int trailing_rows = n - panel_start;
            T* Aii = at(A, lda, inner, inner);
            // diagonal block (rocsolver_potrf at :334); its 64-wide leaves
            // also leave the inverses of their blocks in `inv`
            potrf_diag<T>(h, inner_size, Aii, lda, info, inner, inv);
            // panel solve (:336-354)
            if (trailing_rows > 0) {
                if (use_dbtrsm) {
                    double_blocked_trsm_impl<T>(h, Aii, lda, at(A, lda, panel_start, inner), lda,
                                                trailing_rows, inner_size, inner_size,
                                                imin(kDbTrsmMicroBlock, inner_size));
                } else {
                    trsm_rlt_inv<T>(h, trailing_rows, inner_size, Aii, lda,
                                    at(A, lda, panel_start, inner), lda, inv);
                }
            }
            // next inner panel -= L * L^T over the outer block's columns (:210-249);
            // only its lower trapezoid is needed.
            if (panel_start < outer_end) {
                const int panel_cols = imin(b, outer_end - panel_start);
                const int k = panel_start - outer;
                update_nt<T>(h, prec, kShapeLower, n - panel_start, panel_cols, k,
                             at(A, lda, panel_start, outer), lda, at(A, lda, panel_start, outer),
                             lda, at(A, lda, panel_start, panel_start), lda);
            }
        }
        // next outer block -= L * L^T over all previous columns (:251-292),
        // here as the contribution of this block (SYRK of the diagonal block +
        // GEMM below it, fused in one lower-trapezoid launch); the earlier
        // blocks' contributions came through the side stream.
        const int next_outer = outer_end;
        if (next_outer < n) {
            const int next_size = imin(nb, n - next_outer);
            const int after = next_outer + next_size;   // first column after the next block
            if (side_pending) MIFACT_HIP(h, hipStreamWaitEvent(h->stream, h->side_join, 0));
            side_pending = false;
            update_nt<T>(h, prec, kShapeLower, n - next_outer, next_size, outer_size,
                         at(A, lda, next_outer, outer), lda, at(A, lda, next_outer, outer), lda,
                         at(A, lda, next_outer, next_outer), lda);
            // forked after the launch above, so that the side update runs
            // beside block k + 1's panel work rather than beside that update
            if (after < n) {
                if (h->side_stream != nullptr)
                    MIFACT_HIP(h, hipEventRecord(h->side_fork, h->stream));
                OnSideStream side(h);
                if (h->side_join != nullptr)
                    MIFACT_HIP(h, hipStreamWaitEvent(h->stream, h->side_fork, 0));
                update_nt<T>(h, prec, kShapeLower, n - after, n - after, outer_size,
                             at(A, lda, after, outer), lda, at(A, lda, after, outer), lda,
                             at(A, lda, after, after), lda);
                if (h->side_join != nullptr) {
                    MIFACT_HIP(h, hipEventRecord(h->side_join, h->stream));
                    side_pending = true;
                }
            }
        }
    }
    if (side_pending) MIFACT_HIP(h, hipStreamWaitEvent(h->stream, h->side_join, 0));
}

template void double_blocked_potrf_impl<float>(mifact_handle, float*, int, int, int, int, int*,
                                              bool, LowPrecision, float*);
template void double_blocked_potrf_impl<double>(mifact_handle, double*, int, int, int, int, int*,
                                               bool, LowPrecision, double*);

namespace {

template <typename T>
void syrk_kblocked(mifact_handle h, int n, int k, const T* A, int lda, T* C, int ldc, int k_tile)
{
    if (n <= 0 || k <= 0) return;
    const int step = (k_tile <= 0 || k_tile >= k) ? k : k_tile;
    for (int k0 = 0; k0 < k; k0 += step) {
        const int kc = imin(step, k - k0);
        gemm<T>(h, kOpNT, kShapeLower, n, n, kc, T(-1), at(A, lda, 0, k0), lda,
                at(A, lda, 0, k0), lda, C, ldc);
    }
}

template <typename T>
void gemm_nt_update_impl(mifact_handle h, int rows, int cols, int k, const T* A, int lda,
                         const T* B, int ldb, T* C, int ldc)
{
    if (rows <= 0 || cols <= 0 || k <= 0) return;
    gemm<T>(h, kOpNT, kShapeFull, rows, cols, k, T(-1), A, lda, B, ldb, C, ldc);
}

template <typename T>
void gemm_nn_update_impl(mifact_handle h, int rows, int cols, int k, const T* A, int lda,
                         const T* B, int ldb, T* C, int ldc)
{
    if (rows <= 0 || cols <= 0 || k <= 0) return;
    gemm<T>(h, kOpNN, kShapeFull, rows, cols, k, T(-1), A, lda, B, ldb, C, ldc);
}

template <typename T>
void potrf_block_impl(mifact_handle h, int n, T* A, int lda, int* info)
{
    if (n <= 0) return;
    if (!ensure_workspace(h, n)) return;
    if (info != nullptr) clear_int(h, info);
    potrf_diag<T>(h, n, A, lda, info, 0, invdiag<T>(h));
}

template <typename T>
void getrf_panel_impl(mifact_handle h, int m, int n, T* A, int lda, int* ipiv, int* info)
{
    if (m <= 0 || n <= 0) return;
    if (m < n) {
        report_status(h, mifact_status_invalid_value, "mifact_getrf_panel: needs m >= n");
        return;
    }
    if (!ensure_workspace(h, m)) return;
    Workspace ws = workspace(h);
    if (info != nullptr) clear_int(h, info);
    clear_int(h, h->kernel_status);
    perm_identity(h, ws.orig, m);
    getrf_panel_rec<T>(h, A, lda, m, 0, n, ipiv, ws.orig, ws.inv, h->ws_n, info, 1,
                       panel_inv<T>(h));
}

}  // namespace
}  // namespace mifact

static void potrf_f32(mifact_handle h, float* A, int lda, int n, int nb, int b, int* info,
                      bool dbtrsm)
{
    mifact::double_blocked_potrf_impl<float>(h, A, lda, n, nb, b, info, dbtrsm, mifact::kPrecFP32,
                                             nullptr);
}

static void potrf_f64(mifact_handle h, double* A, int lda, int n, int nb, int b, int* info,
                      bool dbtrsm)
{
    mifact::double_blocked_potrf_impl<double>(h, A, lda, n, nb, b, info, dbtrsm,
                                              mifact::kPrecFP32, nullptr);
}

// ------------------------------------------------------------ allocation-free POTRF
//
// Scratch of double_blocked_potrf_noalloc: the inverses of the 64 x 64
// diagonal blocks of one inner panel, (b + 64) x 64 floats for inner blocks
// b <= kNoAllocMaxB (1 MiB on every GPU the library runs on, held until the
// process ends).  Every call on a GPU shares that GPU's copy: calls that
// overlap on one GPU race.  Nothing in it needs to start at zero (each
// block inverse is written by the diagonal factorization before the panel
// solve reads it).
namespace {
constexpr int kNoAllocMaxB = 4096;
constexpr size_t kScratchFloats = static_cast<size_t>(kNoAllocMaxB + 64) * 64;
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) float scratch[kScratchFloats];

namespace {
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (resolved once per
// device; the array's name is a host-side stand-in that must not be passed
// to kernels).  Returns null on failure.
float* scratch_address(mifact_handle h)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        mifact::report_status(h, mifact_status_hip_error, "noalloc: current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || ptr == nullptr) {
            mifact::report_status(h, mifact_status_hip_error,
                                  "noalloc: hipGetSymbolAddress(scratch) failed");
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<float*>(ptr);
    }
    return g_scratch_ptr[dev];
}

__global__ void set_info_kernel(int* info, int v) { *info = v; }
}  // namespace

namespace mifact {
// Resolves the allocation-free entry point's scratch for the current device
// (called by mifact_create_handle, the library's set-up call, so the entry
// point itself finds the address cached).
bool noalloc_setup(mifact_handle h) { return scratch_address(h) != nullptr; }
}  // namespace mifact

extern "C" {

void double_blocked_potrf(mifact_handle handle, float* A, int lda, int n, int nb, int b, int* info)
{
    potrf_f32(handle, A, lda, n, nb, b, info, false);
}

void double_blocked_potrf_noalloc(mifact_handle handle, float* A, int lda, int n, int nb, int b,
                                  int* info)
{
    if (handle == nullptr) {
        mifact::report_status(nullptr, mifact_status_invalid_handle,
                              "double_blocked_potrf_noalloc: null handle");
        return;
    }
    if (n <= 0) return;
    // the reference's block-size normalization, then the scratch bound
    int nb_n = nb <= 0 ? n : nb;
    int b_n = b <= 0 ? nb_n : b;
    if (b_n > nb_n) b_n = nb_n;
    if (b_n > n) b_n = n;       // no inner panel is wider than n
    if (b_n > kNoAllocMaxB) {
        std::fprintf(stderr,
                     "double_blocked_potrf_noalloc: inner block %d exceeds the %d its "
                     "fixed scratch holds\n", b_n, kNoAllocMaxB);
        mifact::report_status(handle, mifact_status_invalid_value,
                              "double_blocked_potrf_noalloc: inner block too large");
        if (info != nullptr) {
            hipLaunchKernelGGL(set_info_kernel, dim3(1), dim3(1), 0, handle->stream, info, -1);
            (void)hipGetLastError();
        }
        return;
    }
    float* inv = scratch_address(handle);
    if (inv == nullptr) return;
    mifact::double_blocked_potrf_impl<float>(handle, A, lda, n, nb, b, info, false,
                                             mifact::kPrecFP32, inv);
}

void double_blocked_potrf_dbtrsm(mifact_handle handle, float* A, int lda, int n, int nb, int b,
                                 int* info)
{
    potrf_f32(handle, A, lda, n, nb, b, info, true);
}

void double_blocked_potrf_kblocked_syrk(mifact_handle handle, float* A, int lda, int n, int nb,
                                        int b, int syrk_k_tile, int* info)
{
    (void)syrk_k_tile;
    potrf_f32(handle, A, lda, n, nb, b, info, false);
}

void double_blocked_potrf_kblocked_syrk_adaptive_gemm(mifact_handle handle, float* A, int lda,
                                                      int n, int nb, int b, int syrk_k_tile,
                                                      int* info)
{
    (void)syrk_k_tile;
    potrf_f32(handle, A, lda, n, nb, b, info, false);
}

void double_blocked_potrf_kblocked_syrk_adaptive_gemm_configured(
    mifact_handle handle, float* A, int lda, int n, int nb, int b, int syrk_k_tile,
    int gemm_m_tile, int gemm_m_threshold, int gemm_k_threshold, int* info)
{
    (void)syrk_k_tile;
    (void)gemm_m_tile;
    (void)gemm_m_threshold;
    (void)gemm_k_threshold;
    potrf_f32(handle, A, lda, n, nb, b, info, false);
}

void double_blocked_potrf_double(mifact_handle handle, double* A, int lda, int n, int nb, int b,
                                 int* info)
{
    potrf_f64(handle, A, lda, n, nb, b, info, false);
}

void double_blocked_potrf_double_dbtrsm(mifact_handle handle, double* A, int lda, int n, int nb,
                                        int b, int* info)
{
    potrf_f64(handle, A, lda, n, nb, b, info, true);
}

void double_blocked_potrf_double_kblocked_syrk(mifact_handle handle, double* A, int lda, int n,
                                               int nb, int b, int syrk_k_tile, int* info)
{
    (void)syrk_k_tile;
    potrf_f64(handle, A, lda, n, nb, b, info, false);
}

void double_blocked_potrf_double_kblocked_syrk_adaptive_gemm(mifact_handle handle, double* A,
                                                             int lda, int n, int nb, int b,
                                                             int syrk_k_tile, int* info)
{
    (void)syrk_k_tile;
    potrf_f64(handle, A, lda, n, nb, b, info, false);
}

void double_blocked_potrf_double_kblocked_syrk_adaptive_gemm_configured(
    mifact_handle handle, double* A, int lda, int n, int nb, int b, int syrk_k_tile,
    int gemm_m_tile, int gemm_m_threshold, int gemm_k_threshold, int* info)
{
    (void)syrk_k_tile;
    (void)gemm_m_tile;
    (void)gemm_m_threshold;
    (void)gemm_k_threshold;
    potrf_f64(handle, A, lda, n, nb, b, info, false);
}

// ------------------------------------------------------------ BLAS-3 blocks

void gemm_nt_update(mifact_handle handle, int rows, int cols, int k, const float* A, int lda,
                    const float* B, int ldb, float* C, int ldc)
{
    mifact::gemm_nt_update_impl<float>(handle, rows, cols, k, A, lda, B, ldb, C, ldc);
}

void gemm_nt_update_double(mifact_handle handle, int rows, int cols, int k, const double* A,
                           int lda, const double* B, int ldb, double* C, int ldc)
{
    mifact::gemm_nt_update_impl<double>(handle, rows, cols, k, A, lda, B, ldb, C, ldc);
}

void gemm_nt_update_m_tiled(mifact_handle handle, int rows, int cols, int k, const float* A,
                            int lda, const float* B, int ldb, float* C, int ldc, int m_tile)
{
    (void)m_tile;
    mifact::gemm_nt_update_impl<float>(handle, rows, cols, k, A, lda, B, ldb, C, ldc);
}

void gemm_nt_update_m_tiled_double(mifact_handle handle, int rows, int cols, int k,
                                   const double* A, int lda, const double* B, int ldb,
                                   double* C, int ldc, int m_tile)
{
    (void)m_tile;
    mifact::gemm_nt_update_impl<double>(handle, rows, cols, k, A, lda, B, ldb, C, ldc);
}

void gemm_nn_update(mifact_handle handle, int rows, int cols, int k, const float* L, int ldl,
                    const float* U, int ldu, float* C, int ldc)
{
    mifact::gemm_nn_update_impl<float>(handle, rows, cols, k, L, ldl, U, ldu, C, ldc);
}

void gemm_nn_update_double(mifact_handle handle, int rows, int cols, int k, const double* L,
                           int ldl, const double* U, int ldu, double* C, int ldc)
{
    mifact::gemm_nn_update_impl<double>(handle, rows, cols, k, L, ldl, U, ldu, C, ldc);
}

void syrk_lower_update(mifact_handle handle, int n, int k, const float* A, int lda, float* C,
                       int ldc)
{
    mifact::syrk_kblocked<float>(handle, n, k, A, lda, C, ldc, 0);
}

void syrk_lower_update_double(mifact_handle handle, int n, int k, const double* A, int lda,
                              double* C, int ldc)
{
    mifact::syrk_kblocked<double>(handle, n, k, A, lda, C, ldc, 0);
}

void syrk_lower_update_kblocked(mifact_handle handle, int n, int k, const float* A, int lda,
                                float* C, int ldc, int k_tile)
{
    mifact::syrk_kblocked<float>(handle, n, k, A, lda, C, ldc, k_tile);
}

void syrk_lower_update_kblocked_double(mifact_handle handle, int n, int k, const double* A,
                                       int lda, double* C, int ldc, int k_tile)
{
    mifact::syrk_kblocked<double>(handle, n, k, A, lda, C, ldc, k_tile);
}

void trsm_right_lower_transpose(mifact_handle handle, int rows, int cols, const float* A, int lda,
                                float* B, int ldb)
{
    mifact::trsm_rlt<float>(handle, rows, cols, A, lda, B, ldb);
}

void trsm_right_lower_transpose_double(mifact_handle handle, int rows, int cols, const double* A,
                                       int lda, double* B, int ldb)
{
    mifact::trsm_rlt<double>(handle, rows, cols, A, lda, B, ldb);
}

void double_blocked_trsm(mifact_handle handle, float* A, int lda, float* B, int ldb, int n, int m,
                         int nb, int b)
{
    mifact::double_blocked_trsm_impl<float>(handle, A, lda, B, ldb, n, m, nb, b);
}

void double_blocked_trsm_double(mifact_handle handle, double* A, int lda, double* B, int ldb,
                                int n, int m, int nb, int b)
{
    mifact::double_blocked_trsm_impl<double>(handle, A, lda, B, ldb, n, m, nb, b);
}

void trsm_left_lower_unit(mifact_handle handle, int rows, int cols, const float* L, int ldl,
                          float* B, int ldb)
{
    mifact::trsm_llu<float>(handle, rows, cols, L, ldl, B, ldb);
}

void trsm_left_lower_unit_double(mifact_handle handle, int rows, int cols, const double* L,
                                 int ldl, double* B, int ldb)
{
    mifact::trsm_llu<double>(handle, rows, cols, L, ldl, B, ldb);
}

void trsm_left_lower_unit_iterative(mifact_handle handle, int rows, int cols, const float* L,
                                    int ldl, float* B, int ldb, int block_size)
{
    mifact::trsm_llu_iterative<float>(handle, rows, cols, L, ldl, B, ldb, block_size);
}

void trsm_left_lower_unit_iterative_double(mifact_handle handle, int rows, int cols,
                                           const double* L, int ldl, double* B, int ldb,
                                           int block_size)
{
    mifact::trsm_llu_iterative<double>(handle, rows, cols, L, ldl, B, ldb, block_size);
}

// ------------------------------------------------------------ panel kernels

void mifact_potrf_block(mifact_handle handle, int n, float* A, int lda, int* info)
{
    mifact::potrf_block_impl<float>(handle, n, A, lda, info);
}

void mifact_potrf_block_double(mifact_handle handle, int n, double* A, int lda, int* info)
{
    mifact::potrf_block_impl<double>(handle, n, A, lda, info);
}

void mifact_getrf_panel(mifact_handle handle, int m, int n, float* A, int lda, int* ipiv,
                        int* info)
{
    mifact::getrf_panel_impl<float>(handle, m, n, A, lda, ipiv, info);
}

void mifact_getrf_panel_double(mifact_handle handle, int m, int n, double* A, int lda, int* ipiv,
                               int* info)
{
    mifact::getrf_panel_impl<double>(handle, m, n, A, lda, ipiv, info);
}

void mifact_isamax(mifact_handle handle, int n, const float* x, int incx, int* result)
{
    mifact::iamax<float>(handle, n, x, incx, result);
}

void mifact_idamax(mifact_handle handle, int n, const double* x, int incx, int* result)
{
    mifact::iamax<double>(handle, n, x, incx, result);
}

}  // extern "C"

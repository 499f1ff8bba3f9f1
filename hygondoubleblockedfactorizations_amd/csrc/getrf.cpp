// Double-blocked GETRF drivers (immediate and batched pivot application).
//
// Schedule restated from the reference:
//   double_blocked_getrf_impl        src/getrf/double_blocked_getrf.cpp:125-191
//   prepare_next_inner_panel         src/getrf/double_blocked_getrf.cpp:32-71
//   prepare_next_outer_panel         src/getrf/double_blocked_getrf.cpp:73-123
//   batched-pivot double_blocked_getrf
//                                    src/getrf/getrf_batched_pivot_internal.h:445-624
// The panel GETRF (rocsolver_sgetrf), its IAMAX and row swaps, the laswp
// kernels and the rocBLAS TRSM/GEMM calls are replaced by the gfx950
// kernels.  Pivot bookkeeping is device-resident (see pivot.hip).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <type_traits>

#include "blas3.hpp"
#include "drivers.hpp"
#include "getrf/double_blocked_getrf.h"
#include "getrf/double_blocked_getrf_npvt.h"

namespace mifact {
namespace {

constexpr int kLevelPanel = 0;
constexpr int kLevelOuter = kPermLevels - 2;
constexpr int kLevelAll = kPermLevels - 1;

__global__ void set_int_kernel(int* p, int v) { *p = v; }

void set_int(mifact_handle h, int* p, int v)
{
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, h->stream, p, v);
    MIFACT_LAUNCHED(h);
}

// getrf_batched_pivot_internal.h:43-57
bool workspace_is_valid(const getrf_batched_pivot_workspace* w, int n)
{
    return w != nullptr && w->device_sources != nullptr && w->device_targets != nullptr
        && w->device_swap_buffer != nullptr && w->host_pivots != nullptr
        && w->host_permutation != nullptr && w->host_sources != nullptr
        && w->host_targets != nullptr && w->capacity >= n && w->column_tile > 0;
}

// Trailing update C -= L U (NN): FP32/FP64 MFMA, or BF16/FP16 MFMA with FP32
// accumulation for the mixed-precision factorization.
template <typename T>
void update_nn(mifact_handle h, LowPrecision prec, int M, int N, int K, const T* L, long long ldl,
               const T* U, long long ldu, T* C, long long ldc)
{
    if constexpr (std::is_same_v<T, float>) {
        gemm_lowp(h, prec, kOpNN, kShapeFull, M, N, K, -1.0f, L, ldl, U, ldu, C, ldc);
    } else {
        gemm<T>(h, kOpNN, kShapeFull, M, N, K, T(-1), L, ldl, U, ldu, C, ldc);
    }
}

template <typename T>
void outer_update(mifact_handle h, T* A, long long lda, int n, int next_outer, int outer_nb,
                  getrf_outer_trsm_mode mode, int trsm_block, LowPrecision prec)
{
    if (next_outer >= n) return;
    const int next_size = imin(outer_nb, n - next_outer);
    if (mode == getrf_outer_trsm_direct) {
        trsm_llu<T>(h, next_outer, next_size, A, lda, at(A, lda, 0, next_outer), lda);
    } else {
        trsm_llu_iterative<T>(h, next_outer, next_size, A, lda, at(A, lda, 0, next_outer), lda,
                              trsm_block);
    }
    update_nn<T>(h, prec, n - next_outer, next_size, next_outer, at(A, lda, next_outer, 0), lda,
                 at(A, lda, 0, next_outer), lda, at(A, lda, next_outer, next_outer), lda);
}

template <typename T>
void inner_update(mifact_handle h, T* A, long long lda, int n, int outer, int next_inner,
                  int outer_end, int inner_nb, LowPrecision prec)
{
    if (next_inner >= outer_end) return;
    const int next_size = imin(inner_nb, outer_end - next_inner);
    const int history = next_inner - outer;
    trsm_llu<T>(h, history, next_size, at(A, lda, outer, outer), lda,
                at(A, lda, outer, next_inner), lda);
    update_nn<T>(h, prec, n - next_inner, next_size, history, at(A, lda, next_inner, outer), lda,
                 at(A, lda, outer, next_inner), lda, at(A, lda, next_inner, next_inner), lda);
}

}  // namespace

template <typename T>
void double_blocked_getrf_impl(mifact_handle h, T* A, int lda_i, int n, int outer_nb,
                               int inner_nb, getrf_outer_trsm_mode mode, int trsm_block,
                               int* ipiv, int* info, bool batched, T* stage,
                               long long stage_elems, LowPrecision prec)
{
    if (n <= 0) return;
    if (outer_nb <= 0) outer_nb = n;
    if (inner_nb <= 0) inner_nb = outer_nb;
    inner_nb = imin(inner_nb, outer_nb);
    const long long lda = lda_i;
    if (!ensure_workspace(h, n)) return;
    Workspace ws = workspace(h);
    const size_t nws = static_cast<size_t>(h->ws_n);
    int* orig = ws.orig;
    int* inv_panel = ws.inv + kLevelPanel * nws;
    int* inv_outer = ws.inv + kLevelOuter * nws;
    int* inv_all = ws.inv + kLevelAll * nws;
    if (info != nullptr) clear_int(h, info);
    clear_int(h, h->kernel_status);
    perm_identity(h, orig, n);
    if (batched) perm_identity(h, inv_all, n);

    for (int outer = 0; outer < n; outer += outer_nb) {
        const int outer_size = imin(outer_nb, n - outer);
        const int outer_end = outer + outer_size;
        if (batched && outer > 0) perm_snapshot(h, orig, inv_outer, outer, n);
        for (int inner = outer; inner < outer_end; inner += inner_nb) {
            const int isz = imin(inner_nb, outer_end - inner);
            const int next_inner = inner + isz;
            // panel GETRF (rocsolver_sgetrf at double_blocked_getrf.cpp:156-163)
            perm_snapshot(h, orig, inv_panel, inner, n);
            getrf_panel_rec<T>(h, A, lda, n, inner, isz, ipiv, orig, ws.inv,
                               static_cast<int>(nws), info, kLevelPanel + 1, panel_inv<T>(h));
            if (!batched) {
                // apply_panel_pivots (:165-166): every column outside the panel
                apply_row_moves<T>(h, A, lda, 0, inner, inner, isz, ipiv, orig, inv_panel,
                                   stage, stage_elems);
                apply_row_moves<T>(h, A, lda, next_inner, n, inner, isz, ipiv, orig, inv_panel,
                                   stage, stage_elems);
            } else {
                // Inner pivot batch (getrf_batched_pivot_internal.h:500-538)
                apply_row_moves<T>(h, A, lda, outer, inner, inner, isz, ipiv, orig, inv_panel,
                                   stage, stage_elems);
                apply_row_moves<T>(h, A, lda, next_inner, outer_end, inner, isz, ipiv, orig,
                                   inv_panel, stage, stage_elems);
            }
            inner_update<T>(h, A, lda, n, outer, next_inner, outer_end, inner_nb, prec);
        }
        if (batched) {
            // Outer pivot batch (:556-609): this block's exchanges complete
            // the left history; all exchanges so far prepare the next block.
            if (outer > 0) {
                apply_row_moves<T>(h, A, lda, 0, outer, outer, outer_size, ipiv, orig, inv_outer,
                                   stage, stage_elems);
            }
            if (outer_end < n) {
                apply_row_moves<T>(h, A, lda, outer_end, outer_end + imin(outer_nb, n - outer_end),
                                   0, outer_end, ipiv, orig, inv_all, stage, stage_elems);
            }
        }
        outer_update<T>(h, A, lda, n, outer_end, outer_nb, mode, trsm_block, prec);
    }
}

// Unpivoted double-blocked LU (src/getrf/double_blocked_getrf_npvt.cpp:94-148):
// the same two-level schedule with rocsolver_sgetrf_npvt replaced by the
// recursive no-pivot panel and no row exchanges; outer TRSM direct.
template <typename T>
void double_blocked_getrf_npvt_impl(mifact_handle h, T* A, int lda_i, int n, int nb, int b,
                                    int* info)
{
    if (h == nullptr) {
        report_status(nullptr, mifact_status_invalid_handle, "double_blocked_getrf_npvt: null handle");
        return;
    }
    if (n <= 0) return;
    if (nb <= 0) nb = n;
    if (b <= 0) b = nb;
    b = imin(b, nb);
    const long long lda = lda_i;
    if (!ensure_workspace(h, n)) return;
    T* inv = invdiag<T>(h);
    if (info != nullptr) clear_int(h, info);
    for (int outer = 0; outer < n; outer += nb) {
        const int outer_end = outer + imin(nb, n - outer);
        for (int inner = outer; inner < outer_end; inner += b) {
            const int isz = imin(b, outer_end - inner);
            getrf_npvt_panel<T>(h, at(A, lda, inner, inner), lda, n - inner, isz, info, inner, inv);
            inner_update<T>(h, A, lda, n, outer, inner + isz, outer_end, b, kPrecFP32);
        }
        outer_update<T>(h, A, lda, n, outer_end, nb, getrf_outer_trsm_direct, 0, kPrecFP32);
    }
}

template void double_blocked_getrf_impl<float>(mifact_handle, float*, int, int, int, int,
                                              getrf_outer_trsm_mode, int, int*, int*, bool,
                                              float*, long long, LowPrecision);
template void double_blocked_getrf_impl<double>(mifact_handle, double*, int, int, int, int,
                                               getrf_outer_trsm_mode, int, int*, int*, bool,
                                               double*, long long, LowPrecision);

namespace {

template <typename T>
void getrf_entry(mifact_handle h, T* A, int lda, int n, int outer_nb, int inner_nb,
                 getrf_outer_trsm_mode mode, int trsm_block, int* ipiv, int* info)
{
    if (h == nullptr) {
        report_status(nullptr, mifact_status_invalid_handle, "double_blocked_getrf: null handle");
        return;
    }
    double_blocked_getrf_impl<T>(h, A, lda, n, outer_nb, inner_nb, mode, trsm_block, ipiv, info,
                                 false, nullptr, 0, kPrecFP32);
}

template <typename T>
void getrf_batched_entry(mifact_handle h, T* A, int lda, int n, int outer_nb, int inner_nb,
                         getrf_outer_trsm_mode mode, int trsm_block, int* ipiv, int* info,
                         getrf_batched_pivot_workspace* w)
{
    if (h == nullptr) {
        report_status(nullptr, mifact_status_invalid_handle, "double_blocked_getrf: null handle");
        return;
    }
    if (!workspace_is_valid(w, n)) {
        std::fprintf(stderr, "Invalid batched-pivot GETRF workspace\n");
        if (info != nullptr) set_int(h, info, -1);
        return;
    }
    const long long stage_elems =
        static_cast<long long>(w->capacity) * static_cast<long long>(w->column_tile);
    double_blocked_getrf_impl<T>(h, A, lda, n, outer_nb, inner_nb, mode, trsm_block, ipiv, info,
                                 true, static_cast<T*>(w->device_swap_buffer), stage_elems,
                                 kPrecFP32);
}

}  // namespace
}  // namespace mifact

extern "C" {

void double_blocked_getrf(mifact_handle handle, float* A, int lda, int n, int outer_nb,
                          int inner_nb, getrf_outer_trsm_mode outer_trsm_mode, int trsm_block,
                          int* ipiv, int* info)
{
    mifact::getrf_entry<float>(handle, A, lda, n, outer_nb, inner_nb, outer_trsm_mode, trsm_block,
                               ipiv, info);
}

void double_blocked_getrf_double(mifact_handle handle, double* A, int lda, int n, int outer_nb,
                                 int inner_nb, getrf_outer_trsm_mode outer_trsm_mode,
                                 int trsm_block, int* ipiv, int* info)
{
    mifact::getrf_entry<double>(handle, A, lda, n, outer_nb, inner_nb, outer_trsm_mode,
                                trsm_block, ipiv, info);
}

void double_blocked_getrf_batched_pivot(mifact_handle handle, float* A, int lda, int n,
                                        int outer_nb, int inner_nb,
                                        getrf_outer_trsm_mode outer_trsm_mode, int trsm_block,
                                        int inner_gemm_m_tile, int* ipiv, int* info,
                                        getrf_batched_pivot_workspace* workspace)
{
    (void)inner_gemm_m_tile;
    mifact::getrf_batched_entry<float>(handle, A, lda, n, outer_nb, inner_nb, outer_trsm_mode,
                                       trsm_block, ipiv, info, workspace);
}

void double_blocked_getrf_double_batched_pivot(mifact_handle handle, double* A, int lda, int n,
                                               int outer_nb, int inner_nb,
                                               getrf_outer_trsm_mode outer_trsm_mode,
                                               int trsm_block, int inner_gemm_m_tile, int* ipiv,
                                               int* info,
                                               getrf_batched_pivot_workspace* workspace)
{
    (void)inner_gemm_m_tile;
    mifact::getrf_batched_entry<double>(handle, A, lda, n, outer_nb, inner_nb, outer_trsm_mode,
                                        trsm_block, ipiv, info, workspace);
}

void double_blocked_getrf_npvt(mifact_handle handle, float* A, int lda, int n, int nb, int b,
                               int* info)
{
    mifact::double_blocked_getrf_npvt_impl<float>(handle, A, lda, n, nb, b, info);
}

void double_blocked_getrf_npvt_double(mifact_handle handle, double* A, int lda, int n, int nb,
                                      int b, int* info)
{
    mifact::double_blocked_getrf_npvt_impl<double>(handle, A, lda, n, nb, b, info);
}

}  // extern "C"

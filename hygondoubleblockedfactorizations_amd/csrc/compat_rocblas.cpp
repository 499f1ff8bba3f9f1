// rocblas_handle compatibility layer (libmifact_compat.so): the reference's
// exact prototypes (see include/compat/rocblas_compat.h), forwarding to
// libmifact.so.  Not part of the hot path: per call it reads the rocBLAS
// handle's stream, binds it to this thread's mifact handle for the current
// device and calls the native entry point through a pointer resolved once.
#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <string>

#include "compat/rocblas_compat.h"
#include "mifact/handle.h"

namespace {

constexpr int kMaxDevices = 64;

void die(const char* what, const char* detail)
{
    std::fprintf(stderr, "mifact_compat: %s%s%s\n", what, detail ? ": " : "", detail ? detail : "");
    std::exit(EXIT_FAILURE);
}

// libmifact.so from this library's own directory, loaded privately.
void* native_library()
{
    static void* lib = [] {
        Dl_info info;
        if (dladdr(reinterpret_cast<void*>(&native_library), &info) == 0 || info.dli_fname == nullptr)
            die("cannot locate libmifact_compat.so", nullptr);
        std::string path(info.dli_fname);
        const size_t slash = path.rfind('/');
        path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/libmifact.so";
        void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (h == nullptr) die("cannot load libmifact.so", dlerror());
        return h;
    }();
    return lib;
}

void* resolve(const char* name)
{
    void* f = dlsym(native_library(), name);
    if (f == nullptr) die("missing entry point", name);
    return f;
}

// This thread's handle on the current device, bound to the rocBLAS
// handle's stream.
mifact_handle bound_handle(rocblas_handle rh)
{
    using create_t = mifact_status (*)(mifact_handle*);
    using stream_t = mifact_status (*)(mifact_handle, hipStream_t);
    static const create_t create = reinterpret_cast<create_t>(resolve("mifact_create_handle"));
    static const stream_t set_stream = reinterpret_cast<stream_t>(resolve("mifact_set_stream"));
    thread_local mifact_handle handles[kMaxDevices] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        die("no current device", nullptr);
    if (handles[dev] == nullptr && create(&handles[dev]) != mifact_status_success)
        die("mifact_create_handle failed", nullptr);
    hipStream_t s = nullptr;
    if (rocblas_get_stream(rh, &s) != rocblas_status_success) die("rocblas_get_stream failed", nullptr);
    set_stream(handles[dev], s);
    return handles[dev];
}

}  // namespace

extern "C" {

void gemm_nt_update(rocblas_handle handle, int rows, int cols, int k, const float* A, int lda, const float* B, int ldb, float* C, int ldc)
{
    using fn_t = void (*)(mifact_handle, int, int, int, const float*, int, const float*, int, float*, int);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("gemm_nt_update"));
    fn(bound_handle(handle), rows, cols, k, A, lda, B, ldb, C, ldc);
}

void gemm_nt_update_double(rocblas_handle handle, int rows, int cols, int k, const double* A, int lda, const double* B, int ldb, double* C, int ldc)
{
    using fn_t = void (*)(mifact_handle, int, int, int, const double*, int, const double*, int, double*, int);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("gemm_nt_update_double"));
    fn(bound_handle(handle), rows, cols, k, A, lda, B, ldb, C, ldc);
}

void gemm_nt_update_m_tiled(rocblas_handle handle, int rows, int cols, int k, const float* A, int lda, const float* B, int ldb, float* C, int ldc, int m_tile)
{
    using fn_t = void (*)(mifact_handle, int, int, int, const float*, int, const float*, int, float*, int, int);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("gemm_nt_update_m_tiled"));
    fn(bound_handle(handle), rows, cols, k, A, lda, B, ldb, C, ldc, m_tile);
}

void gemm_nt_update_m_tiled_double(rocblas_handle handle, int rows, int cols, int k, const double* A, int lda, const double* B, int ldb, double* C, int ldc, int m_tile)
{
    using fn_t = void (*)(mifact_handle, int, int, int, const double*, int, const double*, int, double*, int, int);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("gemm_nt_update_m_tiled_double"));
    fn(bound_handle(handle), rows, cols, k, A, lda, B, ldb, C, ldc, m_tile);
}

void double_blocked_getrf(rocblas_handle handle, float* A, int lda, int n, int outer_nb, int inner_nb, getrf_outer_trsm_mode outer_trsm_mode, int trsm_block, int* ipiv, int* info)
{
    using fn_t = void (*)(mifact_handle, float*, int, int, int, int, getrf_outer_trsm_mode, int, int*, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_getrf"));
    fn(bound_handle(handle), A, lda, n, outer_nb, inner_nb, outer_trsm_mode, trsm_block, ipiv, info);
}

void double_blocked_getrf_double(rocblas_handle handle, double* A, int lda, int n, int outer_nb, int inner_nb, getrf_outer_trsm_mode outer_trsm_mode, int trsm_block, int* ipiv, int* info)
{
    using fn_t = void (*)(mifact_handle, double*, int, int, int, int, getrf_outer_trsm_mode, int, int*, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_getrf_double"));
    fn(bound_handle(handle), A, lda, n, outer_nb, inner_nb, outer_trsm_mode, trsm_block, ipiv, info);
}

void double_blocked_getrf_batched_pivot(rocblas_handle handle, float* A, int lda, int n, int outer_nb, int inner_nb, getrf_outer_trsm_mode outer_trsm_mode, int trsm_block, int inner_gemm_m_tile, int* ipiv, int* info, getrf_batched_pivot_workspace* workspace)
{
    using fn_t = void (*)(mifact_handle, float*, int, int, int, int, getrf_outer_trsm_mode, int, int, int*, int*, getrf_batched_pivot_workspace*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_getrf_batched_pivot"));
    fn(bound_handle(handle), A, lda, n, outer_nb, inner_nb, outer_trsm_mode, trsm_block, inner_gemm_m_tile, ipiv, info, workspace);
}

void double_blocked_getrf_double_batched_pivot(rocblas_handle handle, double* A, int lda, int n, int outer_nb, int inner_nb, getrf_outer_trsm_mode outer_trsm_mode, int trsm_block, int inner_gemm_m_tile, int* ipiv, int* info, getrf_batched_pivot_workspace* workspace)
{
    using fn_t = void (*)(mifact_handle, double*, int, int, int, int, getrf_outer_trsm_mode, int, int, int*, int*, getrf_batched_pivot_workspace*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_getrf_double_batched_pivot"));
    fn(bound_handle(handle), A, lda, n, outer_nb, inner_nb, outer_trsm_mode, trsm_block, inner_gemm_m_tile, ipiv, info, workspace);
}

void double_blocked_getrf_npvt(rocblas_handle handle, float* A, int lda, int n, int nb, int b, int* info)
{
    using fn_t = void (*)(mifact_handle, float*, int, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_getrf_npvt"));
    fn(bound_handle(handle), A, lda, n, nb, b, info);
}

void double_blocked_getrf_npvt_double(rocblas_handle handle, double* A, int lda, int n, int nb, int b, int* info)
{
    using fn_t = void (*)(mifact_handle, double*, int, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_getrf_npvt_double"));
    fn(bound_handle(handle), A, lda, n, nb, b, info);
}

void double_blocked_potrf(rocblas_handle handle, float* A, int lda, int n, int nb, int b, int* info)
{
    using fn_t = void (*)(mifact_handle, float*, int, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_potrf"));
    fn(bound_handle(handle), A, lda, n, nb, b, info);
}

void double_blocked_potrf_dbtrsm(rocblas_handle handle, float* A, int lda, int n, int nb, int b, int* info)
{
    using fn_t = void (*)(mifact_handle, float*, int, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_potrf_dbtrsm"));
    fn(bound_handle(handle), A, lda, n, nb, b, info);
}

void double_blocked_potrf_kblocked_syrk(rocblas_handle handle, float* A, int lda, int n, int nb, int b, int syrk_k_tile, int* info)
{
    using fn_t = void (*)(mifact_handle, float*, int, int, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_potrf_kblocked_syrk"));
    fn(bound_handle(handle), A, lda, n, nb, b, syrk_k_tile, info);
}

void double_blocked_potrf_kblocked_syrk_adaptive_gemm(rocblas_handle handle, float* A, int lda, int n, int nb, int b, int syrk_k_tile, int* info)
{
    using fn_t = void (*)(mifact_handle, float*, int, int, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_potrf_kblocked_syrk_adaptive_gemm"));
    fn(bound_handle(handle), A, lda, n, nb, b, syrk_k_tile, info);
}

void double_blocked_potrf_kblocked_syrk_adaptive_gemm_configured(rocblas_handle handle, float* A, int lda, int n, int nb, int b, int syrk_k_tile, int gemm_m_tile, int gemm_m_threshold, int gemm_k_threshold, int* info)
{
    using fn_t = void (*)(mifact_handle, float*, int, int, int, int, int, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_potrf_kblocked_syrk_adaptive_gemm_configured"));
    fn(bound_handle(handle), A, lda, n, nb, b, syrk_k_tile, gemm_m_tile, gemm_m_threshold, gemm_k_threshold, info);
}

void double_blocked_potrf_double(rocblas_handle handle, double* A, int lda, int n, int nb, int b, int* info)
{
    using fn_t = void (*)(mifact_handle, double*, int, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_potrf_double"));
    fn(bound_handle(handle), A, lda, n, nb, b, info);
}

void double_blocked_potrf_double_dbtrsm(rocblas_handle handle, double* A, int lda, int n, int nb, int b, int* info)
{
    using fn_t = void (*)(mifact_handle, double*, int, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_potrf_double_dbtrsm"));
    fn(bound_handle(handle), A, lda, n, nb, b, info);
}

void double_blocked_potrf_double_kblocked_syrk(rocblas_handle handle, double* A, int lda, int n, int nb, int b, int syrk_k_tile, int* info)
{
    using fn_t = void (*)(mifact_handle, double*, int, int, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_potrf_double_kblocked_syrk"));
    fn(bound_handle(handle), A, lda, n, nb, b, syrk_k_tile, info);
}

void double_blocked_potrf_double_kblocked_syrk_adaptive_gemm(rocblas_handle handle, double* A, int lda, int n, int nb, int b, int syrk_k_tile, int* info)
{
    using fn_t = void (*)(mifact_handle, double*, int, int, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_potrf_double_kblocked_syrk_adaptive_gemm"));
    fn(bound_handle(handle), A, lda, n, nb, b, syrk_k_tile, info);
}

void double_blocked_potrf_double_kblocked_syrk_adaptive_gemm_configured(rocblas_handle handle, double* A, int lda, int n, int nb, int b, int syrk_k_tile, int gemm_m_tile, int gemm_m_threshold, int gemm_k_threshold, int* info)
{
    using fn_t = void (*)(mifact_handle, double*, int, int, int, int, int, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_potrf_double_kblocked_syrk_adaptive_gemm_configured"));
    fn(bound_handle(handle), A, lda, n, nb, b, syrk_k_tile, gemm_m_tile, gemm_m_threshold, gemm_k_threshold, info);
}

void recursive_potrf(rocblas_handle handle, StreamPool* pool, float* A, int lda, int n, int nb, int* info)
{
    using fn_t = void (*)(mifact_handle, void*, float*, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("recursive_potrf"));
    fn(bound_handle(handle), pool, A, lda, n, nb, info);
}

void recursive_potrf_tiled_trsm(rocblas_handle handle, StreamPool* pool, float* A, int lda, int n, int nb, int* info)
{
    using fn_t = void (*)(mifact_handle, void*, float*, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("recursive_potrf_tiled_trsm"));
    fn(bound_handle(handle), pool, A, lda, n, nb, info);
}

void recursive_potrf_double(rocblas_handle handle, StreamPool* pool, double* A, int lda, int n, int nb, int* info)
{
    using fn_t = void (*)(mifact_handle, void*, double*, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("recursive_potrf_double"));
    fn(bound_handle(handle), pool, A, lda, n, nb, info);
}

void recursive_potrf_mixed_driver(rocblas_handle handle, StreamPool* pool, float* A, int lda, int n, int nb, int* info)
{
    using fn_t = void (*)(mifact_handle, void*, float*, int, int, int, int*);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("recursive_potrf_mixed_driver"));
    fn(bound_handle(handle), pool, A, lda, n, nb, info);
}

void syrk_lower_update(rocblas_handle handle, int n, int k, const float* A, int lda, float* C, int ldc)
{
    using fn_t = void (*)(mifact_handle, int, int, const float*, int, float*, int);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("syrk_lower_update"));
    fn(bound_handle(handle), n, k, A, lda, C, ldc);
}

void syrk_lower_update_double(rocblas_handle handle, int n, int k, const double* A, int lda, double* C, int ldc)
{
    using fn_t = void (*)(mifact_handle, int, int, const double*, int, double*, int);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("syrk_lower_update_double"));
    fn(bound_handle(handle), n, k, A, lda, C, ldc);
}

void syrk_lower_update_kblocked(rocblas_handle handle, int n, int k, const float* A, int lda, float* C, int ldc, int k_tile)
{
    using fn_t = void (*)(mifact_handle, int, int, const float*, int, float*, int, int);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("syrk_lower_update_kblocked"));
    fn(bound_handle(handle), n, k, A, lda, C, ldc, k_tile);
}

void syrk_lower_update_kblocked_double(rocblas_handle handle, int n, int k, const double* A, int lda, double* C, int ldc, int k_tile)
{
    using fn_t = void (*)(mifact_handle, int, int, const double*, int, double*, int, int);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("syrk_lower_update_kblocked_double"));
    fn(bound_handle(handle), n, k, A, lda, C, ldc, k_tile);
}

void double_blocked_trsm(rocblas_handle handle, float* A, int lda, float* B, int ldb, int n, int m, int nb, int b)
{
    using fn_t = void (*)(mifact_handle, float*, int, float*, int, int, int, int, int);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_trsm"));
    fn(bound_handle(handle), A, lda, B, ldb, n, m, nb, b);
}

void double_blocked_trsm_double(rocblas_handle handle, double* A, int lda, double* B, int ldb, int n, int m, int nb, int b)
{
    using fn_t = void (*)(mifact_handle, double*, int, double*, int, int, int, int, int);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("double_blocked_trsm_double"));
    fn(bound_handle(handle), A, lda, B, ldb, n, m, nb, b);
}

void trsm_right_lower_transpose(rocblas_handle handle, int rows, int cols, const float* A, int lda, float* B, int ldb)
{
    using fn_t = void (*)(mifact_handle, int, int, const float*, int, float*, int);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("trsm_right_lower_transpose"));
    fn(bound_handle(handle), rows, cols, A, lda, B, ldb);
}

void trsm_right_lower_transpose_double(rocblas_handle handle, int rows, int cols, const double* A, int lda, double* B, int ldb)
{
    using fn_t = void (*)(mifact_handle, int, int, const double*, int, double*, int);
    static const fn_t fn = reinterpret_cast<fn_t>(resolve("trsm_right_lower_transpose_double"));
    fn(bound_handle(handle), rows, cols, A, lda, B, ldb);
}

}  // extern "C"

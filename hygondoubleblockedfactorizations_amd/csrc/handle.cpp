// Handle, error reporting and workspace arena.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <new>

#include "common.hpp"
#include "kernels.hpp"

namespace mifact {

namespace {
thread_local mifact_status g_unbound_error = mifact_status_success;
}

void report_hip(mifact_handle h, hipError_t e, const char* what,
                const char* file, int line)
{
    std::fprintf(stderr, "mifact: HIP error '%s' (%d) at %s:%d in %s\n",
                 hipGetErrorString(e), static_cast<int>(e), file, line, what);
    if (h == nullptr || h->error_mode == mifact_error_exit) {
        std::exit(EXIT_FAILURE);
    }
    if (h->last_error == mifact_status_success) {
        h->last_error = mifact_status_hip_error;
    }
}

void report_status(mifact_handle h, mifact_status s, const char* what)
{
    std::fprintf(stderr, "mifact: %s (status %d)\n", what, static_cast<int>(s));
    if (h == nullptr) {
        g_unbound_error = s;
        return;
    }
    if (h->error_mode == mifact_error_exit && s == mifact_status_hip_error) {
        std::exit(EXIT_FAILURE);
    }
    if (h->last_error == mifact_status_success) {
        h->last_error = s;
    }
}

static size_t int_slots(size_t n) { return (1 + kPermLevels + 4) * n; }

static size_t invdiag_bytes(size_t n) { return (n + 64) * 64 * sizeof(double); }

// Staging of row moves whose move list does not fit in LDS: 2n entries of
// kStageCols FP64 columns (2 kStageCols FP32 columns).
static size_t stage_bytes(size_t n) { return 2 * n * kStageCols * sizeof(double); }

static size_t workspace_bytes(int n)
{
    const size_t ints = int_slots(static_cast<size_t>(n));
    return ((ints * sizeof(int) + 255) & ~static_cast<size_t>(255)) + kExchangeBytes
         + 2 * invdiag_bytes(static_cast<size_t>(n)) + stage_bytes(static_cast<size_t>(n));
}

bool ensure_workspace(mifact_handle h, int n)
{
    if (n <= h->ws_n && h->ws != nullptr) {
        return true;
    }
    const int want = n < 4096 ? 4096 : n;
    const size_t bytes = workspace_bytes(want);
    if (h->ws != nullptr) {
        // Growing: the old arena may still be read by queued kernels.
        if (hipStreamSynchronize(h->stream) != hipSuccess) {
            report_status(h, mifact_status_hip_error, "stream sync before workspace growth");
            return false;
        }
        (void)hipFree(h->ws);
        h->ws = nullptr;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        report_status(h, mifact_status_allocation_failed, "workspace allocation failed");
        return false;
    }
    h->ws = p;
    h->ws_bytes = bytes;
    h->ws_n = want;
    // Exchange records start with epoch 0, which no leaf column uses.
    Workspace w = workspace(h);
    if (hipMemset(w.exchange, 0, kExchangeBytes) != hipSuccess) {
        report_status(h, mifact_status_hip_error, "workspace initialization failed");
        return false;
    }
    h->leaf_epoch = 1;
    return true;
}

Workspace workspace(mifact_handle h)
{
    Workspace w;
    char* base = static_cast<char*>(h->ws);
    const size_t n = static_cast<size_t>(h->ws_n);
    w.orig = reinterpret_cast<int*>(base);
    w.inv = w.orig + n;
    w.moves = w.inv + static_cast<size_t>(kPermLevels) * n;
    const size_t off_bytes = (int_slots(n) * sizeof(int) + 255) & ~static_cast<size_t>(255);
    w.exchange = reinterpret_cast<float*>(base + off_bytes);
    w.invdiag = base + off_bytes + kExchangeBytes;
    w.panel_inv = static_cast<char*>(w.invdiag) + invdiag_bytes(n);
    w.stage = static_cast<char*>(w.panel_inv) + invdiag_bytes(n);
    // the upper half of the exchange area holds the last leaf's move list
    w.leaf_moves = reinterpret_cast<int*>(base + off_bytes + kExchangeBytes / 2);
    w.stage_bytes = stage_bytes(n);
    return w;
}

static hipEvent_t next_event(mifact_handle h)
{
    if (h->events_used == h->event_pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        h->event_pool.push_back(e);
    }
    return h->event_pool[h->events_used++];
}

StageScope::StageScope(mifact_handle h, Stage stage, double flops, double bytes) : h_(h)
{
    if (h == nullptr || !h->profiling || ((h->profile_mask >> stage) & 1u) == 0) return;
    hipEvent_t a = next_event(h);
    hipEvent_t b = next_event(h);
    if (a == nullptr || b == nullptr) return;
    if (hipEventRecord(a, h->stream) != hipSuccess) return;
    h->records.push_back(ProfileRecord{static_cast<int>(stage), flops, bytes, a, b});
    index_ = static_cast<long>(h->records.size()) - 1;
}

StageScope::~StageScope()
{
    if (index_ < 0) return;
    (void)hipEventRecord(h_->records[static_cast<size_t>(index_)].stop, h_->stream);
}

}  // namespace mifact

extern "C" {

mifact_status mifact_profile_begin(mifact_handle h)
{
    return mifact_profile_begin_stages(h, ~0u);
}

mifact_status mifact_profile_begin_stages(mifact_handle h, unsigned stage_mask)
{
    if (h == nullptr) return mifact_status_invalid_handle;
    h->records.clear();
    h->events_used = 0;
    h->profiling = true;
    h->profile_mask = stage_mask;
    return mifact_status_success;
}

mifact_status mifact_profile_end(mifact_handle h, mifact_stage_stats* stats, int max_stages,
                                 int* count)
{
    if (h == nullptr) return mifact_status_invalid_handle;
    h->profiling = false;
    static const char* names[mifact::kStageCount] = {"gemm", "potf2", "trsm_leaf", "getf2_leaf",
                                                     "row_moves", "perm", "gemm_lowp", "solve",
                                                     "batched", "gemm_update", "gemm_update_256",
                                                     "gemm_update_lower", "gemm_update_256_lower"};
    mifact_stage_stats acc[mifact::kStageCount];
    for (int s = 0; s < mifact::kStageCount; ++s) {
        std::snprintf(acc[s].name, sizeof(acc[s].name), "%s", names[s]);
        acc[s].calls = 0;
        acc[s].flops = 0.0;
        acc[s].bytes = 0.0;
        acc[s].ms = 0.0;
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess) return mifact_status_hip_error;
    for (const auto& r : h->records) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.start, r.stop) != hipSuccess) continue;
        mifact_stage_stats& a = acc[r.stage];
        a.calls += 1;
        a.flops += r.flops;
        a.bytes += r.bytes;
        a.ms += ms;
    }
    h->records.clear();
    h->events_used = 0;
    int n = 0;
    for (int s = 0; s < mifact::kStageCount; ++s) {
        if (acc[s].calls == 0) continue;
        if (stats != nullptr && n < max_stages) stats[n] = acc[s];
        ++n;
    }
    if (count != nullptr) *count = n;
    return mifact_status_success;
}

mifact_status mifact_create_handle(mifact_handle* handle)
{
    if (handle == nullptr) {
        return mifact_status_invalid_value;
    }
    mifact_handle h = new (std::nothrow) mifact_handle_s();
    if (h == nullptr) {
        return mifact_status_allocation_failed;
    }
    if (hipGetDevice(&h->device) != hipSuccess) {
        delete h;
        return mifact_status_hip_error;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) {
        delete h;
        return mifact_status_hip_error;
    }
    h->num_cu = prop.multiProcessorCount;
    void* bar = nullptr;
    if (hipMalloc(&bar, 256) != hipSuccess) {
        delete h;
        return mifact_status_allocation_failed;
    }
    if (hipMemset(bar, 0, 256) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(bar);
        delete h;
        return mifact_status_hip_error;
    }
    h->barrier = static_cast<unsigned long long*>(bar);
    h->kernel_status = reinterpret_cast<int*>(static_cast<char*>(bar) + 64);
    h->barrier_base = 0;
    if (!mifact::noalloc_setup(h)) {
        (void)hipFree(bar);
        delete h;
        return mifact_status_hip_error;
    }
    // the side stream (default priority) and its fork/join events; without
    // them the drivers keep every launch on the main stream
    if (hipStreamCreateWithFlags(&h->side_stream, hipStreamNonBlocking) != hipSuccess) {
        h->side_stream = nullptr;
    } else if (hipEventCreateWithFlags(&h->side_fork, hipEventDisableTiming) != hipSuccess
               || hipEventCreateWithFlags(&h->side_join, hipEventDisableTiming) != hipSuccess) {
        if (h->side_fork != nullptr) (void)hipEventDestroy(h->side_fork);
        (void)hipStreamDestroy(h->side_stream);
        h->side_stream = nullptr;
        h->side_fork = nullptr;
        h->side_join = nullptr;
    }
    *handle = h;
    return mifact_status_success;
}

mifact_status mifact_destroy_handle(mifact_handle h)
{
    if (h == nullptr) {
        return mifact_status_invalid_handle;
    }
    (void)hipDeviceSynchronize();
    if (h->ws != nullptr) {
        (void)hipFree(h->ws);
    }
    if (h->barrier != nullptr) {
        (void)hipFree(h->barrier);
    }
    if (h->solve_ws != nullptr) {
        (void)hipFree(h->solve_ws);
    }
    if (h->host_norms != nullptr) {
        (void)hipHostFree(h->host_norms);
    }
    if (h->batch_ws != nullptr) {
        (void)hipFree(h->batch_ws);
    }
    if (h->pack_ws != nullptr) {
        (void)hipFree(h->pack_ws);
    }
    if (h->capture_stream != nullptr) {
        (void)hipStreamDestroy(h->capture_stream);
    }
    if (h->side_pack_ws != nullptr) {
        (void)hipFree(h->side_pack_ws);
    }
    if (h->side_fork != nullptr) {
        (void)hipEventDestroy(h->side_fork);
    }
    if (h->side_join != nullptr) {
        (void)hipEventDestroy(h->side_join);
    }
    if (h->side_stream != nullptr) {
        (void)hipStreamDestroy(h->side_stream);
    }
    for (hipEvent_t e : h->event_pool) {
        (void)hipEventDestroy(e);
    }
    delete h;
    return mifact_status_success;
}

mifact_status mifact_set_stream(mifact_handle h, hipStream_t stream)
{
    if (h == nullptr) {
        return mifact_status_invalid_handle;
    }
    h->stream = stream;
    return mifact_status_success;
}

mifact_status mifact_get_stream(mifact_handle h, hipStream_t* stream)
{
    if (h == nullptr || stream == nullptr) {
        return mifact_status_invalid_handle;
    }
    *stream = h->stream;
    return mifact_status_success;
}

mifact_status mifact_set_error_mode(mifact_handle h, mifact_error_mode mode)
{
    if (h == nullptr) {
        return mifact_status_invalid_handle;
    }
    h->error_mode = mode;
    return mifact_status_success;
}

mifact_status mifact_get_last_error(mifact_handle h)
{
    if (h == nullptr) {
        return mifact_status_invalid_handle;
    }
    mifact_status s = h->last_error;
    // A cooperative panel kernel whose bounded barrier wait expired leaves a
    // non-zero device status word; surface it here (synchronizes the stream).
    int kst = 0;
    if (hipMemcpyAsync(&kst, h->kernel_status, sizeof(int), hipMemcpyDeviceToHost,
                       h->stream) == hipSuccess
        && hipStreamSynchronize(h->stream) == hipSuccess) {
        if (kst != 0) {
            if (s == mifact_status_success) s = mifact_status_kernel_timeout;
            // read and clear: later panel kernels on this handle run again
            mifact::clear_int(h, h->kernel_status);
            (void)hipStreamSynchronize(h->stream);
        }
    }
    h->last_error = mifact_status_success;
    return s;
}

mifact_status mifact_reserve_workspace(mifact_handle h, int n)
{
    if (h == nullptr) {
        return mifact_status_invalid_handle;
    }
    if (n < 0) {
        return mifact_status_invalid_value;
    }
    return mifact::ensure_workspace(h, n) ? mifact_status_success
                                          : mifact_status_allocation_failed;
}

mifact_status mifact_set_gemm_config(mifact_handle h, int config)
{
    if (h == nullptr) return mifact_status_invalid_handle;
    if (config < 0 || config > 4) return mifact_status_invalid_value;
    h->gemm_config = config;
    return mifact_status_success;
}

mifact_status mifact_set_gemm_policy(mifact_handle h, int policy)
{
    if (h == nullptr) return mifact_status_invalid_handle;
    if (policy < 0 || policy > 3) return mifact_status_invalid_value;
    h->gemm_policy = policy;
    return mifact_status_success;
}

mifact_status mifact_set_batched_tiles(mifact_handle h, int potrf_small_max, int getrf_small_max)
{
    if (h == nullptr) return mifact_status_invalid_handle;
    if (potrf_small_max < 0 || getrf_small_max < 0) return mifact_status_invalid_value;
    h->batched_small_max_chol = potrf_small_max;
    h->batched_small_max_lu = getrf_small_max;
    return mifact_status_success;
}

mifact_status mifact_set_wait_limit(mifact_handle h, long long polls)
{
    if (h == nullptr) return mifact_status_invalid_handle;
    if (polls < 0) return mifact_status_invalid_value;
    h->wait_limit = polls;
    return mifact_status_success;
}

const char* mifact_version(void) { return "mifact 0.1.0 (gfx950)"; }

}  // extern "C"

// Internal definitions shared by the kernels and the host drivers.
#pragma once

#include <utility>

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mifact/handle.h"

namespace mifact {

// Stages of the stage profiler.
enum Stage {
    kStageGemm = 0,
    kStagePotf2,
    kStageTrsmLeaf,
    kStageGetf2Leaf,
    kStageRowMoves,
    kStagePerm,
    kStageGemmLowp,
    kStageSolve,
    kStageBatched,
    kStageGemmUpdate,
    kStageGemmUpdate256,
    kStageGemmUpdateLower,
    kStageGemmUpdate256Lower,
    kStageCount
};

struct ProfileRecord {
    int stage;
    double flops;
    double bytes;
    hipEvent_t start;
    hipEvent_t stop;
};

}  // namespace mifact

struct mifact_handle_s {
    hipStream_t stream = nullptr;
    int device = 0;
    int num_cu = 0;
    int error_mode = mifact_error_exit;
    // FP32 GEMM tile configuration (0 = automatic; see mifact_set_gemm_config).
    int gemm_config = 0;
    // Automatic tile-selection policy (see mifact_set_gemm_policy).
    int gemm_policy = 0;
    mifact_status last_error = mifact_status_success;

    // Device workspace arena: [ints | floats] sized for order ws_n.
    void* ws = nullptr;
    size_t ws_bytes = 0;
    int ws_n = 0;

    // Grid barrier counter of the cooperative panel kernel (device, 64-bit,
    // monotonically increasing) plus the number of arrivals the host has
    // issued so far (each launch waits for targets above this base).
    unsigned long long* barrier = nullptr;
    unsigned long long barrier_base = 0;
    // Epoch of the next GETF2 leaf column (exchange records carry it).
    unsigned int leaf_epoch = 1;
    // Device status word written by kernels whose bounded waits expired
    // (cleared at the start of every pivoting driver call and by
    // mifact_get_last_error).
    int* kernel_status = nullptr;
    // Polls after which a bounded wait gives up (0 = each kernel's default).
    long long wait_limit = 0;

    // Solver arena (triangular solves, residuals, refinement) and the pinned
    // host buffer its convergence tests read back.
    void* solve_ws = nullptr;
    size_t solve_ws_bytes = 0;
    int solve_cap_n = 0;
    int solve_cap_nrhs = 0;
    void* host_norms = nullptr;
    size_t host_norms_bytes = 0;
    // Blocked TRSV: epoch of the last launch (block-ready flags carry it)
    // and the tickets handed out so far (each launch draws nblk of them).
    int trsv_epoch = 0;
    unsigned long long trsv_ticket_base = 0;

    // Low-precision GEMM: packed BF16/FP16/FP8 operand copies and FP8 scales.
    void* pack_ws = nullptr;
    size_t pack_ws_bytes = 0;

    // Second stream of the handle's device (default priority, created with
    // the handle) for updates that overlap the panel work of the main stream
    // inside one driver call; fork/join events; the pack arena its
    // low-precision updates use.  Null when its creation failed: the drivers
    // then run the same launches, in the same order, on the main stream.
    hipStream_t side_stream = nullptr;
    hipEvent_t side_fork = nullptr;
    hipEvent_t side_join = nullptr;
    void* side_pack_ws = nullptr;
    size_t side_pack_ws_bytes = 0;

    // Batched path: work-queue order, counter and FP32 staging; the private
    // stream refinement graphs are captured on.
    void* batch_ws = nullptr;
    size_t batch_ws_bytes = 0;
    hipStream_t capture_stream = nullptr;
    // Size classes of the batched kernels' trailing updates: orders up to
    // these take 32 x 32 tiles (one per wave), larger ones 64 x 64 tiles on
    // four waves (see mifact_set_batched_tiles).
    int batched_small_max_chol = 96;
    int batched_small_max_lu = 96;

    // Stage profiler: event pool and the records of the current session.
    bool profiling = false;
    unsigned profile_mask = ~0u;   // stages recorded while profiling (bit per Stage)
    std::vector<hipEvent_t> event_pool;
    size_t events_used = 0;
    std::vector<mifact::ProfileRecord> records;
};

namespace mifact {

// ---------------------------------------------------------------- errors

void report_hip(mifact_handle h, hipError_t e, const char* what,
                const char* file, int line);
void report_status(mifact_handle h, mifact_status s, const char* what);

#define MIFACT_HIP(h, call)                                                 \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) {                                             \
            ::mifact::report_hip((h), e_, #call, __FILE__, __LINE__);       \
            return;                                                         \
        }                                                                   \
    } while (0)

#define MIFACT_HIP_RET(h, call, ret)                                        \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) {                                             \
            ::mifact::report_hip((h), e_, #call, __FILE__, __LINE__);       \
            return (ret);                                                   \
        }                                                                   \
    } while (0)

// Checks the launch status of the last kernel.
#define MIFACT_LAUNCHED(h) MIFACT_HIP((h), hipGetLastError())

// Launches made while an instance lives go to the handle's side stream and
// use its side pack arena (no-op when the handle has no side stream).
class OnSideStream {
public:
    explicit OnSideStream(mifact_handle h) : h_(h->side_stream != nullptr ? h : nullptr)
    {
        if (h_ != nullptr) swap_();
    }
    ~OnSideStream()
    {
        if (h_ != nullptr) swap_();
    }
    OnSideStream(const OnSideStream&) = delete;
    OnSideStream& operator=(const OnSideStream&) = delete;

private:
    void swap_()
    {
        std::swap(h_->stream, h_->side_stream);
        std::swap(h_->pack_ws, h_->side_pack_ws);
        std::swap(h_->pack_ws_bytes, h_->side_pack_ws_bytes);
    }
    mifact_handle h_;
};

// ---------------------------------------------------------------- indexing

// 64-bit column-major element offset: n = 65536 gives n^2 = 2^32 elements.
__host__ __device__ inline size_t off(long long row, long long col, long long ld)
{
    return static_cast<size_t>(col) * static_cast<size_t>(ld)
         + static_cast<size_t>(row);
}

template <typename T>
__host__ __device__ inline T* at(T* A, long long ld, long long row, long long col)
{
    return A + off(row, col, ld);
}

__host__ __device__ inline int imin(int a, int b) { return a < b ? a : b; }
__host__ __device__ inline int imax(int a, int b) { return a > b ? a : b; }
__host__ __device__ inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

// XCD of the executing workgroup (hardware register XCC_ID, bits 3:0); the
// hardware deals consecutive workgroups of a launch to the XCDs in turn.
constexpr int kNumXcd = 8;
__device__ inline unsigned int xcd_id()
{
    unsigned int v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return v & 0xfu;
}

// ---------------------------------------------------------------- workspace

// Fixed layout of the workspace arena for order n (all offsets in bytes).
struct Workspace {
    int* orig = nullptr;         // n ints: origin row of each row's content
    int* inv = nullptr;          // kPermLevels x n ints: snapshots of orig^-1
    int* moves = nullptr;        // 4 x n ints: move lists (src 2n | dst 2n)
    float* exchange = nullptr;   // cooperative panel exchange records
    void* invdiag = nullptr;     // (n + 64) x 64 elements (double-sized): inverses
                                 // of the 64 x 64 diagonal blocks of a solve
    void* panel_inv = nullptr;   // (n + 64) x 64 elements (double-sized): the
                                 // GETRF panel's diagonal-block inverses, the
                                 // block at global column c at offset 64 c
    int* leaf_moves = nullptr;   // move list of the last GETF2 leaf (2 x 64 ints)
    void* stage = nullptr;       // row-move staging: 2n x kStageCols doubles
    size_t stage_bytes = 0;
};

constexpr int kPermLevels = 24;             // panel/outer + recursion levels
constexpr size_t kExchangeBytes = 1 << 20;  // 2 parities x (G+1) records
constexpr int kStageCols = 16;              // columns per row-move staging chunk (FP64)

bool ensure_workspace(mifact_handle h, int n);
// Set-up of the allocation-free POTRF entry point on the current device.
bool noalloc_setup(mifact_handle h);
Workspace workspace(mifact_handle h);

// Brackets the kernel launches issued in its lifetime with an event pair
// when the handle's profiler is on (no effect otherwise).
class StageScope {
public:
    StageScope(mifact_handle h, Stage stage, double flops, double bytes);
    ~StageScope();
    StageScope(const StageScope&) = delete;
    StageScope& operator=(const StageScope&) = delete;

private:
    mifact_handle h_;
    long index_ = -1;
};

}  // namespace mifact

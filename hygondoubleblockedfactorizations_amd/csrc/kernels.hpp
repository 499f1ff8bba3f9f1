// Host-side launchers of the gfx950 kernels (internal interface).
//
// Every launcher enqueues on h->stream and returns immediately; all matrices
// are column-major with 64-bit leading dimensions.
#pragma once

#include "common.hpp"

namespace mifact {

enum GemmOp { kOpNT = 0, kOpNN = 1, kOpTN = 2 };
enum GemmShape { kShapeFull = 0, kShapeLower = 1 };

// C(M x N) += alpha * A(M x K) * op(B)
//   kOpNT: op(B) = B^T with B stored N x K   (C -= A B^T update: alpha = -1)
//   kOpNN: op(B) = B   with B stored K x N
//   kOpTN: C += alpha * A^T * B with A stored K x M and B stored K x N
// kShapeLower touches only C(i, j) with i >= j (SYRK / trapezoid updates).
template <typename T>
void gemm(mifact_handle h, GemmOp op, GemmShape shape, int M, int N, int K,
          T alpha, const T* A, long long lda, const T* B, long long ldb, T* C,
          long long ldc);

// Arithmetic of the trailing updates of the mixed-precision factorizations.
enum LowPrecision { kPrecFP32 = 0, kPrecBF16 = 1, kPrecFP16 = 2, kPrecFP8 = 3 };

// FP32 C += alpha * A * op(B) with the products on BF16 / FP16 / FP8-e4m3
// MFMA (FP32 accumulation; FP8 with a power-of-two scale per 32 x 32 operand
// tile); kPrecFP32 forwards to gemm<float>.  op: kOpNT or kOpNN.  Uses the
// handle's pack arena (grown on demand).
void gemm_lowp(mifact_handle h, LowPrecision prec, GemmOp op, GemmShape shape, int M, int N,
               int K, float alpha, const float* A, long long lda, const float* B, long long ldb,
               float* C, long long ldc);

// C(M x N) = alpha * A * op(B) (no accumulation into C; C may alias A or B
// when every workgroup's tile reads only the rows/columns it writes, as in
// the triangular-solve leaves).
template <typename T>
void gemm_overwrite(mifact_handle h, GemmOp op, int M, int N, int K, T alpha, const T* A,
                    long long lda, const T* B, long long ldb, T* C, long long ldc);

// Unblocked lower Cholesky of an s x s block (s <= 64) in one wave.  On a
// non-positive pivot at local column j: *info = col_offset + j + 1 if *info
// was 0 (first failure in launch order wins).  When inv != null, inv(L) is
// written column-major with ld 64.
template <typename T>
void potf2_base(mifact_handle h, int s, T* A, long long lda, T* inv, int* info,
                int col_offset);

// Unpivoted LU of an s x s block (s <= 64) in one wave; a zero pivot at
// local column j sets *info = col_offset + j + 1 if it was 0.
template <typename T>
void getf2_npvt_block(mifact_handle h, int s, T* A, long long lda, int* info, int col_offset);

// Inverses of the 64 x 64 diagonal blocks of the n x n lower triangle of L
// (unit diagonal when unit): block b at inv + 4096 * b, ld 64.  upper: the
// blocks of an upper-triangular U are inverted as U^T, giving inv(U)^T.
template <typename T>
void trtri_blocks(mifact_handle h, int n, const T* L, long long ldl, T* inv, bool unit,
                  bool upper = false);

// Width of the GETF2 leaves: 64 columns in FP32, 32 in FP64 (the rows of a
// leaf stay in registers: 2 x 64 FP32 or 2 x 32 FP64 values per thread).
template <typename T>
inline constexpr int kLeafWidth = sizeof(T) == 4 ? 64 : 32;

// Cooperative partial-pivot LU of a leaf panel: rows [c0, n) x columns
// [c0, c0 + w) of the full matrix (w <= kLeafWidth<T>).  Writes ipiv[c0 + j]
// (global, 1-based), carries orig[] (the origin of each row's content) along
// with the row exchanges and records the first zero pivot in *info.
// Optionally:
//   inv      receives the inverse of the leaf's unit-lower w x w diagonal
//            block (kLeafWidth<T> square, ld 64, identity rows past w);
//            inv_zero, when set, names a 32 x 32 block (ld 64) that is
//            cleared (FP64: the off-diagonal block of a 64-column pair);
//   mv_src,  receive the move list of the leaf's own exchanges (2w entries,
//   mv_dst   global rows, src -1 for "no move"), the same list
//            apply_row_moves would derive from a snapshot taken before it.
template <typename T>
void getf2_leaf(mifact_handle h, T* A, long long lda, int n, int c0, int w,
                int* ipiv, int* orig, int* info, T* inv = nullptr, T* inv_zero = nullptr,
                int* mv_src = nullptr, int* mv_dst = nullptr);

// Completes the 64 x 64 inverse at inv (ld 64) of the unit-lower block
// L(0:64, 0:64) whose two 32 x 32 diagonal inverses are already in place.
template <typename T>
void inv_pair(mifact_handle h, const T* L, long long ldl, T* inv);

// inv[orig[p]] = p for p in [r0, n): snapshot of the row-origin map.
void perm_snapshot(mifact_handle h, const int* orig, int* inv, int r0, int n);
// orig[p] = p for p in [0, n).
void perm_identity(mifact_handle h, int* orig, int n);
// *info = 0 with a kernel (no memset nodes in captured graphs).
void clear_int(mifact_handle h, int* p);

// Applies a move list to columns [c_begin, c_end) of A: for each entry t
// with src[t] >= 0, row dst[t] receives the (pre-call) entry of row src[t].
// tmp: optional global staging (>= nm elements per column chunk), used when
// nm values per column do not fit in LDS; else the workspace's staging.
template <typename T>
void apply_moves(mifact_handle h, T* A, long long lda, int c_begin, int c_end, int nm,
                 const int* src, const int* dst, T* tmp, long long tmp_elems);

// Applies the rows exchanged since a snapshot to a block of columns.
// Candidate destination rows: [p0, p0 + K) and ipiv[p0 .. p0+K) - 1.
// For each candidate pos: src = inv[orig[pos]]; if src != pos the column
// entry at pos receives the (pre-call) entry at src.  Columns [c_begin,
// c_end) of A.  tmp as for apply_moves.
template <typename T>
void apply_row_moves(mifact_handle h, T* A, long long lda, int c_begin,
                     int c_end, int p0, int K, const int* ipiv,
                     const int* orig, const int* inv, T* tmp, long long tmp_elems);

// Wave-reduction IAMAX: *result = 1-based index of the first entry of
// maximum |x| (0 when n <= 0).
template <typename T>
void iamax(mifact_handle h, int n, const T* x, int incx, int* result);

}  // namespace mifact

// Recursive BLAS-3 building blocks composed from the gfx950 kernels.
#pragma once

#include "kernels.hpp"

namespace mifact {

// Splits an order into two parts with the first a multiple of `align`.
inline int split_point(int w, int align)
{
    int w1 = ((w / 2 + align - 1) / align) * align;
    if (w1 >= w) w1 = w - align;
    if (w1 <= 0) w1 = w / 2;
    return w1;
}

// Scratch for the inverses of the 64 x 64 diagonal blocks of a solve
// (column-major, ld 64; the leaf starting at local column c sits at
// inv + 64 c).  Sized for the workspace order.
template <typename T>
T* invdiag(mifact_handle h)
{
    return static_cast<T*>(workspace(h).invdiag);
}

// B(rows x cols) := B * L^{-T}, L lower non-unit (cols x cols), given the
// inverses of L's 64 x 64 diagonal blocks at inv (leaf at local column c:
// inv + 64 c).  Recursive halving at multiples of 64; each leaf is an
// in-place MFMA multiply X = B inv(L_cc)^T.
template <typename T>
void trsm_rlt_inv(mifact_handle h, int rows, int cols, const T* L, long long ldl, T* B,
                  long long ldb, const T* inv)
{
    if (rows <= 0 || cols <= 0) return;
    if (cols <= 64) {
        gemm_overwrite<T>(h, kOpNT, rows, cols, cols, T(1), B, ldb, inv, 64, B, ldb);
        return;
    }
    const int c1 = split_point(cols, 64);
    const int c2 = cols - c1;
    trsm_rlt_inv<T>(h, rows, c1, L, ldl, B, ldb, inv);
    gemm<T>(h, kOpNT, kShapeFull, rows, c2, c1, T(-1), B, ldb, at(L, ldl, c1, 0), ldl,
            at(B, ldb, 0, c1), ldb);
    trsm_rlt_inv<T>(h, rows, c2, at(L, ldl, c1, c1), ldl, at(B, ldb, 0, c1), ldb,
                    inv + static_cast<size_t>(c1) * 64);
}

// B(rows x cols) := L^{-1} B, L lower (rows x rows; unit or not as the
// inverses were formed), leaves multiply by inv(L_rr) in place on MFMA.
template <typename T>
void trsm_llu_inv(mifact_handle h, int rows, int cols, const T* L, long long ldl, T* B,
                  long long ldb, const T* inv)
{
    if (rows <= 0 || cols <= 0) return;
    if (rows <= 64) {
        gemm_overwrite<T>(h, kOpNN, rows, cols, rows, T(1), inv, 64, B, ldb, B, ldb);
        return;
    }
    const int r1 = split_point(rows, 64);
    const int r2 = rows - r1;
    trsm_llu_inv<T>(h, r1, cols, L, ldl, B, ldb, inv);
    gemm<T>(h, kOpNN, kShapeFull, r2, cols, r1, T(-1), at(L, ldl, r1, 0), ldl, B, ldb,
            at(B, ldb, r1, 0), ldb);
    trsm_llu_inv<T>(h, r2, cols, at(L, ldl, r1, r1), ldl, at(B, ldb, r1, 0), ldb,
                    inv + static_cast<size_t>(r1) * 64);
}

// B := B * L^{-T} for a standalone lower non-unit L (inverses formed here).
template <typename T>
void trsm_rlt(mifact_handle h, int rows, int cols, const T* L, long long ldl, T* B,
              long long ldb)
{
    if (rows <= 0 || cols <= 0) return;
    if (!ensure_workspace(h, cols)) return;
    T* inv = invdiag<T>(h);
    trtri_blocks<T>(h, cols, L, ldl, inv, false);
    trsm_rlt_inv<T>(h, rows, cols, L, ldl, B, ldb, inv);
}

// B := L^{-1} B for a standalone unit lower L (inverses formed here).
template <typename T>
void trsm_llu(mifact_handle h, int rows, int cols, const T* L, long long ldl, T* B,
              long long ldb)
{
    if (rows <= 0 || cols <= 0) return;
    if (!ensure_workspace(h, rows)) return;
    T* inv = invdiag<T>(h);
    trtri_blocks<T>(h, rows, L, ldl, inv, true);
    trsm_llu_inv<T>(h, rows, cols, L, ldl, B, ldb, inv);
}

// B(rows x cols) := L^{-T} B, L lower (rows x rows), given inverses of L's
// diagonal blocks (leaves multiply by inv(L_jj)^T: TN MFMA GEMM in place).
template <typename T>
void trsm_llt_inv(mifact_handle h, int rows, int cols, const T* L, long long ldl, T* B,
                  long long ldb, const T* inv)
{
    if (rows <= 0 || cols <= 0) return;
    if (rows <= 64) {
        gemm_overwrite<T>(h, kOpTN, rows, cols, rows, T(1), inv, 64, B, ldb, B, ldb);
        return;
    }
    const int r1 = split_point(rows, 64);
    const int r2 = rows - r1;
    trsm_llt_inv<T>(h, r2, cols, at(L, ldl, r1, r1), ldl, at(B, ldb, r1, 0), ldb,
                    inv + static_cast<size_t>(r1) * 64);
    gemm<T>(h, kOpTN, kShapeFull, r1, cols, r2, T(-1), at(L, ldl, r1, 0), ldl,
            at(B, ldb, r1, 0), ldb, B, ldb);
    trsm_llt_inv<T>(h, r1, cols, L, ldl, B, ldb, inv);
}

// B(rows x cols) := U^{-1} B, U upper (rows x rows), given inv(U_jj)^T of
// its diagonal blocks (trtri_blocks with upper = true).
template <typename T>
void trsm_lun_inv(mifact_handle h, int rows, int cols, const T* U, long long ldu, T* B,
                  long long ldb, const T* invt)
{
    if (rows <= 0 || cols <= 0) return;
    if (rows <= 64) {
        gemm_overwrite<T>(h, kOpTN, rows, cols, rows, T(1), invt, 64, B, ldb, B, ldb);
        return;
    }
    const int r1 = split_point(rows, 64);
    const int r2 = rows - r1;
    trsm_lun_inv<T>(h, r2, cols, at(U, ldu, r1, r1), ldu, at(B, ldb, r1, 0), ldb,
                    invt + static_cast<size_t>(r1) * 64);
    gemm<T>(h, kOpNN, kShapeFull, r1, cols, r2, T(-1), at(U, ldu, 0, r1), ldu,
            at(B, ldb, r1, 0), ldb, B, ldb);
    trsm_lun_inv<T>(h, r1, cols, U, ldu, B, ldb, invt);
}

// Row-blocked variant: diagonal solves of `block` rows, each followed by a
// GEMM update of the rows below (getrf_pivot_internal.h:64-110 schedule).
template <typename T>
void trsm_llu_iterative(mifact_handle h, int rows, int cols, const T* L, long long ldl,
                        T* B, long long ldb, int block)
{
    if (rows <= 0 || cols <= 0) return;
    if (block <= 0 || rows <= block) {
        trsm_llu<T>(h, rows, cols, L, ldl, B, ldb);
        return;
    }
    for (int row = 0; row < rows; row += block) {
        const int dr = imin(block, rows - row);
        trsm_llu<T>(h, dr, cols, at(L, ldl, row, row), ldl, B + row, ldb);
        const int tr = rows - row - dr;
        if (tr > 0) {
            gemm<T>(h, kOpNN, kShapeFull, tr, cols, dr, T(-1), at(L, ldl, row + dr, row), ldl,
                    B + row, ldb, B + row + dr, ldb);
        }
    }
}

// Unpivoted LU of an m x w panel (rows below the diagonal block included):
// recursive halving at multiples of 64; each 64-column leaf factors its
// diagonal block in one wave and forms the rows below as A21 inv(U11) with
// an in-place MFMA multiply (inv(U11)^T from trtri_blocks).
template <typename T>
void getrf_npvt_panel(mifact_handle h, T* A, long long lda, int m, int w, int* info,
                      int col_offset, T* inv)
{
    if (w <= 0 || m <= 0) return;
    if (w <= 64) {
        getf2_npvt_block<T>(h, w, A, lda, info, col_offset);
        if (m > w) {
            trtri_blocks<T>(h, w, A, lda, inv, false, true);
            gemm_overwrite<T>(h, kOpNT, m - w, w, w, T(1), A + w, lda, inv, 64, A + w, lda);
        }
        return;
    }
    const int w1 = split_point(w, 64);
    const int w2 = w - w1;
    getrf_npvt_panel<T>(h, A, lda, m, w1, info, col_offset, inv);
    trsm_llu<T>(h, w1, w2, A, lda, at(A, lda, 0, w1), lda);
    gemm<T>(h, kOpNN, kShapeFull, m - w1, w2, w1, T(-1), at(A, lda, w1, 0), lda,
            at(A, lda, 0, w1), lda, at(A, lda, w1, w1), lda);
    getrf_npvt_panel<T>(h, at(A, lda, w1, w1), lda, m - w1, w2, info, col_offset + w1, inv);
}

// In-place lower Cholesky of an s x s diagonal block (any s): recursive
// halving down to 64-wide one-wave leaves, which also leave the inverses of
// their blocks at inv (local column c at inv + 64 c) for the panel solve.
// info: first failing global column.
template <typename T>
void potrf_diag(mifact_handle h, int s, T* A, long long lda, int* info, int col_offset, T* inv)
{
    if (s <= 0) return;
    if (s <= 64) {
        potf2_base<T>(h, s, A, lda, inv, info, col_offset);
        return;
    }
    const int s1 = split_point(s, 64);
    const int s2 = s - s1;
    potrf_diag<T>(h, s1, A, lda, info, col_offset, inv);
    trsm_rlt_inv<T>(h, s2, s1, A, lda, at(A, lda, s1, 0), lda, inv);
    gemm<T>(h, kOpNT, kShapeLower, s2, s2, s1, T(-1), at(A, lda, s1, 0), lda,
            at(A, lda, s1, 0), lda, at(A, lda, s1, s1), lda);
    potrf_diag<T>(h, s2, at(A, lda, s1, s1), lda, info, col_offset + s1,
                  inv + static_cast<size_t>(s1) * 64);
}

inline bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Recursive partial-pivot LU of rows [c0, n) x columns [c0, c0 + w) of A.
// ipiv is written 1-based and global; orig tracks row contents; inv holds
// kPermLevels snapshot arrays of n ints (level 0 belongs to the caller).
//
// Leaves (w <= LW = kLeafWidth<T>: 64 FP32 columns, 32 FP64) leave behind
// the move list of their own exchanges and the inverse of their unit-lower
// diagonal block at pinv + 64 c0, so a node whose left or right child is a
// leaf needs no snapshot, and its triangular solve needs no separate
// inversion.  A left child of power-of-two width >= LW is made of leaves at
// multiples of LW, so its solve uses their blocks too (FP64: a node of width
// exactly 64 pairs its two 32-leaves' inverses into the 64 x 64 inverse,
// first leaf on the top-left, second on the bottom-right, inv_pair filling
// the bottom-left once the block's rows are final); any other width forms
// its inverses with trtri (the generic trsm_llu).
// role (FP64 pairs): 0 lone, 1 first leaf of a 64-node, 2 second leaf.
template <typename T>
void getrf_panel_rec(mifact_handle h, T* A, long long lda, int n, int c0, int w, int* ipiv,
                     int* orig, int* inv, int nws, int* info, int level, T* pinv, int role = 0)
{
    constexpr int LW = kLeafWidth<T>;
    if (w <= 0) return;
    Workspace ws = workspace(h);
    int* mv_src = ws.leaf_moves;
    int* mv_dst = ws.leaf_moves + 2 * LW;
    if (w <= LW) {
        T* ib = (role == 2) ? pinv + 64 * static_cast<size_t>(c0 - 32) + 32 + 64 * 32
                            : pinv + 64 * static_cast<size_t>(c0);
        T* iz = (role == 1) ? pinv + 64 * static_cast<size_t>(c0) + 64 * 32 : nullptr;
        getf2_leaf<T>(h, A, lda, n, c0, w, ipiv, orig, info, ib, iz, mv_src, mv_dst);
        return;
    }
    if (level >= kPermLevels - 2) {
        report_status(h, mifact_status_invalid_value, "getrf panel recursion too deep");
        return;
    }
    const int w1 = split_point(w, LW);
    const int w2 = w - w1;
    const bool pair = (LW == 32 && w == 64);
    int* inv_l = inv + static_cast<size_t>(level) * nws;
    if (w1 > LW) perm_snapshot(h, orig, inv_l, c0, n);
    getrf_panel_rec<T>(h, A, lda, n, c0, w1, ipiv, orig, inv, nws, info, level + 1, pinv,
                       pair ? 1 : 0);
    if (w1 > LW) {
        apply_row_moves<T>(h, A, lda, c0 + w1, c0 + w, c0, w1, ipiv, orig, inv_l, nullptr, 0);
    } else {
        apply_moves<T>(h, A, lda, c0 + w1, c0 + w, 2 * w1, mv_src, mv_dst, nullptr, 0);
    }
    if (w1 <= LW || is_pow2(w1)) {
        trsm_llu_inv<T>(h, w1, w2, at(A, lda, c0, c0), lda, at(A, lda, c0, c0 + w1), lda,
                        pinv + 64 * static_cast<size_t>(c0));
    } else {
        trsm_llu<T>(h, w1, w2, at(A, lda, c0, c0), lda, at(A, lda, c0, c0 + w1), lda);
    }
    gemm<T>(h, kOpNN, kShapeFull, n - c0 - w1, w2, w1, T(-1), at(A, lda, c0 + w1, c0), lda,
            at(A, lda, c0, c0 + w1), lda, at(A, lda, c0 + w1, c0 + w1), lda);
    if (w2 > LW) perm_snapshot(h, orig, inv_l, c0 + w1, n);
    getrf_panel_rec<T>(h, A, lda, n, c0 + w1, w2, ipiv, orig, inv, nws, info, level + 1, pinv,
                       pair ? 2 : 0);
    if (w2 > LW) {
        apply_row_moves<T>(h, A, lda, c0, c0 + w1, c0 + w1, w2, ipiv, orig, inv_l, nullptr, 0);
    } else {
        apply_moves<T>(h, A, lda, c0, c0 + w1, 2 * w2, mv_src, mv_dst, nullptr, 0);
    }
    if (pair) inv_pair<T>(h, at(A, lda, c0, c0), lda, pinv + 64 * static_cast<size_t>(c0));
}

// The panel-inverse region of the workspace (see getrf_panel_rec).
template <typename T>
T* panel_inv(mifact_handle h)
{
    return static_cast<T*>(workspace(h).panel_inv);
}

}  // namespace mifact

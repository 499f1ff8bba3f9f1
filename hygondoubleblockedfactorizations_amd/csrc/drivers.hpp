// Factorization drivers shared by the C-ABI entry points and the
// mixed-precision solvers.
#pragma once

#include "getrf/double_blocked_getrf.h"
#include "kernels.hpp"

namespace mifact {

// Double-blocked POTRF (reference schedule); trailing updates in `prec`.
// inv_scratch: (b + 64) x 64 elements for the diagonal-block inverses, or
// null to use the handle's workspace.
template <typename T>
void double_blocked_potrf_impl(mifact_handle h, T* A, int lda, int n, int nb, int b, int* info,
                               bool use_dbtrsm, LowPrecision prec, T* inv_scratch = nullptr);

// Double-blocked GETRF (reference schedule); trailing updates in `prec`.
// On return the workspace's orig[] holds the final row permutation
// (row p of the factors came from row orig[p] of A).
template <typename T>
void double_blocked_getrf_impl(mifact_handle h, T* A, int lda, int n, int outer_nb, int inner_nb,
                               getrf_outer_trsm_mode mode, int trsm_block, int* ipiv, int* info,
                               bool batched, T* stage, long long stage_elems, LowPrecision prec);

}  // namespace mifact

/*
 * syrk/kblocked_syrk.h -- lower SYRK update C_lower -= A * A^T on the gfx950
 * MFMA kernels (tiles strictly above the diagonal are never launched; the
 * strictly-upper part of diagonal tiles is masked, so only the lower
 * triangle of C is written, as rocblas_ssyrk(lower) does).
 *
 * Same names and argument order as the reference
 * (include/syrk/kblocked_syrk.h:9-41, src/syrk/kblocked_syrk.cpp:17-96).
 * The k_tile of the _kblocked variants splits K into separate passes exactly
 * as the reference does (each chunk is accumulated into C in order).
 */
#pragma once

#include "mifact/handle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference: include/syrk/kblocked_syrk.h:9-15 */
void syrk_lower_update(mifact_handle handle, int n, int k, const float* A,
                       int lda, float* C, int ldc);
/* reference: include/syrk/kblocked_syrk.h:17-23 */
void syrk_lower_update_double(mifact_handle handle, int n, int k,
                              const double* A, int lda, double* C, int ldc);
/* reference: include/syrk/kblocked_syrk.h:25-32 */
void syrk_lower_update_kblocked(mifact_handle handle, int n, int k,
                                const float* A, int lda, float* C, int ldc,
                                int k_tile);
/* reference: include/syrk/kblocked_syrk.h:34-41 */
void syrk_lower_update_kblocked_double(mifact_handle handle, int n, int k,
                                       const double* A, int lda, double* C,
                                       int ldc, int k_tile);

#ifdef __cplusplus
}
#endif

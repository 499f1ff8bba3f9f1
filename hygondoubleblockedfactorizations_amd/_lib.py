"""ctypes binding of ``libmifact.so`` (the C-ABI declared under ``include/``).

The library is the product; this module only loads it and declares argument
types.  Loading fails loudly when the library has not been built: there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libmifact.so")

P = ctypes.c_void_p
I = ctypes.c_int

# name -> argtypes (all functions return void unless listed in _RESTYPE)
SIGNATURES = {
    # mifact/handle.h
    "mifact_create_handle": [ctypes.POINTER(P)],
    "mifact_destroy_handle": [P],
    "mifact_set_stream": [P, P],
    "mifact_get_stream": [P, ctypes.POINTER(P)],
    "mifact_set_error_mode": [P, I],
    "mifact_get_last_error": [P],
    "mifact_reserve_workspace": [P, I],
    "mifact_set_gemm_config": [P, I],
    "mifact_set_gemm_policy": [P, I],
    "mifact_set_batched_tiles": [P, I, I],
    "mifact_set_wait_limit": [P, ctypes.c_longlong],
    "mifact_version": [],
    "mifact_profile_begin": [P],
    "mifact_profile_begin_stages": [P, ctypes.c_uint],
    "mifact_profile_end": [P, P, I, ctypes.POINTER(I)],
    # potrf/double_blocked_potrf.h
    "double_blocked_potrf": [P, P, I, I, I, I, P],
    "double_blocked_potrf_dbtrsm": [P, P, I, I, I, I, P],
    "double_blocked_potrf_noalloc": [P, P, I, I, I, I, P],
    "double_blocked_potrf_kblocked_syrk": [P, P, I, I, I, I, I, P],
    "double_blocked_potrf_kblocked_syrk_adaptive_gemm": [P, P, I, I, I, I, I, P],
    "double_blocked_potrf_kblocked_syrk_adaptive_gemm_configured":
        [P, P, I, I, I, I, I, I, I, I, P],
    "double_blocked_potrf_double": [P, P, I, I, I, I, P],
    "double_blocked_potrf_double_dbtrsm": [P, P, I, I, I, I, P],
    "double_blocked_potrf_double_kblocked_syrk": [P, P, I, I, I, I, I, P],
    "double_blocked_potrf_double_kblocked_syrk_adaptive_gemm": [P, P, I, I, I, I, I, P],
    "double_blocked_potrf_double_kblocked_syrk_adaptive_gemm_configured":
        [P, P, I, I, I, I, I, I, I, I, P],
    # potrf/recursive_potrf.h
    "recursive_potrf": [P, P, P, I, I, I, P],
    "recursive_potrf_tiled_trsm": [P, P, P, I, I, I, P],
    "recursive_potrf_double": [P, P, P, I, I, I, P],
    "recursive_potrf_mixed_driver": [P, P, P, I, I, I, P],
    # getrf/double_blocked_getrf.h
    "double_blocked_getrf": [P, P, I, I, I, I, I, I, P, P],
    "double_blocked_getrf_double": [P, P, I, I, I, I, I, I, P, P],
    "double_blocked_getrf_batched_pivot": [P, P, I, I, I, I, I, I, I, P, P, P],
    "double_blocked_getrf_double_batched_pivot": [P, P, I, I, I, I, I, I, I, P, P, P],
    # getrf/getrf_pivot_kernels.h
    "launch_apply_getrf_panel_pivots": [P, I, I, I, I, P, P],
    "launch_apply_getrf_panel_pivots_double": [P, I, I, I, I, P, P],
    "launch_apply_getrf_permutation": [P, I, I, P, P, I, P, I, P],
    "launch_apply_getrf_permutation_double": [P, I, I, P, P, I, P, I, P],
    # getrf/double_blocked_getrf_npvt.h
    "double_blocked_getrf_npvt": [P, P, I, I, I, I, P],
    "double_blocked_getrf_npvt_double": [P, P, I, I, I, I, P],
    # gemm/update_gemm.h
    "gemm_nt_update": [P, I, I, I, P, I, P, I, P, I],
    "gemm_nt_update_double": [P, I, I, I, P, I, P, I, P, I],
    "gemm_nt_update_m_tiled": [P, I, I, I, P, I, P, I, P, I, I],
    "gemm_nt_update_m_tiled_double": [P, I, I, I, P, I, P, I, P, I, I],
    "gemm_nn_update": [P, I, I, I, P, I, P, I, P, I],
    "gemm_nn_update_double": [P, I, I, I, P, I, P, I, P, I],
    # syrk/kblocked_syrk.h
    "syrk_lower_update": [P, I, I, P, I, P, I],
    "syrk_lower_update_double": [P, I, I, P, I, P, I],
    "syrk_lower_update_kblocked": [P, I, I, P, I, P, I, I],
    "syrk_lower_update_kblocked_double": [P, I, I, P, I, P, I, I],
    # trsm/*.h
    "trsm_right_lower_transpose": [P, I, I, P, I, P, I],
    "trsm_right_lower_transpose_double": [P, I, I, P, I, P, I],
    "double_blocked_trsm": [P, P, I, P, I, I, I, I, I],
    "double_blocked_trsm_double": [P, P, I, P, I, I, I, I, I],
    "trsm_left_lower_unit": [P, I, I, P, I, P, I],
    "trsm_left_lower_unit_double": [P, I, I, P, I, P, I],
    "trsm_left_lower_unit_iterative": [P, I, I, P, I, P, I, I],
    "trsm_left_lower_unit_iterative_double": [P, I, I, P, I, P, I, I],
    # kernels/panel_kernels.h
    "mifact_potrf_block": [P, I, P, I, P],
    "mifact_potrf_block_double": [P, I, P, I, P],
    "mifact_getrf_panel": [P, I, I, P, I, P, P],
    "mifact_getrf_panel_double": [P, I, I, P, I, P, P],
    "mifact_isamax": [P, I, P, I, P],
    "mifact_idamax": [P, I, P, I, P],
    # solve/mixed_refine.h
    "mifact_gemm_update_lowp": [P, I, I, I, I, I, I, ctypes.c_float, P, I, P, I, P, I],
    "double_blocked_potrf_mixed": [P, I, P, I, I, I, I, P],
    "double_blocked_getrf_mixed": [P, I, P, I, I, I, I, I, I, P, P],
    "mifact_potrs": [P, I, I, P, I, P, I],
    "mifact_getrs": [P, I, I, P, I, P, P, I],
    "mifact_residual_fp64": [P, I, I, I, P, I, P, I, P, I, P, I],
    "mifact_posv_refine": [P, I, I, P, I, P, I, P, I, P, I, P, P],
    "mifact_gesv_refine": [P, I, I, P, I, P, I, P, P, I, P, I, P, P],
    # batched/batched_factorizations.h
    "mifact_potrf_batched": [P, I, P, P, P, P, P],
    "mifact_getrf_batched": [P, I, P, P, P, P, P, P, P],
    "mifact_potrf_batched_ex": [P, I, I, P, P, P, P, P],
    "mifact_getrf_batched_ex": [P, I, I, P, P, P, P, P, P, P],
    "mifact_gesv_batched_refine": [P, I, I, I, P, P, P, P, P, P, P, P, P, P,
                                   ctypes.c_longlong, I, ctypes.c_double, P, P],
}

_RESTYPE = {
    "mifact_create_handle": I,
    "mifact_destroy_handle": I,
    "mifact_set_stream": I,
    "mifact_get_stream": I,
    "mifact_set_error_mode": I,
    "mifact_get_last_error": I,
    "mifact_reserve_workspace": I,
    "mifact_set_gemm_config": I,
    "mifact_set_gemm_policy": I,
    "mifact_set_batched_tiles": I,
    "mifact_set_wait_limit": I,
    "mifact_version": ctypes.c_char_p,
    "mifact_profile_begin": I,
    "mifact_profile_begin_stages": I,
    "mifact_profile_end": I,
    "mifact_posv_refine": I,
    "mifact_gesv_refine": I,
    "mifact_gesv_batched_refine": I,
}

# mifact_precision (include/solve/mixed_refine.h)
PRECISION = {"fp32": 0, "bf16": 1, "fp16": 2, "fp8": 3}
# mifact_refine_method
REFINE_METHOD = {"classical": 0, "gmres": 1, "auto": 2}

_lib = None


class LibraryNotBuilt(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Loads libmifact.so and declares every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LibraryNotBuilt(
            f"{path} is missing: build it with "
            "`python -m hygondoubleblockedfactorizations_amd.build`")
    lib = ctypes.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, None)
    _lib = lib
    return lib


class Handle:
    """Owns a mifact_handle; binds it to a HIP stream (torch's by default)."""

    def __init__(self, stream=None, error_mode_return=True):
        lib = load()
        self._lib = lib
        h = P()
        st = lib.mifact_create_handle(ctypes.byref(h))
        if st != 0:
            raise RuntimeError(f"mifact_create_handle failed with status {st}")
        self.ptr = h
        if error_mode_return:
            lib.mifact_set_error_mode(self.ptr, 1)
        if stream is not None:
            self.set_stream(stream)

    def set_stream(self, stream):
        raw = getattr(stream, "cuda_stream", stream)
        self._lib.mifact_set_stream(self.ptr, P(raw))

    def last_error(self):
        return self._lib.mifact_get_last_error(self.ptr)

    def reserve(self, n):
        st = self._lib.mifact_reserve_workspace(self.ptr, n)
        if st != 0:
            raise RuntimeError(f"workspace reservation failed: {st}")

    def close(self):
        if self.ptr:
            self._lib.mifact_destroy_handle(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class StageStats(ctypes.Structure):
    """Layout of mifact_stage_stats (include/mifact/handle.h)."""

    _fields_ = [
        ("name", ctypes.c_char * 32),
        ("calls", I),
        ("flops", ctypes.c_double),
        ("bytes", ctypes.c_double),
        ("ms", ctypes.c_double),
    ]


STAGES = ["gemm", "potf2", "trsm_leaf", "getf2_leaf", "row_moves", "perm", "gemm_lowp",
          "solve", "batched", "gemm_update", "gemm_update_256", "gemm_update_lower",
          "gemm_update_256_lower"]


def profile_begin(handle, stages=None):
    """Starts the stage profiler (all stages, or only the named ones)."""
    if stages is None:
        load().mifact_profile_begin(handle.ptr)
    else:
        mask = 0
        for s_ in stages:
            mask |= 1 << STAGES.index(s_)
        load().mifact_profile_begin_stages(handle.ptr, mask)


def profile_end(handle):
    """Synchronizes and returns {stage: dict(calls, flops, bytes, ms)}."""
    stats = (StageStats * 16)()
    count = I(0)
    st = load().mifact_profile_end(handle.ptr, ctypes.cast(stats, P), 16, ctypes.byref(count))
    if st != 0:
        raise RuntimeError(f"mifact_profile_end failed: {st}")
    return {stats[i].name.decode(): dict(calls=stats[i].calls, flops=stats[i].flops,
                                         bytes=stats[i].bytes, ms=stats[i].ms)
            for i in range(count.value)}


class BatchedPivotWorkspace(ctypes.Structure):
    """Layout of getrf_batched_pivot_workspace (include/getrf/double_blocked_getrf.h)."""

    _fields_ = [
        ("device_sources", P),
        ("device_targets", P),
        ("device_swap_buffer", P),
        ("host_pivots", P),
        ("host_permutation", P),
        ("host_sources", P),
        ("host_targets", P),
        ("capacity", I),
        ("column_tile", I),
    ]


def ptr(t):
    """Device/host address of a torch tensor or numpy array (None -> NULL)."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return P(t.data_ptr())
    return P(t.ctypes.data)


class RefineParams(ctypes.Structure):
    """Layout of mifact_refine_params (include/solve/mixed_refine.h)."""

    _fields_ = [
        ("precision", I),
        ("nb", I),
        ("b", I),
        ("outer_trsm_mode", I),
        ("trsm_block", I),
        ("max_iterations", I),
        ("tolerance", ctypes.c_double),
        ("fallback_fp32", I),
        ("method", I),
    ]


class RefineResult(ctypes.Structure):
    """Layout of mifact_refine_result (include/solve/mixed_refine.h)."""

    _fields_ = [
        ("info", I),
        ("iterations", I),
        ("converged", I),
        ("backward_error", ctypes.c_double),
        ("first_backward_error", ctypes.c_double),
        ("fell_back", I),
        ("factorizations", I),
        ("gmres_iterations", I),
        ("diagonal_shift", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}

"""CPU oracle for the double-blocked POTRF/GETRF path (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this module, and only as the checker.  The product path never
touches it.

Contents
  * ctypes bindings to ``oracle/liboracle.so`` (C restatement, see
    ``residual.c``, ``schedule_impl.h``, ``lapack.c``);
  * the reference's own residual built from its sources into
    ``oracle/_ref/libref_residual.so`` (when that was built);
  * seeded input generators for every benchmark family.

Residual semantics follow the reference's
``include/utils/benchmark/cpu_fp64_residual.h:145-201``.
"""
from __future__ import annotations

import ctypes
import glob
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref_residual.so")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="F_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_c_int = ctypes.c_int
_c_double = ctypes.c_double

_lib = None
_ref = None


def build():
    """Compile liboracle.so (and the reference residual when present)."""
    subprocess.run(["make", "-s", "-B", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/include"):
        subprocess.run(["make", "-s", "-B", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    for suf, fp in (("f32", _f32p), ("f64", _f64p)):
        f = getattr(L, f"oracle_potrf_residual_{suf}")
        f.restype = _c_double
        f.argtypes = [fp, fp, _c_int, _c_int, _c_int]
        f = getattr(L, f"oracle_getrf_residual_{suf}")
        f.restype = _c_double
        f.argtypes = [fp, fp, ctypes.c_void_p, _c_int, _c_int, _c_int, _c_int]
        f = getattr(L, f"oracle_double_blocked_potrf_{suf}")
        f.restype = None
        f.argtypes = [fp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                      ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)]
        f = getattr(L, f"oracle_double_blocked_getrf_{suf}")
        f.restype = _c_int
        f.argtypes = [fp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                      _c_int, _i32p, ctypes.POINTER(_c_int),
                      ctypes.POINTER(_c_int)]
        for name in ("gemm_nt", "gemm_nn"):
            f = getattr(L, f"oracle_{name}_{suf}")
            f.restype = None
            f.argtypes = [_c_int, _c_int, _c_int, fp, _c_int, fp, _c_int,
                          fp, _c_int]
        f = getattr(L, f"oracle_syrk_lower_{suf}")
        f.restype = None
        f.argtypes = [_c_int, _c_int, fp, _c_int, fp, _c_int, _c_int]
        for name in ("trsm_rlt", "trsm_llu"):
            f = getattr(L, f"oracle_{name}_{suf}")
            f.restype = None
            f.argtypes = [_c_int, _c_int, fp, _c_int, fp, _c_int]
        f = getattr(L, f"oracle_potf2_{suf}")
        f.restype = _c_int
        f.argtypes = [_c_int, fp, _c_int]
        f = getattr(L, f"oracle_getf2_{suf}")
        f.restype = _c_int
        f.argtypes = [_c_int, _c_int, fp, _c_int, _i32p]
        f = getattr(L, f"oracle_iamax_{suf}")
        f.restype = _c_int
        f.argtypes = [_c_int, np.ctypeslib.ndpointer(
            dtype=np.float32 if suf == "f32" else np.float64)]
        f = getattr(L, f"oracle_apply_panel_pivots_{suf}")
        f.restype = None
        f.argtypes = [fp, _c_int, _c_int, _c_int, _c_int, _i32p]
    L.oracle_ipiv_to_permutation.restype = _c_int
    L.oracle_ipiv_to_permutation.argtypes = [_i32p, _c_int, _i32p]
    L.oracle_lapack_open.restype = _c_int
    L.oracle_lapack_open.argtypes = [ctypes.c_char_p]
    L.oracle_lapack_path.restype = ctypes.c_char_p
    L.oracle_lapack_config.restype = ctypes.c_char_p
    L.oracle_lapack_set_threads.argtypes = [_c_int]
    L.oracle_lapack_get_threads.restype = _c_int
    L.oracle_spotrf.argtypes = [_c_int, _f32p, _c_int]
    L.oracle_dpotrf.argtypes = [_c_int, _f64p, _c_int]
    L.oracle_sgetrf.argtypes = [_c_int, _c_int, _f32p, _c_int, _i32p]
    L.oracle_dgetrf.argtypes = [_c_int, _c_int, _f64p, _c_int, _i32p]
    L.oracle_dpotrs.argtypes = [_c_int, _c_int, _f64p, _c_int, _f64p, _c_int]
    L.oracle_dgetrs.argtypes = [_c_int, _c_int, _f64p, _c_int, _i32p, _f64p,
                                _c_int]
    L.oracle_residual_use_blas.argtypes = [_c_int]
    L.oracle_residual_use_blas.restype = _c_int
    _lib = L
    return L


def ref_lib():
    """The reference's own residual (compiled from its sources), or None."""
    global _ref
    if _ref is not None:
        return _ref
    if not os.path.exists(REF_LIB_PATH):
        return None
    R = ctypes.CDLL(REF_LIB_PATH)
    for suf, fp in (("f32", _f32p), ("f64", _f64p)):
        f = getattr(R, f"ref_potrf_residual_{suf}")
        f.restype = _c_double
        f.argtypes = [fp, fp, _c_int, _c_int, _c_int]
        f = getattr(R, f"ref_getrf_residual_{suf}")
        f.restype = _c_double
        f.argtypes = [fp, fp, ctypes.c_void_p, _c_int, _c_int, _c_int, _c_int]
    _ref = R
    return R


def _suffix(a):
    if a.dtype == np.float32:
        return "f32"
    if a.dtype == np.float64:
        return "f64"
    raise TypeError(f"unsupported dtype {a.dtype}")


class InvalidArgument(ValueError):
    """Raised where the reference throws std::invalid_argument."""


def _check(r):
    if r < 0:
        raise InvalidArgument("invalid residual inputs")
    return r


# ------------------------------------------------------------------ residual

def potrf_residual(L, A, n=None, lda=None, tile=512, use_ref=False):
    """||A - L L^T||_F / ||A||_F with FP64 products (lower triangle of L)."""
    L = np.asfortranarray(L)
    A = np.asfortranarray(A)
    n = L.shape[1] if n is None else n
    lda = L.shape[0] if lda is None else lda
    suf = _suffix(L)
    mod = ref_lib() if use_ref else lib()
    name = ("ref_" if use_ref else "oracle_") + f"potrf_residual_{suf}"
    return _check(getattr(mod, name)(L, A, n, lda, tile))


def getrf_residual(LU, A, ipiv=None, n=None, lda=None, tile=512,
                   pivoting=True, use_ref=False):
    """||P A - L U||_F / ||A||_F with FP64 products; ipiv 1-based global."""
    LU = np.asfortranarray(LU)
    A = np.asfortranarray(A)
    n = LU.shape[1] if n is None else n
    lda = LU.shape[0] if lda is None else lda
    suf = _suffix(LU)
    piv = None
    if ipiv is not None:
        piv = np.ascontiguousarray(ipiv, dtype=np.int32)
    ptr = piv.ctypes.data_as(ctypes.c_void_p) if piv is not None else None
    mod = ref_lib() if use_ref else lib()
    name = ("ref_" if use_ref else "oracle_") + f"getrf_residual_{suf}"
    return _check(getattr(mod, name)(LU, A, ptr, n, lda, int(pivoting), tile))


def ipiv_to_permutation(ipiv):
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    perm = np.empty_like(ipiv)
    ok = lib().oracle_ipiv_to_permutation(ipiv, len(ipiv), perm)
    return perm if ok else None


def use_blas_for_residual(enable=True):
    """Back the residual's FP64 product with the LAPACK provider's dgemm."""
    lapack_open()
    return lib().oracle_residual_use_blas(int(enable)) == 0


# ------------------------------------------------------------------ schedule

def double_blocked_potrf(A, nb, b, use_dbtrsm=False, syrk_k_tile=0):
    """CPU restatement of double_blocked_potrf_impl (in place).

    Returns (info_reference_semantics, info_first_global)."""
    assert A.flags.f_contiguous
    lda, n = A.shape
    ir, ig = _c_int(0), _c_int(0)
    getattr(lib(), f"oracle_double_blocked_potrf_{_suffix(A)}")(
        A, lda, n, nb, b, int(use_dbtrsm), syrk_k_tile,
        ctypes.byref(ir), ctypes.byref(ig))
    return ir.value, ig.value


def double_blocked_getrf(A, outer_nb, inner_nb, mode=1, trsm_block=0,
                         batched=False):
    """CPU restatement of the double-blocked GETRF (in place).

    Returns (ipiv, info_reference_semantics, info_first_global)."""
    assert A.flags.f_contiguous
    lda, n = A.shape
    ipiv = np.zeros(n, dtype=np.int32)
    ir, ig = _c_int(0), _c_int(0)
    getattr(lib(), f"oracle_double_blocked_getrf_{_suffix(A)}")(
        A, lda, n, outer_nb, inner_nb, mode, trsm_block, int(batched), ipiv,
        ctypes.byref(ir), ctypes.byref(ig))
    return ipiv, ir.value, ig.value


def gemm_nt(C, A, B, rows, cols, k):
    getattr(lib(), f"oracle_gemm_nt_{_suffix(C)}")(
        rows, cols, k, A, A.shape[0], B, B.shape[0], C, C.shape[0])


def gemm_nn(C, A, B, rows, cols, k):
    getattr(lib(), f"oracle_gemm_nn_{_suffix(C)}")(
        rows, cols, k, A, A.shape[0], B, B.shape[0], C, C.shape[0])


def syrk_lower(C, A, n, k, k_tile=0):
    getattr(lib(), f"oracle_syrk_lower_{_suffix(C)}")(
        n, k, A, A.shape[0], C, C.shape[0], k_tile)


def trsm_rlt(B, L, rows, cols):
    getattr(lib(), f"oracle_trsm_rlt_{_suffix(B)}")(
        rows, cols, L, L.shape[0], B, B.shape[0])


def trsm_llu(B, L, rows, cols):
    getattr(lib(), f"oracle_trsm_llu_{_suffix(B)}")(
        rows, cols, L, L.shape[0], B, B.shape[0])


def potf2(A, n=None):
    n = A.shape[1] if n is None else n
    return getattr(lib(), f"oracle_potf2_{_suffix(A)}")(n, A, A.shape[0])


def getf2(A, m=None, n=None):
    m = A.shape[0] if m is None else m
    n = A.shape[1] if n is None else n
    ipiv = np.zeros(min(m, n), dtype=np.int32)
    info = getattr(lib(), f"oracle_getf2_{_suffix(A)}")(m, n, A, A.shape[0], ipiv)
    return ipiv, info


def iamax(x):
    x = np.ascontiguousarray(x)
    return getattr(lib(), f"oracle_iamax_{_suffix(x)}")(len(x), x)


def apply_panel_pivots(A, panel, b, ipiv_local):
    """Reference laswp kernel semantics; returns the globalized pivots."""
    piv = np.ascontiguousarray(ipiv_local, dtype=np.int32).copy()
    getattr(lib(), f"oracle_apply_panel_pivots_{_suffix(A)}")(
        A, A.shape[0], A.shape[1], panel, b, piv)
    return piv


# ------------------------------------------------------------------ LAPACK

def _scipy_openblas():
    try:
        import scipy
    except ImportError:
        return ""
    d = os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs")
    hits = sorted(glob.glob(os.path.join(d, "libscipy_openblas*.so")))
    return os.path.abspath(hits[0]) if hits else ""


def lapack_open():
    L = lib()
    if L.oracle_lapack_open(_scipy_openblas().encode()) != 0:
        raise RuntimeError("no LAPACK provider found")
    return L.oracle_lapack_path().decode(), L.oracle_lapack_config().decode()


def lapack_set_threads(threads):
    lapack_open()
    return lib().oracle_lapack_set_threads(int(threads))


def spotrf(A):
    lapack_open()
    return lib().oracle_spotrf(A.shape[1], A, A.shape[0])


def dpotrf(A):
    lapack_open()
    return lib().oracle_dpotrf(A.shape[1], A, A.shape[0])


def sgetrf(A):
    lapack_open()
    ipiv = np.zeros(min(A.shape), dtype=np.int32)
    info = lib().oracle_sgetrf(A.shape[0], A.shape[1], A, A.shape[0], ipiv)
    return ipiv, info


def dgetrf(A):
    lapack_open()
    ipiv = np.zeros(min(A.shape), dtype=np.int32)
    info = lib().oracle_dgetrf(A.shape[0], A.shape[1], A, A.shape[0], ipiv)
    return ipiv, info


def dpotrs(L, B):
    lapack_open()
    return lib().oracle_dpotrs(L.shape[1], B.shape[1], L, L.shape[0], B, B.shape[0])


def dgetrs(LU, ipiv, B):
    lapack_open()
    return lib().oracle_dgetrs(LU.shape[1], B.shape[1], LU, LU.shape[0],
                               np.ascontiguousarray(ipiv, dtype=np.int32),
                               B, B.shape[0])


# ------------------------------------------------------------------ inputs

def rng(seed):
    """Seeded generator: numpy PCG64."""
    return np.random.Generator(np.random.PCG64(seed))


def spd_matrix(n, seed, lda=None, dtype=np.float32):
    """A = G G^T / n + I, G iid N(0,1) (FP64 product, rounded once, exactly
    symmetric), column-major n x n inside an lda x n array."""
    lda = n if lda is None else lda
    g = rng(seed).standard_normal((n, n))
    a = g @ g.T / n + np.eye(n)
    a = 0.5 * (a + a.T)
    out = np.zeros((lda, n), dtype=dtype, order="F")
    out[:n, :] = a.astype(dtype)
    return out


def uniform_matrix(n, seed, lda=None, lo=-1.0, hi=1.0, dtype=np.float32, m=None):
    """iid Uniform(lo, hi) entries, column-major (m x n in an lda x n array)."""
    m = n if m is None else m
    lda = m if lda is None else lda
    a = rng(seed).uniform(lo, hi, size=(m, n))
    out = np.zeros((lda, n), dtype=dtype, order="F")
    out[:m, :] = a.astype(dtype)
    return out


def spd_condition_matrix(n, kappa, seed, dtype=np.float32):
    """A = Q diag(sigma) Q^T, Q from the Householder QR of a seeded Gaussian,
    sigma_i = kappa^(-i/(n-1)); exactly symmetric after rounding."""
    q, _ = np.linalg.qr(rng(seed).standard_normal((n, n)))
    sig = kappa ** (-np.arange(n) / max(n - 1, 1))
    a = (q * sig) @ q.T
    a = 0.5 * (a + a.T)
    return np.asfortranarray(a.astype(dtype))


def general_condition_matrix(n, kappa, seed, dtype=np.float32):
    """A = U diag(sigma) V^T with seeded orthogonal U, V."""
    r = rng(seed)
    u, _ = np.linalg.qr(r.standard_normal((n, n)))
    v, _ = np.linalg.qr(r.standard_normal((n, n)))
    sig = kappa ** (-np.arange(n) / max(n - 1, 1))
    return np.asfortranarray(((u * sig) @ v.T).astype(dtype))

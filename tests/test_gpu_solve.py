"""GPU parity of the low-precision update GEMM, the triangular solves, the
FP64 residual and the mixed-precision refinement solvers through the C-ABI.

Bar (written per test):
  * BF16/FP16 GEMM: exact on operands whose products and sums are exact in
    FP32 (small integers); on random operands within (k + 2) eps32 |A'||B'|
    of the FP64 product of the operands rounded to BF16/FP16 (A', B');
  * FP32 solves: normwise backward error <= 10 n eps32 and the solution
    within 1e-3 of LAPACK dpotrs/dgetrs on the same FP32 factors;
  * FP64 residual: within 1e-13 (||A|| ||x|| + ||b||) of numpy's FP64 product;
  * refinement: converged, backward error <= the requested tolerance, and
    the solution within kappa * 1e-12 of the FP64 LAPACK solution
    (dpotrf/dpotrs, dgetrf/dgetrs).
"""
import ctypes

import numpy as np
import pytest

from gpu_util import EPS32, Session, dev, dev_int, host, log, sync
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def H():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return Session.get()


def lib():
    from hygondoubleblockedfactorizations_amd import _lib
    return _lib.load()


def P(t):
    from hygondoubleblockedfactorizations_amd._lib import ptr
    return ptr(t)


def round_bf16(x):
    """Round-to-nearest-even FP32 -> BF16, returned as FP32 values."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16 << 16
    return b.astype(np.uint32).view(np.float32)


def round_fp8_tiles(x):
    """The FP8 pack of a row-by-k operand (rows x k, k the contraction index):
    per 32 x 32 tile, e = floor(log2 max|x|) - 7 (the tile's largest
    magnitude scaled into [128, 256)), q = e4m3(x 2^-e) (torch's
    float8_e4m3fn rounding: to nearest, ties to even), returned as q 2^e."""
    import torch
    x = np.asarray(x, np.float64)
    out = np.zeros_like(x)
    for r0 in range(0, x.shape[0], 32):
        for k0 in range(0, x.shape[1], 32):
            t = x[r0:r0 + 32, k0:k0 + 32]
            m = np.abs(t).max()
            e = 0 if m == 0 else int(np.frexp(m)[1]) - 1 - 7
            q = torch.from_numpy(t * 2.0 ** -e).to(torch.float8_e4m3fn).to(torch.float64).numpy()
            out[r0:r0 + 32, k0:k0 + 32] = q * 2.0 ** e
    return out


def round_lowp(x, prec, kaxis=1):
    """x rounded as the GEMM packs it; kaxis: the contraction axis of x."""
    if prec == "fp8":
        return round_fp8_tiles(x) if kaxis == 1 else round_fp8_tiles(np.asarray(x).T).T
    return round_bf16(x) if prec == "bf16" else np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


PREC = {"fp32": 0, "bf16": 1, "fp16": 2, "fp8": 3}


def emulated_mixed_lu(A, prec, nb=128):
    """CPU model of the mixed-precision LU's rounding: nb-column panels
    factored in FP32 with partial pivoting (LAPACK order), the panel's
    interchanges applied to every other column, U12 = L11^{-1} A12, and the
    Schur update A22 -= L21 U12 from operands rounded as the update GEMM packs
    them (round_lowp), products accumulated exactly and stored to FP32.
    Returns (LU, ipiv 1-based)."""
    a = np.array(A, np.float32, order="F")
    n = a.shape[1]
    tiny = np.finfo(np.float32).tiny
    ipiv = np.zeros(n, np.int32)
    for k in range(0, n, nb):
        kb = min(nb, n - k)
        for j in range(k, k + kb):
            p = j + int(np.argmax(np.abs(a[j:, j])))
            ipiv[j] = p + 1
            if p != j:
                a[[j, p], :] = a[[p, j], :]
            piv = a[j, j]
            if piv != 0:
                col = a[j + 1:, j]
                a[j + 1:, j] = col * np.float32(1.0 / piv) if abs(piv) >= tiny else col / piv
                a[j + 1:, j + 1:k + kb] -= np.outer(a[j + 1:, j], a[j, j + 1:k + kb]).astype(np.float32)
        c1 = k + kb
        if c1 < n:
            L11 = np.tril(a[k:c1, k:c1], -1).astype(np.float64) + np.eye(kb)
            a[k:c1, c1:] = np.linalg.solve(L11, a[k:c1, c1:].astype(np.float64)).astype(np.float32)
            L21 = np.asarray(round_lowp(a[c1:, k:c1], prec, kaxis=1), np.float64)
            U12 = np.asarray(round_lowp(a[k:c1, c1:], prec, kaxis=0), np.float64)
            a[c1:, c1:] = (a[c1:, c1:].astype(np.float64) - L21 @ U12).astype(np.float32)
    return a, ipiv


def tile_pow2(rows, cols, seed):
    """Powers of two constant on 32 x 32 tiles (2^-2 .. 2^2): with them the
    FP8 tiles carry different scales (products stay multiples of 2^-4 below
    2^8, so sums of up to 4096 of them are exact in FP32)."""
    r = O.rng(seed)
    e = r.integers(-2, 3, size=((rows + 31) // 32, (cols + 31) // 32))
    return np.kron(2.0 ** e, np.ones((32, 32)))[:rows, :cols].astype(np.float32)


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp8"])
@pytest.mark.parametrize("op", ["nt", "nn"])
@pytest.mark.parametrize("m,n,k,lower", [(256, 256, 64, 0), (300, 130, 77, 0),
                                         (384, 256, 160, 1), (97, 33, 5, 0),
                                         (1100, 600, 300, 0), (1300, 1300, 260, 1),
                                         (4200, 4100, 96, 0), (4352, 4352, 136, 1)])
def test_gemm_lowp_exact(H, prec, op, m, n, k, lower):
    """Small integers: every product and partial sum is exact in FP32, so the
    MFMA result must equal the integer product bit for bit (pins the
    32x32x16 / scaled 32x32x64 operand and accumulator lane layouts, the
    packing, both tile sizes, and for FP8 the per-tile scale: the operands
    are small integers times a power of two per 32 x 32 tile, all exact in
    e4m3 after the tile's scaling)."""
    log("gemm_lowp exact", prec, op, m, n, k, lower)
    r = O.rng(7)
    lda = m + 3
    A = np.zeros((lda, k), np.float32, order="F")
    A[:m] = r.integers(-4, 5, size=(m, k))
    if prec == "fp8":
        A[:m] *= tile_pow2(m, k, 1)
    if op == "nt":
        ldb = n + 1
        B = np.zeros((ldb, k), np.float32, order="F")
        B[:n] = r.integers(-4, 5, size=(n, k))
        if prec == "fp8":
            B[:n] *= tile_pow2(n, k, 2)
        prod = A[:m].astype(np.float64) @ B[:n].astype(np.float64).T
    else:
        ldb = k + 2
        B = np.zeros((ldb, n), np.float32, order="F")
        B[:k] = r.integers(-4, 5, size=(k, n))
        if prec == "fp8":
            B[:k] *= tile_pow2(n, k, 2).T
        prod = A[:m].astype(np.float64) @ B[:k].astype(np.float64)
    ldc = m + 5
    C = np.zeros((ldc, n), np.float32, order="F")
    C[:m] = r.integers(-50, 51, size=(m, n))
    dA, dB, dC = dev(A), dev(B), dev(C)
    lib().mifact_gemm_update_lowp(H.ptr, PREC[prec], 1 if op == "nt" else 0, lower, m, n, k,
                                  -1.0, P(dA), lda, P(dB), ldb, P(dC), ldc)
    sync()
    assert H.last_error() == 0
    out = host(dC)
    want = C.astype(np.float64).copy()
    upd = want[:m] - prod
    if lower:
        mask = np.tril(np.ones((m, n), bool))
        want[:m][mask] = upd[mask]
    else:
        want[:m] = upd
    np.testing.assert_array_equal(out, want.astype(np.float32))


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp8"])
@pytest.mark.parametrize("op", ["nt", "nn"])
def test_gemm_lowp_random(H, prec, op):
    m, n, k = 1000, 520, 700
    log("gemm_lowp random", prec, op, m, n, k)
    r = O.rng(11)
    A = np.asfortranarray(r.uniform(-1, 1, (m, k)).astype(np.float32))
    B = np.asfortranarray(r.uniform(-1, 1, (n, k) if op == "nt" else (k, n)).astype(np.float32))
    C = np.asfortranarray(r.uniform(-1, 1, (m, n)).astype(np.float32))
    dA, dB, dC = dev(A), dev(B), dev(C)
    lib().mifact_gemm_update_lowp(H.ptr, PREC[prec], 1 if op == "nt" else 0, 0, m, n, k, -1.0,
                                  P(dA), m, P(dB), B.shape[0], P(dC), m)
    sync()
    out = host(dC).astype(np.float64)
    Al = round_lowp(A, prec).astype(np.float64)
    Bl = round_lowp(B, prec, kaxis=1 if op == "nt" else 0).astype(np.float64)
    Bop = Bl.T if op == "nt" else Bl
    want = C - Al @ Bop
    bound = (k + 2) * EPS32 * (np.abs(Al) @ np.abs(Bop) + np.abs(C))
    assert np.all(np.abs(out - want) <= bound)


def backward_error(A, x, b):
    r = b - A @ x
    return (np.abs(r).max(axis=0)
            / (np.abs(A).sum(axis=1).max() * np.abs(x).max(axis=0) + np.abs(b).max(axis=0))).max()


@pytest.mark.parametrize("n,nrhs", [(1000, 1), (2048, 3), (64, 2), (130, 17)])
def test_potrs_fp32(H, n, nrhs):
    log("potrs", n, nrhs)
    A = O.spd_matrix(n, 5)
    dA, info = dev(A), dev_int(1, 7)
    L = lib()
    L.double_blocked_potrf(H.ptr, P(dA), n, n, 512, 128, P(info))
    b = np.asfortranarray(O.rng(6).standard_normal((n, nrhs)).astype(np.float32))
    db = dev(b)
    L.mifact_potrs(H.ptr, n, nrhs, P(dA), n, P(db), n)
    sync()
    assert int(info.cpu()[0]) == 0 and H.last_error() == 0
    x = host(db).astype(np.float64)
    A64 = A.astype(np.float64)
    assert backward_error(A64, x, b.astype(np.float64)) <= 10 * n * EPS32
    Lf = np.asfortranarray(np.tril(host(dA)).astype(np.float64))
    xr = np.asfortranarray(b.astype(np.float64))
    assert O.dpotrs(Lf, xr) == 0
    assert np.abs(x - xr).max() <= 1e-3 * np.abs(xr).max()


@pytest.mark.parametrize("n,nrhs", [(1000, 1), (2048, 4), (96, 17)])
def test_getrs_fp32(H, n, nrhs):
    log("getrs", n, nrhs)
    A = O.uniform_matrix(n, 9)
    A[np.arange(n), np.arange(n)] += 4.0 * np.sqrt(n)   # well conditioned
    dA, ipiv, info = dev(A), dev_int(n), dev_int(1, 7)
    L = lib()
    L.double_blocked_getrf(H.ptr, P(dA), n, n, 512, 128, 0, 0, P(ipiv), P(info))
    b = np.asfortranarray(O.rng(10).standard_normal((n, nrhs)).astype(np.float32))
    db = dev(b)
    L.mifact_getrs(H.ptr, n, nrhs, P(dA), n, P(ipiv), P(db), n)
    sync()
    assert int(info.cpu()[0]) == 0 and H.last_error() == 0
    x = host(db).astype(np.float64)
    assert backward_error(A.astype(np.float64), x, b.astype(np.float64)) <= 10 * n * EPS32
    LU = np.asfortranarray(host(dA).astype(np.float64))
    xr = np.asfortranarray(b.astype(np.float64))
    assert O.dgetrs(LU, ipiv.cpu().numpy(), xr) == 0
    assert np.abs(x - xr).max() <= 1e-3 * np.abs(xr).max()


def test_getrs_pivots_general_permutation(H):
    """Long pivot chains through the LDS permutation kernel: A = a row
    permutation of a dominant diagonal, so sgetrf's ipiv is a chain of
    non-trivial exchanges."""
    n = 3000
    r = O.rng(12)
    perm = r.permutation(n)
    A = np.zeros((n, n), np.float32, order="F")
    A[perm, np.arange(n)] = r.uniform(1.0, 2.0, n).astype(np.float32)
    A += np.asfortranarray(r.uniform(-1e-3, 1e-3, (n, n)).astype(np.float32))
    dA, ipiv, info = dev(A), dev_int(n), dev_int(1, 7)
    L = lib()
    L.double_blocked_getrf(H.ptr, P(dA), n, n, 1024, 256, 0, 0, P(ipiv), P(info))
    b = np.asfortranarray(r.standard_normal((n, 2)).astype(np.float32))
    db = dev(b)
    L.mifact_getrs(H.ptr, n, 2, P(dA), n, P(ipiv), P(db), n)
    sync()
    x = host(db).astype(np.float64)
    assert backward_error(A.astype(np.float64), x, b.astype(np.float64)) <= 10 * n * EPS32


@pytest.mark.parametrize("sym", [0, 1])
@pytest.mark.parametrize("n,nrhs", [(1000, 1), (777, 5), (2048, 18)])
def test_residual_fp64(H, sym, n, nrhs):
    log("residual", sym, n, nrhs)
    import torch
    r = O.rng(13)
    if sym:
        A = O.spd_matrix(n, 14)
        Ad = A.copy(order="F")
        Ad[np.triu_indices(n, 1)] = np.nan          # upper triangle must not be read
    else:
        A = O.uniform_matrix(n, 14)
        Ad = A
    x = np.asfortranarray(r.standard_normal((n, nrhs)))
    b = np.asfortranarray(r.standard_normal((n, nrhs)))
    dA = dev(Ad)
    dx = torch.from_numpy(np.ascontiguousarray(x.T)).cuda()
    db = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
    dr = torch.zeros_like(db)
    lib().mifact_residual_fp64(H.ptr, sym, n, nrhs, P(dA), n, P(dx), n, P(db), n, P(dr), n)
    sync()
    got = dr.cpu().numpy().T
    A64 = A.astype(np.float64)
    want = b - A64 @ x
    scale = np.abs(A64).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max()
    assert np.abs(got - want).max() <= 1e-13 * scale


METHOD = {"classical": 0, "gmres": 1, "auto": 2}


def refine(H, sym, A, b, prec, nb, bb, tol=0.0, max_it=30, fallback=0, method="classical"):
    import torch
    from hygondoubleblockedfactorizations_amd import _lib
    n = A.shape[1]
    nrhs = b.shape[1]
    dA = dev(A)
    F = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    db = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
    dx = torch.zeros_like(db)
    prm = _lib.RefineParams(PREC[prec], nb, bb, 0, 0, max_it, tol, fallback, METHOD[method])
    res = _lib.RefineResult()
    L = lib()
    if sym:
        st = L.mifact_posv_refine(H.ptr, n, nrhs, P(dA), n, P(F), n, P(db), n, P(dx), n,
                                  ctypes.byref(prm), ctypes.byref(res))
    else:
        ipiv = dev_int(n)
        st = L.mifact_gesv_refine(H.ptr, n, nrhs, P(dA), n, P(F), n, P(ipiv), P(db), n, P(dx),
                                  n, ctypes.byref(prm), ctypes.byref(res))
    sync()
    assert st == 0
    return dx.cpu().numpy().T.copy(order="F"), res


# Classical refinement converges when kappa * (factorization backward error)
# stays well below one: every kappa here for FP32 updates, kappa <= 1e3 for
# FP16 (u = 2^-11) and kappa <= 1e2 for BF16 (u = 2^-8).  Beyond that
# GMRES-IR ("auto": classical steps while they contract, then FGMRES).
@pytest.mark.parametrize("prec,n,nrhs,kappa,method", [
    ("fp32", 1024, 1, 1e2, "classical"), ("fp32", 2048, 16, 1e4, "classical"),
    ("fp32", 1500, 3, 1e6, "classical"),
    ("fp16", 1024, 1, 1e2, "classical"), ("fp16", 1500, 3, 1e3, "classical"),
    ("fp16", 2048, 16, 1e4, "auto"), ("fp16", 1500, 2, 1e6, "auto"),
    ("bf16", 1024, 1, 1e2, "classical"), ("bf16", 2048, 4, 1e4, "auto"),
    ("bf16", 1024, 1, 1e4, "gmres"), ("fp8", 1024, 1, 1e2, "auto")])
def test_posv_refine(H, prec, n, nrhs, kappa, method):
    log("posv_refine", prec, n, nrhs, kappa, method)
    A = O.spd_condition_matrix(n, kappa, 21)
    A *= 37.0                                    # exercise the equilibration
    xt = O.rng(22).standard_normal((n, nrhs))
    b = np.asfortranarray(A.astype(np.float64) @ xt)
    x, res = refine(H, True, A, b, prec, 512, 128, method=method,
                    fallback=1 if prec == "fp8" else 0)
    log("  ->", res.as_dict())
    assert res.info == 0 and res.converged == 1
    tol = np.sqrt(n) * np.finfo(np.float64).eps
    assert res.backward_error <= tol
    A64 = A.astype(np.float64)
    assert backward_error(A64, x, b) <= 2 * tol
    Lr = np.asfortranarray(A64.copy())
    assert O.dpotrf(Lr) == 0
    xr = np.asfortranarray(b.copy())
    assert O.dpotrs(Lr, xr) == 0
    assert np.abs(x - xr).max() <= kappa * 1e-12 * np.abs(xr).max()


# Uniform(-1, 1) matrices have kappa ~ n: classical refinement converges
# with FP32 updates; FP16 updates need GMRES-IR there; BF16-preconditioned
# FGMRES stalls from n ~ 2048 on, where the FP32 fallback finishes the solve.
@pytest.mark.parametrize("prec,method", [("fp32", "classical"), ("fp16", "auto"),
                                         ("bf16", "auto"), ("fp16", "gmres"), ("fp8", "auto")])
@pytest.mark.parametrize("n,nrhs", [(1024, 1), (2048, 2), (1111, 17)])
def test_gesv_refine_uniform(H, prec, method, n, nrhs):
    log("gesv_refine", prec, method, n, nrhs)
    A = O.uniform_matrix(n, 31)
    A[:, 5] *= 1000.0                            # badly scaled column
    xt = O.rng(32).standard_normal((n, nrhs))
    b = np.asfortranarray(A.astype(np.float64) @ xt)
    tol = np.sqrt(n) * np.finfo(np.float64).eps
    x, res = refine(H, False, A, b, prec, 512, 128, tol=tol, method=method,
                    fallback=1 if prec in ("bf16", "fp8") else 0)
    log("  ->", res.as_dict())
    assert res.info == 0 and res.converged == 1
    A64 = A.astype(np.float64)
    assert backward_error(A64, x, b) <= 2 * tol
    LU = np.asfortranarray(A64.copy())
    ipiv, info = O.dgetrf(LU)
    assert info == 0
    xr = np.asfortranarray(b.copy())
    assert O.dgetrs(LU, ipiv, xr) == 0
    kappa = np.linalg.cond(A64, 1)
    assert np.abs(x - xr).max() <= kappa * 10 * tol * np.abs(xr).max()


@pytest.mark.parametrize("sym", [True, False])
def test_refine_fallback(H, sym):
    """BF16 updates at kappa = 1e5 stagnate; with fallback_fp32 the solver
    refactors with FP32 updates and converges."""
    n = 1024
    if sym:
        A = O.spd_condition_matrix(n, 1e5, 71)
    else:
        A = O.general_condition_matrix(n, 1e5, 71)
    xt = O.rng(72).standard_normal((n, 2))
    b = np.asfortranarray(A.astype(np.float64) @ xt)
    x, res = refine(H, sym, A, b, "bf16", 512, 128, fallback=1, method="classical")
    log("  ->", res.as_dict())
    assert res.info == 0 and res.converged == 1
    # one factorization, plus a shifted retry when the BF16 Cholesky broke
    # down, plus the FP32 refactorization when the refinement stalled
    assert 1 + res.fell_back <= res.factorizations <= 2 + res.fell_back
    assert backward_error(A.astype(np.float64), x, b) <= 2 * np.sqrt(n) * np.finfo(np.float64).eps


def test_gesv_refine_conditioned(H):
    n = 1536
    A = O.general_condition_matrix(n, 1e5, 41)
    xt = O.rng(42).standard_normal((n, 1))
    b = np.asfortranarray(A.astype(np.float64) @ xt)
    # FP16-preconditioned FGMRES contracts slowly at kappa = 1e5; the FP32
    # fallback guarantees the solve.
    x, res = refine(H, False, A, b, "fp16", 512, 128, method="auto", fallback=1)
    log("  ->", res.as_dict())
    assert res.info == 0 and res.converged == 1
    assert backward_error(A.astype(np.float64), x, b) <= 2 * np.sqrt(n) * np.finfo(np.float64).eps


def test_refine_not_spd_reports_info(H):
    n = 512
    A = O.spd_matrix(n, 3)
    A[300, 300] = -5.0
    b = np.ones((n, 1))
    _, res = refine(H, True, A, b, "fp16", 256, 64)
    assert res.info == 301 and res.converged == 0


def test_refine_tolerance_and_limit(H):
    """A loose tolerance stops after the first solve; max_iterations bounds
    the refinement steps."""
    n = 1024
    A = O.uniform_matrix(n, 51)
    b = np.asfortranarray(O.rng(52).standard_normal((n, 1)))
    _, res = refine(H, False, A, b, "bf16", 512, 128, tol=1.0)
    assert res.converged == 1 and res.iterations == 0
    _, res = refine(H, False, A, b, "bf16", 512, 128, tol=1e-300, max_it=2)
    assert res.converged == 0 and res.iterations <= 2


def _growth(F, lower_unit):
    """|| |L| |U| ||_F for the factors packed in F (L unit lower when
    lower_unit, else L = tril(F) and U = L^T), FP64."""
    F = F.astype(np.float64)
    if lower_unit:
        Lf = np.tril(F, -1) + np.eye(F.shape[0])
        Uf = np.triu(F)
    else:
        Lf = np.tril(F)
        Uf = Lf.T
    return float(np.linalg.norm(np.abs(Lf) @ np.abs(Uf)))


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp8"])
def test_mixed_factorizations_residual(H, prec):
    """The mixed drivers factor with a backward error set by the update
    precision.  The trailing updates round both operands to the 16-bit or
    e4m3 format (unit roundoff u = 2^-8, 2^-11, 2^-4; FP8 after the per-tile
    power-of-two scaling, which is exact), so each entry of the computed product carries an error
    of at most about 2u sum_k |l_ik||u_kj|; the rounding errors are independent
    and add like a random walk, so the normwise residual stays below
    u || |L||U| ||_F / ||A||_F, computed here from the returned factors (the
    growth-aware form of the LU backward-error bound |dA| <= g_n |L||U|).  That
    bound is tight for Cholesky (|L||L^T| ~ |A|) and checked there; for LU
    |L||U| grows with n, so the GPU residual is held to 3x the residual of a
    CPU model of the same rounding (emulated_mixed_lu).  A dropped or
    transposed update term gives an O(1) residual, far above either."""
    n = 2048
    u = {"bf16": 2.0 ** -8, "fp16": 2.0 ** -11, "fp8": 2.0 ** -4}[prec]
    L = lib()
    A = O.spd_matrix(n, 61)
    dA, info = dev(A), dev_int(1, 7)
    L.double_blocked_potrf_mixed(H.ptr, PREC[prec], P(dA), n, n, 512, 128, P(info))
    sync()
    assert int(info.cpu()[0]) == 0
    F = host(dA)
    res = O.potrf_residual(F, A)
    bound = u * _growth(F, False) / np.linalg.norm(A.astype(np.float64))
    log("potrf mixed", prec, res, "bound", bound)
    assert res <= bound
    A = O.uniform_matrix(n, 62)
    if prec == "fp8":
        # e4m3 (u = 2^-4) leaves an O(1) residual on a plain uniform matrix at
        # this order (the CPU model gives 2.7); a sqrt(n) diagonal shift
        # (kappa ~ 10) keeps the comparison meaningful
        A[np.arange(n), np.arange(n)] += np.float32(np.sqrt(n))
    dA, ipiv = dev(A), dev_int(n)
    L.double_blocked_getrf_mixed(H.ptr, PREC[prec], P(dA), n, n, 512, 128, 0, 0, P(ipiv), P(info))
    sync()
    assert int(info.cpu()[0]) == 0
    F = host(dA)
    res = O.getrf_residual(F, A, ipiv.cpu().numpy())
    # LU: |L||U| grows with n for general matrices, so the growth-aware bound
    # says little; the residual is held to 3x that of the CPU model of the
    # same rounding (emulated_mixed_lu) instead
    E, pe = emulated_mixed_lu(A, prec)
    emu = O.getrf_residual(np.asfortranarray(E), A, pe)
    log("getrf mixed", prec, res, "emulated", emu)
    assert res <= 3 * emu

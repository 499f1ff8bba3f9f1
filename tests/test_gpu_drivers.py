"""GPU parity of the double-blocked POTRF / GETRF drivers through the C-ABI.

Bar (written per test):
  * exact closed forms: factors and ipiv bit-identical to the construction;
  * random inputs: CPU FP64 residual (the reference's metric) within the
    reference's gates (POTRF 1e-5, GETRF 5e-4; run_paper_experiments.sh:56-57)
    and scaled residual / (n eps32) <= 10; L within 100 n eps of LAPACK
    spotrf; ipiv identical to the oracle schedule / LAPACK sgetrf;
  * immediate and batched pivot GETRF: bit-identical factors and pivots.
"""
import numpy as np
import pytest

from gpu_util import EPS32, Session, dev, dev_int, host, log, pivot_tie_audit, sync
from oracle import oracle as O

pytestmark = pytest.mark.gpu

POTRF_GATE = 1e-5
GETRF_GATE = 5e-4


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


def run_potrf(H, A, nb, b, variant="plain", dtype=np.float32):
    n = A.shape[1]
    lda = A.shape[0]
    dA, info = dev(A.astype(dtype)), dev_int(1, 99)
    L = lib()
    if dtype == np.float64:
        if variant == "dbtrsm":
            L.double_blocked_potrf_double_dbtrsm(H.ptr, P(dA), lda, n, nb, b, P(info))
        else:
            L.double_blocked_potrf_double(H.ptr, P(dA), lda, n, nb, b, P(info))
    elif variant == "plain":
        L.double_blocked_potrf(H.ptr, P(dA), lda, n, nb, b, P(info))
    elif variant == "dbtrsm":
        L.double_blocked_potrf_dbtrsm(H.ptr, P(dA), lda, n, nb, b, P(info))
    elif variant == "kbsyrk":
        L.double_blocked_potrf_kblocked_syrk(H.ptr, P(dA), lda, n, nb, b, 128, P(info))
    elif variant == "adaptive":
        L.double_blocked_potrf_kblocked_syrk_adaptive_gemm(H.ptr, P(dA), lda, n, nb, b, 128,
                                                           P(info))
    elif variant == "configured":
        L.double_blocked_potrf_kblocked_syrk_adaptive_gemm_configured(
            H.ptr, P(dA), lda, n, nb, b, 128, 256, 512, 64, P(info))
    sync()
    assert H.last_error() == 0
    return host(dA), int(info[0])


def run_getrf(H, A, onb, inb, mode=1, tb=0, batched=False, dtype=np.float32):
    import torch
    n = A.shape[1]
    lda = A.shape[0]
    dA, ipiv, info = dev(A.astype(dtype)), dev_int(n), dev_int(1, 99)
    L = lib()
    if not batched:
        fn = L.double_blocked_getrf_double if dtype == np.float64 else L.double_blocked_getrf
        fn(H.ptr, P(dA), lda, n, onb, inb, mode, tb, P(ipiv), P(info))
    else:
        from hygondoubleblockedfactorizations_amd._lib import BatchedPivotWorkspace
        col_tile = 64
        src = torch.zeros(n, dtype=torch.int32, device="cuda")
        dst = torch.zeros(n, dtype=torch.int32, device="cuda")
        swap = torch.zeros(n * col_tile * 2, dtype=torch.float64, device="cuda")
        hp = [np.zeros(n, dtype=np.int32) for _ in range(4)]
        ws = BatchedPivotWorkspace(P(src), P(dst), P(swap), P(hp[0]), P(hp[1]), P(hp[2]),
                                   P(hp[3]), n, col_tile)
        import ctypes
        fn = (L.double_blocked_getrf_double_batched_pivot if dtype == np.float64
              else L.double_blocked_getrf_batched_pivot)
        fn(H.ptr, P(dA), lda, n, onb, inb, mode, tb, 0, P(ipiv), P(info), ctypes.byref(ws))
    sync()
    assert H.last_error() == 0
    return host(dA), ipiv.cpu().numpy()[:n].copy(), int(info[0])


def exact_spd(n, seed):
    """A = L L^T, small-integer L whose entries within 64 of the diagonal are
    zero (every 64-wide diagonal block is diag(1|2), so its inverse -- what
    the triangular-solve leaves multiply by -- is exact too)."""
    r = np.random.default_rng(seed)
    L = np.tril(r.integers(-2, 3, size=(n, n)), -64).astype(np.float64)
    L += np.diag(r.choice([1.0, 2.0], size=n))
    return np.asfortranarray(L.astype(np.float32)), np.asfortranarray((L @ L.T).astype(np.float32))


def exact_lu(n, seed):
    r = np.random.default_rng(seed)
    L = np.tril(r.choice([-0.5, 0.0, 0.5], size=(n, n)), -64) + np.eye(n)
    U = np.triu(r.integers(-3, 4, size=(n, n)), 1).astype(np.float64)
    U += np.diag(r.choice([1.0, -1.0, 2.0, -2.0, 4.0], size=n))
    perm = r.permutation(n)
    A = np.empty((n, n))
    A[perm, :] = L @ U
    return (np.asfortranarray(L.astype(np.float32)), np.asfortranarray(U.astype(np.float32)),
            np.asfortranarray(A.astype(np.float32)))


# ----------------------------------------------------------------- POTRF

def test_potrf_config_n2048(H):
    """The CPU-runnable benchmark case: n=2048, nb/b = 256/64, A = G G^T/n + I."""
    log("potrf n=2048 256/64")
    n = 2048
    A = O.spd_matrix(n, 20260716)
    Lg, info = run_potrf(H, A, 256, 64)
    assert info == 0
    res = O.potrf_residual(Lg, A)
    log("residual", res, "scaled", res / (n * EPS32))
    assert res <= POTRF_GATE
    assert res / (n * EPS32) <= 10
    Y = A.copy(order="F")
    assert O.spotrf(Y) == 0
    assert np.max(np.abs(np.tril(Lg) - np.tril(Y))) <= 100 * n * EPS32


@pytest.mark.parametrize("n,nb,b", [(300, 128, 64), (257, 100, 30), (1000, 0, 0), (513, 512, 512)])
@pytest.mark.parametrize("variant", ["plain", "dbtrsm", "kbsyrk", "adaptive", "configured"])
def test_potrf_closed_form_exact(H, n, nb, b, variant):
    log("potrf exact", n, nb, b, variant)
    L, A = exact_spd(n, n + nb)
    Lg, info = run_potrf(H, A, nb, b, variant)
    assert info == 0
    assert np.array_equal(np.tril(Lg), L)


@pytest.mark.parametrize("n,nb,b,pad", [(1024, 256, 64, 0), (1500, 512, 128, 5), (4096, 1024, 256, 0)])
def test_potrf_random_vs_oracle(H, n, nb, b, pad):
    log("potrf random", n, nb, b, pad)
    A = O.spd_matrix(n, n, lda=n + pad)
    Lg, info = run_potrf(H, A, nb, b)
    assert info == 0
    res = O.potrf_residual(Lg, A, n=n, lda=n + pad)
    assert res <= POTRF_GATE and res / (n * EPS32) <= 10
    assert np.array_equal(Lg[n:, :], A[n:, :])          # padding rows untouched
    assert np.array_equal(np.triu(Lg[:n], 1), np.triu(A[:n], 1))  # upper untouched


def test_potrf_info_first_global_column(H):
    log("potrf info")
    n = 600
    A = O.spd_matrix(n, 9)
    A[333, 333] = -10.0
    _, info = run_potrf(H, A, 256, 64)
    assert info == 334


def test_potrf_n_nonpositive_noop(H):
    log("potrf n=0")
    A = O.spd_matrix(8, 1)
    dA, infod = dev(A), dev_int(1, 42)
    lib().double_blocked_potrf(H.ptr, P(dA), 8, 0, 4, 2, P(infod))
    sync()
    assert np.array_equal(host(dA), A) and int(infod[0]) == 42


def test_potrf_double(H):
    log("potrf double")
    n = 700
    A = O.spd_matrix(n, 2, dtype=np.float64)
    Lg, info = run_potrf(H, A, 256, 64, dtype=np.float64)
    assert info == 0
    D = A.copy(order="F")
    assert O.dpotrf(D) == 0
    assert np.max(np.abs(np.tril(Lg) - np.tril(D))) <= 1e-12
    Lg2, _ = run_potrf(H, A, 300, 100, "dbtrsm", dtype=np.float64)
    assert np.max(np.abs(np.tril(Lg2) - np.tril(D))) <= 1e-12


# ----------------------------------------------------------------- GETRF

@pytest.mark.parametrize("n,onb,inb,mode,tb", [(256, 128, 64, 1, 64), (300, 96, 40, 0, 0),
                                               (512, 0, 0, 1, 0), (1000, 512, 128, 1, 200)])
@pytest.mark.parametrize("batched", [False, True])
def test_getrf_closed_form_exact(H, n, onb, inb, mode, tb, batched):
    log("getrf exact", n, onb, inb, mode, tb, batched)
    L, U, A = exact_lu(n, n + onb)
    X = A.copy(order="F")
    ipiv_ref, _, _ = O.double_blocked_getrf(X, onb, inb, mode, tb)
    LU, ipiv, info = run_getrf(H, A, onb, inb, mode, tb, batched)
    assert info == 0
    assert np.array_equal(ipiv, ipiv_ref)
    assert np.array_equal(np.tril(LU, -1), np.tril(L, -1))
    assert np.array_equal(np.triu(LU), U)


@pytest.mark.parametrize("n,onb,inb", [(512, 256, 64), (1024, 512, 256), (2048, 1024, 512)])
def test_getrf_random_vs_oracle(H, n, onb, inb):
    log("getrf random", n, onb, inb)
    A = O.uniform_matrix(n, 1234567890 % (n + 7))
    LU, ipiv, info = run_getrf(H, A, onb, inb, 1, 128)
    assert info == 0
    X = A.copy(order="F")
    ipiv_ref, info_ref = O.sgetrf(X)
    ok, k, mult = pivot_tie_audit(LU, ipiv, ipiv_ref, n, tol=4 * n * EPS32)
    log("pivot audit", ok, k, mult)
    assert ok, (k, mult)
    res = O.getrf_residual(LU, A, ipiv)
    log("residual", res, "scaled", res / (n * EPS32))
    assert res <= GETRF_GATE and res / (n * EPS32) <= 10


@pytest.mark.parametrize("n,onb,inb", [(777, 300, 100), (1024, 512, 128)])
def test_getrf_batched_bit_identical_to_immediate(H, n, onb, inb):
    log("getrf batched==immediate", n, onb, inb)
    A = O.uniform_matrix(n, 77)
    LU1, p1, i1 = run_getrf(H, A, onb, inb, 1, 128, batched=False)
    LU2, p2, i2 = run_getrf(H, A, onb, inb, 1, 128, batched=True)
    assert i1 == i2 == 0
    assert np.array_equal(p1, p2)
    assert np.array_equal(LU1, LU2)


def test_getrf_zero_column_info(H):
    log("getrf info")
    n = 400
    A = O.uniform_matrix(n, 8)
    A[:, 201] = 0.0
    _, ipiv, info = run_getrf(H, A, 200, 64, 1, 64)
    assert info == 202


def test_getrf_invalid_workspace(H):
    log("getrf invalid workspace")
    import ctypes
    from hygondoubleblockedfactorizations_amd._lib import BatchedPivotWorkspace
    n = 64
    A = O.uniform_matrix(n, 1)
    dA, ipiv, info = dev(A), dev_int(n), dev_int(1, 0)
    ws = BatchedPivotWorkspace()       # all NULL, capacity 0
    lib().double_blocked_getrf_batched_pivot(H.ptr, P(dA), n, n, 32, 16, 1, 16, 0, P(ipiv),
                                             P(info), ctypes.byref(ws))
    sync()
    assert int(info[0]) == -1
    assert np.array_equal(host(dA), A)


def test_getrf_double(H):
    log("getrf double")
    n = 600
    A = O.uniform_matrix(n, 31, dtype=np.float64)
    LU, ipiv, info = run_getrf(H, A, 256, 64, 1, 64, dtype=np.float64)
    assert info == 0
    D = A.copy(order="F")
    ipiv_ref, _ = O.dgetrf(D)
    assert np.array_equal(ipiv, ipiv_ref)
    assert np.max(np.abs(LU - D)) <= 1e-10
    LU2, ipiv2, _ = run_getrf(H, A, 256, 64, 1, 64, batched=True, dtype=np.float64)
    assert np.array_equal(ipiv2, ipiv) and np.array_equal(LU2, LU)

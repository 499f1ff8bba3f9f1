"""GPU parity of the gfx950 building blocks, called through the C-ABI.

Floating-point tolerances are written per test: products are checked against
an FP64 reference within the componentwise FP32 accumulation bound
(k + 2) eps |A| |B|; exact closed forms must come back bit for bit; integer
outputs (IAMAX, pivots) must be identical.
"""
import numpy as np
import pytest

from gpu_util import EPS32, EPS64, Session, dev, dev_int, gemm_bound, host, log, sync
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


GEMM_SHAPES = [(1, 1, 1, 0), (64, 64, 16, 0), (130, 70, 33, 0), (257, 129, 64, 3),
               (1000, 300, 200, 0), (555, 555, 1024, 1), (4096, 256, 512, 0)]


@pytest.mark.parametrize("M,N,K,pad", GEMM_SHAPES)
@pytest.mark.parametrize("op", ["nt", "nn"])
def test_gemm_update(H, M, N, K, pad, op):
    log("gemm", op, M, N, K, pad)
    r = np.random.default_rng(M + N + K)
    lda = M + pad
    A = np.asfortranarray(r.standard_normal((lda, K)).astype(np.float32))
    if op == "nt":
        ldb = N + pad
        B = np.asfortranarray(r.standard_normal((ldb, K)).astype(np.float32))
        Bop = B[:N, :K].astype(np.float64).T
    else:
        ldb = K + pad
        B = np.asfortranarray(r.standard_normal((ldb, N)).astype(np.float32))
        Bop = B[:K, :N].astype(np.float64)
    C = np.asfortranarray(r.standard_normal((lda, N)).astype(np.float32))
    dA, dB, dC = dev(A), dev(B), dev(C)
    fn = lib().gemm_nt_update if op == "nt" else lib().gemm_nn_update
    fn(H.ptr, M, N, K, P(dA), lda, P(dB), ldb, P(dC), lda)
    sync()
    assert H.last_error() == 0
    got = host(dC)
    ref = C[:M, :N].astype(np.float64) - A[:M, :K].astype(np.float64) @ Bop
    bound = gemm_bound(A[:M, :K], np.abs(Bop), K) + 2 * EPS32 * np.abs(C[:M, :N])
    assert np.all(np.abs(got[:M, :N] - ref) <= bound)
    assert np.array_equal(got[M:, :], C[M:, :])     # padding rows untouched


@pytest.mark.parametrize("n,k,kt", [(64, 64, 0), (300, 100, 0), (513, 257, 100), (1024, 96, 0)])
def test_syrk_lower(H, n, k, kt):
    log("syrk", n, k, kt)
    r = np.random.default_rng(n)
    A = np.asfortranarray(r.standard_normal((n, k)).astype(np.float32))
    C = np.asfortranarray(r.standard_normal((n, n)).astype(np.float32))
    dA, dC = dev(A), dev(C)
    lib().syrk_lower_update_kblocked(H.ptr, n, k, P(dA), n, P(dC), n, kt)
    sync()
    got = host(dC)
    ref = C.astype(np.float64) - A.astype(np.float64) @ A.astype(np.float64).T
    bound = gemm_bound(A, np.abs(A.T), k) * (2 if kt else 1) + 2 * EPS32 * np.abs(C)
    low = np.tril_indices(n)
    up = np.triu_indices(n, 1)
    assert np.all(np.abs(got[low] - ref[low]) <= bound[low])
    assert np.array_equal(got[up], C[up])


@pytest.mark.parametrize("rows,cols", [(100, 64), (37, 1), (1000, 200), (5000, 1024), (64, 130)])
def test_trsm_right_lower_transpose(H, rows, cols):
    log("trsm_rlt", rows, cols)
    r = np.random.default_rng(rows * cols)
    L = np.tril(r.standard_normal((cols, cols))) * 0.1 + np.eye(cols) * 2
    L = np.asfortranarray(L.astype(np.float32))
    B = np.asfortranarray(r.standard_normal((rows, cols)).astype(np.float32))
    dL, dB = dev(L), dev(B)
    lib().trsm_right_lower_transpose(H.ptr, rows, cols, P(dL), cols, P(dB), rows)
    sync()
    X = host(dB).astype(np.float64)
    Ld = np.tril(L).astype(np.float64)
    resid = np.abs(X @ Ld.T - B)
    bound = 4 * (cols + 2) * EPS32 * (np.abs(X) @ np.abs(Ld.T) + np.abs(B))
    assert np.all(resid <= bound)


@pytest.mark.parametrize("rows,cols,block", [(64, 100, 0), (1, 7, 0), (200, 1000, 0),
                                             (1024, 300, 0), (700, 50, 256)])
def test_trsm_left_lower_unit(H, rows, cols, block):
    log("trsm_llu", rows, cols, block)
    r = np.random.default_rng(rows + cols)
    L = np.tril(r.uniform(-1, 1, (rows, rows)), -1) * (2.0 / np.sqrt(rows)) + np.eye(rows)
    L = np.asfortranarray(L.astype(np.float32))
    B = np.asfortranarray(r.standard_normal((rows, cols)).astype(np.float32))
    dL, dB = dev(L), dev(B)
    if block:
        lib().trsm_left_lower_unit_iterative(H.ptr, rows, cols, P(dL), rows, P(dB), rows, block)
    else:
        lib().trsm_left_lower_unit(H.ptr, rows, cols, P(dL), rows, P(dB), rows)
    sync()
    X = host(dB).astype(np.float64)
    Ld = (np.tril(L, -1) + np.eye(rows)).astype(np.float64)
    resid = np.abs(Ld @ X - B)
    bound = 4 * (rows + 2) * EPS32 * (np.abs(Ld) @ np.abs(X) + np.abs(B))
    assert np.all(resid <= bound)


def test_double_blocked_trsm(H):
    log("double_blocked_trsm")
    r = np.random.default_rng(9)
    n, m = 900, 700
    L = np.asfortranarray((np.tril(r.standard_normal((m, m))) * 0.05 + 2 * np.eye(m)).astype(np.float32))
    B = np.asfortranarray(r.standard_normal((n, m)).astype(np.float32))
    dL, dB = dev(L), dev(B)
    lib().double_blocked_trsm(H.ptr, P(dL), m, P(dB), n, n, m, 300, 128)
    sync()
    X = host(dB).astype(np.float64)
    Ld = np.tril(L).astype(np.float64)
    resid = np.abs(X @ Ld.T - B)
    assert np.all(resid <= 4 * (m + 2) * EPS32 * (np.abs(X) @ np.abs(Ld.T) + np.abs(B)))


def exact_spd(n, seed):
    """A = L L^T, small-integer L whose entries within 64 of the diagonal are
    zero (every 64-wide diagonal block is diag(1|2), so its inverse -- what
    the triangular-solve leaves multiply by -- is exact too)."""
    r = np.random.default_rng(seed)
    L = np.tril(r.integers(-2, 3, size=(n, n)), -64).astype(np.float64)
    L += np.diag(r.choice([1.0, 2.0], size=n))
    return np.asfortranarray(L.astype(np.float32)), np.asfortranarray((L @ L.T).astype(np.float32))


@pytest.mark.parametrize("n", [1, 5, 64, 65, 200, 1024])
def test_potrf_block_closed_form_exact(H, n):
    log("potrf_block exact", n)
    L, A = exact_spd(n, n)
    dA, info = dev(A), dev_int(1, 7)
    lib().mifact_potrf_block(H.ptr, n, P(dA), n, P(info))
    sync()
    assert int(info[0]) == 0
    assert np.array_equal(np.tril(host(dA)), L)


@pytest.mark.parametrize("n", [5, 33, 64])
def test_potf2_leaf_elimination_exact(H, n):
    """One 64-wide leaf (no inverse applied): integer L with entries next to
    the diagonal must come back bit for bit."""
    log("potf2 leaf exact", n)
    r = np.random.default_rng(n)
    L = np.tril(r.integers(-2, 3, size=(n, n)), -1).astype(np.float64)
    L += np.diag(r.choice([1.0, 2.0], size=n))
    A = np.asfortranarray((L @ L.T).astype(np.float32))
    dA, info = dev(A), dev_int(1, 7)
    lib().mifact_potrf_block(H.ptr, n, P(dA), n, P(info))
    sync()
    assert int(info[0]) == 0
    assert np.array_equal(np.tril(host(dA)), L.astype(np.float32))


@pytest.mark.parametrize("n,bad", [(64, 10), (300, 150), (1000, 999)])
def test_potrf_block_info(H, n, bad):
    log("potrf_block info", n, bad)
    A = O.spd_matrix(n, 5)
    A[bad, bad] = -1.0
    dA, info = dev(A), dev_int(1)
    lib().mifact_potrf_block(H.ptr, n, P(dA), n, P(info))
    sync()
    assert int(info[0]) == bad + 1


@pytest.mark.parametrize("n", [200, 777])
def test_potrf_block_vs_lapack(H, n):
    log("potrf_block lapack", n)
    A = O.spd_matrix(n, 6)
    dA, info = dev(A), dev_int(1)
    lib().mifact_potrf_block(H.ptr, n, P(dA), n, P(info))
    sync()
    Lg = host(dA)
    Y = A.copy(order="F")
    assert O.spotrf(Y) == 0
    assert np.max(np.abs(np.tril(Lg) - np.tril(Y))) <= 50 * n * EPS32
    assert O.potrf_residual(Lg, A) <= 1e-6


@pytest.mark.parametrize("x,expect", [
    ([1, -3, 3, 2], 2), ([0, 0, 0], 1), ([-5, 5], 1), ([float("nan"), 1, 2], 3), ([], 0)])
def test_iamax_semantics(H, x, expect):
    log("iamax", x)
    import torch
    xs = torch.tensor(x if x else [0.0], dtype=torch.float32, device="cuda")
    res = dev_int(1, -9)
    lib().mifact_isamax(H.ptr, len(x), P(xs), 1, P(res))
    sync()
    assert int(res[0]) == expect


@pytest.mark.parametrize("n", [1000, 100000])
def test_iamax_large_with_ties(H, n):
    log("iamax large", n)
    import torch
    r = np.random.default_rng(n)
    x = r.uniform(-1, 1, n).astype(np.float32)
    x[n // 3] = 7.0
    x[n // 2] = -7.0
    xs = torch.from_numpy(x).cuda()
    res = dev_int(1)
    lib().mifact_isamax(H.ptr, n, P(xs), 1, P(res))
    sync()
    assert int(res[0]) == n // 3 + 1


def exact_lu(m, n, seed, band=64):
    """P A = L U, |l_ij| <= 1/2, power-of-two diag(U): exact in FP32.  L is
    zero within `band` of the diagonal (exact 64-block inverses); band=1 is
    used for single-leaf panels, where no inverse is applied."""
    r = np.random.default_rng(seed)
    L = np.tril(r.choice([-0.5, 0.0, 0.5], size=(m, n)), -band)
    L[:n, :n] += np.eye(n)
    U = np.triu(r.integers(-3, 4, size=(n, n)), 1).astype(np.float64)
    U += np.diag(r.choice([1.0, -1.0, 2.0, -2.0, 4.0], size=n))
    perm = r.permutation(m)
    A = np.empty((m, n))
    A[perm, :] = L @ U
    return np.asfortranarray(A.astype(np.float32))


@pytest.mark.parametrize("m,n,band", [(32, 32, 1), (100, 20, 1), (3000, 32, 1), (700, 64, 64),
                                      (5000, 96, 64), (40000, 40, 64)])
def test_getrf_panel_closed_form(H, m, n, band):
    log("getrf_panel exact", m, n, band)
    A = exact_lu(m, n, m + n, band)
    X = A.copy(order="F")
    ipiv_ref, info_ref = O.getf2(X)
    dA, ipiv, info = dev(A), dev_int(n), dev_int(1, 5)
    lib().mifact_getrf_panel(H.ptr, m, n, P(dA), m, P(ipiv), P(info))
    sync()
    assert H.last_error() == 0
    assert int(info[0]) == 0 == info_ref
    assert np.array_equal(ipiv.cpu().numpy()[:n], ipiv_ref)
    assert np.array_equal(host(dA), X)


@pytest.mark.parametrize("m,n", [(300, 33), (2000, 128), (20000, 64)])
def test_getrf_panel_random_vs_oracle(H, m, n):
    log("getrf_panel random", m, n)
    A = O.uniform_matrix(n, m, m=m)
    X = A.copy(order="F")
    ipiv_ref, _ = O.getf2(X)
    dA, ipiv, info = dev(A), dev_int(n), dev_int(1)
    lib().mifact_getrf_panel(H.ptr, m, n, P(dA), m, P(ipiv), P(info))
    sync()
    assert np.array_equal(ipiv.cpu().numpy()[:n], ipiv_ref)
    got = host(dA)
    assert np.max(np.abs(got - X)) <= 1e-4


def test_getrf_panel_zero_column_info(H):
    log("getrf_panel zero column")
    m, n = 500, 50
    A = O.uniform_matrix(n, 3, m=m)
    A[:, 17] = 0.0
    X = A.copy(order="F")
    ipiv_ref, info_ref = O.getf2(X)
    dA, ipiv, info = dev(A), dev_int(n), dev_int(1)
    lib().mifact_getrf_panel(H.ptr, m, n, P(dA), m, P(ipiv), P(info))
    sync()
    assert int(info[0]) == info_ref == 18
    assert np.array_equal(ipiv.cpu().numpy()[:n], ipiv_ref)


@pytest.mark.parametrize("M,N,K", [(100, 60, 40), (513, 257, 129)])
def test_gemm_double(H, M, N, K):
    log("gemm double", M, N, K)
    r = np.random.default_rng(M)
    A = np.asfortranarray(r.standard_normal((M, K)))
    B = np.asfortranarray(r.standard_normal((N, K)))
    C = np.asfortranarray(r.standard_normal((M, N)))
    dA, dB, dC = dev(A), dev(B), dev(C)
    lib().gemm_nt_update_double(H.ptr, M, N, K, P(dA), M, P(dB), N, P(dC), M)
    sync()
    ref = C - A @ B.T
    bound = (K + 2) * EPS64 * (np.abs(A) @ np.abs(B.T) + np.abs(C))
    assert np.all(np.abs(host(dC) - ref) <= bound)

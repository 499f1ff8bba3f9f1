"""Pins the CPU restatement of the double-blocked schedules (oracle/schedule_impl.h).

Each check is one a plausible mistake would fail:
  * building blocks vs an FP64 numpy product / solve (transposed operand,
    wrong triangle, dropped term);
  * exact closed forms: A = L L^T and P A = L U built from dyadic factors so
    every FP32 operation is exact -> factors and ipiv must come back bit for
    bit (wrong index, wrong pivot rule, P vs P^-1);
  * LAPACK spotrf / sgetrf / dpotrf / dgetrf (OpenBLAS) on the same bytes;
  * immediate and batched pivot application give bit-identical factors
    (the reference's own evidence: identical residuals in
    results/.../accuracy/getrf_N2048_timings.csv:3-4);
  * info semantics (reference: last block's local info; global first).
"""
import numpy as np
import pytest

from oracle import oracle as O


def f32(a):
    return np.asfortranarray(a, dtype=np.float32)


# ----------------------------------------------------------- building blocks

@pytest.mark.parametrize("rows,cols,k", [(7, 5, 3), (33, 17, 9), (1, 1, 1)])
def test_gemm_blocks_vs_numpy(rows, cols, k):
    r = np.random.default_rng(rows)
    A = f32(r.standard_normal((rows, k)))
    Bt = f32(r.standard_normal((cols, k)))
    Bn = f32(r.standard_normal((k, cols)))
    C = f32(r.standard_normal((rows, cols)))
    C1 = C.copy(order="F")
    O.gemm_nt(C1, A, Bt, rows, cols, k)
    ref = C.astype(np.float64) - A.astype(np.float64) @ Bt.astype(np.float64).T
    assert np.allclose(C1, ref, rtol=1e-5, atol=1e-5)
    C2 = C.copy(order="F")
    O.gemm_nn(C2, A, Bn, rows, cols, k)
    ref = C.astype(np.float64) - A.astype(np.float64) @ Bn.astype(np.float64)
    assert np.allclose(C2, ref, rtol=1e-5, atol=1e-5)


def test_syrk_lower_only_and_ktile():
    r = np.random.default_rng(3)
    n, k = 13, 11
    A = f32(r.standard_normal((n, k)))
    C = f32(r.standard_normal((n, n)))
    ref = C.astype(np.float64) - A.astype(np.float64) @ A.astype(np.float64).T
    for kt in (0, 4):
        C1 = C.copy(order="F")
        O.syrk_lower(C1, A, n, k, kt)
        low = np.tril_indices(n)
        up = np.triu_indices(n, 1)
        assert np.allclose(C1[low], ref[low], rtol=1e-5, atol=1e-5)
        assert np.array_equal(C1[up], C[up])


def test_trsm_blocks_vs_numpy():
    r = np.random.default_rng(4)
    s, rows = 9, 15
    L = np.tril(r.standard_normal((s, s))) + 4 * np.eye(s)
    B = r.standard_normal((rows, s))
    X = B.copy()
    Bf = f32(B)
    O.trsm_rlt(Bf, f32(L), rows, s)
    ref = np.linalg.solve(L.astype(np.float32).astype(np.float64),
                          B.astype(np.float32).astype(np.float64).T).T
    assert np.allclose(Bf, ref, rtol=1e-4, atol=1e-5)
    Lu = np.tril(r.standard_normal((s, s)), -1) * 0.3 + np.eye(s)
    Bl = r.standard_normal((s, rows))
    Bf = f32(Bl)
    O.trsm_llu(Bf, f32(Lu), s, rows)
    ref = np.linalg.solve(Lu.astype(np.float32).astype(np.float64),
                          Bl.astype(np.float32).astype(np.float64))
    assert np.allclose(Bf, ref, rtol=1e-4, atol=1e-5)
    del X


def test_iamax_first_of_ties():
    assert O.iamax(np.array([1, -3, 3, 2], dtype=np.float32)) == 1
    assert O.iamax(np.array([0, 0, 0], dtype=np.float32)) == 0
    assert O.iamax(np.array([-5, 5], dtype=np.float64)) == 0


def test_apply_panel_pivots_brute_force():
    r = np.random.default_rng(5)
    n, panel, b = 12, 4, 3
    A = f32(r.standard_normal((n, n)))
    piv_local = np.array([3, 1 + 4, 3 + 2], dtype=np.int32)  # 1-based, local
    ref = A.copy()
    for c in range(n):
        if panel <= c < panel + b:
            continue
        for k in range(b):
            i, p = panel + k, panel + piv_local[k] - 1
            ref[[i, p], c] = ref[[p, i], c]
    got = A.copy(order="F")
    glob = O.apply_panel_pivots(got, panel, b, piv_local)
    assert np.array_equal(got, ref)
    assert list(glob) == list(piv_local + panel)


# ----------------------------------------------------------- closed forms

def exact_spd(n, seed):
    """A = L L^T with small-integer L (diagonal 1 or 2): exact in FP32."""
    r = np.random.default_rng(seed)
    L = np.tril(r.integers(-2, 3, size=(n, n)), -1).astype(np.float64)
    L += np.diag(r.choice([1.0, 2.0], size=n))
    return L, f32(L @ L.T)


def exact_lu(n, seed):
    """P A = L U with |l_ij| <= 1/2 (unique pivots) and power-of-two U
    diagonal: every FP32 operation of the factorization is exact."""
    r = np.random.default_rng(seed)
    L = np.tril(r.choice([-0.5, 0.0, 0.5], size=(n, n)), -1) + np.eye(n)
    U = np.triu(r.integers(-3, 4, size=(n, n)), 1).astype(np.float64)
    U += np.diag(r.choice([1.0, -1.0, 2.0, -2.0, 4.0], size=n))
    perm = r.permutation(n)          # row i of P A is row perm[i] of A
    PA = L @ U
    A = np.empty_like(PA)
    A[perm, :] = PA
    # LAPACK ipiv encoding of the same permutation
    cur = list(range(n))             # cur[i] = original row now at position i
    pos = {v: i for i, v in enumerate(cur)}
    ipiv = np.zeros(n, dtype=np.int32)
    for k in range(n):
        p = pos[perm[k]]
        ipiv[k] = p + 1
        a, b = cur[k], cur[p]
        cur[k], cur[p] = b, a
        pos[a], pos[b] = p, k
    return L, U, f32(A), ipiv


@pytest.mark.parametrize("n,nb,b", [(64, 32, 8), (100, 40, 12), (97, 97, 97)])
def test_potrf_closed_form_exact(n, nb, b):
    L, A = exact_spd(n, n)
    X = A.copy(order="F")
    ir, ig = O.double_blocked_potrf(X, nb, b)
    assert (ir, ig) == (0, 0)
    assert np.array_equal(np.tril(X), f32(L))


@pytest.mark.parametrize("n,onb,inb,batched", [(64, 32, 8, False), (64, 32, 8, True),
                                                (90, 50, 16, False), (90, 50, 16, True)])
def test_getrf_closed_form_exact(n, onb, inb, batched):
    L, U, A, ipiv_ref = exact_lu(n, n + int(batched))
    X = A.copy(order="F")
    ipiv, ir, ig = O.double_blocked_getrf(X, onb, inb, 1, 16, batched)
    assert (ir, ig) == (0, 0)
    assert np.array_equal(ipiv, ipiv_ref)
    assert np.array_equal(np.tril(X, -1), f32(np.tril(L, -1)))
    assert np.array_equal(np.triu(X), f32(U))


# ----------------------------------------------------------- LAPACK parity

@pytest.mark.parametrize("n,nb,b,db", [(256, 64, 16, False), (300, 128, 50, True),
                                       (257, 0, 0, False)])
def test_potrf_vs_lapack(n, nb, b, db):
    A = O.spd_matrix(n, 7)
    X = A.copy(order="F")
    assert O.double_blocked_potrf(X, nb, b, use_dbtrsm=db) == (0, 0)
    Y = A.copy(order="F")
    assert O.spotrf(Y) == 0
    assert np.max(np.abs(np.tril(X) - np.tril(Y))) < 1e-5 * np.max(np.abs(Y))
    assert O.potrf_residual(X, A) < 1e-6
    D = A.astype(np.float64)
    Xd = D.copy(order="F")
    assert O.double_blocked_potrf(Xd, nb, b) == (0, 0)
    assert O.dpotrf(D) == 0
    assert np.max(np.abs(np.tril(Xd) - np.tril(D))) < 1e-12


@pytest.mark.parametrize("n,onb,inb,mode,tb", [(256, 128, 32, 1, 64), (200, 64, 24, 0, 0),
                                               (131, 0, 0, 1, 16)])
def test_getrf_vs_lapack(n, onb, inb, mode, tb):
    A = O.uniform_matrix(n, 11)
    X = A.copy(order="F")
    ipiv, ir, ig = O.double_blocked_getrf(X, onb, inb, mode, tb)
    assert ig == 0
    Y = A.copy(order="F")
    ipiv_l, info = O.sgetrf(Y)
    assert info == 0
    assert np.array_equal(ipiv, ipiv_l)
    assert O.getrf_residual(X, A, ipiv) < 1e-5
    D = A.astype(np.float64)
    Xd = D.copy(order="F")
    ipiv_d, _, _ = O.double_blocked_getrf(Xd, onb, inb, mode, tb)
    ipiv_dl, _ = O.dgetrf(D)
    assert np.array_equal(ipiv_d, ipiv_dl)
    assert np.max(np.abs(Xd - D)) < 1e-11


@pytest.mark.parametrize("n,onb,inb", [(160, 64, 16), (128, 128, 32)])
def test_getrf_batched_equals_immediate(n, onb, inb):
    A = O.uniform_matrix(n, 21)
    X1, X2 = A.copy(order="F"), A.copy(order="F")
    p1, _, _ = O.double_blocked_getrf(X1, onb, inb, 1, 32, batched=False)
    p2, _, _ = O.double_blocked_getrf(X2, onb, inb, 1, 32, batched=True)
    assert np.array_equal(p1, p2)
    assert np.array_equal(X1, X2)


# ----------------------------------------------------------- info semantics

def test_potrf_info_semantics():
    n = 48
    A = O.spd_matrix(n, 3)
    A[20, 20] = -50.0                     # first non-positive pivot at column 20
    X = A.copy(order="F")
    ir, ig = O.double_blocked_potrf(X, 16, 8)
    assert ig == 21                       # 1-based global
    # the reference keeps only the last diagonal call's local status
    assert ir != 21


def test_getrf_zero_pivot_info():
    n = 40
    A = O.uniform_matrix(n, 4)
    A[:, 13] = 0.0                        # exactly singular column 13
    X = A.copy(order="F")
    ipiv, ir, ig = O.double_blocked_getrf(X, 16, 8, 1, 8)
    assert ig == 14
    Y = A.copy(order="F")
    _, info = O.sgetrf(Y)
    assert info == 14

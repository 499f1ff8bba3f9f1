"""Pins the CPU FP64 residual oracle (oracle/residual.c).

Restates, as behaviour specs, the checks of the reference's own unit test
tests/cpu_fp64_residual/test_cpu_fp64_residual.cpp (170 checks):
  * :51-107  layouts n in {1,2,3,5,11}, lda padding {0,3}, LU / Cholesky,
             with and without pivoting, a 3-cycle permutation (catches P vs
             P^-1), NaN-filled ignored triangles, tiles {1,2,3,4,512},
             against an independent unblocked long-double residual (:24-49);
  * :109-147 FP64 product resolves FP32 rounding of (1+2^-23)^2, zero and
             non-finite handling, invalid pivot -> inf, invalid tile rejected,
             scaled norms at 1e+-160/200;
and cross-checks the restatement against the reference's header compiled from
its own sources (oracle/_ref/libref_residual.so) on random inputs.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _longdouble_reference(factor, original, permutation, n, lda, cholesky):
    """Unblocked long-double residual (test_cpu_fp64_residual.cpp:24-49)."""
    ld = np.longdouble
    f = factor.astype(ld)
    num = ld(0)
    den = ld(0)
    for j in range(n):
        for i in range(n):
            prod = ld(0)
            for k in range(n):
                if i < k:
                    left = ld(0)
                elif (not cholesky) and i == k:
                    left = ld(1)
                else:
                    left = f[k * lda + i]
                if cholesky:
                    right = ld(0) if j < k else f[k * lda + j]
                else:
                    right = ld(0) if k > j else f[j * lda + k]
                prod += left * right
            a = ld(original[j * lda + permutation[i]])
            d = prod - a
            num += d * d
            den += a * a
    return float(np.sqrt(num / den))


def _close(actual, expected):
    tol = max(1e-18, abs(expected) * 1e-7)
    return math.isfinite(actual) and abs(actual - expected) <= tol


CASES = [(n, pad, chol, piv)
         for n in (1, 2, 3, 5, 11)
         for pad in (0, 3)
         for chol in (False, True)
         for piv in (False, True)
         if not (chol and piv)]


def _build_case(n, pad, cholesky, pivoting):
    lda = n + pad
    factor = np.full(lda * n, np.nan, dtype=np.float32)
    original = np.zeros(lda * n, dtype=np.float32)
    pivots = np.arange(1, n + 1, dtype=np.int32)
    permutation = list(range(n))
    if pivoting and n >= 3:
        pivots[0] = 3
        pivots[1] = 3
        for k in range(n):
            p = pivots[k] - 1
            permutation[k], permutation[p] = permutation[p], permutation[k]
    for j in range(n):
        for i in range(n):
            if (not cholesky) or i >= j:
                factor[j * lda + i] = np.float32(
                    np.float32(0.13) * np.float32(i + 1)
                    - np.float32(0.07) * np.float32(j + 1)
                    + (np.float32(2.0) if i == j else np.float32(0.0)))
    ld = np.longdouble
    for j in range(n):
        for i in range(n):
            prod = ld(0)
            for k in range(min(i, j) + 1):
                l = ld(1) if ((not cholesky) and i == k) else ld(factor[k * lda + i])
                r = ld(factor[k * lda + j]) if cholesky else ld(factor[j * lda + k])
                prod += l * r
            original[j * lda + permutation[i]] = np.float32(prod)
    return lda, factor, original, pivots, permutation


@pytest.mark.parametrize("n,pad,cholesky,pivoting", CASES)
def test_layout_tile_permutation(n, pad, cholesky, pivoting):
    lda, factor, original, pivots, permutation = _build_case(n, pad, cholesky, pivoting)
    expected = _longdouble_reference(factor, original, permutation, n, lda, cholesky)
    F = factor.reshape(n, lda).T.copy(order="F")
    A = original.reshape(n, lda).T.copy(order="F")
    for tile in (1, 2, 3, 4, 512):
        if cholesky:
            got = O.potrf_residual(F, A, n=n, lda=lda, tile=tile)
        else:
            got = O.getrf_residual(F, A, pivots if pivoting else None, n=n, lda=lda,
                                   tile=tile, pivoting=pivoting)
        assert _close(got, expected), (tile, got, expected)
        if O.ref_lib() is not None:
            if cholesky:
                ref = O.potrf_residual(F, A, n=n, lda=lda, tile=tile, use_ref=True)
            else:
                ref = O.getrf_residual(F, A, pivots if pivoting else None, n=n, lda=lda,
                                       tile=tile, pivoting=pivoting, use_ref=True)
            assert got == ref


def test_fp64_product_resolves_fp32_rounding():
    f = np.nextafter(np.float32(1.0), np.float32(2.0))
    orig = np.float32(f * f)
    exact = np.longdouble(f) * np.longdouble(f)
    F = np.array([[f]], dtype=np.float32, order="F")
    A = np.array([[orig]], dtype=np.float32, order="F")
    got = O.potrf_residual(F, A)
    assert _close(got, float(abs(exact - np.longdouble(orig)) / np.longdouble(orig)))
    assert got > 0


def test_zero_and_nonfinite():
    z = np.zeros((1, 1), dtype=np.float32, order="F")
    one = np.ones((1, 1), dtype=np.float32, order="F")
    nan = np.full((1, 1), np.nan, dtype=np.float32, order="F")
    assert O.potrf_residual(z, z) == 0.0
    assert math.isinf(O.potrf_residual(one, z))
    assert math.isinf(O.potrf_residual(nan, one))
    bad = np.array([0], dtype=np.int32)
    assert math.isinf(O.getrf_residual(one, one, bad))


@pytest.mark.parametrize("tile", [0, -1])
def test_invalid_tile_rejected(tile):
    one = np.ones((1, 1), dtype=np.float32, order="F")
    with pytest.raises(O.InvalidArgument):
        O.potrf_residual(one, one, tile=tile)


def test_scaled_norm_overflow_underflow():
    big = np.array([[1e160]], order="F")
    big2 = np.array([[2e160]], order="F")
    assert _close(O.getrf_residual(big, big2, None, pivoting=False), 0.5)
    small = np.array([[1e-200]], order="F")
    small2 = np.array([[2e-200]], order="F")
    assert _close(O.getrf_residual(small, small2, None, pivoting=False), 0.5)


def test_invalid_pivots_give_inf():
    lu = np.eye(3, dtype=np.float32, order="F")
    for piv in ([1, 1, 3], [2, 4, 3], [0, 2, 3]):
        assert math.isinf(O.getrf_residual(lu, lu, np.array(piv, dtype=np.int32)))


def test_ipiv_to_permutation_three_cycle():
    # pivots (3, 3, 3): swap(0,2) then swap(1,2) -> [2, 0, 1]
    perm = O.ipiv_to_permutation(np.array([3, 3, 3], dtype=np.int32))
    assert list(perm) == [2, 0, 1]
    assert O.ipiv_to_permutation(np.array([2, 1], dtype=np.int32)) is None


@pytest.mark.skipif(O.ref_lib() is None, reason="reference residual not built here")
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_matches_reference_build_random(seed):
    r = np.random.default_rng(seed)
    n, lda = 37, 41
    F = np.asfortranarray(r.standard_normal((lda, n)).astype(np.float32))
    A = np.asfortranarray(r.standard_normal((lda, n)).astype(np.float32))
    piv = np.array([r.integers(k + 1, n + 1) for k in range(n)], dtype=np.int32)
    for tile in (5, 16, 512):
        assert O.getrf_residual(F, A, piv, n=n, lda=lda, tile=tile) == \
            O.getrf_residual(F, A, piv, n=n, lda=lda, tile=tile, use_ref=True)
        assert O.potrf_residual(F, A, n=n, lda=lda, tile=tile) == \
            O.potrf_residual(F, A, n=n, lda=lda, tile=tile, use_ref=True)
        F64, A64 = F.astype(np.float64), A.astype(np.float64)
        assert O.getrf_residual(F64, A64, piv, n=n, lda=lda, tile=tile) == \
            O.getrf_residual(F64, A64, piv, n=n, lda=lda, tile=tile, use_ref=True)


@pytest.mark.skipif(O.ref_lib() is None, reason="reference test binary not built here")
def test_reference_unit_test_passes():
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(O.__file__), "_ref", "test_cpu_fp64_residual")
    if not os.path.exists(exe):
        pytest.skip("reference test binary not built")
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    assert out.startswith("PASS: 170 checks")

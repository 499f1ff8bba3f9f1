"""GPU parity of the unpivoted double-blocked LU (double_blocked_getrf_npvt)
through the C-ABI.

Inputs are column diagonally dominant (iid U(-1,1) plus n I, the reference's
no-pivot generator, getrf_benchmark_common.cpp:45-49), so LAPACK sgetrf/dgetrf
chooses no row exchange (asserted: ipiv = 1..n) and computes the same
unpivoted factors.  Bar: factors within 50 n eps of LAPACK's, and the CPU
FP64 residual ||A - L U|| / ||A|| <= 10 n eps (the reference's metric).
"""
import numpy as np
import pytest

from gpu_util import EPS32, EPS64, Session, dev, dev_int, host, log, sync
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


def dominant(n, seed, dtype=np.float32, lda=None):
    A = O.uniform_matrix(n, seed, lda=lda, dtype=np.float64)
    A[np.arange(n), np.arange(n)] += n
    return np.asfortranarray(A.astype(dtype))


@pytest.mark.parametrize("n,nb,b,pad", [(2048, 256, 64, 0), (3000, 1024, 128, 3), (777, 0, 0, 1),
                                        (130, 64, 32, 0)])
def test_npvt_vs_lapack(H, n, nb, b, pad):
    log("getrf_npvt", n, nb, b, pad)
    A = dominant(n, 90 + n, lda=n + pad)
    dA, info = dev(A), dev_int(1, 7)
    lib().double_blocked_getrf_npvt(H.ptr, P(dA), n + pad, n, nb, b, P(info))
    sync()
    assert int(info.cpu()[0]) == 0 and H.last_error() == 0
    LU = host(dA)
    assert np.array_equal(LU[n:], A[n:]), "padding rows modified"
    ref = np.asfortranarray(A[:n].copy())
    ipiv, inf = O.sgetrf(ref)
    assert inf == 0 and np.array_equal(ipiv, np.arange(1, n + 1)), "input must need no pivoting"
    d = np.abs(LU[:n].astype(np.float64) - ref).max()
    assert d <= 50 * n * EPS32 * np.abs(ref).max(), d
    res = O.getrf_residual(np.asfortranarray(LU[:n]), np.asfortranarray(A[:n]),
                           np.arange(1, n + 1, dtype=np.int32))
    assert res <= 10 * n * EPS32, res


def test_npvt_double(H):
    n = 1024
    A = dominant(n, 5, dtype=np.float64)
    dA, info = dev(A), dev_int(1, 7)
    lib().double_blocked_getrf_npvt_double(H.ptr, P(dA), n, n, 512, 128, P(info))
    sync()
    assert int(info.cpu()[0]) == 0
    ref = np.asfortranarray(A.copy())
    ipiv, inf = O.dgetrf(ref)
    assert inf == 0 and np.array_equal(ipiv, np.arange(1, n + 1))
    assert np.abs(host(dA) - ref).max() <= 50 * n * EPS64 * np.abs(ref).max()


def test_npvt_zero_pivot_info(H):
    """A zero pivot reports its global column (LAPACK info) and the call
    returns normally; the columns before it are factored."""
    n = 300
    A = dominant(n, 6)
    A[200, :] = 0.0
    A[:, 200] = 0.0
    dA, info = dev(A), dev_int(1, 7)
    lib().double_blocked_getrf_npvt(H.ptr, P(dA), n, n, 128, 64, P(info))
    sync()
    assert int(info.cpu()[0]) == 201
    assert H.last_error() == 0


def test_npvt_empty(H):
    info = dev_int(1, 7)
    lib().double_blocked_getrf_npvt(H.ptr, None, 1, 0, 64, 32, P(info))
    sync()
    assert int(info.cpu()[0]) == 7 and H.last_error() == 0

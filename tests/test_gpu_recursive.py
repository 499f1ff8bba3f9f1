"""GPU parity of the recursive POTRF family (recursive_potrf,
recursive_potrf_tiled_trsm, recursive_potrf_double,
recursive_potrf_mixed_driver) through the C-ABI.

Bars (per test docstring): exact closed forms bit for bit; random SPD
inputs within the reference's POTRF gate (1e-5, run_paper_experiments.sh:
56-57) and 10 n eps scaled, L within 100 n eps of LAPACK spotrf; FP64
within 1e-12 of dpotrf; the mixed driver (FP16 products) within the
growth-aware bound u16 || |L||L^T| ||_F / ||A||_F, its damping path
succeeding on a marginally indefinite matrix and reporting a strongly
indefinite one.
"""
import numpy as np
import pytest

from gpu_util import EPS32, Session, dev, dev_int, host, log, sync
from oracle import oracle as O
from test_gpu_drivers import exact_spd

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


def run(H, name, A, nb):
    n = A.shape[1]
    dA, info = dev(A), dev_int(1, 99)
    getattr(lib(), name)(H.ptr, None, P(dA), A.shape[0], n, nb, P(info))
    sync()
    return host(dA), int(info[0])


@pytest.mark.parametrize("name", ["recursive_potrf", "recursive_potrf_tiled_trsm"])
@pytest.mark.parametrize("n,nb", [(300, 64), (513, 100), (1000, 0), (777, 300)])
def test_recursive_closed_form_exact(H, name, n, nb):
    """A = L L^T with small-integer L (64-wide diagonal blocks diagonal):
    every product and block inverse is exact, so L comes back bit for bit."""
    log("recursive exact", name, n, nb)
    L, A = exact_spd(n, n + nb + 5)
    Lg, info = run(H, name, A, nb)
    assert H.last_error() == 0 and info == 0
    assert np.array_equal(np.tril(Lg), L)


@pytest.mark.parametrize("n,nb,pad", [(2048, 256, 0), (1500, 128, 7)])
def test_recursive_random_vs_oracle(H, n, nb, pad):
    """A = G G^T / n + I: oracle FP64 residual within the POTRF gate, L within
    100 n eps of LAPACK spotrf, the strictly upper triangle untouched."""
    log("recursive random", n, nb, pad)
    A = O.spd_matrix(n, 31 + n, lda=n + pad)
    Lg, info = run(H, "recursive_potrf", A, nb)
    assert H.last_error() == 0 and info == 0
    res = O.potrf_residual(Lg, A, n=n, lda=n + pad)
    log("residual", res)
    assert res <= 1e-5 and res / (n * EPS32) <= 10
    Y = A[:n].copy(order="F")
    assert O.spotrf(Y) == 0
    assert np.max(np.abs(np.tril(Lg[:n]) - np.tril(Y))) <= 100 * n * EPS32
    assert np.array_equal(np.triu(Lg[:n], 1), np.triu(A[:n], 1))


def test_recursive_double(H):
    """FP64 recursion against LAPACK dpotrf."""
    n = 900
    A = O.spd_matrix(n, 4, dtype=np.float64)
    Lg, info = run(H, "recursive_potrf_double", A, 128)
    assert H.last_error() == 0 and info == 0
    D = A.copy(order="F")
    assert O.dpotrf(D) == 0
    assert np.max(np.abs(np.tril(Lg) - np.tril(D))) <= 1e-12


def test_recursive_info_first_global_column(H):
    """A non-positive pivot at column 411 (0-based) is reported as 412."""
    n = 700
    A = O.spd_matrix(n, 12)
    A[411, 411] = -5.0
    _, info = run(H, "recursive_potrf", A, 128)
    assert info == 412


def _growth(Lf):
    Lf = np.tril(Lf.astype(np.float64))
    return float(np.linalg.norm(np.abs(Lf) @ np.abs(Lf.T)))


def test_mixed_driver_well_conditioned(H):
    """FP16 products after scaling: the residual stays within
    u16 || |L||L^T| || / ||A|| (u16 = 2^-11), no damping needed."""
    n = 1024
    A = O.spd_matrix(n, 77)
    Lg, info = run(H, "recursive_potrf_mixed_driver", A, 128)
    assert H.last_error() == 0 and info == 0
    res = O.potrf_residual(Lg, A)
    bound = 2.0 ** -11 * _growth(Lg) / np.linalg.norm(A.astype(np.float64))
    log("mixed residual", res, "bound", bound)
    assert res <= bound


def test_mixed_driver_damping(H):
    """A marginally indefinite matrix (smallest eigenvalue -1e-4 of a
    spectrum in [~-1e-4, ~5]) is factored after diagonal damping (info 0, A +
    D = L L^T with a small D); A = -I cannot be rescued: info = 1."""
    n = 512
    rng = O.rng(5)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.linspace(-1e-4, 5.0, n)
    A = np.asfortranarray(((q * ev) @ q.T).astype(np.float32))
    A = np.asfortranarray(0.5 * (A + A.T))
    Lg, info = run(H, "recursive_potrf_mixed_driver", A, 128)
    log("damped info", info)
    assert H.last_error() == 0 and info == 0
    res = O.potrf_residual(Lg, A)
    log("damped residual", res)
    assert res <= 1e-2
    Neg = np.asfortranarray(-np.eye(256, dtype=np.float32))
    _, info2 = run(H, "recursive_potrf_mixed_driver", Neg, 64)
    assert info2 == 1

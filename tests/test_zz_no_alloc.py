"""The allocation-free POTRF entry point (double_blocked_potrf_noalloc)
against double_blocked_potrf through the C-ABI.

Bar: bit-identical factors and info on two calls in a row, the second on a
smaller matrix with a smaller inner block (the scratch keeps the first
call's contents); an inner block beyond the scratch bound is rejected with
info = -1 and A untouched."""
import numpy as np
import pytest

from gpu_util import Session, dev, dev_int, host, log, sync
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


def _run(H, fn, A, nb, b):
    n = A.shape[1]
    dA, info = dev(A), dev_int(1, 99)
    fn(H.ptr, P(dA), A.shape[0], n, nb, b, P(info))
    sync()
    return host(dA), int(info[0])


def test_noalloc_matches_allocating_entry_point(H):
    import torch
    L = lib()
    cases = [(2048, 512, 256, 3), (700, 300, 128, 0)]      # second call: smaller
    for n, nb, b, pad in cases:
        log("noalloc potrf", n, nb, b, pad)
        A = O.spd_matrix(n, 9000 + n, lda=n + pad)
        ref, info_ref = _run(H, L.double_blocked_potrf, A, nb, b)
        assert H.last_error() == 0
        dA, info = dev(A), dev_int(1, 99)
        torch.cuda.synchronize()
        L.double_blocked_potrf_noalloc(H.ptr, P(dA), n + pad, n, nb, b, P(info))
        sync()
        assert H.last_error() == 0
        got = host(dA)
        assert int(info[0]) == info_ref == 0
        assert np.array_equal(got, ref)


def test_noalloc_rejects_oversized_inner_block(H):
    L = lib()
    # nb <= 0 -> n, b = min(b, nb, n): a 256-order matrix with b = 5000 runs
    A = O.spd_matrix(256, 5)
    dA, info = dev(A), dev_int(1, 7)
    L.double_blocked_potrf_noalloc(H.ptr, P(dA), 256, 256, 0, 5000, P(info))
    sync()
    assert H.last_error() == 0 and int(info[0]) == 0
    ref, _ = _run(H, L.double_blocked_potrf, A, 0, 5000)
    assert np.array_equal(host(dA), ref)
    # an inner block of 4097 columns exceeds the scratch: rejected, A untouched
    big = 4097
    A3 = np.asfortranarray(np.eye(big, dtype=np.float32))
    dA3, info3 = dev(A3), dev_int(1, 0)
    L.double_blocked_potrf_noalloc(H.ptr, P(dA3), big, big, big, big, P(info3))
    sync()
    assert H.last_error() == 2                  # mifact_status_invalid_value
    assert int(info3[0]) == -1
    assert np.array_equal(host(dA3), A3)

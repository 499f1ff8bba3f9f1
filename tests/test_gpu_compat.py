"""GPU parity of the reference-compatible surface: the stand-alone pivot
launchers (include/getrf/getrf_pivot_kernels.h) and the rocblas_handle
compatibility library (libmifact_compat.so, include/compat/rocblas_compat.h).

Bars: row exchanges are data movement, so results are bit-identical to the
oracle's LAPACK-laswp restatement (oracle.apply_panel_pivots, which follows
src/getrf/getrf_pivot_kernels.hip:14-105) and to an exact gather/scatter;
the compatibility entry points give bit-identical factors, pivots and info
to the native ones on the same input.
"""
import ctypes
import os

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


def stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def random_local_pivots(rng, panel, ps, n):
    """LAPACK-valid local pivots of a panel starting at `panel`: step k
    exchanges panel rows k and piv[k] - 1 >= k."""
    return np.array([rng.integers(k, n - panel) + 1 for k in range(ps)], dtype=np.int32)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n,panel,ps,pad", [(600, 128, 200, 0), (1000, 256, 600, 3),
                                            (300, 0, 300, 1), (513, 100, 1, 0)])
def test_panel_pivots_launcher(H, dtype, n, panel, ps, pad):
    """launch_apply_getrf_panel_pivots == the oracle's sequential laswp on
    every column outside the panel (multi-chunk at ps > 256), pivots made
    global; the panel's own columns untouched."""
    log("panel pivots", np.dtype(dtype).name, n, panel, ps, pad)
    rng = np.random.default_rng(n + panel + ps)
    A = np.asfortranarray(rng.standard_normal((n + pad, n)).astype(dtype))
    piv = random_local_pivots(rng, panel, ps, n)
    ref = A.copy(order="F")
    gpiv = O.apply_panel_pivots(ref, panel, ps, piv)
    dA = dev(A)
    import torch
    dp = torch.from_numpy(piv.copy()).cuda()
    fn = (lib().launch_apply_getrf_panel_pivots_double if dtype == np.float64
          else lib().launch_apply_getrf_panel_pivots)
    fn(P(dA), n + pad, n, panel, ps, P(dp), stream())
    sync()
    assert np.array_equal(host(dA), ref)
    assert np.array_equal(dp.cpu().numpy(), gpiv)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("rows,cols,moves,tile", [(3000, 40, 700, 8), (40000, 5, 20000, 2)])
def test_permutation_launcher(H, dtype, rows, cols, moves, tile):
    """launch_apply_getrf_permutation: row targets[m] receives the pre-call
    row sources[m] in the first `cols` columns (LDS path; the temporary in
    column tiles when one column of the list exceeds 64 KiB)."""
    import torch
    log("permutation", np.dtype(dtype).name, rows, cols, moves, tile)
    rng = np.random.default_rng(rows + moves)
    A = np.asfortranarray(rng.standard_normal((rows, cols + 2)).astype(dtype))
    rows_sel = rng.choice(rows, size=moves, replace=False).astype(np.int32)
    perm = rng.permutation(moves)
    sources, targets = rows_sel, rows_sel[perm].copy()
    ref = A.copy(order="F")
    ref[targets, :cols] = A[sources, :cols]
    dA = dev(A)
    ds = torch.from_numpy(sources).cuda()
    dt = torch.from_numpy(targets).cuda()
    tmp = torch.zeros(tile * moves, dtype=torch.float64 if dtype == np.float64 else torch.float32,
                      device="cuda")
    fn = (lib().launch_apply_getrf_permutation_double if dtype == np.float64
          else lib().launch_apply_getrf_permutation)
    fn(P(dA), rows, cols, P(ds), P(dt), moves, P(tmp), tile, stream())
    sync()
    assert np.array_equal(host(dA), ref)


# ------------------------------------------------------------------ rocblas_handle layer

@pytest.fixture(scope="module")
def compat():
    from hygondoubleblockedfactorizations_amd import _lib
    rb = ctypes.CDLL("/opt/rocm/lib/librocblas.so")
    rb.rocblas_create_handle.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    rb.rocblas_set_stream.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    rb.rocblas_destroy_handle.argtypes = [ctypes.c_void_p]
    h = ctypes.c_void_p()
    assert rb.rocblas_create_handle(ctypes.byref(h)) == 0
    assert rb.rocblas_set_stream(h, stream()) == 0
    cl = ctypes.CDLL(os.path.join(os.path.dirname(_lib.LIB_PATH), "libmifact_compat.so"))
    for name in ("double_blocked_potrf", "recursive_potrf"):
        getattr(cl, name).restype = None
    yield rb, h, cl
    rb.rocblas_destroy_handle(h)


def test_compat_potrf_and_getrf_match_native(H, compat):
    """The reference's prototypes called with a real rocblas_handle give the
    native entry points' results bit for bit (POTRF, GETRF with pivots,
    the recursive POTRF with a null StreamPool)."""
    rb, rh, cl = compat
    L = lib()
    n = 1024
    A = O.spd_matrix(n, 4321)
    dA, info = dev(A), dev_int(1, 9)
    cl.double_blocked_potrf(rh, P(dA), n, n, 256, 64, P(info))
    sync()
    dB, info2 = dev(A), dev_int(1, 9)
    L.double_blocked_potrf(H.ptr, P(dB), n, n, 256, 64, P(info2))
    sync()
    assert int(info[0]) == int(info2[0]) == 0
    assert np.array_equal(host(dA), host(dB))
    G = O.uniform_matrix(n, 99)
    dG, ipiv, info = dev(G), dev_int(n), dev_int(1, 9)
    cl.double_blocked_getrf(rh, P(dG), n, n, 512, 128, 1, 128, P(ipiv), P(info))
    sync()
    dG2, ipiv2, info2 = dev(G), dev_int(n), dev_int(1, 9)
    L.double_blocked_getrf(H.ptr, P(dG2), n, n, 512, 128, 1, 128, P(ipiv2), P(info2))
    sync()
    assert int(info[0]) == int(info2[0]) == 0
    assert np.array_equal(ipiv.cpu().numpy(), ipiv2.cpu().numpy())
    assert np.array_equal(host(dG), host(dG2))
    dR, info = dev(A), dev_int(1, 9)
    cl.recursive_potrf(rh, None, P(dR), n, n, 128, P(info))
    sync()
    dR2, info2 = dev(A), dev_int(1, 9)
    L.recursive_potrf(H.ptr, None, P(dR2), n, n, 128, P(info2))
    sync()
    assert int(info[0]) == int(info2[0]) == 0
    assert np.array_equal(host(dR), host(dR2))

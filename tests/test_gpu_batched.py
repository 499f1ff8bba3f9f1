"""GPU parity of the batched (uneven, one persistent launch) Cholesky / LU
and the graph-captured batched refinement through the C-ABI.

Bar (written per test):
  * every matrix: CPU FP64 residual (the reference's metric) within the
    reference's gates (POTRF 1e-5, GETRF 5e-4) and <= 10 n eps32 scaled;
    L within 100 n eps32 of LAPACK spotrf; ipiv identical to LAPACK sgetrf
    (tie audit at the first difference); info = LAPACK's;
  * strictly-upper triangles of POTRF inputs untouched; other matrices of the
    batch unaffected by a failing one;
  * refinement: every matrix's backward error <= the tolerance.
"""
import ctypes

import numpy as np
import pytest

from gpu_util import (EPS32, Session, dev_int, emulated_lowp_lu, log, pivot_tie_audit,
                      sync)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PREC = {"fp32": 0, "bf16": 1, "fp16": 2}



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


class Packed:
    """Uneven batch packed into one FP32 buffer (column-major, per-matrix
    offset and leading dimension)."""

    def __init__(self, mats, pads):
        import torch
        self.n = np.array([m.shape[1] for m in mats], np.int32)
        self.lda = np.array([m.shape[1] + p for m, p in zip(mats, pads)], np.int32)
        sizes = self.lda.astype(np.int64) * np.maximum(self.n, 1)
        self.off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        flat = np.full(int(sizes.sum()), 7.25, np.float32)
        for m, o, ld in zip(mats, self.off, self.lda):
            n = m.shape[1]
            if n == 0:
                continue
            blk = np.full((ld, n), 7.25, np.float32, order="F")
            blk[:n, :] = m
            flat[o:o + ld * n] = blk.ravel(order="F")
        self.flat = flat
        self.dA = torch.from_numpy(flat.copy()).cuda()
        self.d_off = torch.from_numpy(self.off).cuda()
        self.d_n = torch.from_numpy(self.n).cuda()
        self.d_lda = torch.from_numpy(self.lda).cuda()
        self.piv_off = np.concatenate([[0], np.cumsum(np.maximum(self.n, 1))[:-1]]).astype(np.int64)
        self.d_piv_off = torch.from_numpy(self.piv_off).cuda()
        self.batch = len(mats)

    def get(self, flat, i):
        n, ld, o = int(self.n[i]), int(self.lda[i]), int(self.off[i])
        return flat[o:o + ld * n].reshape((ld, n), order="F")


def random_sizes(batch, lo, hi, seed):
    return O.rng(seed).integers(lo, hi + 1, size=batch)


@pytest.fixture
def tiles(H):
    """Sets the batched size classes for one test (None: the defaults) and
    restores the defaults afterwards."""
    def set_(small):
        if small is not None:
            assert lib().mifact_set_batched_tiles(H.ptr, small, small) == 0
    yield set_
    assert lib().mifact_set_batched_tiles(H.ptr, 96, 96) == 0


# small = largest trailing order on the 32 x 32 per-wave tiles: 0 keeps every
# update on the 64 x 64 tiles, 4096 puts every update on the small tiles
SIZE_CLASSES = [None, 0, 4096]


@pytest.mark.parametrize("small", SIZE_CLASSES)
@pytest.mark.parametrize("lo,hi,batch", [(1, 96, 40), (32, 512, 24), (600, 1100, 3)])
def test_potrf_batched(H, tiles, lo, hi, batch, small):
    log("potrf_batched", lo, hi, batch, small)
    tiles(small)
    sizes = random_sizes(batch, lo, hi, 100 + hi)
    mats = [O.spd_matrix(int(n), 1000 + i)[:n, :n] for i, n in enumerate(sizes)]
    pk = Packed(mats, [i % 3 for i in range(batch)])
    info = dev_int(batch, 99)
    lib().mifact_potrf_batched(H.ptr, batch, P(pk.dA), P(pk.d_off), P(pk.d_n), P(pk.d_lda),
                               P(info))
    sync()
    assert H.last_error() == 0
    out = pk.dA.cpu().numpy()
    infos = info.cpu().numpy()
    for i, A in enumerate(mats):
        n = A.shape[1]
        assert infos[i] == 0, (i, n)
        F = pk.get(out, i)
        orig = pk.get(pk.flat, i)
        iu = np.triu_indices(n, 1)
        assert np.array_equal(F[:n][iu], orig[:n][iu]), "upper triangle modified"
        assert np.all(F[n:] == orig[n:]), "padding rows modified"
        res = O.potrf_residual(np.asfortranarray(F[:n]), np.asfortranarray(A))
        assert res <= 1e-5 and res <= 10 * n * EPS32, (i, n, res)
        ref = np.asfortranarray(A.copy())
        assert O.spotrf(ref) == 0
        d = np.abs(np.tril(F[:n]) - np.tril(ref)).max()
        assert d <= 100 * n * EPS32 * np.abs(np.tril(ref)).max(), (i, n, d)


def test_potrf_batched_info(H):
    """A matrix that is not positive definite reports LAPACK's info; the
    rest of the batch factors normally."""
    sizes = [50, 130, 77, 200]
    mats = [O.spd_matrix(n, 2000 + i)[:n, :n] for i, n in enumerate(sizes)]
    mats[1][90, 90] = -1.0
    mats[3][0, 0] = 0.0
    pk = Packed(mats, [0, 1, 2, 0])
    info = dev_int(4, 99)
    lib().mifact_potrf_batched(H.ptr, 4, P(pk.dA), P(pk.d_off), P(pk.d_n), P(pk.d_lda), P(info))
    sync()
    infos = info.cpu().numpy()
    want = []
    for A in mats:
        ref = np.asfortranarray(A.copy())
        want.append(O.spotrf(ref))
    assert list(infos) == want, (list(infos), want)
    out = pk.dA.cpu().numpy()
    for i in (0, 2):
        F = pk.get(out, i)
        n = sizes[i]
        assert O.potrf_residual(np.asfortranarray(F[:n]), np.asfortranarray(mats[i])) <= 1e-5


@pytest.mark.parametrize("small", SIZE_CLASSES)
@pytest.mark.parametrize("lo,hi,batch", [(1, 96, 40), (64, 512, 16), (700, 1024, 3)])
def test_getrf_batched(H, tiles, lo, hi, batch, small):
    log("getrf_batched", lo, hi, batch, small)
    tiles(small)
    sizes = random_sizes(batch, lo, hi, 200 + hi)
    mats = [O.uniform_matrix(int(n), 3000 + i) for i, n in enumerate(sizes)]
    pk = Packed(mats, [(i * 5) % 4 for i in range(batch)])
    ipiv = dev_int(int(np.maximum(pk.n, 1).sum()), 0)
    info = dev_int(batch, 99)
    lib().mifact_getrf_batched(H.ptr, batch, P(pk.dA), P(pk.d_off), P(pk.d_n), P(pk.d_lda),
                               P(ipiv), P(pk.d_piv_off), P(info))
    sync()
    assert H.last_error() == 0
    out = pk.dA.cpu().numpy()
    piv = ipiv.cpu().numpy()
    infos = info.cpu().numpy()
    for i, A in enumerate(mats):
        n = A.shape[1]
        assert infos[i] == 0
        F = np.asfortranarray(pk.get(out, i)[:n])
        p = piv[pk.piv_off[i]:pk.piv_off[i] + n]
        res = O.getrf_residual(F, np.asfortranarray(A), p)
        assert res <= 5e-4 and res <= 10 * n * EPS32, (i, n, res)
        ref = np.asfortranarray(A.copy())
        pref, inf = O.sgetrf(ref)
        assert inf == 0
        ok, k, mult = pivot_tie_audit(F, p, pref, n, 4 * n * EPS32)
        assert ok, (i, n, k, mult)


def test_getrf_batched_zero_column(H):
    sizes = [40, 70]
    mats = [O.uniform_matrix(n, 4000 + i) for i, n in enumerate(sizes)]
    mats[1][:, 33] = 0.0
    pk = Packed(mats, [0, 0])
    ipiv = dev_int(int(pk.n.sum()), 0)
    info = dev_int(2, 99)
    lib().mifact_getrf_batched(H.ptr, 2, P(pk.dA), P(pk.d_off), P(pk.d_n), P(pk.d_lda), P(ipiv),
                               P(pk.d_piv_off), P(info))
    sync()
    assert list(info.cpu().numpy()) == [0, 34]


def test_getrf_batched_order_limit(H):
    """Orders above the LU path's 1024 limit get info = -1 and stay
    untouched; the other matrices of the batch factor normally."""
    sizes = [1026, 300]
    mats = [O.uniform_matrix(n, 4100 + i) for i, n in enumerate(sizes)]
    pk = Packed(mats, [0, 0])
    ipiv = dev_int(int(pk.n.sum()), 0)
    info = dev_int(2, 99)
    lib().mifact_getrf_batched(H.ptr, 2, P(pk.dA), P(pk.d_off), P(pk.d_n), P(pk.d_lda), P(ipiv),
                               P(pk.d_piv_off), P(info))
    sync()
    assert H.last_error() == 0
    assert list(info.cpu().numpy()) == [-1, 0]
    out = pk.dA.cpu().numpy()
    assert np.array_equal(pk.get(out, 0), pk.get(pk.flat, 0))
    F = np.asfortranarray(pk.get(out, 1)[:300])
    p = ipiv.cpu().numpy()[pk.piv_off[1]:pk.piv_off[1] + 300]
    assert O.getrf_residual(F, np.asfortranarray(mats[1]), p) <= 5e-4


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_batched_lowp_residual(H, prec):
    u = 2.0 ** -8 if prec == "bf16" else 2.0 ** -11
    sizes = random_sizes(12, 64, 400, 7)
    mats = [O.spd_matrix(int(n), 5000 + i)[:n, :n] for i, n in enumerate(sizes)]
    pk = Packed(mats, [0] * 12)
    info = dev_int(12, 99)
    lib().mifact_potrf_batched_ex(H.ptr, PREC[prec], 12, P(pk.dA), P(pk.d_off), P(pk.d_n),
                                  P(pk.d_lda), P(info))
    sync()
    out = pk.dA.cpu().numpy()
    for i, A in enumerate(mats):
        assert int(info.cpu()[i]) == 0
        res = O.potrf_residual(np.asfortranarray(pk.get(out, i)), np.asfortranarray(A))
        assert res <= 4 * u, (i, res)

@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_batched_lowp_residual_lu(H, prec):
    """BF16/FP16 trailing updates: info 0 and the FP64 residual within 2x
    of a CPU emulation of the same arithmetic (FP32 panels, update operands
    rounded to BF16/FP16, FP32 accumulation; tests/gpu_util.py)."""
    u = 2.0 ** -8 if prec == "bf16" else 2.0 ** -11
    sizes = random_sizes(12, 64, 400, 7)
    info = dev_int(12, 99)
    mats = [O.uniform_matrix(int(n), 6000 + i) for i, n in enumerate(sizes)]
    pk = Packed(mats, [0] * 12)
    ipiv = dev_int(int(pk.n.sum()), 0)
    lib().mifact_getrf_batched_ex(H.ptr, PREC[prec], 12, P(pk.dA), P(pk.d_off), P(pk.d_n),
                                  P(pk.d_lda), P(ipiv), P(pk.d_piv_off), P(info))
    sync()
    out = pk.dA.cpu().numpy()
    piv = ipiv.cpu().numpy()
    infos = info.cpu().numpy()
    for i, A in enumerate(mats):
        n = A.shape[1]
        assert infos[i] == 0, (i, infos[i])
        p = piv[pk.piv_off[i]:pk.piv_off[i] + n]
        F = np.asfortranarray(pk.get(out, i))
        res = O.getrf_residual(F, np.asfortranarray(A), p)
        # the same rounding model on the CPU: the GPU residual within 2x of
        # the emulation's (plus u / 4 for residuals near zero)
        E, pe = emulated_lowp_lu(A, prec)
        emu = O.getrf_residual(np.asfortranarray(E), np.asfortranarray(A), pe)
        assert res <= 2 * emu + u / 4, (i, n, res, emu)


@pytest.mark.parametrize("sym", [1, 0])
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_batched_refine(H, sym, prec):
    import torch
    batch = 64
    sizes = random_sizes(batch, 32, 300, 8 + sym)
    if sym:
        mats = [O.spd_matrix(int(n), 7000 + i)[:n, :n] for i, n in enumerate(sizes)]
    else:
        mats = []
        for i, n in enumerate(sizes):
            a = O.uniform_matrix(int(n), 8000 + i)
            a[np.arange(n), np.arange(n)] += np.float32(np.sqrt(n))   # kappa ~ 10
            mats.append(a)
    pk = Packed(mats, [i % 2 for i in range(batch)])
    F = torch.zeros_like(pk.dA)
    voff = np.concatenate([[0], np.cumsum(pk.n)[:-1]]).astype(np.int64)
    total = int(pk.n.sum())
    r = O.rng(9)
    xt = r.standard_normal(total)
    b = np.zeros(total)
    for i, A in enumerate(mats):
        n = A.shape[1]
        b[voff[i]:voff[i] + n] = A.astype(np.float64) @ xt[voff[i]:voff[i] + n]
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros_like(db)
    dvo = torch.from_numpy(voff).cuda()
    berr = torch.zeros(batch, dtype=torch.float64, device="cuda")
    info = dev_int(batch, 99)
    ipiv = dev_int(int(pk.n.sum()), 0)
    tol = 1e-13
    st = lib().mifact_gesv_batched_refine(H.ptr, sym, PREC[prec], batch, P(pk.dA), P(F),
                                          P(pk.d_off), P(pk.d_n), P(pk.d_lda),
                                          None if sym else P(ipiv),
                                          None if sym else P(pk.d_piv_off), P(db), P(dx),
                                          P(dvo), total, 12, tol, P(berr), P(info))
    sync()
    assert st == 0
    assert np.all(info.cpu().numpy() == 0)
    be = berr.cpu().numpy()
    log("batched refine", sym, prec, "max berr", be.max())
    assert np.all(be <= tol), be.max()
    x = dx.cpu().numpy()
    for i, A in enumerate(mats):
        n = A.shape[1]
        xi = x[voff[i]:voff[i] + n]
        bi = b[voff[i]:voff[i] + n]
        A64 = A.astype(np.float64)
        rr = np.abs(bi - A64 @ xi).max() / (np.abs(A64).sum(1).max() * np.abs(xi).max()
                                            + np.abs(bi).max())
        assert rr <= 2 * tol, (i, n, rr)

"""GPU parity at the benchmark sizes: the large-tile MFMA GEMM configurations
and the FP32 drivers at n = 8192 / 16384 (the benchmark sweep sizes, autotuned
block sizes), plus size-independent properties at n = 32768.

Bars (per test docstring):
  * GEMM: every entry within the componentwise FP32 accumulation bound
    (k + 2) eps |A||op(B)| (+ 2 eps |C|) of the FP64 product; rows and columns
    outside the update untouched;
  * drivers: the CPU FP64 residual of the oracle (BLAS-backed product) within
    the reference's gates (POTRF 1e-5, GETRF 5e-4; run_paper_experiments.sh:
    56-57) and <= 10 n eps32 scaled; ipiv identical to LAPACK sgetrf up to the
    first accepted near-tie (gpu_util.pivot_tie_audit);
  * n = 32768: the GPU FP64 residual within the same gates, info = 0,
    |multipliers| <= 1 (partial pivoting), positive Cholesky diagonal.
"""
import ctypes

import numpy as np
import pytest

from gpu_util import (EPS32, Session, dev, dev_int, gemm_bound, getrf_residual_gpu, host, log,
                      pivot_tie_audit, potrf_residual_gpu, spd_gpu, sync, uniform_gpu)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

POTRF_GATE = 1e-5
GETRF_GATE = 5e-4


@pytest.fixture(scope="module")
def H():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    O.use_blas_for_residual(True)
    return Session.get()


def lib():
    from hygondoubleblockedfactorizations_amd import _lib
    return _lib.load()


def P(t):
    from hygondoubleblockedfactorizations_amd._lib import ptr
    return ptr(t)


class gemm_config:
    """Forces an FP32 GEMM tile configuration on the handle for a block."""

    def __init__(self, H, cfg):
        self.H, self.cfg = H, cfg

    def __enter__(self):
        assert lib().mifact_set_gemm_config(self.H.ptr, self.cfg) == 0

    def __exit__(self, *exc):
        lib().mifact_set_gemm_config(self.H.ptr, 0)


# ------------------------------------------------------------------ GEMM configs

LARGE_SHAPES = [(1100, 1030, 300, 5), (2048, 1024, 520, 0), (1537, 1281, 129, 3)]


@pytest.mark.parametrize("M,N,K,pad", LARGE_SHAPES)
@pytest.mark.parametrize("op", ["nt", "nn"])
@pytest.mark.parametrize("cfg", [1, 2, 4])
def test_gemm_forced_config(H, cfg, op, M, N, K, pad):
    """Config 1 (128x128x16), 2 (128x128x32), 4 (256x256x16) on full NT/NN
    updates with ragged tails and padded leading dimensions."""
    log("gemm cfg", cfg, op, M, N, K, pad)
    r = np.random.default_rng(cfg * 1000 + M + N + K)
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
    ldc = M + 2 * pad + 1
    C = np.asfortranarray(r.standard_normal((ldc, N + 1)).astype(np.float32))
    dA, dB, dC = dev(A), dev(B), dev(C)
    fn = lib().gemm_nt_update if op == "nt" else lib().gemm_nn_update
    with gemm_config(H, cfg):
        fn(H.ptr, M, N, K, P(dA), lda, P(dB), ldb, P(dC), ldc)
        sync()
    assert H.last_error() == 0
    got = host(dC)
    ref = C[:M, :N].astype(np.float64) - A[:M, :K].astype(np.float64) @ Bop
    bound = gemm_bound(A[:M, :K], np.abs(Bop), K) + 2 * EPS32 * np.abs(C[:M, :N])
    assert np.all(np.abs(got[:M, :N] - ref) <= bound)
    assert np.array_equal(got[M:, :], C[M:, :])
    assert np.array_equal(got[:, N:], C[:, N:])


@pytest.mark.parametrize("cfg", [1, 2, 4])
@pytest.mark.parametrize("n,k,pad", [(1100, 300, 5), (2049, 700, 0)])
def test_syrk_forced_config(H, cfg, n, k, pad):
    """Lower-triangle (SYRK) tiles of each large config: diagonal tiles are
    masked, the strictly upper triangle is left untouched."""
    log("syrk cfg", cfg, n, k, pad)
    r = np.random.default_rng(cfg + n + k)
    lda = n + pad
    A = np.asfortranarray(r.standard_normal((lda, k)).astype(np.float32))
    C = np.asfortranarray(r.standard_normal((lda, n)).astype(np.float32))
    dA, dC = dev(A), dev(C)
    with gemm_config(H, cfg):
        lib().syrk_lower_update(H.ptr, n, k, P(dA), lda, P(dC), lda)
        sync()
    assert H.last_error() == 0
    got = host(dC)
    Ad = A[:n].astype(np.float64)
    ref = C[:n].astype(np.float64) - Ad @ Ad.T
    bound = gemm_bound(A[:n], np.abs(Ad.T), k) + 2 * EPS32 * np.abs(C[:n])
    low = np.tril_indices(n)
    up = np.triu_indices(n, 1)
    assert np.all(np.abs(got[:n][low] - ref[low]) <= bound[low])
    assert np.array_equal(got[:n][up], C[:n][up])
    assert np.array_equal(got[n:], C[n:])


def test_gemm_auto_large_tiles(H):
    """Automatic selection: 4096 x 4096 updates take the 128x128 tiles
    (>= 2 tiles per CU); checked NN and NT against FP64."""
    M = N = 4096
    K = 256
    r = np.random.default_rng(4096)
    A = np.asfortranarray(r.standard_normal((M, K)).astype(np.float32))
    B = np.asfortranarray(r.standard_normal((K, N)).astype(np.float32))
    C = np.asfortranarray(r.standard_normal((M, N)).astype(np.float32))
    for op in ("nn", "nt"):
        log("gemm auto", op)
        Bm = B if op == "nn" else np.asfortranarray(B.T)
        Bop = B.astype(np.float64)
        dA, dB, dC = dev(A), dev(Bm), dev(C)
        fn = lib().gemm_nn_update if op == "nn" else lib().gemm_nt_update
        fn(H.ptr, M, N, K, P(dA), M, P(dB), Bm.shape[0], P(dC), M)
        sync()
        got = host(dC)
        ref = C.astype(np.float64) - A.astype(np.float64) @ Bop
        bound = gemm_bound(A, np.abs(Bop), K) + 2 * EPS32 * np.abs(C)
        assert np.all(np.abs(got - ref) <= bound)


# ------------------------------------------------------------------ drivers at 8192 / 16384

def tuned(n):
    from hygondoubleblockedfactorizations_amd import tuning
    return tuning.lookup(n)


def potrf_dev(H, A_buf, nb, b):
    import torch
    work = A_buf.clone()
    info = torch.full((1,), 99, dtype=torch.int32, device="cuda")
    n = A_buf.shape[0]
    lib().double_blocked_potrf(H.ptr, P(work), n, n, nb, b, P(info))
    sync()
    assert H.last_error() == 0
    return work, int(info[0])


def getrf_dev(H, A_buf, nb, b, tb, mode=1):
    import torch
    work = A_buf.clone()
    n = A_buf.shape[0]
    info = torch.full((1,), 99, dtype=torch.int32, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    lib().double_blocked_getrf(H.ptr, P(work), n, n, nb, b, mode, tb, P(ipiv), P(info))
    sync()
    assert H.last_error() == 0
    return work, ipiv.cpu().numpy().copy(), int(info[0])


def to_host(buf):
    """Column-major device buffer -> numpy Fortran array."""
    return np.asfortranarray(buf.cpu().numpy().T)


@pytest.mark.parametrize("nb,b,policy", [(512, 128, 0), (1024, 256, 1), (1024, 256, 2)])
def test_potrf_lookahead_deterministic(H, nb, b, policy):
    """POTRF with many outer blocks, whose later blocks are updated on the
    handle's side stream while the next block's panel runs on the main
    stream: two runs on one handle and one on a fresh handle bound to
    another stream give the same bits (every block receives its updates in
    block order whatever the overlap), and the oracle residual is within the
    1e-5 gate at n = 8192."""
    import torch
    from hygondoubleblockedfactorizations_amd import _lib
    n = 8192
    log("potrf lookahead", n, nb, b, policy)
    assert lib().mifact_set_gemm_policy(H.ptr, policy) == 0
    try:
        Ab = spd_gpu(n, 31337)
        L1, i1 = potrf_dev(H, Ab, nb, b)
        L2, i2 = potrf_dev(H, Ab, nb, b)
        s = torch.cuda.Stream()
        h2 = _lib.Handle(stream=s)
        assert lib().mifact_set_gemm_policy(h2.ptr, policy) == 0
        L3 = Ab.clone()
        info3 = torch.full((1,), 99, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        lib().double_blocked_potrf(h2.ptr, P(L3), n, n, nb, b, P(info3))
        torch.cuda.synchronize()
        assert h2.last_error() == 0
        h2.close()
        assert i1 == i2 == int(info3[0]) == 0
        assert torch.equal(L1, L2) and torch.equal(L1, L3)
        res = O.potrf_residual(to_host(L1), to_host(Ab))
        assert res <= POTRF_GATE and res / (n * EPS32) <= 10, res
    finally:
        lib().mifact_set_gemm_policy(H.ptr, 0)


@pytest.mark.parametrize("n", [8192, 16384])
def test_potrf_sweep_sizes(H, n):
    """Benchmark-sweep POTRF at n = 8192 / 16384 with the autotuned blocks:
    oracle residual (FP64, BLAS product) within 1e-5 and <= 10 n eps
    scaled; at 8192 also L within 100 n eps of LAPACK spotrf."""
    t = tuned(n)
    log("potrf sweep", n, t["potrf_nb"], t["potrf_b"])
    Ab = spd_gpu(n, 20260716 + n)
    Lb, info = potrf_dev(H, Ab, t["potrf_nb"], t["potrf_b"])
    assert info == 0
    A = to_host(Ab)
    Lh = to_host(Lb)
    res = O.potrf_residual(Lh, A)
    log("residual", res, "scaled", res / (n * EPS32))
    assert res <= POTRF_GATE and res / (n * EPS32) <= 10
    if n == 8192:
        Y = A.copy(order="F")
        assert O.spotrf(Y) == 0
        assert np.max(np.abs(np.tril(Lh) - np.tril(Y))) <= 100 * n * EPS32


@pytest.mark.parametrize("n", [8192, 16384])
def test_getrf_sweep_sizes(H, n):
    """Benchmark-sweep GETRF (iid U(-1,1)) at n = 8192 / 16384 with the
    autotuned blocks and the iterative outer TRSM: oracle residual within
    5e-4 and <= 10 n eps scaled; ipiv against LAPACK sgetrf with the tie
    audit (tolerance 4 n eps on the runner-up multiplier)."""
    t = tuned(n)
    log("getrf sweep", n, t["getrf_nb"], t["getrf_b"], t["trsm_block"])
    Ab = uniform_gpu(n, 1234567890 + n)
    LUb, ipiv, info = getrf_dev(H, Ab, t["getrf_nb"], t["getrf_b"], t["trsm_block"])
    assert info == 0
    A = to_host(Ab)
    LU = to_host(LUb)
    res = O.getrf_residual(LU, A, ipiv)
    log("residual", res, "scaled", res / (n * EPS32))
    assert res <= GETRF_GATE and res / (n * EPS32) <= 10
    X = A.copy(order="F")
    ipiv_ref, info_ref = O.sgetrf(X)
    assert info_ref == 0
    ok, k, mult = pivot_tie_audit(LU, ipiv, ipiv_ref, n, tol=4 * n * EPS32)
    log("pivot audit", ok, k, mult)
    assert ok, (k, mult)


@pytest.mark.parametrize("cfg", [1, 2, 4])
def test_drivers_forced_gemm_config(H, cfg):
    """Both drivers at n = 8192 with each large GEMM configuration forced for
    every update: residual gates as above; immediate and batched-pivot GETRF
    bit-identical."""
    n = 8192
    t = tuned(n)
    log("drivers cfg", cfg)
    Ab = spd_gpu(n, 77 + cfg)
    with gemm_config(H, cfg):
        Lb, info = potrf_dev(H, Ab, t["potrf_nb"], t["potrf_b"])
    assert info == 0
    res = O.potrf_residual(to_host(Lb), to_host(Ab))
    log("potrf residual", res)
    assert res <= POTRF_GATE and res / (n * EPS32) <= 10
    Gb = uniform_gpu(n, 99 + cfg)
    with gemm_config(H, cfg):
        LUb, ipiv, info = getrf_dev(H, Gb, t["getrf_nb"], t["getrf_b"], t["trsm_block"])
    assert info == 0
    res = O.getrf_residual(to_host(LUb), to_host(Gb), ipiv)
    log("getrf residual", res)
    assert res <= GETRF_GATE and res / (n * EPS32) <= 10


# ------------------------------------------------------------------ row-move staging

def _batched_ws(n, col_tile, elem_bytes):
    import torch
    from hygondoubleblockedfactorizations_amd._lib import BatchedPivotWorkspace
    src = torch.zeros(n, dtype=torch.int32, device="cuda")
    dst = torch.zeros(n, dtype=torch.int32, device="cuda")
    swap = torch.zeros((n * col_tile * elem_bytes + 7) // 8, dtype=torch.float64, device="cuda")
    hp = [np.zeros(n, dtype=np.int32) for _ in range(4)]
    ws = BatchedPivotWorkspace(P(src), P(dst), P(swap), P(hp[0]), P(hp[1]), P(hp[2]),
                               P(hp[3]), n, col_tile)
    return ws, (src, dst, swap, hp)


@pytest.mark.parametrize("n,nb,b,dtype", [(12288, 12288, 10240, np.float32),
                                          (8192, 6144, 6144, np.float64)])
def test_getrf_row_moves_beyond_lds(H, n, nb, b, dtype):
    """Inner blocks whose move lists exceed the LDS budget (2 b elements per
    column > 64 KiB) are staged through the workspace: the immediate-pivot
    driver must still apply every exchange (residual gate, ipiv audit), and
    the batched-pivot driver with column_tile = 1, whose swap buffer holds
    less than one column of those move lists, must give the same factors
    and pivots bit for bit."""
    import torch
    log("getrf staging", n, nb, b, np.dtype(dtype).name)
    Ab = uniform_gpu(n, 4242 + n)
    if dtype == np.float64:
        Ab = Ab.double()
    A = to_host(Ab)
    L = lib()
    work = Ab.clone()
    info = torch.full((1,), 99, dtype=torch.int32, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    fn = L.double_blocked_getrf_double if dtype == np.float64 else L.double_blocked_getrf
    fn(H.ptr, P(work), n, n, nb, b, 1, 2048, P(ipiv), P(info))
    sync()
    assert H.last_error() == 0 and int(info[0]) == 0
    LU = to_host(work)
    piv = ipiv.cpu().numpy().copy()
    res = O.getrf_residual(LU, A, piv)
    eps = float(np.finfo(dtype).eps)
    log("residual", res, "scaled", res / (n * eps))
    assert res / (n * eps) <= 10
    X = A.copy(order="F")
    ipiv_ref, _ = O.dgetrf(X) if dtype == np.float64 else O.sgetrf(X)
    ok, k, mult = pivot_tie_audit(LU, piv, ipiv_ref, n, tol=4 * n * eps)
    log("pivot audit", ok, k, mult)
    assert ok, (k, mult)
    # batched pivots, one staging column
    work2 = Ab.clone()
    ipiv2 = torch.zeros(n, dtype=torch.int32, device="cuda")
    ws, keep = _batched_ws(n, 1, np.dtype(dtype).itemsize)
    fnb = (L.double_blocked_getrf_double_batched_pivot if dtype == np.float64
           else L.double_blocked_getrf_batched_pivot)
    fnb(H.ptr, P(work2), n, n, nb, b, 1, 2048, 0, P(ipiv2), P(info), ctypes.byref(ws))
    sync()
    assert H.last_error() == 0 and int(info[0]) == 0
    assert torch.equal(ipiv2, ipiv)
    assert torch.equal(work2, work)


def test_leaf_wait_expiry_reported_and_cleared(H):
    """Failure path of the cooperative GETF2 leaf, forced from the host with a
    one-poll wait limit: every kernel still exits; an expiry, when it
    happens, shows as info = -2 and as mifact_status_kernel_timeout from
    mifact_get_last_error, which clears it; the next GETRF on the same
    handle (default limit) is correct."""
    import torch
    n = 8192
    L = lib()
    Ab = uniform_gpu(n, 5150)
    work = Ab.clone()
    info = torch.full((1,), 99, dtype=torch.int32, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    assert L.mifact_set_wait_limit(H.ptr, 1) == 0
    try:
        L.double_blocked_getrf(H.ptr, P(work), n, n, 4096, 2048, 1, 2048, P(ipiv), P(info))
        sync()
    finally:
        L.mifact_set_wait_limit(H.ptr, 0)
    st = H.last_error()
    inf = int(info[0])
    log("forced limit: status", st, "info", inf)
    assert st in (0, 5)
    if st == 5:
        assert inf == -2
    assert H.last_error() == 0          # cleared by the previous read
    LUb, piv, inf2 = getrf_dev(H, Ab, 4096, 2048, 2048)
    assert inf2 == 0
    res = O.getrf_residual(to_host(LUb), to_host(Ab), piv)
    assert res <= GETRF_GATE


# ------------------------------------------------------------------ n = 32768 properties

@pytest.mark.slow
def test_potrf_n32768_properties(H):
    """POTRF at the headline size (autotuned blocks): info = 0, positive
    diagonal, GPU FP64 residual within 1e-5 and <= 10 n eps scaled."""
    import torch
    n = 32768
    t = tuned(n)
    log("potrf 32768", t["potrf_nb"], t["potrf_b"])
    Ab = spd_gpu(n, 20260716)
    Lb, info = potrf_dev(H, Ab, t["potrf_nb"], t["potrf_b"])
    assert info == 0
    assert bool((torch.diagonal(Lb) > 0).all())
    res = potrf_residual_gpu(Ab, Lb)
    log("residual", res, "scaled", res / (n * EPS32))
    assert res <= POTRF_GATE and res / (n * EPS32) <= 10


@pytest.mark.slow
def test_getrf_n32768_properties(H):
    """GETRF at the headline size: info = 0, ipiv a valid LAPACK pivot
    sequence (k < ipiv[k] <= n, 1-based), every multiplier |l_ik| <= 1 + 2 eps
    (partial pivoting), GPU FP64 residual within 5e-4 and <= 10 n eps."""
    import torch
    n = 32768
    t = tuned(n)
    log("getrf 32768", t["getrf_nb"], t["getrf_b"], t["trsm_block"])
    Ab = uniform_gpu(n, 1234567891)
    LUb, ipiv, info = getrf_dev(H, Ab, t["getrf_nb"], t["getrf_b"], t["trsm_block"])
    assert info == 0
    k = np.arange(n)
    assert np.all(ipiv >= k + 1) and np.all(ipiv <= n)
    lmax = float(torch.tril(LUb.T, -1).abs().max())
    log("max multiplier", lmax)
    assert lmax <= 1.0 + 2 * EPS32
    res = getrf_residual_gpu(Ab, LUb, ipiv)
    log("residual", res, "scaled", res / (n * EPS32))
    assert res <= GETRF_GATE and res / (n * EPS32) <= 10

"""CPU replay of the batched LU kernel's address arithmetic on the exact
inputs of the GPU batched-LU tests (tests/replay/batched_lu_replay.cpp).

The replay runs the kernel's phases thread by thread with bounds-checked
accessors on the same packed batches (sizes, padded leading dimensions,
offsets, ipiv offsets) and checks:
  * no access leaves the buffer, the matrix's lda x n block, its n x n part,
    its ipiv segment or the workgroup arrays;
  * the replay computes what the kernel must compute: ipiv identical to
    LAPACK sgetrf and the FP64 residual within the GETRF gate, so the pivot
    sequence that decides the row-exchange addresses is the real one.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "replay", "batched_lu_replay.cpp")


@pytest.fixture(scope="module")
def replay(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("replay") / "batched_lu_replay")
    subprocess.run([gxx, "-O2", "-std=c++17", SRC, "-o", exe], check=True)
    return exe


def packed_case(lo, hi, batch):
    """The packing of tests/test_gpu_batched.py::test_getrf_batched."""
    sizes = O.rng(200 + hi).integers(lo, hi + 1, size=batch)
    mats = [O.uniform_matrix(int(n), 3000 + i) for i, n in enumerate(sizes)]
    pads = [(i * 5) % 4 for i in range(batch)]
    n = np.array([m.shape[1] for m in mats], np.int32)
    lda = np.array([m.shape[1] + p for m, p in zip(mats, pads)], np.int32)
    sizes_el = lda.astype(np.int64) * np.maximum(n, 1)
    off = np.concatenate([[0], np.cumsum(sizes_el)[:-1]]).astype(np.int64)
    flat = np.full(int(sizes_el.sum()), 7.25, np.float32)
    for m, o, ld in zip(mats, off, lda):
        k = m.shape[1]
        blk = np.full((ld, k), 7.25, np.float32, order="F")
        blk[:k, :] = m
        flat[o:o + ld * k] = blk.ravel(order="F")
    piv_off = np.concatenate([[0], np.cumsum(np.maximum(n, 1))[:-1]]).astype(np.int64)
    return mats, n, lda, off, piv_off, flat


@pytest.mark.parametrize("small_max", [0, 96, 1024])
@pytest.mark.parametrize("lo,hi,batch", [(1, 96, 40), (64, 512, 16), (700, 1024, 3)])
def test_batched_lu_addresses_in_range(replay, tmp_path, lo, hi, batch, small_max):
    """small_max: largest trailing order updated by the small size class's
    32 x 32 per-wave tiles (0: none, 96: the default, 1024: every update)."""
    mats, n, lda, off, piv_off, flat = packed_case(lo, hi, batch)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(struct.pack("<iq", batch, flat.size))
        for a in (n, lda, off, piv_off, flat):
            f.write(a.tobytes())
    r = subprocess.run([replay, str(inp), str(out), str(small_max)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "out-of-range accesses 0" in r.stdout
    raw = open(out, "rb").read()
    F = np.frombuffer(raw, np.float32, flat.size, 0)
    ptot = int(np.maximum(n, 1).sum())
    piv = np.frombuffer(raw, np.int32, ptot, 4 * flat.size)
    info = np.frombuffer(raw, np.int32, batch, 4 * (flat.size + ptot))
    for i, A in enumerate(mats):
        k, ld, o = int(n[i]), int(lda[i]), int(off[i])
        Fi = np.asfortranarray(F[o:o + ld * k].reshape((ld, k), order="F")[:k])
        blk = flat[o:o + ld * k].reshape((ld, k), order="F")
        assert np.all(F[o:o + ld * k].reshape((ld, k), order="F")[k:] == blk[k:])
        p = piv[piv_off[i]:piv_off[i] + k].copy()
        ref = np.asfortranarray(A.copy())
        pref, inf = O.sgetrf(ref)
        assert inf == 0 and info[i] == 0
        assert np.array_equal(p, pref), i
        assert O.getrf_residual(Fi, np.asfortranarray(A), p) <= 5e-4


def test_batched_lu_order_limit(replay, tmp_path):
    """Orders above the LU path's 1024 limit report info = -1 and leave the
    matrix untouched; the rest of the batch factors normally."""
    mats, n, lda, off, piv_off, flat = packed_case(1020, 1030, 3)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(struct.pack("<iq", 3, flat.size))
        for a in (n, lda, off, piv_off, flat):
            f.write(a.tobytes())
    r = subprocess.run([replay, str(inp), str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(out, "rb").read()
    F = np.frombuffer(raw, np.float32, flat.size, 0)
    ptot = int(np.maximum(n, 1).sum())
    info = np.frombuffer(raw, np.int32, 3, 4 * (flat.size + ptot))
    for i in range(3):
        o, k, ld = int(off[i]), int(n[i]), int(lda[i])
        if k > 1024:
            assert info[i] == -1
            assert np.array_equal(F[o:o + ld * k], flat[o:o + ld * k])
        else:
            assert info[i] == 0

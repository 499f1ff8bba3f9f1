"""CPU replay of the batched Cholesky kernel's address arithmetic on the exact
inputs of the GPU batched-Cholesky tests (tests/replay/batched_chol_replay.cpp).

The replay runs the kernel's phases thread by thread (POTF2 lane by lane on
wave 0, the row-per-thread TRSM, the 64 x 64 tiles with their k slices, the
small size class's per-wave tiles) with bounds-checked accessors on the same
packed batches (sizes, padded leading dimensions, offsets) and checks:
  * no access leaves the buffer, the matrix's lda x n block or its n x n
    part, no write lands above the diagonal, no workgroup-array index leaves
    its array;
  * the replay computes what the kernel must compute: the FP64 residual
    within the POTRF gate and LAPACK spotrf's info, so the schedule (which
    columns each step touches) is the real one.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "replay", "batched_chol_replay.cpp")
EPS32 = float(np.finfo(np.float32).eps)


@pytest.fixture(scope="module")
def replay(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("replay") / "batched_chol_replay")
    subprocess.run([gxx, "-O2", "-std=c++17", SRC, "-o", exe], check=True)
    return exe


def packed_case(mats, pads):
    """The packing of tests/test_gpu_batched.py::Packed."""
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
    return n, lda, off, flat


def run(replay, tmp_path, mats, pads, small_max):
    n, lda, off, flat = packed_case(mats, pads)
    batch = len(mats)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(struct.pack("<iq", batch, flat.size))
        for a in (n, lda, off, flat):
            f.write(a.tobytes())
    r = subprocess.run([replay, str(inp), str(out), str(small_max)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "out-of-range accesses 0" in r.stdout
    raw = open(out, "rb").read()
    F = np.frombuffer(raw, np.float32, flat.size, 0)
    info = np.frombuffer(raw, np.int32, batch, 4 * flat.size)
    return n, lda, off, flat, F, info


@pytest.mark.parametrize("small_max", [0, 96, 4096])
@pytest.mark.parametrize("lo,hi,batch", [(1, 96, 40), (32, 512, 24), (600, 1100, 3)])
def test_batched_chol_addresses_in_range(replay, tmp_path, lo, hi, batch, small_max):
    """small_max: largest trailing order updated by the small size class's
    32 x 32 per-wave tiles (0: none, 96, 4096: every update).  The sizes and
    pads are those of tests/test_gpu_batched.py::test_potrf_batched."""
    sizes = O.rng(100 + hi).integers(lo, hi + 1, size=batch)
    mats = [O.spd_matrix(int(k), 1000 + i)[:k, :k] for i, k in enumerate(sizes)]
    n, lda, off, flat, F, info = run(replay, tmp_path, mats, [i % 3 for i in range(batch)],
                                     small_max)
    for i, A in enumerate(mats):
        k, ld, o = int(n[i]), int(lda[i]), int(off[i])
        blk_in = flat[o:o + ld * k].reshape((ld, k), order="F")
        blk = F[o:o + ld * k].reshape((ld, k), order="F")
        assert info[i] == 0, (i, k)
        iu = np.triu_indices(k, 1)
        assert np.array_equal(blk[:k][iu], blk_in[:k][iu])
        assert np.all(blk[k:] == blk_in[k:])
        res = O.potrf_residual(np.asfortranarray(blk[:k]), np.asfortranarray(A))
        assert res <= 1e-5 and res <= 10 * k * EPS32, (i, k, res)


def test_batched_chol_info(replay, tmp_path):
    """Non-positive pivots (in the first and in the second panel of an
    outer block, and at column 0) give LAPACK's info; the rest of the
    batch factors normally (tests/test_gpu_batched.py::test_potrf_batched_info
    plus a pivot in a second panel)."""
    sizes = [50, 130, 77, 200, 100]
    mats = [O.spd_matrix(k, 2000 + i)[:k, :k] for i, k in enumerate(sizes)]
    mats[1][90, 90] = -1.0
    mats[3][0, 0] = 0.0
    mats[4][40, 40] = -1.0
    n, lda, off, flat, F, info = run(replay, tmp_path, mats, [0, 1, 2, 0, 1], 0)
    want = []
    for A in mats:
        ref = np.asfortranarray(A.copy())
        want.append(O.spotrf(ref))
    assert list(info) == want, (list(info), want)
    for i in (0, 2):
        k, ld, o = int(n[i]), int(lda[i]), int(off[i])
        blk = F[o:o + ld * k].reshape((ld, k), order="F")
        assert O.potrf_residual(np.asfortranarray(blk[:k]), np.asfortranarray(mats[i])) <= 1e-5

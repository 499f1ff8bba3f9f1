"""Helpers for the GPU tests: column-major host <-> device copies and a
session handle.  torch supplies device memory and the stream only."""
import sys

import numpy as np

from hygondoubleblockedfactorizations_amd import _lib

EPS32 = float(np.finfo(np.float32).eps)
EPS64 = float(np.finfo(np.float64).eps)


def log(*args):
    print("[gpu]", *args, flush=True)
    sys.stdout.flush()


def dev(a):
    """numpy column-major (lda x n) -> torch CUDA tensor with the same layout."""
    import torch
    a = np.asfortranarray(a)
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


def host(t):
    """torch CUDA tensor made by dev() -> numpy column-major array."""
    return np.asfortranarray(t.cpu().numpy().T)


def dev_int(n, fill=0):
    import torch
    return torch.full((max(n, 1),), fill, dtype=torch.int32, device="cuda")


def sync():
    import torch
    torch.cuda.synchronize()


class Session:
    handle = None

    @classmethod
    def get(cls):
        import torch
        if cls.handle is None:
            cls.handle = _lib.Handle(stream=torch.cuda.current_stream())
        return cls.handle


def gemm_bound(A, B, k):
    """Componentwise FP32 accumulation bound (k + 2) eps |A| |B|."""
    return (k + 2) * EPS32 * (np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64)))


def pivot_tie_audit(LU, ipiv, ipiv_ref, n, tol):
    """ipiv parity with the tie rule: identical, or at the FIRST mismatch k
    the row the reference chose must sit, in our factorization, at a
    position i > k whose multiplier |L(i, k)| >= 1 - tol, i.e. its candidate
    magnitude ties ours within the accumulated rounding bound.  After an
    accepted tie the trajectories legitimately differ (checked by residual).
    Returns (ok, k or None, multiplier or None)."""
    mism = np.nonzero(np.asarray(ipiv[:n]) != np.asarray(ipiv_ref[:n]))[0]
    if len(mism) == 0:
        return True, None, None
    k = int(mism[0])
    q = int(ipiv_ref[k]) - 1            # reference's row position before step k
    pos = q
    for s_ in range(k, n):              # follow that row through our swaps
        p = int(ipiv[s_]) - 1
        if pos == s_:
            pos = p
        elif pos == p:
            pos = s_
    mult = abs(float(LU[pos, k])) if pos > k else 0.0
    return mult >= 1.0 - tol, k, mult

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

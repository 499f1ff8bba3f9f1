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


def spd_gpu(n, seed):
    """A = G G^T / n + I (G iid N(0,1)) built in FP64 on the GPU and rounded
    once to FP32: the same family as the oracle's spd_matrix, for orders
    where a host FP64 product would dominate a test.  Returns the torch
    tensor t with A = t.T (column-major buffer), exactly symmetric."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    G = torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64)
    A = G @ G.T
    del G
    A.div_(float(n))
    A.diagonal().add_(1.0)
    A = 0.5 * (A + A.T)
    return A.float().contiguous()


def uniform_gpu(n, seed):
    """iid Uniform(-1, 1) FP32 on the GPU (column-major buffer t, A = t.T)."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    A = torch.rand(n, n, generator=g, device="cuda", dtype=torch.float32)
    return A.mul_(2.0).sub_(1.0)


def potrf_residual_gpu(A_buf, L_buf, bs=8192):
    """||A - L L^T||_F / ||A||_F with FP64 products on the GPU; A_buf, L_buf
    are column-major buffers (torch tensors t with A = t.T).  A property
    check for orders the host oracle cannot reach in seconds."""
    import torch
    n = A_buf.shape[0]
    num = 0.0
    den = 0.0
    for c0 in range(0, n, bs):
        c1 = min(n, c0 + bs)
        acc = A_buf[c0:c1, :].T.double()           # A(:, c0:c1)
        den += float((acc ** 2).sum())
        Lr = torch.tril(L_buf.T[:, :c1].double())  # L(:, 0:c1)
        Lc = torch.tril(L_buf.T[c0:c1, :c1].double(), diagonal=c0)
        acc -= Lr @ Lc.T
        num += float((acc ** 2).sum())
        del acc, Lr, Lc
    return float(np.sqrt(num / den))


def getrf_residual_gpu(A_buf, LU_buf, ipiv, bs=8192):
    """||P A - L U||_F / ||A||_F with FP64 products on the GPU (blocked over
    column and k blocks); ipiv 1-based global (numpy)."""
    import torch
    n = A_buf.shape[0]
    perm = np.arange(n)
    piv = np.asarray(ipiv)[:n] - 1
    for k in range(n):
        p = piv[k]
        if p != k:
            perm[k], perm[p] = perm[p], perm[k]
    permd = torch.from_numpy(perm).cuda()
    X = LU_buf.T
    M = A_buf.T
    num = 0.0
    den = 0.0
    for c0 in range(0, n, bs):
        c1 = min(n, c0 + bs)
        acc = M[permd, c0:c1].double()
        den += float((M[:, c0:c1].double() ** 2).sum())
        for k0 in range(0, c1, bs):
            k1 = min(c1, k0 + bs)
            Lk = X[:, k0:k1].double()
            rows = torch.arange(n, device=Lk.device).unsqueeze(1)
            cols = torch.arange(k0, k1, device=Lk.device).unsqueeze(0)
            Lk = torch.where(rows > cols, Lk, torch.where(rows == cols, 1.0, 0.0))
            Uk = X[k0:k1, c0:c1].double()
            ur = torch.arange(k0, k1, device=Uk.device).unsqueeze(1)
            uc = torch.arange(c0, c1, device=Uk.device).unsqueeze(0)
            Uk = torch.where(ur <= uc, Uk, 0.0)
            acc -= Lk @ Uk
            del Lk, Uk
        num += float((acc ** 2).sum())
        del acc
    return float(np.sqrt(num / den))


def round_to(x, prec):
    """x (FP32 array) rounded to BF16 or FP16 and back (nearest even)."""
    import torch
    t = torch.from_numpy(np.ascontiguousarray(x, np.float32))
    dt = torch.bfloat16 if prec == "bf16" else torch.float16
    return t.to(dt).to(torch.float32).numpy()


def emulated_lowp_lu(A, prec, nb=32):
    """CPU emulation of the batched kernel's mixed-precision LU: 32-column
    panels factored in FP32 with partial pivoting (LAPACK order), the
    panel's interchanges applied to the other columns, U12 = L11^{-1} A12 in
    FP32, and A22 -= L21 U12 with both operands rounded to `prec` and the
    products accumulated in FP32.  Returns (LU, ipiv 1-based)."""
    a = np.array(A, np.float32, order="F")
    n = a.shape[1]
    ipiv = np.zeros(n, np.int32)
    for k in range(0, n, nb):
        kb = min(nb, n - k)
        for j in range(k, k + kb):
            p = j + int(np.argmax(np.abs(a[j:, j])))
            ipiv[j] = p + 1
            if p != j:
                a[[j, p], :] = a[[p, j], :]
            piv = a[j, j]
            if piv != 0:
                a[j + 1:, j] = a[j + 1:, j] * np.float32(1.0 / piv) if abs(piv) >= np.finfo(np.float32).tiny else a[j + 1:, j] / piv
                a[j + 1:, j + 1:k + kb] -= np.outer(a[j + 1:, j], a[j, j + 1:k + kb]).astype(np.float32)
        c1 = k + kb
        if c1 < n:
            L11 = np.tril(a[k:c1, k:c1], -1).astype(np.float64) + np.eye(kb)
            a[k:c1, c1:] = np.linalg.solve(L11, a[k:c1, c1:].astype(np.float64)).astype(np.float32)
            L21 = round_to(a[c1:, k:c1], prec).astype(np.float64)
            U12 = round_to(a[k:c1, c1:], prec).astype(np.float64)
            a[c1:, c1:] = (a[c1:, c1:].astype(np.float64) - L21 @ U12).astype(np.float32)
    return a, ipiv

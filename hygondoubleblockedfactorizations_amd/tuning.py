"""Autotuner for the double-blocked drivers' block sizes and update tiles.

The reference picks nb/b by hand per run (its benchmark scripts pass them on
the command line; potrf_double_blocked_cases.cpp:12-114,
getrf_double_blocked_cases.cpp:13-109).  Here the outer/inner block sizes of
POTRF and GETRF, the GETRF outer-TRSM block and, at the chosen blocks, the
MFMA tile-selection policy of the update GEMMs (mifact_set_gemm_policy:
which updates take 256 x 256, 128 x 128 or 64 x 64 tiles) are measured on
the device: every candidate is timed (HIP events, median of `reps` runs
after one warm-up) on the same seeded inputs as the benchmark, and the
fastest is stored in ``tuning_table.json`` next to this file, keyed by n.
``lookup(n)`` returns the entry of the nearest tuned order (block sizes
clipped to n), or the built-in defaults when no table exists.

The batched factorizations' size classes (mifact_set_batched_tiles: the
largest trailing order updated in 32 x 32 per-wave tiles rather than 64 x 64
tiles on four waves) are tuned the same way on the batched workloads --
4096 SPD matrices with n ~ U[32, 512] and 2048 general matrices with
n ~ U[64, 1024] -- and stored in ``tuning_batched.json``
(``lookup_batched()``).

    python -m hygondoubleblockedfactorizations_amd.tuning --n 8192 16384 32768
    python -m hygondoubleblockedfactorizations_amd.tuning --batched
"""
from __future__ import annotations

import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning_table.json")
BATCHED_TABLE = os.path.join(HERE, "tuning_batched.json")

# Candidate grids (outer, inner) per routine; the outer-TRSM block of GETRF.
POTRF_GRID = [(1024, 256), (1536, 256), (2048, 128), (2048, 256), (4096, 256), (4096, 512),
              (8192, 256), (8192, 512), (8192, 1024), (12288, 512), (16384, 512), (16384, 1024)]
GETRF_GRID = [(4096, 2048), (4096, 4096), (6144, 3072), (8192, 2048), (8192, 4096),
              (8192, 8192), (12288, 4096), (16384, 4096), (16384, 8192)]
TRSM_BLOCKS = [2048, 4096]
GEMM_POLICIES = [0, 1, 2, 3]   # mifact_set_gemm_policy
# largest trailing order of the small size class (mifact_set_batched_tiles)
BATCHED_SMALL = [0, 64, 96, 128, 192, 256, 384]
BATCHED_DEFAULT = {"potrf_small_max": 96, "getrf_small_max": 96}
# (routine, batch, smallest order, largest order) of the batched workloads
BATCHED_WORKLOADS = [("potrf", 4096, 32, 512), ("getrf", 2048, 64, 1024)]


def defaults(n: int) -> dict:
    """Block sizes used when nothing is tuned for n."""
    return {
        "potrf_nb": min(8192, n), "potrf_b": min(1024, n),
        "getrf_nb": min(max(n // 2, 1), 20480), "getrf_b": min(max(n // 4, 1), 4096),
        "trsm_block": 2048, "potrf_gemm_policy": 0, "getrf_gemm_policy": 0,
    }


def load_table(path: str = TABLE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return {int(k): v for k, v in json.load(f).items()}


def lookup(n: int, path: str = TABLE) -> dict:
    """Tuned block sizes for order n (nearest tuned order, clipped to n)."""
    table = load_table(path)
    if not table:
        return defaults(n)
    key = min(table, key=lambda m: abs(m - n))
    e = dict(table[key])
    for k in ("potrf_nb", "potrf_b", "getrf_nb", "getrf_b"):
        e[k] = max(1, min(int(e[k]), n))
    e["potrf_b"] = min(e["potrf_b"], e["potrf_nb"])
    e["getrf_b"] = min(e["getrf_b"], e["getrf_nb"])
    e.setdefault("potrf_gemm_policy", 0)
    e.setdefault("getrf_gemm_policy", 0)
    return {k: e[k] for k in ("potrf_nb", "potrf_b", "getrf_nb", "getrf_b", "trsm_block",
                              "potrf_gemm_policy", "getrf_gemm_policy")}


def _time(torch, fn, reset, reps):
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    out = []
    for i in range(reps + 1):
        reset()
        torch.cuda.synchronize()
        ev0.record()
        fn()
        ev1.record()
        torch.cuda.synchronize()
        if i > 0:
            out.append(ev0.elapsed_time(ev1))
    out.sort()
    return out[len(out) // 2]


def tune(n: int, reps: int = 2, log=print) -> dict:
    """Times the candidate grids at order n on the current GPU."""
    import torch
    from . import _lib
    lib = _lib.load()
    h = _lib.Handle(stream=torch.cuda.current_stream())
    h.reserve(n)
    P = _lib.ptr
    g = torch.Generator(device="cuda")
    g.manual_seed(20260716)
    G = torch.randn(n, n, device="cuda", generator=g)
    spd = torch.addmm(torch.eye(n, device="cuda"), G, G.t(), alpha=1.0 / n)
    del G
    gen = torch.rand(n, n, device="cuda", generator=g) * 2 - 1
    work = torch.empty_like(spd)
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    best = {}
    res = []
    for nb, b in POTRF_GRID:
        if nb > n or b > nb:
            continue
        ms = _time(torch, lambda: lib.double_blocked_potrf(h.ptr, P(work), n, n, nb, b, P(info)),
                   lambda: work.copy_(spd), reps)
        res.append((ms, nb, b))
        log(f"  potrf n={n} nb={nb} b={b}: {ms:.2f} ms")
    ms, nb, b = min(res)
    best.update(potrf_nb=nb, potrf_b=b, potrf_ms=ms)
    res = []
    for nb, b in GETRF_GRID:
        if nb > n or b > nb:
            continue
        for tb in TRSM_BLOCKS:
            if tb > nb:
                continue
            ms = _time(torch, lambda: lib.double_blocked_getrf(h.ptr, P(work), n, n, nb, b, 1, tb,
                                                               P(ipiv), P(info)),
                       lambda: work.copy_(gen), reps)
            res.append((ms, nb, b, tb))
            log(f"  getrf n={n} nb={nb} b={b} tb={tb}: {ms:.2f} ms")
    if res:
        ms, nb, b, tb = min(res)
        best.update(getrf_nb=nb, getrf_b=b, trsm_block=tb, getrf_ms=ms)
    else:
        d = defaults(n)
        best.update(getrf_nb=d["getrf_nb"], getrf_b=d["getrf_b"], trsm_block=d["trsm_block"])
    # update-tile policy at the chosen blocks
    for routine in ("potrf", "getrf"):
        res = []
        for pol in GEMM_POLICIES:
            lib.mifact_set_gemm_policy(h.ptr, pol)
            if routine == "potrf":
                nb, b = best["potrf_nb"], best["potrf_b"]
                ms = _time(torch, lambda: lib.double_blocked_potrf(h.ptr, P(work), n, n, nb, b,
                                                                   P(info)),
                           lambda: work.copy_(spd), reps)
            else:
                nb, b, tb = best["getrf_nb"], best["getrf_b"], best["trsm_block"]
                ms = _time(torch, lambda: lib.double_blocked_getrf(h.ptr, P(work), n, n, nb, b, 1,
                                                                   tb, P(ipiv), P(info)),
                           lambda: work.copy_(gen), reps)
            res.append((ms, pol))
            log(f"  {routine} n={n} gemm policy {pol}: {ms:.2f} ms")
        lib.mifact_set_gemm_policy(h.ptr, 0)
        ms, pol = min(res)
        best[f"{routine}_gemm_policy"] = pol
        best[f"{routine}_ms"] = ms
    if h.last_error() != 0:
        raise RuntimeError("library reported an error while tuning")
    return best


def lookup_batched(path: str = BATCHED_TABLE) -> dict:
    """Tuned size classes of the batched factorizations (the library's
    defaults when nothing is tuned)."""
    if not os.path.exists(path):
        return dict(BATCHED_DEFAULT)
    with open(path) as f:
        e = json.load(f)
    return {k: int(e.get(k, v)) for k, v in BATCHED_DEFAULT.items()}


def batched_inputs(torch, routine, batch, lo, hi, seed):
    """Packed column-major batch (lda = n) of the batched workloads: SPD
    A = G G^T / n + I, or iid U(-1, 1); returns (flat, offsets, n, ipiv
    offsets) as device tensors and the flop count."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    ns = rng.integers(lo, hi + 1, size=batch).astype(np.int32)
    sizes = ns.astype(np.int64) ** 2
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    if routine == "potrf":
        flat = torch.empty(int(sizes.sum()), dtype=torch.float32, device="cuda")
        for i, n in enumerate(ns):
            n = int(n)
            G = torch.randn(n, n, device="cuda", generator=g, dtype=torch.float64)
            a = (G @ G.t() / n + torch.eye(n, device="cuda", dtype=torch.float64)).float()
            flat[offs[i]:offs[i] + n * n] = a.reshape(-1)
        flops = float((ns.astype(np.float64) ** 3).sum() / 3.0)
    else:
        flat = torch.rand(int(sizes.sum()), dtype=torch.float32, device="cuda",
                          generator=g).mul_(2).sub_(1)
        flops = float((2.0 * ns.astype(np.float64) ** 3).sum() / 3.0)
    poff = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
    return (flat, torch.from_numpy(offs).cuda(), torch.from_numpy(ns).cuda(),
            torch.from_numpy(poff).cuda(), flops)


def tune_batched(reps: int = 3, log=print) -> dict:
    """Times every size-class threshold on the batched workloads."""
    import torch
    from . import _lib
    lib = _lib.load()
    h = _lib.Handle(stream=torch.cuda.current_stream())
    P = _lib.ptr
    best = {}
    for routine, batch, lo, hi in BATCHED_WORKLOADS:
        flat, d_off, d_n, d_poff, flops = batched_inputs(torch, routine, batch, lo, hi,
                                                         20260716 + (0 if routine == "potrf" else 1))
        work = torch.empty_like(flat)
        info = torch.zeros(batch, dtype=torch.int32, device="cuda")
        ipiv = torch.zeros(int(d_n.sum().item()), dtype=torch.int32, device="cuda")
        res = []
        for small in BATCHED_SMALL:
            other = BATCHED_DEFAULT["getrf_small_max" if routine == "potrf" else "potrf_small_max"]
            if routine == "potrf":
                lib.mifact_set_batched_tiles(h.ptr, small, other)
                fn = lambda: lib.mifact_potrf_batched(h.ptr, batch, P(work), P(d_off), P(d_n),
                                                      P(d_n), P(info))
            else:
                lib.mifact_set_batched_tiles(h.ptr, other, small)
                fn = lambda: lib.mifact_getrf_batched(h.ptr, batch, P(work), P(d_off), P(d_n),
                                                      P(d_n), P(ipiv), P(d_poff), P(info))
            ms = _time(torch, fn, lambda: work.copy_(flat), reps)
            res.append((ms, small))
            log(f"  batched {routine} small_max={small}: {ms:.3f} ms "
                f"({flops / ms / 1e9:.2f} TFLOP/s)")
        ms, small = min(res)
        best[f"{routine}_small_max"] = small
        best[f"{routine}_ms"] = ms
        best[f"{routine}_candidates_ms"] = {str(s): m for m, s in res}
    if h.last_error() != 0:
        raise RuntimeError("library reported an error while tuning")
    best["workload"] = ("4096 SPD n~U[32,512] (A = G G^T/n + I); 2048 general n~U[64,1024] "
                        "(iid U(-1,1)); FP32, lda = n")
    return best


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, nargs="+", default=[8192, 16384, 32768])
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--batched", action="store_true",
                    help="tune the batched size classes instead (tuning_batched.json)")
    args = ap.parse_args(argv)
    if args.batched:
        best = tune_batched(max(args.reps, 3), log=lambda s: print(s, flush=True))
        print(f"  best: {best}", flush=True)
        with open(BATCHED_TABLE, "w") as f:
            json.dump(best, f, indent=1)
        return 0
    table = load_table(args.out)
    for n in args.n:
        print(f"tuning n={n}", flush=True)
        table[n] = tune(n, args.reps, log=lambda s: print(s, flush=True))
        print(f"  best: {table[n]}", flush=True)
        with open(args.out, "w") as f:
            json.dump({str(k): v for k, v in sorted(table.items())}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())

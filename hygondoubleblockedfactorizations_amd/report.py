"""Raw timing CSV in the reference's layout.

The reference's benchmark runner writes one row per timed iteration to the
file named by ``BENCHMARK_CSV_PATH`` (src/utils/benchmark/benchmark_printer.cpp:
26-68) with the columns below; archived results use that layout.  Same
columns, quoting and environment variable here, so both can be compared
directly.  Residual metric names follow
include/utils/benchmark/residual_metric.h:11-23, plus
``gpu_fp64_full_relative_residual`` for the same full FP64 relative residual
computed on the GPU.
"""
from __future__ import annotations

import os
import statistics

COLUMNS = ("benchmark,method,parameters,iteration,time_ms,mean_ms,median_ms,stddev_ms,"
           "min_ms,max_ms,gflops,speedup,forward_error,residual_metric,residual_value")


def _quote(s: str) -> str:
    return '"' + str(s).replace('"', '""') + '"'


def timing_rows(title, results):
    """results: list of dicts with method, parameters, times_ms (list),
    flops, and optional speedup, forward_error, residual_metric,
    residual_value.  Returns the CSV lines (header first)."""
    out = [COLUMNS]
    for r in results:
        t = list(r["times_ms"])
        mean = statistics.fmean(t)
        med = statistics.median(t)
        sd = statistics.pstdev(t) if len(t) > 1 else 0.0
        gflops = r["flops"] / (mean * 1e-3) / 1e9 if mean > 0 else 0.0
        for i, ms in enumerate(t):
            out.append(",".join([
                _quote(title), _quote(r["method"]), _quote(r["parameters"]), str(i),
                repr(float(ms)), repr(mean), repr(med), repr(sd), repr(min(t)), repr(max(t)),
                repr(gflops), repr(float(r.get("speedup", 0.0))),
                repr(float(r.get("forward_error", 0.0))), r.get("residual_metric", "none"),
                repr(float(r.get("residual_value", 0.0)))]))
    return out


def write_timing_csv(title, results, path=None):
    """Writes the CSV to ``path`` or to $BENCHMARK_CSV_PATH; no-op if unset."""
    path = path or os.environ.get("BENCHMARK_CSV_PATH", "")
    if not path:
        return None
    with open(path, "w") as f:
        f.write("\n".join(timing_rows(title, results)) + "\n")
    return path

"""The raw timing CSV keeps the reference's columns, quoting and per-row
statistics (src/utils/benchmark/benchmark_printer.cpp:41-63)."""
import csv
import io

from hygondoubleblockedfactorizations_amd import report


def test_columns_and_rows(tmp_path):
    res = [{"method": 'double_blocked "potrf"', "parameters": "nb=8192,b=256",
            "times_ms": [2.0, 4.0, 3.0], "flops": 3e9, "residual_metric":
            "gpu_fp64_full_relative_residual", "residual_value": 1e-7}]
    path = report.write_timing_csv("POTRF n=1024", res, path=str(tmp_path / "t.csv"))
    rows = list(csv.reader(open(path)))
    assert rows[0] == report.COLUMNS.split(",")
    assert len(rows) == 4
    r = dict(zip(rows[0], rows[2]))
    assert r["method"] == 'double_blocked "potrf"' and r["iteration"] == "1"
    assert float(r["time_ms"]) == 4.0 and float(r["mean_ms"]) == 3.0
    assert float(r["median_ms"]) == 3.0 and float(r["min_ms"]) == 2.0
    assert abs(float(r["gflops"]) - 1000.0) < 1e-9
    assert r["residual_metric"] == "gpu_fp64_full_relative_residual"


def test_env_unset_is_noop(monkeypatch):
    monkeypatch.delenv("BENCHMARK_CSV_PATH", raising=False)
    assert report.write_timing_csv("x", []) is None

"""Host logic of the block-size autotuner (no GPU): table lookup, nearest
order, clipping to n, defaults."""
import json

from hygondoubleblockedfactorizations_amd import tuning


def test_defaults_without_table(tmp_path):
    d = tuning.lookup(2048, path=str(tmp_path / "none.json"))
    assert d == tuning.defaults(2048)
    assert d["potrf_nb"] == 2048 and d["potrf_b"] == 1024


def test_nearest_entry_and_clipping(tmp_path):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({
        "8192": {"potrf_nb": 4096, "potrf_b": 512, "getrf_nb": 8192, "getrf_b": 2048,
                 "trsm_block": 1024},
        "32768": {"potrf_nb": 8192, "potrf_b": 1024, "getrf_nb": 16384, "getrf_b": 4096,
                  "trsm_block": 2048}}))
    assert tuning.lookup(30000, path=str(p))["getrf_nb"] == 16384
    e = tuning.lookup(1000, path=str(p))
    assert e["potrf_nb"] == 1000 and e["potrf_b"] == 512 and e["getrf_b"] == 1000
    assert e["trsm_block"] == 1024


def test_committed_table_is_valid():
    t = tuning.load_table()
    for n, e in t.items():
        for k in ("potrf_nb", "potrf_b", "getrf_nb", "getrf_b", "trsm_block"):
            assert int(e[k]) > 0
        assert e["potrf_b"] <= e["potrf_nb"] <= n
        assert e["getrf_b"] <= e["getrf_nb"] <= n

"""The autotuner's search and cache logic on the CPU (simulated timings)."""
import pytest

from sleek_amd import autotune as at


def fake_measure(costs, digests=None):
    calls = []

    def measure(cfg):
        calls.append(dict(cfg))
        e = 1.0 - 0.1 * (cfg["encode_variant"] == 3) - 0.05 * (cfg["lb_window"] == 8)
        d = 1.0 - 0.2 * (cfg["decode_grid_mult"] == 2.0 and cfg["decoder"] == 0)
        dig = "same" if digests is None else digests(cfg)
        return e * costs, d * costs, dig
    return measure, calls


def test_search_picks_fastest_per_coordinate():
    measure, calls = fake_measure(1e-3)
    best, trace = at.search(measure)
    assert best["encode_variant"] == 3 and best["lb_window"] == 8 and best["decode_grid_mult"] == 2.0
    assert calls[0] == at.DEFAULT and len(trace) == len(calls)


def test_search_raises_when_bytes_change():
    measure, _ = fake_measure(1e-3, digests=lambda c: "other" if c["lb_group"] == 16 else "same")
    with pytest.raises(RuntimeError):
        at.search(measure)


def test_cache_roundtrip(tmp_path):
    p = str(tmp_path / "t.json")
    k = at.key("MI355X", 4, True, 512 << 20)
    assert at.load_cache(p) == {}
    at.save_cache({k: dict(at.DEFAULT)}, p)
    assert at.load_cache(p)[k] == at.DEFAULT
    assert at.size_bucket(1) == 4 ** 7 and at.size_bucket(512 << 20) == 4 ** 15
    assert at.key("d", 8, False, 100) == "d|f64|lossless|16384"

"""bench.py's host-side arithmetic, without a GPU: the VALU roofline of the
batched line and the lookup of committed counter profiles."""
import importlib.util
import json
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_valu_roofline_fractions():
    """floor = ops / rate; frac = floor / measured time, for the step and for
    each phase at the autotuner's best time of that phase."""
    b = load_bench()
    rate = 256 * 4 * 16 * 2.4e9          # lanes x clock, one multiply or add each
    assert b.VALU_PEAK_OPS == rate
    report = {"phase_a": [{"ms": 70.0}, {"ms": 60.0}], "phase_b": [{"ms": 40.0}]}
    r = b.valu_roofline(ops=rate * 0.05, ops_a=rate * 0.03, kms=100.0, report=report)
    assert r["floor_ms"] == pytest.approx(50.0) and r["frac"] == pytest.approx(0.5)
    assert r["phase_a"]["best_ms"] == 60.0
    assert r["phase_a"]["floor_ms"] == pytest.approx(30.0)
    assert r["phase_a"]["frac"] == pytest.approx(0.5)
    assert r["phase_b"]["ops"] == pytest.approx(rate * 0.02)
    assert r["phase_b"]["frac"] == pytest.approx(0.5)
    # without an autotuner report: the step only
    r = b.valu_roofline(ops=1.0, ops_a=0.5, kms=1.0, report=None)
    assert "phase_a" not in r and "phase_b" not in r


def test_committed_traffic_entries():
    """The counter profiles bench.py reads: each entry's bytes are its scaled
    fetch plus write, and an entry is used only for the kernel it was
    measured on."""
    b = load_bench()
    t = json.load(open(b.PROFILE_TRAFFIC))
    cal = t["_calibration"]
    assert cal["write_scale"] == 1.0 and 1.9 < cal["fetch_scale"] < 2.1
    c3 = t["cfg3"]
    assert c3["bytes_per_step"] == c3["fetch_bytes_per_step"] + c3["write_bytes_per_step"]
    # per-kernel figures are rounded to whole bytes each
    per_kernel = sum(k["fetch_bytes"] + k["write_bytes"] for k in c3["kernels"].values())
    assert abs(c3["bytes_per_step"] - per_kernel) <= 2 * len(c3["kernels"])
    assert c3["ratio"] == pytest.approx(c3["bytes_per_step"] / c3["algorithmic_bytes"])
    for key in ("cfg1_limit", "cfg1_refine"):
        e = t[key]
        assert b.traffic_for(key, e["kernel_short"]) == e["bytes_per_launch"]
        assert b.traffic_for(key, e["kernel_short"] + "x") is None
        assert e["ratio"] == pytest.approx(e["bytes_per_launch"] / e["algorithmic_bytes"])

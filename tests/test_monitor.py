"""Utilisation summary and run qualification of the benchmark harness
(nkr-zkp_amd/monitor.py) on hand-worked sample sets: the reference harness's
trimming, percentile, gate and m / T_max definitions
(scripts/bench/monitor/base.py:59-103, scripts/bench/metrics.py:60-78,
169-248)."""
import os
import sys
import time

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "nkr-zkp_amd"))
import monitor  # noqa: E402


def samples(values, t0=100.0, dt=0.1):
    return [monitor.Sample(t0 + i * dt, util=v) for i, v in enumerate(values)]


def test_trim_and_percentile_twenty_values():
    # 0..19: int(20 * 0.1) = 2 dropped from each end -> 2..17 (16 values);
    # the 10th percentile sits at rank 1.5 of 16 -> between 3 and 4
    u = monitor.summarize(samples(list(range(20))))
    assert u["sample_count"] == 16
    assert u["util_min"] == 2 and u["util_max"] == 17
    assert u["util_mean"] == pytest.approx(9.5)
    assert u["util_p10"] == pytest.approx(3.5)
    assert u["util_ge_target_ratio"] == 0.0


def test_no_trim_below_five_or_when_cut_rounds_to_zero():
    u = monitor.summarize(samples([10, 20, 30, 40]))
    assert u["sample_count"] == 4 and u["util_min"] == 10
    # five values: int(5 * 0.1) = 0, nothing dropped
    u = monitor.summarize(samples([0, 100, 100, 100, 100]))
    assert u["sample_count"] == 5
    assert u["util_mean"] == pytest.approx(80.0)
    assert u["util_ge_target_ratio"] == pytest.approx(0.8)
    # p10 of five values: rank 0.4 between 0 and 100
    assert u["util_p10"] == pytest.approx(40.0)


def test_trim_removes_a_start_up_dip():
    # ten values, one low start-up sample: one dropped from each end
    u = monitor.summarize(samples([50] + [95] * 9))
    assert u["sample_count"] == 8 and u["util_min"] == 95
    rec = monitor.qualify([10.0, 11.0], samples([50] + [95] * 9), (0.0, 1e9))
    assert rec["status"] == "ok"


def test_missing_values_are_ignored():
    s = samples([None, 91.0, None, 93.0])
    u = monitor.summarize(s)
    assert u["sample_count"] == 2 and u["util_mean"] == pytest.approx(92.0)
    assert monitor.summarize([])["util_mean"] is None


def test_qualify_window_status_and_throughput():
    s = samples([0.0] * 5 + [99.0] * 10, t0=0.0, dt=1.0)  # t = 0 .. 14
    # the window [5, 14] holds only the busy samples
    rec = monitor.qualify([100.0, 200.0, 150.0], s, (5.0, 14.0))
    assert rec["status"] == "ok"
    assert rec["util"]["sample_count"] == 8  # 10 inside, one trimmed from each end
    assert rec["m"] == 3 and rec["t_max_ms"] == 200.0 and rec["t_median_ms"] == 150.0
    assert rec["throughput_proofs_s"] == pytest.approx(15.0)  # 3 / 0.2 s
    # the whole range: 5 x 0 and 10 x 99, one of each trimmed -> 891 / 13, below 90
    rec = monitor.qualify([100.0], s, (0.0, 14.0))
    assert rec["status"] == "util_unqualified"
    # m given explicitly: one call of 16 proofs
    rec = monitor.qualify([400.0], s, (5.0, 14.0), m=16)
    assert rec["throughput_proofs_s"] == pytest.approx(40.0)
    # no sample inside the window, or a failed proof
    assert monitor.qualify([1.0], s, (20.0, 30.0))["status"] == "util_unknown"
    assert monitor.qualify([1.0], s, (5.0, 14.0), ok=False)["status"] == "invalid"
    assert monitor.qualify([], s, (5.0, 14.0))["throughput_proofs_s"] is None


def test_sampler_collects_and_survives_errors():
    calls = []

    def fn():
        calls.append(1)
        if len(calls) == 2:
            raise RuntimeError("sensor read failed")
        return monitor.Sample(time.time(), util=42.0)

    sm = monitor.Sampler(fn, interval_s=0.01)
    sm.start()
    time.sleep(0.2)
    got = sm.stop()
    assert len(calls) >= 3
    assert len(got) == len(calls) - 1 and all(x.util == 42.0 for x in got)
    assert sm.error == "sensor read failed"


def test_no_source_for_an_absent_device():
    # no device has this address; neither source may pretend to sample it
    name, fn = monitor.source_for("ffff:ff:1f.7")
    assert name is None and fn is None

"""GPU utilisation sampling and run qualification for the benchmark harness.

The reference harness samples each device's utilisation on a background
thread while its prove processes run (scripts/bench/monitor/base.py:28-56,
scripts/bench/monitor/amd.py:110-140), keeps the samples that fall inside the
window in which every prove was online (scripts/bench/metrics.py:157-166),
summarises them after trimming both tails (scripts/bench/monitor/base.py:73-103)
and marks an otherwise good repetition "util_unqualified" when the mean is
below the target, "util_unknown" when no sample was taken
(scripts/bench/metrics.py:196-245).  Throughput is m / T_max, m proofs that ran
together over the slowest one's prove time (scripts/bench/metrics.py:60-78).
Defaults are the reference's: 100 ms between samples, 10 % trimmed from each
end, target 0.90 (scripts/bench/config.py:66-71).

This module keeps those definitions.  Samples come from the AMD SMI library
(graphics engine activity of the device whose PCI address matches), falling
back to the kernel driver's gpu_busy_percent file in sysfs; with neither, a
run is "util_unknown", as in the reference.
"""
from __future__ import annotations

import glob
import os
import statistics
import threading
import time
from dataclasses import dataclass
from typing import Callable, Optional

SAMPLE_S = 0.100
TRIM_RATIO = 0.10
TARGET = 0.90


@dataclass
class Sample:
    timestamp: float
    util: Optional[float] = None
    power_w: Optional[float] = None
    temp_c: Optional[float] = None
    clock_mhz: Optional[float] = None


def percentile(values: list, fraction: float) -> Optional[float]:
    """Linear interpolation between closest ranks (scripts/bench/monitor/base.py:59-70)."""
    if not values:
        return None
    ordered = sorted(values)
    if len(ordered) == 1:
        return ordered[0]
    pos = (len(ordered) - 1) * fraction
    lo = int(pos)
    hi = min(lo + 1, len(ordered) - 1)
    w = pos - lo
    return ordered[lo] * (1.0 - w) + ordered[hi] * w


def summarize(samples: list, trim_ratio: float = TRIM_RATIO, target: float = TARGET) -> dict:
    """Utilisation summary of one device's samples: with 5 or more values,
    int(n * trim_ratio) values are dropped from each end (when at least 2
    remain); then mean, 10th percentile, min, max and the share of values at
    or above target (scripts/bench/monitor/base.py:73-103)."""
    vals = sorted(s.util for s in samples if s.util is not None)
    if not vals:
        return {"util_mean": None, "util_p10": None, "util_min": None, "util_max": None,
                "util_ge_target_ratio": None, "sample_count": 0}
    if trim_ratio > 0 and len(vals) >= 5:
        cut = int(len(vals) * trim_ratio)
        if cut > 0 and len(vals) - 2 * cut >= 2:
            vals = vals[cut:len(vals) - cut]
    pct = target * 100.0
    return {"util_mean": statistics.fmean(vals), "util_p10": percentile(vals, 0.10),
            "util_min": min(vals), "util_max": max(vals),
            "util_ge_target_ratio": sum(1 for v in vals if v >= pct) / len(vals),
            "sample_count": len(vals)}


def qualify(prove_ms: list, samples: list, window: tuple, m: Optional[int] = None,
            ok: bool = True, trim_ratio: float = TRIM_RATIO, target: float = TARGET) -> dict:
    """One repetition's record: status, m / T_max throughput and the
    utilisation of the samples inside window = (start, end) in the samples'
    clock (scripts/bench/metrics.py:60-78, 169-248).  `ok` is whether every
    proof of the repetition succeeded and verified; `m` defaults to the number
    of prove times."""
    m = len(prove_ms) if m is None else m
    t_max = max(prove_ms) if prove_ms else None
    start, end = window
    inside = [s for s in samples if start <= s.timestamp <= end]
    util = summarize(inside, trim_ratio, target)
    if not ok:
        status = "invalid"
    elif util["util_mean"] is None:
        status = "util_unknown"
    elif util["util_mean"] < target * 100.0:
        status = "util_unqualified"
    else:
        status = "ok"
    return {"status": status, "m": m,
            "t_max_ms": None if t_max is None else round(t_max, 4),
            "t_median_ms": round(statistics.median(prove_ms), 4) if prove_ms else None,
            "throughput_proofs_s": None if not t_max else round(m / (t_max / 1000.0), 3),
            "util": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in util.items()},
            "target": target, "trim_ratio": trim_ratio, "window_s": round(end - start, 4)}


class Sampler(threading.Thread):
    """Calls sample_fn every interval_s on a daemon thread until stop()
    (scripts/bench/monitor/base.py:28-56); an exception in sample_fn is kept
    and ends nothing."""

    def __init__(self, sample_fn: Callable[[], Optional[Sample]], interval_s: float = SAMPLE_S):
        super().__init__(daemon=True)
        self.sample_fn = sample_fn
        self.interval_s = interval_s
        self.samples: list = []
        self.error: Optional[str] = None
        self._halt = threading.Event()

    def run(self) -> None:
        while not self._halt.is_set():
            t0 = time.time()
            try:
                s = self.sample_fn()
                if s is not None:
                    self.samples.append(s)
            except Exception as exc:  # keep the benchmark alive
                self.error = str(exc)
            self._halt.wait(max(0.0, self.interval_s - (time.time() - t0)))

    def stop(self) -> list:
        self._halt.set()
        self.join(timeout=5)
        return self.samples


def _amdsmi_source(pci: str):
    """A sample function over the AMD SMI library for the device at `pci`, or
    None when the library or the device is not available."""
    try:
        import amdsmi  # noqa: PLC0415
        amdsmi.amdsmi_init(amdsmi.AmdSmiInitFlags.INIT_AMD_GPUS)
    except Exception:
        return None
    handle = None
    try:
        for h in amdsmi.amdsmi_get_processor_handles():
            if str(amdsmi.amdsmi_get_gpu_device_bdf(h)).lower() == pci:
                handle = h
                break
    except Exception:
        handle = None
    if handle is None:
        try:
            amdsmi.amdsmi_shut_down()
        except Exception:
            pass
        return None

    def number(v):
        return float(v) if isinstance(v, (int, float)) else None

    def sample() -> Sample:
        act = amdsmi.amdsmi_get_gpu_activity(handle)
        s = Sample(time.time(), util=number(act.get("gfx_activity")))
        try:
            pw = amdsmi.amdsmi_get_power_info(handle)
            s.power_w = number(pw.get("socket_power", pw.get("current_socket_power")))
        except Exception:
            pass
        try:
            s.temp_c = number(amdsmi.amdsmi_get_temp_metric(handle, amdsmi.AmdSmiTemperatureType.HOTSPOT,
                                                            amdsmi.AmdSmiTemperatureMetric.CURRENT))
        except Exception:
            pass
        try:
            s.clock_mhz = number(amdsmi.amdsmi_get_clock_info(handle, amdsmi.AmdSmiClkType.GFX).get("clk"))
        except Exception:
            pass
        return s

    try:
        sample()
    except Exception:
        return None
    return sample


def _sysfs_source(pci: str):
    """A sample function over the driver's gpu_busy_percent file of the device
    at `pci`, or None."""
    for dev in glob.glob("/sys/class/drm/card*/device"):
        try:
            if os.path.basename(os.path.realpath(dev)).lower() != pci:
                continue
            path = os.path.join(dev, "gpu_busy_percent")
            with open(path) as f:
                float(f.read().strip())
        except (OSError, ValueError):
            continue

        def sample(path=path) -> Sample:
            with open(path) as f:
                return Sample(time.time(), util=float(f.read().strip()))
        return sample
    return None


def source_for(pci: str):
    """(name, sample function) for the device at PCI address `pci`: the AMD
    SMI library first, then sysfs; (None, None) when neither works."""
    pci = pci.lower()
    fn = _amdsmi_source(pci)
    if fn is not None:
        return "amdsmi gfx_activity", fn
    fn = _sysfs_source(pci)
    if fn is not None:
        return "sysfs gpu_busy_percent", fn
    return None, None

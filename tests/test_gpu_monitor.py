"""The benchmark's utilisation sampling on a real device: the library
reports the device's PCI address, the sampler finds a source for that device
(the AMD SMI library, or the driver's sysfs file), and its samples are
activity percentages taken at the requested interval."""
import re
import time

import pytest

import monitor
import nkr_g16 as N

pytestmark = pytest.mark.gpu


def test_pci_address_and_samples():
    pci = N.device_pci_bus_id(0)
    assert re.fullmatch(r"[0-9a-f]{4}:[0-9a-f]{2}:[0-9a-f]{2}\.[0-7]", pci), pci
    name, fn = monitor.source_for(pci)
    if fn is None:
        pytest.skip("no utilisation source on this host (neither AMD SMI nor sysfs)")
    sm = monitor.Sampler(fn, interval_s=0.05)
    t0 = time.time()
    sm.start()
    time.sleep(0.5)
    got = sm.stop()
    t1 = time.time()
    assert sm.error is None, sm.error
    assert len(got) >= 4, (name, len(got))
    for s in got:
        assert t0 <= s.timestamp <= t1
        assert s.util is not None and 0.0 <= s.util <= 100.0, (name, s)
    rec = monitor.qualify([1.0], got, (t0, t1))
    assert rec["util"]["sample_count"] > 0 and rec["status"] in ("ok", "util_unqualified")

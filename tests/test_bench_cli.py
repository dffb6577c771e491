"""bench.py's host-side contract, on the CPU: more than one GPU is refused with
a message before any device work, the committed PMC traffic summary parses,
the Montgomery constant the operator leg uses is 1, and the options list the
legs."""
import json
import os
import subprocess
import sys

import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True,
                          text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = run_bench("--gpus", "2")
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_measured_traffic_reads_the_committed_summary():
    with open(bench.TRAFFIC_PROFILE) as f:
        text = f.read()
    for kernel in ("g1", "g2"):
        t = bench.measured_traffic(kernel)
        assert isinstance(t, int) and t > 0
        assert f"traffic_bytes_{kernel}_accumulate: {t}" in text


def test_montgomery_one():
    assert [int(x) for x in bench.ONE_MONT] == O.int_to_words(O.to_mont_int(1, O.R))


def test_options_list_the_legs():
    r = run_bench("--help")
    assert r.returncode == 0
    for opt in ("--operators", "--circuit", "--tuned", "--no-batch", "--no-serve", "--pmc-workload"):
        assert opt in r.stdout, opt

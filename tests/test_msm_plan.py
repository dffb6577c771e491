"""The MSM sizing arithmetic the device buffers are allocated from
(nkr-zkp_amd/csrc/msm_plan.h), compiled for the CPU: window choice, the
level-0 piece bound, the fold levels and the partial-sum capacity, including
schedules with zero entries and invalid piece bounds, which must raise an
error instead of dividing by zero (SIGFPE)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def plan_bin(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("needs g++")
    out = str(tmp_path_factory.mktemp("plan") / "msm_plan_check")
    subprocess.run([cxx, "-O2", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "nkr-zkp_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host", "msm_plan_check.cpp"), "-o", out], check=True)
    return out


def ask(binary, lines):
    res = subprocess.run([binary], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    return res.stdout.strip().splitlines()


def test_window_choice_covers_the_scalar(plan_bin):
    ns = sorted({1, 2, 3, 5, 7, 100, 1000, 4097} | {1 << k for k in range(0, 25)} | {(1 << k) + 1 for k in range(1, 24)})
    out = ask(plan_bin, [f"plan {n} 0 19 4 64" for n in ns])
    for n, line in zip(ns, out):
        c, W, B, pieces, cap, levels = map(int, line.split())
        assert 4 <= c <= 16 and B == 1 << (c - 1)
        assert c * W >= 255 and c * (W - 1) < 255          # W signed digits cover 255 bits, none wasted
        assert 254 - c * (W - 1) >= 4 or c in (4, 16)        # the top window keeps at least 4 bits
        if n >= 1 << 20:
            assert c == 16
        assert pieces >= B + 1 and cap >= pieces


def test_piece_bound_covers_every_split(plan_bin):
    """The device picks the largest power of two p <= E / 2^target within
    [lo, hi] (and longer while buckets would keep more than 64 partials);
    every bucket's run of entries then splits into ceil(count / p) pieces.
    Worst case over the entry count E and the bucket spread: at most
    E / p + B pieces, which max_pieces must bound."""
    rng = np.random.default_rng(7)
    for n, target in ((1 << 10, 19), (1 << 16, 19), (1 << 20, 19), (1 << 20, 12), (5000, 24)):
        line = ask(plan_bin, [f"plan {n} 0 {target} 4 64"])[0]
        c, W, B, pieces, cap, levels = map(int, line.split())
        for E in [0, 1, B, n * W // 3, n * W] + list(rng.integers(0, n * W + 1, 5)):
            p = 4
            while p < 64 and 2 * p <= E >> target:
                p *= 2
            assert -(-E // p) + B <= pieces, (n, target, E)


def test_fold_levels_and_capacity(plan_bin):
    """Zero entries need no fold; the partial-sum capacity covers every fold
    level of any distribution of the level-0 pieces over the buckets (each
    level writes ceil(k / 4) partials per bucket with k > 64)."""
    assert ask(plan_bin, ["folds 0", "folds 1", "folds 64", "folds 65", "folds 256", "folds 257", "folds 4096"]) == \
        ["0", "0", "0", "1", "1", "2", "3"]
    rng = np.random.default_rng(11)
    for n in (1 << 12, 1 << 16, 1 << 20):
        c, W, B, pieces, cap, levels = map(int, ask(plan_bin, [f"plan {n} 0 19 4 64"])[0].split())
        for shape in ("one", "uniform", "skewed"):
            total = pieces - B - 1
            if shape == "one":
                counts = np.zeros(B, dtype=np.int64)
                counts[0] = total
            elif shape == "uniform":
                counts = np.full(B, total // B, dtype=np.int64)
            else:
                counts = rng.multinomial(total, rng.dirichlet(np.full(B, 0.05)))
            used, level = int(counts.sum()), 0
            while counts.max() > 64:
                big = counts > 64
                counts = np.where(big, -(-counts // 4), counts)
                used += int(counts[big].sum())
                level += 1
            assert level <= levels and used <= cap, (n, shape)


def test_zero_divisors_raise_instead_of_crashing(plan_bin):
    out = ask(plan_bin, ["plan 1024 0 19 0 64", "plan 1024 0 19 8 4", "divup 10 0", "divup 10 3",
                         "plan 0 0 19 4 64", "plan 1024 17 19 4 64"])
    assert out[0].startswith("error MSM piece bounds")
    assert out[1].startswith("error MSM piece bounds")
    assert out[2].startswith("error internal sizing error")
    assert out[3] == "4"
    assert out[4].startswith("error MSM needs at least one term")
    assert out[5].startswith("error MSM window")

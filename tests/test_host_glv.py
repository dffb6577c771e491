"""The host GLV scalar multiplication used for the proof's blinding terms
(nkr-zkp_amd/csrc/glv.h), compiled for the CPU and checked against the
oracle's scalar multiplication (libff scalar_mul restated) on random and
edge-case scalars: 0, 1, r - 1, lambda and its neighbours, 2^127, 2^128,
and scalars whose GLV halves are negative."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LAMBDA = 0xB3C4D79D41A917585BFC41088D8DAAA78B17EA66B99C90DD


@pytest.fixture(scope="module")
def glv_bin(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("needs g++ and the HIP headers")
    out = str(tmp_path_factory.mktemp("glv") / "glv_check")
    subprocess.run([cxx, "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I" + os.path.join(ROOT, "nkr-zkp_amd", "csrc"), "-w",
                    os.path.join(ROOT, "tests", "host", "glv_check.cpp"), "-o", out], check=True)
    return out


def run(binary, points, scalars):
    lines = []
    for p, k in zip(points, scalars):
        lines.append(" ".join(str(int(v)) for v in list(p) + list(k)))
    res = subprocess.run([binary], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    return np.array([[int(v) for v in l.split()] for l in res.stdout.strip().splitlines()], dtype=np.uint64)


def test_glv_matches_oracle(glv_bin):
    g = O.g1_generator()
    ints = [0, 1, 2, O.R - 1, O.R - 2, LAMBDA, LAMBDA - 1, LAMBDA + 1, (LAMBDA * LAMBDA) % O.R,
            1 << 127, (1 << 128) - 1, 1 << 128, (1 << 253) + 12345, O.R // 2, O.R // 3]
    ks = np.concatenate([O.fr_array(ints), O.fr_random(0x61F, 200)])
    # a generator multiple with Z != 1 as the second base (non-normalised input)
    q = O.g1_mul(g, O.fr_array([987654321]))[0]
    q = O.g1_add(O.g1_add(q, g), O.g1_neg(g))
    assert O.fq_from_mont(q[8:12])[0].tolist() != [1, 0, 0, 0]  # Z != 1
    for base in (g, q):
        want = O.g1_to_affine(O.g1_mul(base, ks))
        got = run(glv_bin, [base] * len(ks), ks)
        for i in range(len(ks)):
            if not got[i].any():  # infinity
                assert not want[i][8:12].any(), i
            else:
                assert np.array_equal(got[i], want[i]), i


def test_joint_glv_and_batched_normalisation(glv_bin):
    """k P + l Q as one multiplication of four GLV half-scalars
    (g1_mul2_glv), and k P, l Q on their own, normalised together with one
    shared inversion (xyzz_to_jacobian_affine_batch): all three equal the
    oracle's, including sums and terms at infinity."""
    g = O.g1_generator()
    p = O.g1_mul(g, O.fr_array([123456789]))[0]
    p = O.g1_add(O.g1_add(p, g), O.g1_neg(g))  # Z != 1
    q = O.g1_mul(g, O.fr_array([987654321]))[0]
    rng = np.random.default_rng(0x2B1)
    cases = []
    for _ in range(40):
        k, l = (int(rng.integers(1, 1 << 62)) * int(rng.integers(1, 1 << 62)) * 7919) % O.R, \
               (int(rng.integers(1, 1 << 62)) * int(rng.integers(1, 1 << 62)) * 104729) % O.R
        cases.append((p, k, q, l))
    k = 0xDEADBEEF12345
    cases += [(p, 0, q, 5), (p, 7, q, 0), (p, 0, q, 0), (p, k, p, O.R - k), (q, LAMBDA, q, 1),
              (p, O.R - 1, q, O.R - 1), (p, 1 << 128, q, (1 << 127) + 3)]
    lines = []
    for (a, ka, b, kb) in cases:
        kw, lw = O.fr_array([ka]), O.fr_array([kb])
        lines.append(" ".join(str(int(v)) for v in list(a) + list(kw[0]) + list(b) + list(lw[0])))
    res = subprocess.run([glv_bin, "pair"], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         check=True)
    got = np.array([[int(v) for v in l.split()] for l in res.stdout.strip().splitlines()], dtype=np.uint64)
    assert got.shape == (len(cases), 36)
    for i, (a, ka, b, kb) in enumerate(cases):
        pa = O.g1_mul(a, O.fr_array([ka]))[0]
        pb = O.g1_mul(b, O.fr_array([kb]))[0]
        for j, want in enumerate((O.g1_add(pa, pb), pa, pb)):
            w = O.g1_to_affine(np.array([want]))[0]
            g_ = got[i, 12 * j:12 * j + 12]
            if not g_.any():
                assert not w[8:12].any(), (i, j)
            else:
                assert np.array_equal(g_, w), (i, j)

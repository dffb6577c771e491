"""The C++ seeded generator (nkr-g16-gen, the reference's gen driver with
seeded values) against the Python generators of circuits.py: identical CSR
arrays, public input counts and witnesses for every family, density and
option; the witnesses satisfy their R1CS (oracle SpMV); file names and
errors follow the reference's driver."""
import os
import subprocess

import numpy as np
import pytest

import circuits
import pyoracle as O

GEN = circuits.GEN_BIN


def same(a, b):
    assert (a.rows, a.cols, a.public_inputs) == (b.rows, b.cols, b.public_inputs)
    for m, n in ((a.a, b.a), (a.b, b.b), (a.c, b.c)):
        for x, y in zip(m, n):
            assert np.array_equal(np.asarray(x), np.asarray(y))
    assert np.array_equal(a.witness, b.witness)


def satisfied(c):
    az, bz, cz = (O.spmv(c.rows, c.cols, *m, c.witness) for m in (c.a, c.b, c.c))
    return np.array_equal(O.fr_mul(az, bz), cz)


@pytest.mark.parametrize("s,density,wide", [(2, "full", False), (5, "full", True), (6, "sparse", True),
                                             (7, "hybrid", False), (10, "sparse", True)])
def test_matmul_equals_python(s, density, wide):
    c = circuits.generate("matmul", s, 0x4E4B52 + s, density=density, wide=wide)
    same(c, circuits.matmul(s, 0x4E4B52 + s, density=density, wide=wide))
    assert satisfied(c)


def test_fibonacci_equals_python():
    for steps in (2, 3, 100, 1000):
        c = circuits.generate("fibonacci", steps, 77 + steps)
        same(c, circuits.fibonacci(steps, 77 + steps))
        assert satisfied(c)


@pytest.mark.parametrize("count,density,folded", [(1, "full", False), (3, "full", True), (4, "hybrid", False),
                                                  (5, "sparse", True)])
def test_poseidon_equals_python(count, density, folded):
    c = circuits.generate("poseidon", count, 0x9A + count, density=density, folded=folded)
    same(c, circuits.poseidon(count, 0x9A + count, density=density, folded=folded))
    assert satisfied(c)


def test_reference_counts_and_stems(tmp_path):
    """One unfolded hash has the reference circuit's 439 constraints and 442
    variables in a 512 x 512 R1CS (poseidon_test.cpp:40-48); outputs carry
    the reference's stems."""
    res = subprocess.run([GEN, "--path", str(tmp_path), "poseidon", "1"], capture_output=True, text=True, check=True)
    assert "Density: full (inputs: 2/2 nonzero)" in res.stdout
    c = circuits.read_r1cs(str(tmp_path / "poseidon_1.bin"), str(tmp_path / "poseidon_1.witness.bin"))
    assert (c.rows, c.cols) == (512, 512)
    assert int(np.count_nonzero(np.diff(c.a[0]) | np.diff(c.b[0]) | np.diff(c.c[0]))) == 439
    assert int(np.count_nonzero(c.witness.any(axis=1))) <= 442
    subprocess.run([GEN, "--path", str(tmp_path), "--density", "sparse", "matmul", "4"], check=True,
                   capture_output=True)
    subprocess.run([GEN, "--path", str(tmp_path), "--folded", "--density=hybrid", "poseidon", "2"], check=True,
                   capture_output=True)
    names = sorted(os.listdir(tmp_path))
    for stem in ("poseidon_1", "matmul_4_sparse", "poseidon_2_hybrid_folded"):
        assert stem + ".bin" in names and stem + ".witness.bin" in names


def test_errors(tmp_path):
    for args, code, text in ((["matmul"], 1, "matmul needs"), (["fibonacci", "1"], 1, "at least 2"),
                             (["--density", "fibonacci", "fibonacci", "4"], 1, "unknown density"),
                             (["--density", "sparse", "fibonacci", "4"], 1, "only for matmul and poseidon"),
                             (["sql-aggregate", "4"], 1, "unknown case"), (["--bogus"], 2, "Usage"), ([], 2, "Usage")):
        res = subprocess.run([GEN, "--path", str(tmp_path)] + args, capture_output=True, text=True)
        assert res.returncode == code and text in res.stderr, (args, res.returncode, res.stderr)

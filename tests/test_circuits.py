"""The product's seeded circuit generator reproduces the golden generator's
circuits exactly, and its witnesses satisfy the R1CS (check_witness,
protocols/groth16/common/src/protocol.cpp:69-83)."""
import numpy as np
import pytest

import circuits
import gen_golden as GG
import pyoracle as O


@pytest.mark.parametrize("s,seed,density,wide", [(2, 0x4E4B52, "full", False), (16, 0x4E4B52, "full", False),
                                                 (10, 5, "sparse", True), (7, 9, "hybrid", True)])
def test_matmul_matches_golden_generator(s, seed, density, wide):
    got = circuits.matmul(s, seed, density=density, wide=wide)
    ref = GG.matmul_circuit(s, seed, density=density, wide=wide)
    assert (got.rows, got.cols, got.public_inputs) == (ref["rows"], ref["cols"], ref["public_inputs"])
    for k in ("a", "b", "c"):
        rp, ci, va = getattr(got, k)
        assert rp.tolist() == ref[k][0] and ci.tolist() == ref[k][1]
        assert np.array_equal(va, O.fr_array(ref[k][2]))
    assert np.array_equal(got.witness, O.fr_array(ref["z"]))


def test_matmul16_shape_matches_reference_counts():
    """4,352 live rows in a 2^13 domain; nnz A = 4,096 + 256*17, B = C = 4,352."""
    c = circuits.matmul(16, 1)
    assert c.rows == 8192 and c.cols == 8192
    assert c.a[1].size == 4096 + 256 * 17 and c.b[1].size == 4352 and c.c[1].size == 4352


def test_witness_satisfies():
    c = circuits.matmul(6, 3, density="sparse", wide=True)
    az = O.spmv(c.rows, c.cols, *c.a, c.witness)
    bz = O.spmv(c.rows, c.cols, *c.b, c.witness)
    cz = O.spmv(c.rows, c.cols, *c.c, c.witness)
    assert np.array_equal(O.fr_mul(az, bz), cz)


def test_random_scalars_canonical():
    s = circuits.random_scalars(10000, 1)
    ints = [O.words_to_int(r) for r in s]
    assert all(v < O.R for v in ints) and len(set(ints)) == 10000

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


def satisfied(c) -> bool:
    az = O.spmv(c.rows, c.cols, *c.a, c.witness)
    bz = O.spmv(c.rows, c.cols, *c.b, c.witness)
    cz = O.spmv(c.rows, c.cols, *c.c, c.witness)
    return np.array_equal(O.fr_mul(az, bz), cz)


def test_poseidon_counts_match_reference_test():
    """protocols/groth16/tests/poseidon_test.cpp:40-48: one hash gives 439
    constraints, 442 variables with the constant column, one public input,
    and a 512 x 512 padded R1CS; the witness satisfies it and a changed digest
    does not."""
    c = circuits.poseidon(1, 7)
    assert c.public_inputs == 1 and c.rows == 512 and c.cols == 512
    assert int(c.a[0][439]) == int(c.a[0][-1]) and int(c.a[0][438]) < int(c.a[0][439])
    live_cols = int(max(c.a[1].max(), c.b[1].max(), c.c[1].max())) + 1
    assert live_cols == 442
    assert satisfied(c)
    w = c.witness.copy()
    w[1] = O.fr_add(w[1:2], O.fr_array([1]))[0]
    c.witness = w
    assert not satisfied(c)


def test_poseidon_hash_matches_permutation_model():
    """The digest column holds hash2(left, right) = permute([0, left, right])[0]
    of an independent restatement of the round schedule (4 full, 57 partial, 4 full)."""
    prm = circuits.PoseidonParams()
    left, right = 1, 2
    s = [0, left, right]
    k = 0
    for rnd in range(65):
        s = [(x + prm.rc[k + i]) % O.R for i, x in enumerate(s)]
        k += 3
        if 4 <= rnd < 61:
            s[0] = pow(s[0], 5, O.R)
        else:
            s = [pow(x, 5, O.R) for x in s]
        s = [sum(prm.mds[r][cc] * s[cc] for cc in range(3)) % O.R for r in range(3)]
    assert prm.hash2(left, right) == s[0]


@pytest.mark.parametrize("folded", [False, True])
def test_poseidon_batch_satisfies(folded):
    c = circuits.poseidon(5, 11, density="hybrid", folded=folded)
    per_hash = 244 if folded else 439
    assert int(c.a[0][5 * per_hash]) == int(c.a[0][-1])
    assert c.public_inputs == 5 and satisfied(c)


def test_fibonacci():
    c = circuits.fibonacci(1024, 3)
    assert c.rows == 1024 and c.cols == 2048 and c.public_inputs == 3
    assert satisfied(c)
    f = O.fr_ints(c.witness[:4])
    assert f[0] == 1
    # F_steps (public column 3) is the last private term
    last = O.fr_ints(c.witness[3 + 1023:3 + 1024])[0]
    assert f[3] == last
    a, b = f[1], f[2]
    for _ in range(1023):
        a, b = b, (a + b) % O.R
    assert f[3] == b

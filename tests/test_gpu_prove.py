"""End-to-end parity of the HIP Groth16 backend (Backend::prove through the
C ABI) with the CPU oracle and the golden fixtures: every intermediate
(A*z, B*z, C*z, H), every MSM result in the raw proof, and the blinded proof
with fixed r, s are bit-identical in the reference's affine form."""
import numpy as np
import pytest

import circuits
import gen_golden as GG
import goldens
import nkr_g16 as N
import pyoracle as O

pytestmark = pytest.mark.gpu


def affine_raw(raw):
    return np.concatenate([O.g1_to_affine(raw[:48].reshape(4, 12)).reshape(-1), O.g2_to_affine(raw[48:])[0]])


def affine_proof(p):
    return np.concatenate([O.g1_to_affine(p[:12])[0], O.g2_to_affine(p[12:36])[0], O.g1_to_affine(p[36:])[0]])


def test_prove_golden_matmul2():
    e = goldens.E2E()
    pv = N.Prover(e, e)
    raw, (az, bz, cz, h) = pv.prove(e.witness, intermediates=True)
    assert np.array_equal(az, goldens.fe(e.d["az"]))
    assert np.array_equal(bz, goldens.fe(e.d["bz"]))
    assert np.array_equal(cz, goldens.fe(e.d["cz"]))
    assert np.array_equal(h, goldens.fe(e.d["h"]))
    assert np.array_equal(raw, e.raw)
    assert np.array_equal(pv.blind(raw, e.r, e.s), e.proof)
    # proving twice with the same context gives the same bytes
    assert np.array_equal(pv.prove(e.witness), raw)
    # fused prove + blind (comb tables, GLV scalar products) gives the same proof
    raw2, proof2 = pv.prove_blind(e.witness, e.r, e.s)
    assert np.array_equal(raw2, e.raw) and np.array_equal(proof2, e.proof)
    pv.close()


class Circ:
    pass


def matmul_inputs(s, seed, density="full", wide=False):
    c = GG.matmul_circuit(s, seed, density=density, wide=wide)
    circ = Circ()
    circ.rows, circ.cols, circ.public_inputs = c["rows"], c["cols"], c["public_inputs"]
    for k in ("a", "b", "c"):
        rp, ci, va = c[k]
        setattr(circ, k, (np.array(rp, dtype=np.uint32), np.array(ci, dtype=np.uint32), O.fr_array(va)))
    witness = O.fr_array(c["z"])
    return circ, witness


def synthetic_pk(rows, cols, seed):
    """Random multiples of the generators in the reference layout (a valid
    proving key shape; prove parity does not need a trapdoor setup)."""
    g1, g2 = O.g1_generator(), O.g2_generator()
    pk = Circ()
    pk.a_query = N.fixed_base_mul(g1, O.fr_random(seed + 1, cols))
    pk.b_g1_query = N.fixed_base_mul(g1, O.fr_random(seed + 2, cols))
    pk.b_g2_query = N.fixed_base_mul(g2, O.fr_random(seed + 3, cols), g2=True)
    lk = O.fr_random(seed + 4, cols)
    lk[:5] = 0  # public columns carry the point at infinity, as in the CPU setup
    pk.l_query = N.fixed_base_mul(g1, lk)
    hk = O.fr_random(seed + 5, rows)
    hk[-1] = 0
    pk.h_query = N.fixed_base_mul(g1, hk)
    for k, g, sd in (("alpha_g1", g1, 6), ("beta_g1", g1, 7), ("delta_g1", g1, 8)):
        setattr(pk, k, N.fixed_base_mul(g, O.fr_random(seed + sd, 1))[0])
    for k, sd in (("beta_g2", 9), ("delta_g2", 10)):
        setattr(pk, k, N.fixed_base_mul(g2, O.fr_random(seed + sd, 1), g2=True)[0])
    return pk


def test_prove_matmul16_domain_2_13_vs_oracle():
    """The 16x16 matmul circuit (4096 products + 256 sums, domain 2^13) with
    uniform 32-bit A, B entries; r, s fixed from the seed."""
    circ, w = matmul_inputs(16, seed=0x4E4B52)
    assert circ.rows == 1 << 13 and circ.cols == 1 << 13
    pk = synthetic_pk(circ.rows, circ.cols, 77)
    pv = N.Prover(circ, pk)
    raw, (az, bz, cz, h) = pv.prove(w, intermediates=True)
    want_raw, (waz, wbz, wcz, wh) = O.prove(circ, pk, w, intermediates=True)
    assert np.array_equal(az, waz) and np.array_equal(bz, wbz) and np.array_equal(cz, wcz)
    assert np.array_equal(h, wh)
    assert O.fr_ints(h[-1:])[0] == 0  # deg H <= N - 2 for a satisfying witness
    assert np.array_equal(raw, affine_raw(want_raw))
    rs = O.fr_random(0x4E4B52 ^ 0x72, 2)
    proof = pv.blind(raw, rs[0], rs[1])
    want_proof = affine_proof(O.blind(pk, want_raw, rs[0], rs[1]))
    assert np.array_equal(proof, want_proof)
    raw2, proof2 = pv.prove_blind(w, rs[0], rs[1])
    assert np.array_equal(raw2, raw) and np.array_equal(proof2, want_proof)
    pv.close()


def test_prove_sparse_matmul_vs_oracle():
    circ, w = matmul_inputs(10, seed=5, density="sparse", wide=True)
    pk = synthetic_pk(circ.rows, circ.cols, 91)
    pv = N.Prover(circ, pk)
    raw = pv.prove(w)
    assert np.array_equal(raw, affine_raw(O.prove(circ, pk, w)))
    pv.close()


def test_prove_rejects_bad_inputs():
    e = goldens.E2E()
    circ = Circ()
    circ.rows, circ.cols, circ.public_inputs = e.rows, e.cols, e.public_inputs
    circ.a, circ.b = e.a, e.b
    bad = (e.c[0], e.c[1].copy(), e.c[2])
    bad[1][0] = e.cols + 5  # out-of-range column
    circ.c = bad
    with pytest.raises(N.NkrError, match="out-of-range column"):
        N.Prover(circ, e)
    circ.c = e.c
    circ.rows = e.rows - 1
    with pytest.raises(N.NkrError, match="power of two"):
        N.Prover(circ, e)


@pytest.mark.parametrize("count,folded", [(6, False), (16, True)])
def test_prove_poseidon_vs_oracle(count, folded):
    """Poseidon-style hash batches (the second circuit family of the large-proof
    configuration, at small counts; dense witnesses with uniform preimages):
    the raw proof equals the oracle's."""
    circ = circuits.poseidon(count, 0x9A + count, folded=folded)
    pk = synthetic_pk(circ.rows, circ.cols, 93 + count)
    pv = N.Prover(circ, pk)
    raw = pv.prove(circ.witness)
    assert np.array_equal(raw, affine_raw(O.prove(circ, pk, circ.witness)))
    pv.close()

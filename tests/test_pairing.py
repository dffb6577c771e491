"""BN254 reduced pairing and the Groth16 verifier (host code of libnkr_g16.so,
no GPU needed), against the independent flat-Fq12 model's vectors
(tests/golden/gen_pairing.py) and the pairing's defining identities."""
import numpy as np
import pytest

import gen_golden as GG
import gen_pairing as GP
import goldens
import nkr_g16 as N

G1 = goldens.point(GG.g1_json(GG.G1_GEN))
G2 = goldens.point(GG.g2_json(GG.G2_GEN))
GT_ONE = np.zeros(48, dtype=np.uint64)
GT_ONE[:4] = goldens.hex_words(GG.fe_mont_hex(1, GG.Q))


def gt(hs):
    return goldens.fe(hs).reshape(48)


def g1(k):
    return goldens.point(GG.g1_json(GG.smul(GG.G1, GG.G1_GEN, k)))


def g2(k):
    return goldens.point(GG.g2_json(GG.smul(GG.G2, GG.G2_GEN, k)))


@pytest.fixture(scope="module")
def gold():
    return goldens.load("bn254_pairing.json")


def test_pairing_of_generators_matches_model(gold):
    e = N.pairing(G1, G2)
    assert np.array_equal(e, gt(gold["e_g1_g2"]))
    assert not np.array_equal(e, GT_ONE)  # non-degenerate


def test_miller_loop_and_final_exponentiation_match_model(gold):
    f = N.miller_loop(G1, G2)
    assert np.array_equal(f, gt(gold["miller_g1_g2"]))
    assert np.array_equal(N.final_exponentiation(f), gt(gold["e_g1_g2"]))


def test_pairing_cases_match_model(gold):
    for case in gold["cases"]:
        e = N.pairing(goldens.point(case["g1"]), goldens.point(case["g2"]))
        assert np.array_equal(e, gt(case["gt"]))


def test_model_gt_has_order_r(gold):
    """e(G1, G2)^r = 1 and e^k != 1 for small k, in the model, on the library's bytes."""
    words = N.pairing(G1, G2).reshape(12, 4)
    hexes = ["%064x" % sum(int(w) << (64 * i) for i, w in enumerate(row)) for row in words]
    flat = GP.gt_from_hex(hexes)
    assert GP.k_pow(flat, GG.R) == GP.K_ONE
    assert GP.k_pow(flat, 2) != GP.K_ONE and GP.k_pow(flat, 3) != GP.K_ONE
    assert GP.gt_hex(flat) == hexes == gold["e_g1_g2"]  # layout round trip


def test_bilinearity():
    a, b = 0x1234567, 0x89ABCDEF
    e_ab = N.pairing(g1(a), g2(b))
    assert np.array_equal(e_ab, N.pairing(g1(a * b), G2))
    assert np.array_equal(e_ab, N.pairing(G1, g2(a * b)))
    # e(P + P', Q) = e(P, Q) e(P', Q), via the product with one final exponentiation
    assert np.array_equal(N.pairing(g1(a + b), G2), N.pairing_product(np.stack([g1(a), g1(b)]),
                                                                      np.stack([G2, G2])))
    # e(P, Q) e(-P, Q) = 1
    neg = goldens.point(GG.g1_json(GG.G1.neg(GG.smul(GG.G1, GG.G1_GEN, a))))
    assert np.array_equal(N.pairing_product(np.stack([g1(a), neg]), np.stack([G2, G2])), GT_ONE)


def test_infinity_and_empty_product():
    inf1 = goldens.point(GG.g1_json(None))
    inf2 = goldens.point(GG.g2_json(None))
    assert np.array_equal(N.pairing(inf1, G2), GT_ONE)
    assert np.array_equal(N.pairing(G1, inf2), GT_ONE)
    assert np.array_equal(N.pairing_product(np.zeros((0, 12)), np.zeros((0, 24))), GT_ONE)


def test_non_normalized_jacobian_inputs():
    """Jacobian (X Z^2, Y Z^3, Z) pairs like the affine form (the reference's
    proving-key points are not guaranteed affine)."""
    q = GG.Q
    z = 0x5EED
    x, y = GG.smul(GG.G1, GG.G1_GEN, 99)
    jac = [GG.fe_mont_hex(x * z * z, q), GG.fe_mont_hex(y * z ** 3, q), GG.fe_mont_hex(z, q)]
    assert np.array_equal(N.pairing(goldens.point(jac), G2), N.pairing(g1(99), G2))


class Vk:
    pass


@pytest.fixture(scope="module")
def matmul2(gold):
    e = goldens.E2E()
    vk = Vk()
    v = gold["matmul2_vk"]
    vk.alpha_g1_beta_g2 = gt(v["alpha_g1_beta_g2"])
    vk.gamma_g2 = goldens.point(v["gamma_g2"])
    vk.delta_g2 = goldens.point(v["delta_g2"])
    vk.ic = goldens.points(e.d["vk_ic"])
    x = e.witness[1:1 + e.public_inputs]
    return e, vk, x


def test_verify_accepts_golden_proof(matmul2):
    e, vk, x = matmul2
    assert N.verify(vk, x, e.proof)


def test_verify_rejects(matmul2):
    e, vk, x = matmul2
    # C replaced by A
    bad = e.proof.copy()
    bad[36:] = e.proof[:12]
    assert not N.verify(vk, x, bad)
    # a changed public input
    x2 = x.copy()
    x2[0], x2[1] = x[1].copy(), x[0].copy()
    assert not np.array_equal(x2, x) and not N.verify(vk, x2, e.proof)
    # public input count that does not match ic (the reference returns false)
    assert not N.verify(vk, x[:-1], e.proof)
    # B off the twist
    bad = e.proof.copy()
    bad[12] ^= np.uint64(1)
    assert not N.verify(vk, x, bad)
    # a different alpha_g1_beta_g2
    vk2 = Vk()
    vk2.__dict__.update(vk.__dict__)
    vk2.alpha_g1_beta_g2 = N.pairing(G1, G2)
    assert not N.verify(vk2, x, e.proof)

"""Key generation on the GPU (nkr_g16_setup, Backend::setup with given
trapdoors) against the golden key of the matmul2 fixture (made from its
trapdoors by the Python restatement of backend.cpp:41-154), against the
restated query scalars times the generators at a larger size, and end to end:
the keys it writes prove and verify."""
import numpy as np
import pytest

import gen_golden as GG
import goldens
import nkr_g16 as N
import pyoracle as O
from test_gpu_prove import matmul_inputs

pytestmark = pytest.mark.gpu


def mont_words(values):
    return O.fr_array([v % GG.R for v in values])


def test_setup_matches_golden_key_matmul2():
    e = goldens.E2E()
    td = mont_words(int(v) for v in e.d["trapdoors"])
    k = N.setup(e, td)
    for name in ("a_query", "b_g1_query", "b_g2_query", "l_query", "h_query",
                 "alpha_g1", "beta_g1", "beta_g2", "delta_g1", "delta_g2"):
        assert np.array_equal(getattr(k, name).reshape(-1), np.asarray(getattr(e, name)).reshape(-1)), name
    assert np.array_equal(k.ic, goldens.points(e.d["vk_ic"]))
    v = goldens.load("bn254_pairing.json")["matmul2_vk"]
    assert np.array_equal(k.alpha_g1_beta_g2, goldens.fe(v["alpha_g1_beta_g2"]).reshape(-1))
    assert np.array_equal(k.gamma_g2, goldens.point(v["gamma_g2"]))
    assert np.array_equal(k.delta_g2, goldens.point(v["delta_g2"]))


def test_setup_scalars_matmul4():
    """Every query equals (restated scalar) * generator, by the fixed-base kernel."""
    c = GG.matmul_circuit(4, 11)
    circ, _ = matmul_inputs(4, 11)
    tr = O.fr_ints(O.fr_random(0x5E7, 5))
    sc = GG.setup_scalars(c, tr)
    k = N.setup(circ, mont_words(tr))
    g1, g2 = O.g1_generator(), O.g2_generator()
    for name, key, g2_ in (("a_query", "at", False), ("b_g1_query", "bt", False), ("l_query", "l", False),
                           ("h_query", "h", False), ("ic", "ic", False), ("b_g2_query", "bt", True)):
        want = N.fixed_base_mul(g2 if g2_ else g1, mont_words(sc[key]), g2=g2_)
        assert np.array_equal(getattr(k, name), want.reshape(getattr(k, name).shape)), name


def test_setup_prove_verify_matmul16():
    circ, w = matmul_inputs(16, seed=0x4E4B52)
    k = N.setup(circ, O.fr_random(0x7D, 5))
    pv = N.Prover(circ, k)
    rs = O.fr_random(0x4E4B52 ^ 0x72, 2)
    raw, proof = pv.prove_blind(w, rs[0], rs[1])
    pv.close()
    x = w[1:1 + circ.public_inputs]
    assert N.verify(k, x, proof)
    x2 = x.copy()
    x2[0] = x[1]
    assert not N.verify(k, x2, proof)


def test_setup_rejects_bad_trapdoors():
    e = goldens.E2E()
    td = mont_words(int(v) for v in e.d["trapdoors"])
    bad = td.copy()
    bad[0] = O.fr_array([1])[0]  # t = 1 is a root of unity
    with pytest.raises(N.NkrError, match="evaluation domain"):
        N.setup(e, bad)
    bad = td.copy()
    bad[3] = 0  # gamma = 0
    with pytest.raises(N.NkrError, match="invertible"):
        N.setup(e, bad)

"""Pins the CPU oracle: golden vectors from the independent Python model, the C++
standard's mt19937_64 known answer, and closed-form identities at sizes the
fixtures do not reach.  CPU only."""
import numpy as np
import pytest

import goldens
import pyoracle as O

G = goldens.load("bn254_golden.json")


def test_mt19937_64_known_answer():
    d = G["mt19937_64"]
    out = O.mt19937_64(d["seed"], 10000)
    assert [str(int(v)) for v in out[:10]] == d["first"]
    assert str(int(out[9999])) == d["n10000"] == "9981545732273789042"


def test_fr_random_rule():
    d = G["fr_random"]
    got = O.fr_random(d["seed"], len(d["mont"]))
    assert np.array_equal(got, goldens.fe(d["mont"]))
    assert all(O.words_to_int(r) < O.R for r in got)


@pytest.mark.parametrize("field", ["fr", "fq"])
def test_field_ops(field):
    d = G[field]
    a, b = goldens.fe(d["a"]), goldens.fe(d["b"])
    mul, add, sub = (O.fr_mul, O.fr_add, O.fr_sub) if field == "fr" else (O.fq_mul, O.fq_add, O.fq_sub)
    assert np.array_equal(mul(a, b), goldens.fe(d["mul"]))
    assert np.array_equal(add(a, b), goldens.fe(d["add"]))
    assert np.array_equal(sub(a, b), goldens.fe(d["sub"]))
    if field == "fr":
        assert np.array_equal(O.fr_inv(a), goldens.fe(d["inv"]))


def test_fq2_ops():
    d = G["fq2"]
    a = np.array([goldens.fe(x).reshape(-1) for x in d["a"]])
    b = np.array([goldens.fe(x).reshape(-1) for x in d["b"]])
    assert np.array_equal(O.fq2_mul(a, b), np.array([goldens.fe(x).reshape(-1) for x in d["mul"]]))
    assert np.array_equal(O.fq2_sqr(a), np.array([goldens.fe(x).reshape(-1) for x in d["sqr"]]))


def test_montgomery_round_trip():
    x = np.array([O.int_to_words(v) for v in (0, 1, 5, O.R - 1)], dtype=np.uint64)
    m = O.fr_to_mont(x)
    assert [O.words_to_int(v) for v in m] == [O.to_mont_int(v, O.R) for v in (0, 1, 5, O.R - 1)]
    assert np.array_equal(O.fr_from_mont(m), x)


def test_generators_and_scalar_mul():
    g1, g2 = O.g1_generator(), O.g2_generator()
    assert np.array_equal(g1, goldens.point(G["g1"]["gen"]))
    assert np.array_equal(g2, goldens.point(G["g2"]["gen"]))
    assert O.g1_on_curve(g1) and O.g2_on_curve(g2)
    ks = O.fr_array([int(k) for k in G["g1"]["k"]])
    got1 = O.g1_to_affine(O.g1_mul(g1, ks))
    got2 = O.g2_to_affine(O.g2_mul(g2, ks))
    assert np.array_equal(got1, goldens.points(G["g1"]["kG"]))
    assert np.array_equal(got2, goldens.points(G["g2"]["kG"]))
    # group order: (r - 1) G + G is the point at infinity, (0, 1, 0) after to_affine
    last = O.g1_mul(g1, O.fr_array([O.R - 1]))[0]
    inf = goldens.point(["0" * 64, "%064x" % O.to_mont_int(1, O.Q), "0" * 64])
    assert np.array_equal(O.g1_to_affine(O.g1_add(last, g1))[0], inf)


def test_group_law_edge_cases():
    g1 = O.g1_generator()
    two = O.g1_to_affine(O.g1_add(g1, g1))[0]
    assert np.array_equal(two, goldens.point(G["g1"]["kG"][1]))
    zero = O.g1_add(g1, O.g1_neg(g1))
    assert O.words_to_int(zero[8:12]) == 0  # Z == 0


def test_spmv():
    d = G["spmv"]
    y = O.spmv(d["rows"], d["cols"], np.array(d["row_ptr"]), np.array(d["col_idx"]),
               goldens.fe(d["values"]), goldens.fe(d["z"]))
    assert np.array_equal(y, goldens.fe(d["y"]))


@pytest.mark.parametrize("kind,key", [(O.FFT, "fft"), (O.IFFT, "ifft"), (O.COSET_FFT, "coset_fft"),
                                      (O.ICOSET_FFT, "icoset_fft")])
def test_fft_kinds(kind, key):
    for case in G["fft"]:
        got = O.fft(goldens.fe(case["a"]), kind)
        assert np.array_equal(got, goldens.fe(case[key])), (key, case["n"])


def test_quotient_chain():
    for case in G["quotient"]:
        h = O.quotient(goldens.fe(case["az"]), goldens.fe(case["bz"]), goldens.fe(case["cz"]))
        assert np.array_equal(h, goldens.fe(case["h"])), case["n"]


def test_msm_fixtures():
    for case in G["msm"]:
        bases = goldens.points(case["bases"])
        sc = goldens.fe(case["scalars"])
        if case["group"] == "g1":
            got = O.g1_to_affine(O.msm_g1(bases, sc))[0]
        else:
            got = O.g2_to_affine(O.msm_g2(bases, sc))[0]
        assert np.array_equal(got, goldens.point(case["result"])), (case["group"], case["n"], case["flavour"])


def test_msm_empty_and_zero_scalars():
    g1 = O.g1_generator()
    bases = np.repeat(g1[None, :], 5, axis=0)
    out = O.msm_g1(bases, np.zeros((5, 4), dtype=np.uint64))
    assert O.words_to_int(out[8:12]) == 0


# ----------------------------------------------------- closed-form identities
def _horner(coeffs, x, p=O.R):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + c) % p
    return acc


def test_fft_round_trip_and_delta_large():
    n = 1 << 12
    x = O.fr_random(7, n)
    assert np.array_equal(O.fft(O.fft(x, O.FFT), O.IFFT), x)
    assert np.array_equal(O.fft(O.fft(x, O.COSET_FFT), O.ICOSET_FFT), x)
    delta = np.zeros((n, 4), dtype=np.uint64)
    delta[0] = O.int_to_words(O.to_mont_int(1, O.R))
    ones = O.fft(delta, O.FFT)
    assert all(O.words_to_int(v) == O.to_mont_int(1, O.R) for v in ones)


def test_fft_evaluates_polynomial_large():
    """FFT(a)[i] = a(omega^i), checked at a few indices by Horner at N = 2^13."""
    n = 1 << 13
    a = O.fr_random(11, n)
    ev = O.fft(a, O.FFT)
    coeffs = O.fr_ints(a)
    w = 19103219067921713944291392827692070036145651957329286315305642004821462161904
    for _ in range(28 - 13):
        w = w * w % O.R
    assert pow(w, n, O.R) == 1 and pow(w, n // 2, O.R) != 1
    evi = O.fr_ints(ev)
    for i in (0, 1, 2, 777, n - 1):
        assert evi[i] == _horner(coeffs, pow(w, i, O.R))


def test_quotient_identity_large():
    """For a satisfying triple (c = a*b on the domain) H(x)(x^N - 1) = A(x)B(x) - C(x)."""
    n = 1 << 13
    az, bz = O.fr_random(21, n), O.fr_random(22, n)
    cz = O.fr_mul(az, bz)
    h = O.quotient(az, bz, cz)
    A, B, C = (O.fr_ints(O.fft(v, O.IFFT)) for v in (az, bz, cz))
    H = O.fr_ints(h)
    assert H[-1] == 0
    tau = 0x1234567890ABCDEF1234567890ABCDEF % O.R
    lhs = _horner(H, tau) * (pow(tau, n, O.R) - 1) % O.R
    rhs = (_horner(A, tau) * _horner(B, tau) - _horner(C, tau)) % O.R
    assert lhs == rhs


def test_msm_closed_form_large():
    """Bases k_i G: sum s_i (k_i G) = (sum k_i s_i) G, n = 2^11 and a witness-like mix."""
    n = 1 << 11
    k = O.fr_random(31, n)
    s = O.fr_random(32, n)
    s[::3] = 0
    s[1::3] = O.int_to_words(O.to_mont_int(1, O.R))
    g1 = O.g1_generator()
    bases = O.g1_mul(g1, k)
    tot = sum(a * b for a, b in zip(O.fr_ints(k), O.fr_ints(s))) % O.R
    expect = O.g1_to_affine(O.g1_mul(g1, O.fr_array([tot])))[0]
    got = O.g1_to_affine(O.msm_g1(bases, s))[0]
    assert np.array_equal(got, expect)


def test_prove_and_blind_end_to_end():
    e = goldens.E2E()
    raw, (az, bz, cz, h) = O.prove(e, e, e.witness, intermediates=True)
    assert np.array_equal(az, goldens.fe(e.d["az"]))
    assert np.array_equal(bz, goldens.fe(e.d["bz"]))
    assert np.array_equal(cz, goldens.fe(e.d["cz"]))
    assert np.array_equal(h, goldens.fe(e.d["h"]))
    aff = np.concatenate([O.g1_to_affine(raw[:48].reshape(4, 12)).reshape(-1),
                          O.g2_to_affine(raw[48:])[0]])
    assert np.array_equal(aff, e.raw)
    proof = O.blind(e, raw, e.r, e.s)
    aff = np.concatenate([O.g1_to_affine(proof[:12])[0], O.g2_to_affine(proof[12:36])[0],
                          O.g1_to_affine(proof[36:])[0]])
    assert np.array_equal(aff, e.proof)


# ------------------------------------------------------ parallel (all-core) oracle
def test_msm_mt_fixtures_and_closed_form():
    """The chunked MSM (variable_base_msm) gives the golden fixtures' points
    for every thread count (chunks of one term, ragged last chunks, more
    threads than terms), and the closed form at 2^11."""
    for case in G["msm"]:
        bases = goldens.points(case["bases"])
        sc = goldens.fe(case["scalars"])
        want = goldens.point(case["result"])
        for threads in (1, 2, 3, 7, 64):
            if case["group"] == "g1":
                got = O.g1_to_affine(O.msm_g1_mt(bases, sc, threads))[0]
            else:
                got = O.g2_to_affine(O.msm_g2_mt(bases, sc, threads))[0]
            assert np.array_equal(got, want), (case["group"], case["n"], case["flavour"], threads)
    n = 1 << 11
    k = O.fr_random(31, n)
    s = O.fr_random(32, n)
    s[::3] = 0
    g1 = O.g1_generator()
    bases = O.g1_mul(g1, k)
    tot = sum(a * b for a, b in zip(O.fr_ints(k), O.fr_ints(s))) % O.R
    expect = O.g1_to_affine(O.g1_mul(g1, O.fr_array([tot])))[0]
    assert np.array_equal(O.g1_to_affine(O.msm_g1_mt(bases, s, 5))[0], expect)
    empty = O.msm_g1_mt(np.zeros((0, 12), dtype=np.uint64), np.zeros((0, 4), dtype=np.uint64), 4)
    assert O.words_to_int(empty[8:12]) == 0


def test_prove_mt_equals_prove():
    """or_prove_mt gives or_prove's intermediates and (affine) raw proof: the
    golden matmul 2x2 proof, and a 2^10 Fibonacci chain with a random key."""
    import circuits
    e = goldens.E2E()
    for threads in (1, 4):
        raw, inter = O.prove_mt(e, e, e.witness, intermediates=True, threads=threads)
        for got, key in zip(inter, ("az", "bz", "cz", "h")):
            assert np.array_equal(got, goldens.fe(e.d[key]))
        aff = np.concatenate([O.g1_to_affine(raw[:48].reshape(4, 12)).reshape(-1), O.g2_to_affine(raw[48:])[0]])
        assert np.array_equal(aff, e.raw)
    circ = circuits.fibonacci(1000, 0x4E4B52)
    g1, g2 = O.g1_generator(), O.g2_generator()

    class Pk:
        pass
    pk = Pk()
    pk.a_query = O.g1_mul(g1, O.fr_random(1, circ.cols))
    pk.b_g1_query = O.g1_mul(g1, O.fr_random(2, circ.cols))
    pk.b_g2_query = O.g2_mul(g2, O.fr_random(3, circ.cols))
    pk.l_query = O.g1_mul(g1, O.fr_random(4, circ.cols))
    pk.h_query = O.g1_mul(g1, O.fr_random(5, circ.rows))
    for k in ("alpha_g1", "beta_g1", "delta_g1"):
        setattr(pk, k, g1)
    for k in ("beta_g2", "delta_g2"):
        setattr(pk, k, g2)
    want_raw, want_inter = O.prove(circ, pk, circ.witness, intermediates=True)
    got_raw, got_inter = O.prove_mt(circ, pk, circ.witness, intermediates=True, threads=6)
    for a, b in zip(got_inter, want_inter):
        assert np.array_equal(a, b)
    aff = lambda r: np.concatenate([O.g1_to_affine(r[:48].reshape(4, 12)).reshape(-1), O.g2_to_affine(r[48:])[0]])
    assert np.array_equal(aff(got_raw), aff(want_raw))

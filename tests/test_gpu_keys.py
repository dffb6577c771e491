"""Proving keys with non-normalised points, as the reference CPU setup writes
them: libff's windowed exponentiation leaves Jacobian points with Z != 1
(operators/include/nkr/operators/fixed_base_msm.hpp:23-33), and the point at
infinity is any (X, Y, 0) (alt_bn128_g1.cpp:93-96 tests Z == 0 only).  The
reference's HIP load_bases copies X, Y and assumes Z = 1
(operators/hip/include/msm/msm.hip:61-73); this backend normalises on the
device at prepare.  Every key point is rescaled to (X l^2, Y l^3, Z l) with
its own random l (Fq for G1, Fq2 for G2), and some points become infinities
with random X, Y.  Proofs and MSM results must equal those of the normalised
key and the oracle's."""
import numpy as np
import pytest

import nkr_g16 as N
import pyoracle as O
from test_gpu_prove import affine_proof, affine_raw, matmul_inputs, synthetic_pk

pytestmark = pytest.mark.gpu


def nonzero_fq(seed, n, width=4):
    x = O.fr_random(seed, n * width // 4).reshape(n, width)
    x[~x.any(axis=1), 0] = 1  # a zero draw would give Z = 0
    return x


def rescale_g1(p, seed, inf_garbage=()):
    p = np.ascontiguousarray(p, dtype=np.uint64).reshape(-1, 12)
    lam = nonzero_fq(seed, p.shape[0])
    l2 = O.fq_mul(lam, lam)
    l3 = O.fq_mul(l2, lam)
    out = np.concatenate([O.fq_mul(p[:, 0:4], l2), O.fq_mul(p[:, 4:8], l3), O.fq_mul(p[:, 8:12], lam)], axis=1)
    for i in inf_garbage:  # infinity with arbitrary X, Y
        assert not p[i, 8:12].any()
        out[i, 0:8] = O.fr_random(seed ^ (i + 1), 2).reshape(-1)
    return out


def rescale_g2(p, seed, inf_garbage=()):
    p = np.ascontiguousarray(p, dtype=np.uint64).reshape(-1, 24)
    lam = nonzero_fq(seed, p.shape[0], 8)
    l2 = O.fq2_mul(lam, lam)
    l3 = O.fq2_mul(l2, lam)
    out = np.concatenate([O.fq2_mul(p[:, 0:8], l2), O.fq2_mul(p[:, 8:16], l3), O.fq2_mul(p[:, 16:24], lam)],
                         axis=1)
    for i in inf_garbage:
        assert not p[i, 16:24].any()
        out[i, 0:16] = O.fr_random(seed ^ (i + 1), 4).reshape(-1)
    return out


class Pk:
    pass


def rescaled_key(pk, rows, cols, seed):
    inf_cols = [i for i in range(cols) if not np.asarray(pk.l_query)[i, 8:12].any()][:3]
    q = Pk()
    q.a_query = rescale_g1(pk.a_query, seed + 1)
    q.b_g1_query = rescale_g1(pk.b_g1_query, seed + 2)
    q.b_g2_query = rescale_g2(pk.b_g2_query, seed + 3)
    q.l_query = rescale_g1(pk.l_query, seed + 4, inf_garbage=inf_cols)
    q.h_query = rescale_g1(pk.h_query, seed + 5, inf_garbage=[rows - 1])
    for k, sd in (("alpha_g1", 6), ("beta_g1", 7), ("delta_g1", 8)):
        setattr(q, k, rescale_g1(getattr(pk, k), seed + sd)[0])
    for k, sd in (("beta_g2", 9), ("delta_g2", 10)):
        setattr(q, k, rescale_g2(getattr(pk, k), seed + sd)[0])
    return q


def test_rescale_helpers_keep_the_points():
    g1, g2 = O.g1_generator(), O.g2_generator()
    pts = O.g1_mul(np.tile(g1, (5, 1)), O.fr_random(3, 5))
    r = rescale_g1(pts, 11)
    assert not np.array_equal(r, pts)
    assert np.array_equal(O.g1_to_affine(r), O.g1_to_affine(pts))
    pts2 = O.g2_mul(np.tile(g2, (3, 1)), O.fr_random(4, 3))
    r2 = rescale_g2(pts2, 12)
    assert np.array_equal(O.g2_to_affine(r2), O.g2_to_affine(pts2))


@pytest.mark.parametrize("s", [8, 16])
def test_prove_with_non_normalised_key(s):
    """matmul s x s (domains 2^10 and 2^13): the rescaled key gives the same
    raw proof, blinded proof and intermediates as the normalised key, and the
    oracle agrees."""
    circ, w = matmul_inputs(s, seed=0x4E4B52 + s)
    pk = synthetic_pk(circ.rows, circ.cols, 500 + s)
    bad = rescaled_key(pk, circ.rows, circ.cols, 600 + s)
    assert np.asarray(bad.a_query)[:, 8:12].any(axis=1).all()  # Z != 0 ...
    one = np.array(O.int_to_words(O.to_mont_int(1, O.Q)), dtype=np.uint64)
    assert not (np.asarray(bad.a_query)[:, 8:12] == one).all(axis=1).any()  # ... and never 1
    rs = O.fr_random(0x5A5A + s, 2)
    pv_ok, pv_bad = N.Prover(circ, pk), N.Prover(circ, bad)
    raw, inter = pv_ok.prove(w, intermediates=True)
    raw_b, inter_b = pv_bad.prove(w, intermediates=True)
    for x, y in zip(inter, inter_b):
        assert np.array_equal(x, y)
    assert np.array_equal(raw_b, raw)
    want_raw = O.prove(circ, bad, w)
    assert np.array_equal(raw_b, affine_raw(want_raw))
    proof_b = pv_bad.blind(raw_b, rs[0], rs[1])
    assert np.array_equal(proof_b, pv_ok.blind(raw, rs[0], rs[1]))
    assert np.array_equal(proof_b, affine_proof(O.blind(bad, want_raw, rs[0], rs[1])))
    raw2, proof2 = pv_bad.prove_blind(w, rs[0], rs[1])
    assert np.array_equal(raw2, raw) and np.array_equal(proof2, proof_b)
    pv_ok.close()
    pv_bad.close()


@pytest.mark.parametrize("n", [1000, 1 << 14])
def test_msm_with_non_normalised_bases(n):
    """nkr_g1_msm_create / nkr_g2_msm_create with rescaled bases and scaled
    infinities equal the normalised bases' results and the oracle's."""
    g1, g2 = O.g1_generator(), O.g2_generator()
    k = O.fr_random(700 + n, n)
    k[::97] = 0  # infinities
    inf = np.nonzero(~k.any(axis=1))[0].tolist()
    b1 = N.fixed_base_mul(g1, k)
    s = O.fr_random(701 + n, n)
    want1 = O.g1_to_affine(O.msm_g1(b1, s))[0]
    got1 = N.msm(rescale_g1(b1, 702 + n, inf_garbage=inf[::2]), s)
    assert np.array_equal(got1, want1) and np.array_equal(N.msm(b1, s), want1)
    m = min(n, 4096)
    b2 = N.fixed_base_mul(g2, k[:m], g2=True)
    want2 = O.g2_to_affine(O.msm_g2(b2, s[:m]))[0]
    got2 = N.msm(rescale_g2(b2, 703 + n, inf_garbage=[i for i in inf if i < m][::2]), s[:m], g2=True)
    assert np.array_equal(got2, want2)

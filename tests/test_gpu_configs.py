"""Parity at the benchmark configurations' full sizes, through the C ABI,
against the CPU oracle (the C restatement of the reference CPU backend,
protocols/groth16/backends/cpu/src/backend.cpp:156-201, with the MSMs chunked
over host threads as variable_base_msm does):

  * the 100x100 matmul R1CS with 10% non-zero A/B entries (domain 2^20) and
    4096 folded Poseidon-style hashes (domain 2^20): A*z, B*z, C*z, H, the
    five MSM results (ar, bs1, zk, qz, bs2) and the blinded proof with seeded
    r, s are bit-identical; the fused prove + blind gives the same bytes and
    the verifier accepts -- the checks of the reference's own 2^20 GPU test
    (protocols/groth16/backends/hip/test/groth16-test.hip:182-195,266-283);
  * the 64-proof heterogeneous batch (16 circuit shapes x 4 witnesses,
    domains 2^9 .. 2^18) proved in one nkr_g16_batch_prove call: every raw
    and blinded proof equals the oracle's;
  * a seeded Poisson request stream at 400 requests/s drawn from that mix
    plus 10% 2^20 matmul requests, served with continuous batching: every
    result equals the plain one-at-a-time path and the oracle.

Points are compared in the reference's normalised affine form (Z = 1, or
(0, 1, 0) for infinity)."""
import time

import numpy as np
import pytest

import circuits
import nkr_g16 as N
import pyoracle as O

pytestmark = pytest.mark.gpu

SEED = 0x4E4B52


def affine_raw(raw):
    return np.concatenate([O.g1_to_affine(raw[:48].reshape(4, 12)).reshape(-1), O.g2_to_affine(raw[48:])[0]])


def affine_proof(p):
    return np.concatenate([O.g1_to_affine(p[:12])[0], O.g2_to_affine(p[12:36])[0], O.g1_to_affine(p[36:])[0]])


def full_prove_parity(circ, keys, rs):
    """Every intermediate, every MSM and the blinded proof against the oracle."""
    w = np.ascontiguousarray(circ.witness)
    pv = N.Prover(circ, keys)
    try:
        raw, (az, bz, cz, h) = pv.prove(w, intermediates=True)
        t0 = time.perf_counter()
        want_raw, (waz, wbz, wcz, wh) = O.prove_mt(circ, keys, w, intermediates=True)
        print(f"[oracle] {circ.rows} rows: {time.perf_counter() - t0:.1f} s on {O.host_threads()} threads",
              flush=True)
        for got, want, name in ((az, waz, "A*z"), (bz, wbz, "B*z"), (cz, wcz, "C*z"), (h, wh, "H")):
            assert np.array_equal(got, want), name
        assert O.fr_ints(h[-1:])[0] == 0  # deg H <= N - 2
        want_aff = affine_raw(want_raw)
        for k, name in enumerate(("ar", "bs1", "zk", "qz")):
            assert np.array_equal(raw[12 * k:12 * k + 12], want_aff[12 * k:12 * k + 12]), name
        assert np.array_equal(raw[48:], want_aff[48:]), "bs2"
        proof = pv.blind(raw, rs[0], rs[1])
        want_proof = affine_proof(O.blind(keys, want_raw, rs[0], rs[1]))
        assert np.array_equal(proof, want_proof)
        raw2, proof2 = pv.prove_blind(w, rs[0], rs[1])
        assert np.array_equal(raw2, raw) and np.array_equal(proof2, want_proof)
        assert N.verify(keys, w[1:1 + circ.public_inputs], proof)
    finally:
        pv.close()
    return want_raw


@pytest.fixture(scope="module")
def big():
    """The 2^20 matmul configuration: circuit, keys, context and the oracle's raw proof."""
    circ = circuits.generate("matmul", 100, SEED, density="sparse", wide=True)
    assert circ.rows == 1 << 20 and circ.cols == 1 << 20
    keys = N.setup(circ, circuits.random_scalars(5, SEED ^ 0x5E7))
    pv = N.Prover(circ, keys)
    yield circ, keys, pv, {}
    pv.close()


def test_c2a_matmul100_sparse_2p20_vs_oracle(big):
    circ, keys, _, memo = big
    assert (circ.a[1].size, circ.b[1].size, circ.c[1].size) == (2010000, 1010000, 1010000)
    memo["oracle_raw"] = full_prove_parity(circ, keys, circuits.random_scalars(2, SEED ^ 0x72))


def test_c2b_poseidon4096_folded_2p20_vs_oracle():
    circ = circuits.generate("poseidon", 4096, SEED, folded=True)
    assert circ.rows == 1 << 20 and circ.cols == 1 << 20
    keys = N.setup(circ, circuits.random_scalars(5, SEED ^ 0x9E7))
    full_prove_parity(circ, keys, circuits.random_scalars(2, SEED ^ 0x73))


@pytest.fixture(scope="module")
def mix():
    """The 64-proof mix: [(name, circ, keys, prover, witness, r, s, oracle raw)]."""
    shapes = circuits.batch_mix(SEED, 4)
    items, provers = [], []
    t0 = time.perf_counter()
    for k, (name, circ, wits) in enumerate(shapes):
        keys = N.setup(circ, circuits.random_scalars(5, SEED + 17 * k))
        pv = N.Prover(circ, keys)
        provers.append(pv)
        for j, w in enumerate(wits):
            rs = circuits.random_scalars(2, SEED ^ (k << 8) ^ j)
            w = np.ascontiguousarray(w)
            items.append((name, circ, keys, pv, w, rs[0].copy(), rs[1].copy(), O.prove_mt(circ, keys, w)))
    print(f"[oracle] 64 mix proofs: {time.perf_counter() - t0:.1f} s", flush=True)
    yield items
    for pv in provers:
        pv.close()


def test_c3_batch_mix_64_vs_oracle(mix):
    assert len(mix) == 64
    rows = sorted({it[1].rows for it in mix})
    assert rows[0] <= 1 << 10 and rows[-1] == 1 << 18
    b = N.Batch()
    try:
        outs = b.prove([(pv, w, r, s) for (_, _, _, pv, w, r, s, _) in mix])
    finally:
        b.close()
    for (name, circ, keys, _, w, r, s, want_raw), (raw, proof) in zip(mix, outs):
        assert np.array_equal(raw, affine_raw(want_raw)), name
        assert np.array_equal(proof, affine_proof(O.blind(keys, want_raw, r, s))), name
    for (name, circ, keys, _, w, *_), (_, proof) in list(zip(mix, outs))[::4]:
        assert N.verify(keys, w[1:1 + circ.public_inputs], proof), name


def test_c4_poisson_serving_with_2p20_requests(big, mix):
    """120 requests at 400/s (seeded exponential gaps), about 10% of them the
    2^20 matmul proof, through 3 workers with batches of up to 32; results
    equal the plain path's and the oracle's."""
    circ, keys, pv, memo = big
    w_big = np.ascontiguousarray(circ.witness)
    raw_big = memo.get("oracle_raw")
    if raw_big is None:
        raw_big = O.prove_mt(circ, keys, w_big)
    rng = np.random.default_rng(SEED ^ 0xC4)
    n_req, qps = 120, 400.0
    arrivals = np.cumsum(rng.exponential(1.0 / qps, n_req))
    reqs = []
    for i in range(n_req):
        rs = circuits.random_scalars(2, (SEED << 5) ^ i)
        if rng.random() < 0.1:
            reqs.append(("matmul100_sparse", circ, keys, pv, w_big, rs[0].copy(), rs[1].copy(), raw_big))
        else:
            name, c, k, p, w, _, _, want = mix[int(rng.integers(len(mix)))]
            reqs.append((name, c, k, p, w, rs[0].copy(), rs[1].copy(), want))
    n_big = sum(1 for q in reqs if q[0] == "matmul100_sparse")
    assert n_big >= 5
    srv = N.Server(workers=3, max_batch=32, budget_bytes=32 << 30)
    try:
        tickets = []
        t0 = time.perf_counter()
        for q, at in zip(reqs, arrivals):
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(q[3], q[4], q[5], q[6]))
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    assert st["proofs"] == n_req and st["max_batch"] >= 2
    for (name, c, k, p, w, r, s, want_raw), (raw, proof, ms) in zip(reqs, got):
        plain_raw, plain_proof = p.prove_blind(w, r, s)
        assert np.array_equal(raw, plain_raw) and np.array_equal(proof, plain_proof), name
        assert np.array_equal(raw, affine_raw(want_raw)), name
        assert np.array_equal(proof, affine_proof(O.blind(k, want_raw, r, s))), name
        assert ms > 0
    i_big = next(i for i, q in enumerate(reqs) if q[0] == "matmul100_sparse")
    assert N.verify(keys, w_big[1:1 + circ.public_inputs], got[i_big][1])

"""Serving (nkr_g16_server_*): requests submitted while others are on the GPU
are admitted and batched as workers and memory free up; every result equals
the proof of the same request made alone, and verifies.  The work memory the
workers hold stays within the budget, the budget is clamped to the device's
free memory, and a request that cannot be proved fails alone."""
import threading
import time

import numpy as np
import pytest

import circuits
import nkr_g16 as N

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def contexts():
    out = []
    specs = [("matmul8", circuits.matmul(8, 1), circuits.matmul(8, 2)),
             ("fib2048", circuits.fibonacci(2048, 3), circuits.fibonacci(2048, 4)),
             ("poseidon4", circuits.poseidon(4, 5), circuits.poseidon(4, 6)),
             ("matmul16s", circuits.matmul(16, 7, density="sparse", wide=True),
              circuits.matmul(16, 8, density="sparse", wide=True))]
    for i, (name, c, c2) in enumerate(specs):
        k = N.setup(c, circuits.random_scalars(5, 300 + i))
        out.append((name, c, k, N.Prover(c, k), [c.witness, c2.witness]))
    yield out
    for _, _, _, pv, _ in out:
        pv.close()


def requests(contexts, n, seed):
    rs = circuits.random_scalars(2 * n, seed)
    reqs = []
    for i in range(n):
        name, c, k, pv, wits = contexts[(i * 7) % len(contexts)]
        reqs.append((name, c, k, pv, wits[i % 2], rs[2 * i], rs[2 * i + 1]))
    return reqs


def test_serving_matches_alone_and_batches(contexts):
    reqs = requests(contexts, 24, 91)
    alone = [pv.prove_blind(w, r, s) for (_, _, _, pv, w, r, s) in reqs]
    srv = N.Server(workers=2, max_batch=16)
    tickets = []
    # a first burst, then the rest trickling in while the first batches run
    for q in reqs[:12]:
        tickets.append(srv.submit(q[3], q[4], q[5], q[6]))
    for q in reqs[12:]:
        time.sleep(0.0005)
        tickets.append(srv.submit(q[3], q[4], q[5], q[6]))
    got = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    for (name, c, k, _, w, _, _), (raw, proof, ms), (raw1, proof1) in zip(reqs, got, alone):
        assert np.array_equal(raw, raw1) and np.array_equal(proof, proof1), name
        assert ms > 0
        assert N.verify(k, w[1:1 + c.public_inputs], proof), name
    assert st["proofs"] == len(reqs)
    assert st["batches"] < len(reqs) and st["max_batch"] >= 2


def test_concurrent_submitters(contexts):
    reqs = requests(contexts, 16, 17)
    alone = [pv.prove_blind(w, r, s)[1] for (_, _, _, pv, w, r, s) in reqs]
    srv = N.Server(workers=3, max_batch=8)
    results = [None] * len(reqs)

    def client(lo, hi):
        ts = [(i, srv.submit(reqs[i][3], reqs[i][4], reqs[i][5], reqs[i][6])) for i in range(lo, hi)]
        for i, t in ts:
            results[i] = srv.wait(t)[1]

    threads = [threading.Thread(target=client, args=(0, 8)), threading.Thread(target=client, args=(8, 16))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    srv.close()
    for got, want in zip(results, alone):
        assert np.array_equal(got, want)


def test_memory_admission(contexts):
    """Each of the workers gets an equal share of the budget: a request larger
    than a share is rejected at submission; with a share that fits exactly one
    request, every batch holds one request and everything still completes."""
    name, c, k, pv, wits = contexts[0]
    need = N.Server.request_bytes(pv)
    assert need > 0
    small = N.Server(workers=2, max_batch=8, budget_bytes=2 * need - 2)
    with pytest.raises(N.NkrError, match="share"):
        small.submit(pv, wits[0], *circuits.random_scalars(2, 1))
    small.close()
    one = N.Server(workers=2, max_batch=8, budget_bytes=2 * need)
    rs = circuits.random_scalars(8, 5)
    ts = [one.submit(pv, wits[i % 2], rs[2 * i], rs[2 * i + 1]) for i in range(4)]
    outs = [one.wait(t) for t in ts]
    st = one.stats()
    mem = one.memory()
    one.close()
    assert st["max_batch"] == 1 and need <= st["peak_work_bytes"] <= 2 * need
    assert mem["share"] == need and mem["peak_held"] <= 2 * need
    for i, (raw, proof, _) in enumerate(outs):
        assert np.array_equal(proof, pv.prove_blind(wits[i % 2], rs[2 * i], rs[2 * i + 1])[1])


def test_work_memory_stays_within_the_budget(contexts):
    """Three workers with a budget that fits a little more than one of the
    largest requests per worker: the work buffers the workers hold never add
    up to more than the budget (the workers release their grow-only buffers
    when keeping them would exceed their share), and every result still
    equals the proof made alone."""
    big = max(N.Server.request_bytes(pv) for (_, _, _, pv, _) in contexts)
    budget = 3 * (big + big // 2)
    reqs = requests(contexts, 30, 23)
    alone = [pv.prove_blind(w, r, s)[1] for (_, _, _, pv, w, r, s) in reqs]
    srv = N.Server(workers=3, max_batch=16, budget_bytes=budget)
    try:
        mem0 = srv.memory()
        assert mem0["budget"] == budget and mem0["share"] == budget // 3
        tickets = [srv.submit(q[3], q[4], q[5], q[6]) for q in reqs]
        got = [srv.wait(t)[1] for t in tickets]
        mem = srv.memory()
        st = srv.stats()
    finally:
        srv.close()
    for g, want in zip(got, alone):
        assert np.array_equal(g, want)
    assert 0 < mem["peak_held"] <= mem["budget"], mem
    assert mem["held"] <= mem["budget"]
    assert st["proofs"] == len(reqs)


def test_budget_is_clamped_to_free_device_memory(contexts):
    srv = N.Server(workers=2, max_batch=4, budget_bytes=1 << 60)
    try:
        mem = srv.memory()
    finally:
        srv.close()
    assert 0 < mem["budget"] < 1 << 50 and mem["share"] == mem["budget"] // 2


def test_request_without_blinding_keys_fails_alone(contexts):
    """A context prepared without alpha/beta/delta cannot blind: its request
    fails with that error while the requests batched with it complete."""
    name, c, k, pv, wits = contexts[0]

    class NoBlind:
        pass

    nb = NoBlind()
    for attr in ("a_query", "b_g1_query", "b_g2_query", "l_query", "h_query"):
        setattr(nb, attr, getattr(k, attr))
    bad = N.Prover(c, nb)
    reqs = requests(contexts, 8, 31)
    alone = [pv_.prove_blind(w, r, s)[1] for (_, _, _, pv_, w, r, s) in reqs]
    rs = circuits.random_scalars(2, 41)
    srv = N.Server(workers=1, max_batch=16)
    try:
        tickets = [srv.submit(q[3], q[4], q[5], q[6]) for q in reqs[:4]]
        t_bad = srv.submit(bad, wits[0], rs[0], rs[1])
        tickets += [srv.submit(q[3], q[4], q[5], q[6]) for q in reqs[4:]]
        got = [srv.wait(t)[1] for t in tickets]
        with pytest.raises(N.NkrError, match="alpha/beta/delta"):
            srv.wait(t_bad)
        st = srv.stats()
    finally:
        srv.close()
        bad.close()
    for g, want in zip(got, alone):
        assert np.array_equal(g, want)
    assert st["proofs"] == len(reqs) + 1


def test_warm_up_sizes_every_worker(contexts):
    """nkr_g16_server_warm: every worker proves the warm-up items once (the
    outputs equal the proofs made alone), each worker then holds buffers for
    them (within its share), requests submitted during the warm-up wait for
    it and still equal their proofs made alone, and invalid warm-ups are
    rejected without stopping the server."""
    reqs = requests(contexts, 12, 57)
    alone = [pv.prove_blind(w, r, s) for (_, _, _, pv, w, r, s) in reqs]
    srv = N.Server(workers=3, max_batch=4)
    try:
        assert srv.memory()["held"] == 0
        items = [(q[3], q[4], q[5], q[6]) for q in reqs[:3]]
        outs = srv.warm(items)
        for (raw, proof), (raw1, proof1) in zip(outs, alone[:3]):
            assert np.array_equal(raw, raw1) and np.array_equal(proof, proof1)
        mem = srv.memory()
        # each of the 3 workers holds at least the items' witness buffers
        witness_bytes = sum(q[3].cols * 32 for q in reqs[:3])
        assert mem["held"] >= 3 * witness_bytes, mem
        assert mem["held"] <= mem["budget"] and mem["peak_held"] <= mem["budget"], mem
        # a second warm-up while requests arrive: they wait, then complete
        tickets = []
        done = threading.Event()

        def submit_during():
            for q in reqs[3:]:
                tickets.append(srv.submit(q[3], q[4], q[5], q[6]))
            done.set()

        th = threading.Thread(target=submit_during)
        th.start()
        srv.warm(items[:1])
        th.join()
        assert done.is_set()
        got = [srv.wait(t) for t in tickets]
        for (raw, proof, _), (raw1, proof1) in zip(got, alone[3:]):
            assert np.array_equal(raw, raw1) and np.array_equal(proof, proof1)
        with pytest.raises(N.NkrError):
            srv.warm(items * 2)  # more than max_batch
        # the server still serves after a rejected warm-up
        t = srv.submit(reqs[0][3], reqs[0][4], reqs[0][5], reqs[0][6])
        assert np.array_equal(srv.wait(t)[1], alone[0][1])
    finally:
        srv.close()

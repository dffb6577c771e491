"""Batched proving (nkr_g16_batch_prove): a mix of circuits and sizes, with
one context used several times with different witnesses, gives for every
item exactly the proof that proving it alone gives, and every proof
verifies.  The golden matmul2 item is also checked against its fixture."""
import numpy as np
import pytest

import circuits
import goldens
import nkr_g16 as N

pytestmark = pytest.mark.gpu


def keyed(c, seed):
    return N.setup(c, circuits.random_scalars(5, seed))


@pytest.fixture(scope="module")
def mix():
    e = goldens.E2E()
    cases = []
    specs = [("matmul4", circuits.matmul(4, 1), [circuits.matmul(4, 2), circuits.matmul(4, 3)]),
             ("matmul8s", circuits.matmul(8, 5, density="sparse", wide=True), []),
             ("fib300", circuits.fibonacci(300, 1), [circuits.fibonacci(300, 9)]),
             ("poseidon2", circuits.poseidon(2, 4), []),
             ("poseidon3f", circuits.poseidon(3, 6, folded=True), [])]
    for i, (name, c, others) in enumerate(specs):
        k = keyed(c, 100 + i)
        pv = N.Prover(c, k)
        for w in [c.witness] + [o.witness for o in others]:
            cases.append((name, pv, k, c, w))
    gold = N.Prover(e, e)
    yield e, gold, cases
    for _, pv, *_ in cases:
        pv.close()
    gold.close()


def test_batch_equals_individual(mix):
    e, gold, cases = mix
    rs = circuits.random_scalars(2 * len(cases), 77)
    items = [(pv, w, rs[2 * i], rs[2 * i + 1]) for i, (_, pv, _, _, w) in enumerate(cases)]
    items.append((gold, e.witness, e.r, e.s))
    b = N.Batch()
    outs = b.prove(items)
    for i, (name, pv, k, c, w) in enumerate(cases):
        raw, proof = pv.prove_blind(w, rs[2 * i], rs[2 * i + 1])
        assert np.array_equal(outs[i][0], raw), name
        assert np.array_equal(outs[i][1], proof), name
        assert N.verify(k, w[1:1 + c.public_inputs], outs[i][1]), name
    assert np.array_equal(outs[-1][0], e.raw) and np.array_equal(outs[-1][1], e.proof)
    # a second, smaller batch on the same handle (buffers are reused)
    sub = [items[3], items[0]]
    outs2 = b.prove(sub)
    assert np.array_equal(outs2[0][1], outs[3][1]) and np.array_equal(outs2[1][1], outs[0][1])
    b.close()


def test_batch_rejects_bad_items(mix):
    e, gold, cases = mix
    b = N.Batch()
    with pytest.raises(N.NkrError, match="null"):
        N._check(N.lib().nkr_g16_batch_prove(b.h, None, 1))
    assert b.prove([]) == []
    b.close()


def test_batch_graph_replays_match_plain(mix):
    """A batch repeating the previous call's contexts is captured into a HIP
    graph on its second call and replayed after; a batch of other contexts
    runs eagerly and drops the graph.  Every call gets fresh witnesses (the
    witnesses of one context rotate among its items) and fresh r, s, and
    every proof of every call -- eager, captured and replayed, with eager
    proofs and host synchronisations in between -- equals the plain path's."""
    e, gold, cases = mix
    b = N.Batch()
    by_pv = {}
    for i, (_, pv, _, _, w) in enumerate(cases):
        by_pv.setdefault(id(pv), []).append(i)

    def items_for(call, idx):
        rs = circuits.random_scalars(2 * len(idx), 900 + call)
        out = []
        for n, i in enumerate(idx):
            name, pv, k, c, _ = cases[i]
            group = by_pv[id(pv)]
            w = cases[group[(group.index(i) + call) % len(group)]][4]
            out.append((name, pv, k, c, w, rs[2 * n], rs[2 * n + 1]))
        return out

    def check(call, idx):
        its = items_for(call, idx)
        outs = b.prove([(pv, w, r, s) for (_, pv, _, _, w, r, s) in its])
        for (name, pv, k, c, w, r, s), (raw, proof) in zip(its, outs):
            raw1, proof1 = pv.prove_blind(w, r, s)
            assert np.array_equal(raw, raw1) and np.array_equal(proof, proof1), (call, name)
            assert N.verify(k, w[1:1 + c.public_inputs], proof), (call, name)
        return b.graph_active()

    every = list(range(len(cases)))
    sub = every[3:6]
    assert not check(0, every)      # eager
    assert check(1, every)          # captured, then replayed
    assert check(2, every)          # replayed
    assert check(3, every)          # replayed
    assert not check(4, sub)        # another composition: eager, graph dropped
    assert not check(5, every)      # eager again
    assert check(6, every)          # captured again
    b.set_graphs(False)
    assert not check(7, every) and not check(8, every)  # capture off: eager
    b.close()


def test_batch_upload_chunk_edges():
    """The batch's packed upload lays witnesses end to end from multiples of
    64 and cuts them into chunks of 2^15 elements: witnesses of 35841 columns
    (two chunks, the second partial, a length that is not a multiple of 64)
    between small ones, one of them dense, one with a chunk of zeros and one
    with a run of ones, each give the proof the single-witness upload gives
    (the valid one also verifies)."""
    import pyoracle as O

    base = circuits.matmul(32, 11)
    live = 1 + 3 * 32 * 32 + 32 * 32 * 32
    assert live % 64 and live > 1 << 15 and not base.witness[live:].any()
    c = circuits.R1cs(base.rows, live, base.public_inputs, base.a, base.b, base.c,
                      np.ascontiguousarray(base.witness[:live]))
    small = circuits.fibonacci(300, 1)
    k, ks = keyed(c, 301), keyed(small, 302)
    pv, pvs = N.Prover(c, k), N.Prover(small, ks)
    one = np.array(O.int_to_words(O.to_mont_int(1, O.R)), dtype=np.uint64)
    w_zero = c.witness.copy()
    w_zero[1 << 14:(1 << 14) + (1 << 15)] = 0
    w_ones = c.witness.copy()
    w_ones[100:30000] = one
    ws = [(pv, c.witness), (pvs, small.witness), (pv, w_zero), (pv, w_ones), (pvs, small.witness)]
    rs = circuits.random_scalars(2 * len(ws), 303)
    items = [(p, w, rs[2 * i], rs[2 * i + 1]) for i, (p, w) in enumerate(ws)]
    b = N.Batch()
    for rep in range(2):  # the second call reuses the grown buffers
        outs = b.prove(items)
        for i, (p, w, r, s) in enumerate(items):
            raw, proof = p.prove_blind(w, r, s)
            assert np.array_equal(outs[i][0], raw), (rep, i)
            assert np.array_equal(outs[i][1], proof), (rep, i)
    assert N.verify(k, c.witness[1:1 + c.public_inputs], outs[0][1])
    b.close()
    pv.close()
    pvs.close()

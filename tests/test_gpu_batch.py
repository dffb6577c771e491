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

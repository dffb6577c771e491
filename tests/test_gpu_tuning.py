"""Per-context tuning (nkr_g16_ctx_create_tuned) and the autotuner: a proof
does not depend on the tuning, so contexts built with non-default MSM windows,
transform shapes, piece targets and piece minima must give the same raw and
blinded proofs as the default context, bit for bit; invalid tunings are
rejected; the autotuner's search returns a tuning of the same proof."""
import numpy as np
import pytest

import autotune as A
import nkr_g16 as N
import pyoracle as O
from test_gpu_prove import affine_raw, matmul_inputs, synthetic_pk

pytestmark = pytest.mark.gpu

TUNINGS = [
    {"witness_window_bits": 9, "h_window_bits": 8},
    {"ntt_lds_log": 10, "ntt_radix_log": 3, "ntt_split_policy": 1},
    {"msm_target_pieces_log": 12, "witness_g1_piece_min": 64},
    {"witness_window_bits": 13, "h_window_bits": 14, "ntt_split_policy": 2, "msm_target_pieces_log": 20,
     "witness_g1_piece_min": 1},
] + [t for t in A.TuningTable.shipped().entries.values() if t]  # every tuning the package ships


def test_tuned_contexts_prove_the_same():
    circ, w = matmul_inputs(16, seed=0x7E5)  # domain 2^13: two transform passes
    pk = synthetic_pk(circ.rows, circ.cols, 55)
    rs = O.fr_random(0x7E6, 2)
    base = N.Prover(circ, pk)
    want_raw, want_proof = base.prove_blind(w, rs[0], rs[1])
    assert np.array_equal(want_raw, affine_raw(O.prove(circ, pk, w)))
    base.close()
    for t in TUNINGS:
        pv = N.Prover(circ, pk, tuning=t)
        raw, proof = pv.prove_blind(w, rs[0], rs[1])
        assert np.array_equal(raw, want_raw), t
        assert np.array_equal(proof, want_proof), t
        pv.close()


@pytest.mark.parametrize("bad,msg", [({"witness_window_bits": 16}, "witness window"),
                                     ({"h_window_bits": 1}, "H window"),
                                     ({"ntt_lds_log": 12}, "LDS elements"),
                                     ({"ntt_radix_log": 1}, "register radix"),
                                     ({"ntt_split_policy": 3}, "split policy"),
                                     ({"msm_target_pieces_log": 30}, "target pieces"),
                                     ({"witness_g1_piece_min": 65}, "piece minimum")])
def test_invalid_tunings_are_rejected(bad, msg):
    circ, _ = matmul_inputs(4, seed=3)
    pk = synthetic_pk(circ.rows, circ.cols, 5)
    with pytest.raises(N.NkrError, match=msg):
        N.Prover(circ, pk, tuning=bad)


def test_autotune_search_keeps_the_proof():
    circ, w = matmul_inputs(8, seed=0x7E7)
    pk = synthetic_pk(circ.rows, circ.cols, 57)
    rs = O.fr_random(0x7E8, 2)
    knobs = {"h_window_bits": [0, 6], "witness_g1_piece_min": [0, 8]}
    best, report = A.autotune(circ, pk, w, rs[0], rs[1], proves=3, knobs=knobs)
    assert len(report) >= 3 and all(ms > 0 for _, _, ms in report)
    assert report[0][0] == "default" and not any(report[0][1].values())
    assert set(best) <= {"h_window_bits", "witness_g1_piece_min"}
    pv = N.Prover(circ, pk, tuning=best)
    raw = pv.prove(w)
    assert np.array_equal(raw, affine_raw(O.prove(circ, pk, w)))
    pv.close()

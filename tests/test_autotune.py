"""Host logic of the autotuner (nkr-zkp_amd/autotune.py): density classes,
table keys, the JSON table and its nearest-size lookup, and the tuning
struct's field checks.  The measured search itself runs in the GPU tests."""
import numpy as np
import pytest

import autotune as A
import nkr_g16 as N


def witness(n, nonzero):
    w = np.zeros((n, 4), dtype=np.uint64)
    w[: int(n * nonzero), 0] = 7
    return w


def test_density_classes_and_keys():
    assert A.density_class(witness(1000, 0.02)) == "sparse"
    assert A.density_class(witness(1000, 0.2)) == "hybrid"
    assert A.density_class(witness(1000, 0.5)) == "full"
    assert A.density_class(witness(1000, 1.0)) == "full"
    assert A.table_key(1 << 20, witness(64, 0.03)) == "2^20/sparse"


def test_table_round_trip_and_lookup(tmp_path):
    t = A.TuningTable()
    t.put(1 << 20, witness(64, 0.03), {"witness_g1_piece_min": 64})
    t.put(1 << 13, witness(64, 1.0), {"ntt_lds_log": 10, "h_window_bits": 9})
    p = tmp_path / "t.json"
    t.save(str(p))
    u = A.TuningTable.load(str(p))
    assert u.lookup(1 << 20, witness(64, 0.03)) == {"witness_g1_piece_min": 64}
    # nearest size with the same density
    assert u.lookup(1 << 16, witness(64, 1.0)) == {"ntt_lds_log": 10, "h_window_bits": 9}
    assert u.lookup(1 << 18, witness(64, 0.02)) == {"witness_g1_piece_min": 64}
    assert u.lookup(1 << 16, witness(64, 0.2)) is None


def test_tuning_fields_are_checked(tmp_path):
    t = N.Tuning.of({"ntt_lds_log": 10, "msm_target_pieces_log": 18})
    assert t.ntt_lds_log == 10 and t.msm_target_pieces_log == 18 and t.h_window_bits == 0
    with pytest.raises(ValueError):
        N.Tuning.of({"window": 12})
    p = tmp_path / "bad.json"
    p.write_text('{"2^10/full": {"bogus": 1}}')
    with pytest.raises(ValueError):
        A.TuningTable.load(str(p))


def test_knob_alternatives_stay_in_range():
    for rows, cols in ((1 << 10, 1 << 10), (1 << 13, 1 << 13), (1 << 20, 1 << 20)):
        k = A.knob_alternatives(rows, cols)
        assert set(k) == set(N.Tuning.FIELDS)
        assert all(0 in v for v in k.values())
        assert all(2 <= c <= 15 for c in k["witness_window_bits"] if c)
        assert all(2 <= c <= 16 for c in k["h_window_bits"] if c)


def test_shipped_table():
    """The table measured on MI355X loads, names only known fields, and keeps
    the library defaults for the 2^20 sparse workload it was tuned on."""
    t = A.TuningTable.shipped()
    assert "2^20/sparse" in t.entries and t.entries["2^20/sparse"] == {}
    for key, tuning in t.entries.items():
        size, dens = key.split("/")
        assert size.startswith("2^") and dens in ("sparse", "hybrid", "full")
        N.Tuning.of(tuning)

    class Circ:
        rows = 1 << 20

    assert A.tuning_for(Circ, witness(64, 0.01)) == {}
    assert A.tuning_for(Circ, witness(64, 1.0)) == t.entries["2^20/full"]

"""Autotuner for the per-context choices of the Groth16 backend
(nkr_g16_tuning in include/nkr_g16.h): MSM window widths of the witness and H
MSMs, the transform tile, radix and pass split, the MSM piece target and the
witness G1 piece minimum.  None of them changes a proof, only its time, so
the tuner checks that every candidate's raw proof equals the default one and
keeps the fastest.

The search is one knob at a time from the defaults (each knob's alternatives
are tried with the best values found so far for the others), timing plain
proofs with the witness resident in HBM, as bench.py does.  A candidate
replaces the best only when it is faster by more than `min_gain` (timing
noise is about 0.5 %), and the winner is confirmed at the end by timing it
and the defaults in alternation on fresh contexts: it is kept only if it
still wins by `min_gain`, so a slow first measurement cannot promote it.
Results are kept per (domain size, witness density) in a JSON table that
`TuningTable.lookup` serves to `nkr_g16.Prover(..., tuning=...)`; the table
measured on MI355X ships as `tuning_mi355x.json` beside this file.

    python autotune.py --size 100 --density sparse --out tuning.json
    python autotune.py --circuit poseidon --hashes 4096 --out tuning.json
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import nkr_g16 as N  # noqa: E402

FIELDS = N.Tuning.FIELDS


def density_class(witness) -> str:
    """Class of the fraction f of non-zero scalars in the witness vector the MSMs
    read (zero padding to the power-of-two column count included): f < 0.06
    sparse, f < 0.45 hybrid, else full.  The matmul workloads of those names
    land at f = 0.011-0.017, 0.14-0.26 and 0.50-0.98 for domains 2^13..2^20
    (products of sparse inputs are sparser than the inputs)."""
    w = np.ascontiguousarray(np.asarray(witness, dtype=np.uint64)).reshape(-1, 4)
    frac = float(np.count_nonzero(w.any(axis=1))) / max(1, w.shape[0])
    return "sparse" if frac < 0.06 else ("hybrid" if frac < 0.45 else "full")


def table_key(rows: int, witness) -> str:
    return f"2^{int(math.log2(rows))}/{density_class(witness)}"


def default_window(n: int) -> int:
    """The library's window rule (MsmPlan::choose) without the top-window nudge."""
    lg = max(1, math.ceil(math.log2(max(2, n))))
    return 16 if lg >= 20 else max(4, min(15, (lg * 3 + 2) // 4))


def knob_alternatives(rows: int, cols: int) -> dict[str, list[int]]:
    """Values tried per knob (0 = the library default)."""
    cw = min(15, default_window(cols))
    ch = default_window(rows)
    return {
        "witness_window_bits": [0] + [c for c in (cw - 1, cw + 1) if 2 <= c <= 15 and c != cw],
        "h_window_bits": [0] + [c for c in (ch - 1, ch + 1) if 2 <= c <= 16 and c != ch],
        "ntt_lds_log": [0, 10],
        "ntt_radix_log": [0, 3],
        "ntt_split_policy": [0, 1],
        "msm_target_pieces_log": [0, 18, 20],
        "witness_g1_piece_min": [0, 16, 64],
    }


class _Timer:
    """Plain proofs of one witness, resident in HBM, on contexts of one circuit."""

    def __init__(self, circ, pk, witness, r, s, device: int = 0):
        self.circ, self.pk, self.device = circ, pk, device
        self.w = np.ascontiguousarray(np.asarray(witness, dtype=np.uint64))
        self.r = np.ascontiguousarray(np.asarray(r, dtype=np.uint64).reshape(4))
        self.s = np.ascontiguousarray(np.asarray(s, dtype=np.uint64).reshape(4))
        self.dptr = ctypes.c_void_p()
        N._check(N.lib().nkr_dev_alloc(self.w.nbytes, ctypes.byref(self.dptr)))
        N._check(N.lib().nkr_memcpy_htod(self.dptr, N._p(self.w), self.w.nbytes))

    def close(self) -> None:
        if self.dptr:
            N._check(N.lib().nkr_dev_free(self.dptr))
            self.dptr = ctypes.c_void_p()

    def measure(self, tuning: dict, proves: int, warmup: int = 2):
        """(median ms per proof, raw proof, proof) for one candidate."""
        pv = N.Prover(self.circ, self.pk, device=self.device, tuning={k: v for k, v in tuning.items() if v})
        raw = np.zeros(72, dtype=np.uint64)
        proof = np.zeros(48, dtype=np.uint64)
        try:
            def step():
                N._check(N.lib().nkr_g16_prove_blind_dev(pv.h, self.dptr, N._p(self.r), N._p(self.s), N._p(raw),
                                                         N._p(proof)))
            for _ in range(warmup):
                step()
            N._check(N.lib().nkr_device_synchronize())
            ts = []
            for _ in range(proves):
                t0 = time.perf_counter()
                step()
                ts.append(1e3 * (time.perf_counter() - t0))
            return sorted(ts)[len(ts) // 2], raw.copy(), proof.copy()
        finally:
            pv.close()


def autotune(circ, pk, witness, r, s, proves: int = 10, knobs: dict | None = None, device: int = 0,
             log=None, min_gain: float = 0.01, confirm_rounds: int = 3) -> tuple[dict, list]:
    """Best tuning (dict of non-default fields) for one circuit and witness, and
    the report [(label, tuning, median ms)] of every measurement."""
    knobs = knobs or knob_alternatives(circ.rows, circ.cols)
    timer = _Timer(circ, pk, witness, r, s, device)
    say = log or (lambda m: None)
    try:
        default = {k: 0 for k in FIELDS}
        best = dict(default)
        best_ms, ref_raw, ref_proof = timer.measure(best, proves)
        report = [("default", dict(best), best_ms)]
        say(f"default: {best_ms:.3f} ms")

        def timed(label, cand):
            ms, raw, proof = timer.measure(cand, proves)
            if not (np.array_equal(raw, ref_raw) and np.array_equal(proof, ref_proof)):
                raise RuntimeError(f"tuning {cand} changed the proof")
            report.append((label, dict(cand), ms))
            return ms

        for knob, values in knobs.items():
            for v in values:
                if v == best[knob]:
                    continue
                cand = dict(best)
                cand[knob] = v
                ms = timed(f"{knob}={v}", cand)
                say(f"{knob}={v}: {ms:.3f} ms")
                if ms < best_ms * (1.0 - min_gain):
                    best, best_ms = cand, ms
        if best != default and confirm_rounds > 0:
            d_ms, b_ms = [], []
            for i in range(confirm_rounds):
                d_ms.append(timed(f"confirm default #{i}", default))
                b_ms.append(timed(f"confirm best #{i}", best))
            d, b = float(np.median(d_ms)), float(np.median(b_ms))
            keep = b < d * (1.0 - min_gain)
            say(f"confirm: default {d:.3f} ms, best {b:.3f} ms -> {'kept' if keep else 'defaults'}")
            if not keep:
                best = default
        return {k: v for k, v in best.items() if v}, report
    finally:
        timer.close()


SHIPPED_TABLE = os.path.join(HERE, "tuning_mi355x.json")


class TuningTable:
    """Tunings per (domain size, witness density), stored as JSON."""

    @classmethod
    def shipped(cls) -> "TuningTable":
        """The table measured on MI355X that ships with the package."""
        return cls.load(SHIPPED_TABLE)

    def __init__(self, entries: dict | None = None):
        self.entries = dict(entries or {})

    @classmethod
    def load(cls, path: str) -> "TuningTable":
        with open(path) as f:
            data = json.load(f)
        for key, t in data.items():
            N.Tuning.of(t)  # rejects unknown fields
        return cls(data)

    def save(self, path: str) -> None:
        with open(path, "w") as f:
            json.dump(self.entries, f, indent=1, sort_keys=True)

    def put(self, rows: int, witness, tuning: dict) -> None:
        self.entries[table_key(rows, witness)] = dict(tuning)

    def lookup(self, rows: int, witness) -> dict | None:
        """The tuning of this size and density, else of the nearest size with
        the same density, else None (library defaults)."""
        key = table_key(rows, witness)
        if key in self.entries:
            return dict(self.entries[key])
        dens = key.split("/")[1]
        lg = int(math.log2(rows))
        near = [(abs(int(k.split("/")[0][2:]) - lg), k) for k in self.entries if k.endswith("/" + dens)]
        return dict(self.entries[min(near)[1]]) if near else None


def tuning_for(circ, witness) -> dict | None:
    """The shipped table's tuning for this circuit and witness (None: the
    library defaults), ready for nkr_g16.Prover(..., tuning=...)."""
    return TuningTable.shipped().lookup(circ.rows, witness)


def main() -> int:
    import circuits
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--circuit", choices=["matmul", "poseidon"], default="matmul")
    ap.add_argument("--size", type=int, default=100, help="matmul dimension s")
    ap.add_argument("--density", default="sparse")
    ap.add_argument("--hashes", type=int, default=4096, help="Poseidon-style hashes (folded rounds)")
    ap.add_argument("--proves", type=int, default=10)
    ap.add_argument("--seed", type=lambda x: int(x, 0), default=0x4E4B52)
    ap.add_argument("--min-gain", type=float, default=0.01)
    ap.add_argument("--out", default="tuning.json")
    args = ap.parse_args()
    if args.circuit == "poseidon":
        circ = circuits.poseidon(args.hashes, args.seed, folded=True)
    else:
        circ = circuits.matmul(args.size, args.seed, density=args.density, wide=True)
    pk = N.setup(circ, circuits.random_scalars(5, args.seed ^ 0x5E7))
    rs = circuits.random_scalars(2, args.seed ^ 0x72)
    best, _ = autotune(circ, pk, circ.witness, rs[0], rs[1], proves=args.proves, min_gain=args.min_gain,
                       log=lambda m: print(f"[autotune] {m}", flush=True))
    table = TuningTable.load(args.out) if os.path.exists(args.out) else TuningTable()
    table.put(circ.rows, circ.witness, best)
    table.save(args.out)
    print(json.dumps({"key": table_key(circ.rows, circ.witness), "tuning": best}))
    return 0


if __name__ == "__main__":
    sys.exit(main())

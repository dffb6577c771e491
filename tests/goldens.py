"""Loaders for the committed golden fixtures (tests/golden/*.json) as numpy word arrays."""
import json
import os
from functools import lru_cache

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@lru_cache(maxsize=None)
def load(name: str):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def hex_words(h: str) -> list:
    x = int(h, 16)
    return [(x >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)]


def fe(hs) -> np.ndarray:
    """list of hex Montgomery residues -> (n, 4) uint64"""
    return np.array([hex_words(h) for h in hs], dtype=np.uint64).reshape(-1, 4)


def point(p) -> np.ndarray:
    """G1 [X, Y, Z] or G2 [x0, x1, y0, y1, z0, z1] hex list -> flat uint64 words"""
    return np.array([w for h in p for w in hex_words(h)], dtype=np.uint64)


def points(ps) -> np.ndarray:
    return np.stack([point(p) for p in ps]) if ps else np.zeros((0, 12), dtype=np.uint64)


class E2E:
    """The tiny end-to-end Groth16 fixture as arrays the oracle and the backend take."""

    def __init__(self, name="groth16_matmul2.json"):
        d = load(name)
        self.d = d
        self.rows, self.cols, self.public_inputs = d["rows"], d["cols"], d["public_inputs"]

        def mat(m):
            return (np.array(m[0], dtype=np.uint32), np.array(m[1], dtype=np.uint32), fe(m[2]))
        self.a, self.b, self.c = mat(d["a"]), mat(d["b"]), mat(d["c"])
        self.witness = fe(d["witness"])
        pk = d["pk"]
        self.a_query = points(pk["a_query"])
        self.b_g1_query = points(pk["b_g1_query"])
        self.b_g2_query = points(pk["b_g2_query"])
        self.l_query = points(pk["l_query"])
        self.h_query = points(pk["h_query"])
        for k in ("alpha_g1", "beta_g1", "beta_g2", "delta_g1", "delta_g2"):
            setattr(self, k, point(pk[k]))
        self.r = fe([d["r"]])[0]
        self.s = fe([d["s"]])[0]
        self.raw = np.concatenate([point(d["raw"][k]) for k in ("ar", "bs1", "zk", "qz", "bs2")])
        self.proof = np.concatenate([point(d["proof"][k]) for k in ("a", "b", "c")])

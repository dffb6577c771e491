"""Seeded synthetic Groth16 inputs: R1CS circuits with witnesses, and
synthetic proving keys.

The reference generates circuits with std::random_device
(protocols/groth16/app/gen.cpp:42-84, common/include/density.hpp:72-103), so
they cannot be reproduced; these generators keep the reference's constraint
structure and draw every value from mt19937_64 streams of one seed:

  matmul(s): for i, j, l < s one multiplication constraint
      P[i][j][l] = A[i][l] * B[l][j]       (A row {A_il: 1}, B row {B_lj: 1}, C row {P_ijl: 1})
  then for i, j < s one linear constraint
      C[i][j] = sum_l P[i][j][l]           (A row {0: 0, P_ij0..: 1}, B row {0: 1}, C row {C_ij: 1})
  Columns: the constant 1, A, B, C (public, s^2 each), P (private, s^3);
  rows and columns padded to powers of two (r1cs.hpp:64-99).
  A and B entries: uniform u32 (wide=False) or uniform Fr (wide=True, the
  libff random_element rule), drawn from mt19937_64(seed ^ 0xA11CE) and
  mt19937_64(seed ^ 0xB0B); zero positions for hybrid / sparse density come
  from a Fisher-Yates shuffle driven by mt19937_64(seed ^ 0x3A5C0 / 0x3A5C1),
  with ceil(count * density) non-zeros (density.hpp:55-64).
"""
from __future__ import annotations

import numpy as np

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
Q = 21888242871839275222246405745257275088696311157297823662689037894645226208583
MONT = 1 << 256
_M64 = 0xFFFFFFFFFFFFFFFF


class MT64:
    """mt19937_64 (the C++ standard engine, [rand.predef])."""

    def __init__(self, seed: int):
        mt = [0] * 312
        mt[0] = seed & _M64
        for i in range(1, 312):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & _M64
        self.mt, self.idx = mt, 312

    def next(self) -> int:
        if self.idx >= 312:
            mt = self.mt
            for i in range(312):
                y = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
                v = mt[(i + 156) % 312] ^ (y >> 1)
                mt[i] = v ^ 0xB5026F5AA96619E9 if y & 1 else v
            self.idx = 0
        x = self.mt[self.idx]
        self.idx += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        return (x ^ (x >> 43)) & _M64


def fr_random_mont(rng: MT64) -> int:
    """libff Fp_model::random_element: 4 words, bits 254-255 cleared, reject >= r."""
    while True:
        x = rng.next() | (rng.next() << 64) | (rng.next() << 128) | ((rng.next() & 0x3FFFFFFFFFFFFFFF) << 192)
        if x < R:
            return x


def to_words(values_mont: list[int]) -> np.ndarray:
    """Montgomery residues (ints) -> (n, 4) uint64."""
    out = np.zeros((len(values_mont), 4), dtype=np.uint64)
    for k in range(4):
        out[:, k] = [(v >> (64 * k)) & _M64 for v in values_mont]
    return out


def fr_mont(values: list[int]) -> np.ndarray:
    return to_words([v % R * MONT % R for v in values])


def density_nonzero(count: int, density: str) -> int:
    return {"full": count, "hybrid": count - count // 2, "sparse": count // 10 + (count % 10 != 0)}[density]


def seeded_mask(count: int, nonzero: int, seed: int) -> list[int]:
    mask = [1] * nonzero + [0] * (count - nonzero)
    rng = MT64(seed)
    for i in range(count - 1, 0, -1):
        j = rng.next() % (i + 1)
        mask[i], mask[j] = mask[j], mask[i]
    return mask


class R1cs:
    """Circuit arrays in the reference layout plus its witness."""

    def __init__(self, rows, cols, public_inputs, a, b, c, witness):
        self.rows, self.cols, self.public_inputs = rows, cols, public_inputs
        self.a, self.b, self.c = a, b, c      # (row_ptr u32, col_idx u32, values (nnz, 4) u64)
        self.witness = witness                # (cols, 4) u64, Montgomery


def matmul(s: int, seed: int, density: str = "full", wide: bool = False) -> R1cs:
    ss = s * s
    nz = density_nonzero(ss, density)
    vals = []
    for tag, mtag in ((0xA11CE, 0x3A5C0), (0xB0B, 0x3A5C1)):
        mask = seeded_mask(ss, nz, seed ^ mtag) if nz < ss else [1] * ss
        rng = MT64(seed ^ tag)
        out = []
        for i in range(ss):
            if not mask[i]:
                out.append(0)
                continue
            while True:
                v = (fr_random_mont(rng) * pow(MONT, -1, R) % R) if wide else rng.next() & 0xFFFFFFFF
                if v:
                    break
            out.append(v)
        vals.append(out)
    av, bv = vals
    live_rows = s * ss + ss
    live_cols = 1 + 3 * ss + s * ss
    rows = 1 << (live_rows - 1).bit_length()
    cols = 1 << (live_cols - 1).bit_length()
    colA, colB, colC, colP = 1, 1 + ss, 1 + 2 * ss, 1 + 3 * ss

    one = np.array([(MONT % R) >> (64 * k) & _M64 for k in range(4)], dtype=np.uint64)
    i, j, l = np.meshgrid(np.arange(s), np.arange(s), np.arange(s), indexing="ij")
    i, j, l = i.ravel(), j.ravel(), l.ravel()            # product rows in (i, j, l) order
    prod_rows = s * ss
    # A: products {A_il}, then sums {0: 0, P_ij0 .. P_ij(s-1): 1}
    a_cols = np.concatenate([colA + i * s + l,
                             np.stack([np.zeros(ss, dtype=np.int64)] +
                                      [colP + (np.arange(ss) // s) * ss + (np.arange(ss) % s) * s + ll
                                       for ll in range(s)], axis=1).ravel()])
    a_vals = np.zeros((a_cols.size, 4), dtype=np.uint64)
    a_vals[:prod_rows] = one
    sum_part = a_vals[prod_rows:].reshape(ss, s + 1, 4)
    sum_part[:, 1:] = one
    a_rp = np.concatenate([np.arange(prod_rows + 1), prod_rows + (s + 1) * np.arange(1, ss + 1)])
    # B: products {B_lj}, sums {0: 1}
    b_cols = np.concatenate([colB + l * s + j, np.zeros(ss, dtype=np.int64)])
    b_vals = np.tile(one, (b_cols.size, 1))
    # C: products {P_ijl}, sums {C_ij}
    c_cols = np.concatenate([colP + np.arange(prod_rows), colC + np.arange(ss)])
    c_vals = np.tile(one, (c_cols.size, 1))
    simple_rp = np.arange(live_rows + 1)

    def pad(rp):
        return np.concatenate([rp, np.full(rows - live_rows, rp[-1])]).astype(np.uint32)

    # witness
    z = [0] * cols
    z[0] = 1
    for ii in range(s):
        for jj in range(s):
            acc = 0
            for ll in range(s):
                p = av[ii * s + ll] * bv[ll * s + jj] % R
                z[colP + ii * ss + jj * s + ll] = p
                acc += p
            z[colC + ii * s + jj] = acc % R
    for k in range(ss):
        z[colA + k] = av[k]
        z[colB + k] = bv[k]
    witness = fr_mont(z)
    return R1cs(rows, cols, 3 * ss,
                (pad(a_rp), a_cols.astype(np.uint32), a_vals),
                (pad(simple_rp), b_cols.astype(np.uint32), b_vals),
                (pad(simple_rp), c_cols.astype(np.uint32), c_vals),
                witness)


def random_scalars(n: int, seed: int) -> np.ndarray:
    """Uniform Fr Montgomery residues (libff rejection rule) from numpy's PCG64,
    for synthetic proving keys and operator inputs."""
    rng = np.random.default_rng(seed)
    r_words = np.array([(R >> (64 * k)) & _M64 for k in range(4)], dtype=np.uint64)
    out = np.zeros((n, 4), dtype=np.uint64)
    filled = 0
    while filled < n:
        m = int((n - filled) * 1.3) + 16
        w = rng.integers(0, 2**64, size=(m, 4), dtype=np.uint64)
        w[:, 3] &= np.uint64(0x3FFFFFFFFFFFFFFF)
        lt = np.zeros(m, dtype=bool)
        eq = np.ones(m, dtype=bool)
        for k in (3, 2, 1, 0):
            lt |= eq & (w[:, k] < r_words[k])
            eq &= w[:, k] == r_words[k]
        good = w[lt]
        take = min(n - filled, good.shape[0])
        out[filled:filled + take] = good[:take]
        filled += take
    return out

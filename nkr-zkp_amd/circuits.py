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

The same generators exist in C++ (nkr-zkp_amd/host/gen.cpp, the nkr-g16-gen
program, which writes the reference's .bin / .witness.bin files); generate()
runs it and reads its files back, which is how the 2^20-constraint workloads
are built quickly.  Both produce identical arrays.
"""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile

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


class R1csBuilder:
    """The reference's R1CSManager semantics (protocols/groth16/common/include/
    r1cs/r1cs.hpp:26-99,104-160) on canonical integers mod r: public ids
    count up from 1, private ids from 2^63; a linear combination is a dict
    id -> coefficient whose terms are emitted in id order, so public columns
    precede private ones; LinearCombination() holds {0: 0} and
    LinearCombination(c) holds {0: c}; add_constraint(out, in) is
    in * 1 = out; columns map to 1 + public..., then private...; rows and
    columns are padded to powers of two."""

    HALF = 1 << 63

    def __init__(self):
        self.n_public = 1
        self.n_private = self.HALF
        self.rows = []          # (a, b, c) dicts
        self.values = {}        # id -> witness value

    def public(self, value=0):
        i = self.n_public
        self.n_public += 1
        self.values[i] = value % R
        return i

    def private(self, value=0):
        i = self.n_private
        self.n_private += 1
        self.values[i] = value % R
        return i

    @staticmethod
    def lc(*terms, const=None):
        """terms: (id, coeff) pairs; const: None for a Variable-style LC (no
        constant term), else the constant at id 0."""
        d = {} if const is None else {0: const % R}
        for i, c in terms:
            d[i] = (d.get(i, 0) + c) % R
        return d

    @staticmethod
    def add(a, b):
        d = dict(a)
        for i, c in b.items():
            d[i] = (d.get(i, 0) + c) % R
        return d

    @staticmethod
    def scale(a, s):
        return {i: c * s % R for i, c in a.items()}

    def constrain(self, out, a, b=None):
        self.rows.append((a, {0: 1} if b is None else b, out))

    def dense(self, i):
        return i if i < self.HALF else i - self.HALF + self.n_public

    def finish(self) -> R1cs:
        live_rows = len(self.rows)
        live_cols = self.n_public + (self.n_private - self.HALF)
        rows = 1 << max(live_rows - 1, 0).bit_length()
        cols = 1 << max(live_cols - 1, 0).bit_length()
        mats = []
        for k in range(3):
            rp, ci, va = [0], [], []
            for row in self.rows:
                for i in sorted(row[k]):
                    ci.append(self.dense(i))
                    va.append(row[k][i])
                rp.append(len(ci))
            rp += [rp[-1]] * (rows - live_rows)
            mats.append((np.array(rp, dtype=np.uint32), np.array(ci, dtype=np.uint32).reshape(-1),
                         fr_mont(va) if va else np.zeros((0, 4), dtype=np.uint64)))
        z = [0] * cols
        z[0] = 1
        for i, v in self.values.items():
            z[self.dense(i)] = v
        return R1cs(rows, cols, self.n_public - 1, mats[0], mats[1], mats[2], fr_mont(z))


def fibonacci(steps: int, seed: int) -> R1cs:
    """build_fibonacci (protocols/groth16/app/gen.cpp:87-120): public F0, F1 and
    F_steps, one private term per step with F_i * 1 = F_(i-2) + F_(i-1), and a
    final F_steps * 1 = current.  F0, F1 uniform Fr from mt19937_64(seed ^ 0xF1B)."""
    if steps < 2:
        raise ValueError("fibonacci steps must be at least 2")
    rng = MT64(seed ^ 0xF1B)
    rinv = pow(MONT, -1, R)
    f0 = fr_random_mont(rng) * rinv % R
    f1 = fr_random_mont(rng) * rinv % R
    b = R1csBuilder()
    prev, cur = b.public(f0), b.public(f1)
    expected = b.public()
    pv, cv = f0, f1
    for _ in range(2, steps + 1):
        nv = (pv + cv) % R
        nxt = b.private(nv)
        b.constrain(b.lc((nxt, 1)), b.lc((prev, 1), (cur, 1)))
        pv, cv = cv, nv
        prev, cur = cur, nxt
    b.values[expected] = cv
    b.constrain(b.lc((expected, 1)), b.lc((cur, 1)))
    return b.finish()


class PoseidonParams:
    """Width 3, x^5 S-box, 8 full and 57 partial rounds (the reference's shape,
    common/include/crypto/poseidon.hpp:19-23) with seeded constants: 195 round
    constants and a 3x3 MDS matrix drawn as uniform Fr from mt19937_64(seed).
    (The reference's own constants come from its Grain LFSR generator; these
    are a stand-in of the same shape for synthetic workloads.)"""
    WIDTH, FULL, PARTIAL = 3, 8, 57

    def __init__(self, seed: int = 0x9053):
        rng = MT64(seed)
        rinv = pow(MONT, -1, R)
        draw = lambda: fr_random_mont(rng) * rinv % R
        self.rc = [draw() for _ in range(self.WIDTH * (self.FULL + self.PARTIAL))]
        self.mds = [[draw() for _ in range(3)] for _ in range(3)]

    def permute(self, state):
        s = list(state)
        k = 0
        for rnd in range(self.FULL + self.PARTIAL):
            s = [(x + self.rc[k + i]) % R for i, x in enumerate(s)]
            k += 3
            full = rnd < self.FULL // 2 or rnd >= self.FULL // 2 + self.PARTIAL
            s = [pow(x, 5, R) if (full or i == 0) else x for i, x in enumerate(s)]
            s = [sum(self.mds[r][c] * s[c] for c in range(3)) % R for r in range(3)]
        return s

    def hash2(self, left, right):
        return self.permute([0, left, right])[0]


def _poseidon_hash_constraints(b: R1csBuilder, prm: PoseidonParams, left, lv, right, rv, digest, folded: bool):
    """add_hash2_constraints (protocols/groth16/common/include/r1cs/poseidon.hpp:
    33-105): per round, state += constants; S-boxes as squared = s*s,
    fourth = squared^2, fifth = fourth*s; then one private output and one
    linear constraint per MDS row.  folded=True keeps the MDS rows as linear
    combinations instead (243 + 1 constraints per hash instead of 439)."""
    state = [b.lc(const=0), b.lc((left, 1)), b.lc((right, 1))]
    vals = [0, lv, rv]
    k = 0
    for rnd in range(prm.FULL + prm.PARTIAL):
        sboxed, sv = list(state), list(vals)
        for i in range(3):
            state[i] = b.add(state[i], b.lc(const=prm.rc[k]))
            vals[i] = (vals[i] + prm.rc[k]) % R
            k += 1
        full = rnd < prm.FULL // 2 or rnd >= prm.FULL // 2 + prm.PARTIAL
        nsb = 3 if full else 1
        for i in range(nsb):
            x = vals[i]
            x2 = x * x % R
            x4 = x2 * x2 % R
            x5 = x4 * x % R
            sq, fo, fi = b.private(x2), b.private(x4), b.private(x5)
            b.constrain(b.lc((sq, 1)), state[i], state[i])
            b.constrain(b.lc((fo, 1)), b.lc((sq, 1)), b.lc((sq, 1)))
            b.constrain(b.lc((fi, 1)), b.lc((fo, 1)), state[i])
            sboxed[i], sv[i] = b.lc((fi, 1)), x5
        for i in range(nsb, 3):
            sboxed[i], sv[i] = state[i], vals[i]
        nxt, nv = [], []
        for r in range(3):
            mixed = b.lc(const=0)
            for c in range(3):
                mixed = b.add(mixed, b.scale(sboxed[c], prm.mds[r][c]))
            v = sum(prm.mds[r][c] * sv[c] for c in range(3)) % R
            if folded:
                nxt.append(mixed)
            else:
                out = b.private(v)
                b.constrain(b.lc((out, 1)), mixed)
                nxt.append(b.lc((out, 1)))
            nv.append(v)
        state, vals = nxt, nv
    b.constrain(b.lc((digest, 1)), state[0])


def poseidon(count: int, seed: int, density: str = "full", folded: bool = False,
             params: PoseidonParams | None = None) -> R1cs:
    """build_poseidon (protocols/groth16/app/gen.cpp:136-162): count
    independent hashes with a public digest and private preimage (left, right),
    allocated per hash in the order digest, left, right, then the hash's
    intermediates.  Preimages uniform Fr from mt19937_64(seed ^ 0x9E5), with
    the density's zero positions from seeded_mask(2 * count, ..., seed ^ 0x3A5C2)."""
    prm = params or PoseidonParams()
    n_in = 2 * count
    nz = density_nonzero(n_in, density)
    mask = seeded_mask(n_in, nz, seed ^ 0x3A5C2) if nz < n_in else [1] * n_in
    rng = MT64(seed ^ 0x9E5)
    rinv = pow(MONT, -1, R)
    inputs = []
    for i in range(n_in):
        v = 0
        if mask[i]:
            while v == 0:
                v = fr_random_mont(rng) * rinv % R
        inputs.append(v)
    b = R1csBuilder()
    for h in range(count):
        lv, rv = inputs[2 * h], inputs[2 * h + 1]
        digest = b.public(prm.hash2(lv, rv))
        left, right = b.private(lv), b.private(rv)
        _poseidon_hash_constraints(b, prm, left, lv, right, rv, digest, folded)
    return b.finish()


def batch_mix(seed: int = 0x4E4B52, per_shape: int = 4, use_cpp: bool | None = None):
    """The heterogeneous batch: 16 circuit shapes with domains 2^9 .. 2^18 --
    matmul of size 8 .. 48 (full and sparse), Fibonacci chains of 2^10 .. 2^16
    steps, Poseidon batches of 1 .. 512 hashes (one folded) -- each proved for
    `per_shape` different witnesses (the circuit structure does not depend on
    the inputs, so one proving key serves all of them).  use_cpp: generate
    with the C++ generator (the same arrays, much faster); None uses it when
    it is built.
    Returns [(name, circuit, [witness, ...])]."""
    cpp = os.path.exists(GEN_BIN) if use_cpp is None else use_cpp

    def mm(s, sd, **kw):
        if cpp:
            return generate("matmul", s, sd, density=kw.get("density", "full"), wide=kw.get("wide", False))
        return matmul(s, sd, **kw)

    def fib(steps, sd):
        return generate("fibonacci", steps, sd) if cpp else fibonacci(steps, sd)

    def pos(count, sd, **kw):
        if cpp:
            return generate("poseidon", count, sd, density=kw.get("density", "full"), folded=kw.get("folded", False))
        return poseidon(count, sd, **kw)

    shapes = [("matmul8", lambda s: mm(8, s)), ("matmul16", lambda s: mm(16, s)),
              ("matmul24_sparse", lambda s: mm(24, s, density="sparse", wide=True)),
              ("matmul32", lambda s: mm(32, s)),
              ("matmul40_sparse", lambda s: mm(40, s, density="sparse", wide=True)),
              ("matmul48", lambda s: mm(48, s)),
              ("fibonacci1024", lambda s: fib(1 << 10, s)), ("fibonacci4096", lambda s: fib(1 << 12, s)),
              ("fibonacci16384", lambda s: fib(1 << 14, s)), ("fibonacci65536", lambda s: fib(1 << 16, s)),
              ("poseidon1", lambda s: pos(1, s)), ("poseidon8", lambda s: pos(8, s)),
              ("poseidon64_hybrid", lambda s: pos(64, s, density="hybrid")),
              ("poseidon128_folded", lambda s: pos(128, s, folded=True)),
              ("poseidon256", lambda s: pos(256, s)), ("poseidon512", lambda s: pos(512, s))]
    out = []
    for k, (name, make) in enumerate(shapes):
        circ = make(seed + 1000 * k)
        wits = [circ.witness]
        for j in range(1, per_shape):
            other = make(seed + 1000 * k + j)
            assert np.array_equal(other.a[1], circ.a[1]) and other.rows == circ.rows
            wits.append(other.witness)
        out.append((name, circ, wits))
    return out


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


# ------------------------------------------------- reference-format files
HERE = os.path.dirname(os.path.abspath(__file__))
GEN_BIN = os.path.join(HERE, "bin", "nkr-g16-gen")
_MAGIC = b"LKZ1"


class _Reader:
    """BinaryArchive reader (common/include/binary_archive.hpp:17-97): magic
    "LKZ1", u32 version 1, then values and u64-byte-length-prefixed vectors."""

    def __init__(self, path: str):
        self.data = np.fromfile(path, dtype=np.uint8)
        if self.data.size < 8 or self.data[:4].tobytes() != _MAGIC or \
                struct.unpack("<I", self.data[4:8].tobytes())[0] != 1:
            raise ValueError(f"{path}: not a version-1 LKZ1 archive")
        self.pos = 8

    def u64(self) -> int:
        v = struct.unpack("<Q", self.data[self.pos:self.pos + 8].tobytes())[0]
        self.pos += 8
        return v

    def vector(self, dtype, width: int = 1) -> np.ndarray:
        n = self.u64()
        if self.pos + n > self.data.size:
            raise ValueError("truncated archive")
        a = self.data[self.pos:self.pos + n].view(dtype)
        self.pos += n
        return a.reshape(-1, width) if width > 1 else a


def read_r1cs(path: str, witness_path: str | None = None) -> R1cs:
    """A circuit file (io.cpp write_circuit layout: public input count,
    columns, then row_ptr, col_idx and values of A, B and C) and, when given,
    its witness file."""
    r = _Reader(path)
    public_inputs = r.u64()
    cols = r.u64()
    mats = []
    for _ in range(3):
        rp = r.vector(np.uint32)
        ci = r.vector(np.uint32)
        va = r.vector(np.uint64, 4)
        mats.append((rp, ci, va))
    rows = mats[0][0].size - 1
    witness = None
    if witness_path:
        w = _Reader(witness_path)
        witness = w.vector(np.uint64, 4)
    return R1cs(rows, cols, public_inputs, mats[0], mats[1], mats[2], witness)


def generate(case: str, param: int, seed: int, density: str = "full", wide: bool = True,
             folded: bool = False) -> R1cs:
    """Runs the C++ generator (nkr-g16-gen) and reads its files back.
    case: matmul (param = size; wide=False for non-zero 32-bit entries),
    fibonacci (param = steps) or poseidon (param = hash count)."""
    if not os.path.exists(GEN_BIN):
        raise RuntimeError(f"{GEN_BIN} is not built (make -C nkr-zkp_amd)")
    args = [GEN_BIN, "--seed", str(seed), "--density", density]
    if case == "matmul" and not wide:
        args.append("--u32")
    if case == "poseidon" and folded:
        args.append("--folded")
    with tempfile.TemporaryDirectory() as d:
        res = subprocess.run(args + ["--path", d, case, str(param)], capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"nkr-g16-gen failed: {res.stderr.strip()}")
        stem = [f for f in os.listdir(d) if f.endswith(".bin") and not f.endswith(".witness.bin")][0][:-4]
        return read_r1cs(os.path.join(d, stem + ".bin"), os.path.join(d, stem + ".witness.bin"))

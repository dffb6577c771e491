#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from an independent pure-Python
big-integer model of BN254 (affine group law, naive DFTs, closed-form MSMs).

It shares no code with the C oracle or the HIP backend: field elements are
Python ints, points are affine tuples, transforms are O(N^2) sums and each MSM
is evaluated as (sum k_i s_i mod r) * G for bases P_i = k_i * G.  The results
pin the oracle (and through it the HIP backend) on:

  * mt19937_64 (C++ [rand.predef]: 10000th output of the default seed),
  * Fr / Fq / Fq2 products, sums, differences, inverses,
  * G1 / G2 generators, scalar multiples, the group order,
  * CSR SpMV (protocols/groth16/backends/cpu/src/backend.cpp:25-35),
  * FFT / iFFT / cosetFFT / icosetFFT (libfqfft basic_radix2_domain.tcc:41-74),
  * the quotient chain of CpuBackend::prove (backend.cpp:171-179),
  * G1 / G2 MSMs including the edge cases (zeros, duplicates, P and -P, infinity),
  * a tiny end-to-end Groth16 prove + blind (matmul 2x2, domain 16) with its
    trapdoor setup, per protocols/groth16/backends/cpu/src/backend.cpp:41-201
    and protocols/groth16/common/src/protocol.cpp:85-94.

Run:  python3 tests/golden/gen_golden.py      (writes tests/golden/*.json)
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

Q = 21888242871839275222246405745257275088696311157297823662689037894645226208583
R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
MONT = 1 << 256
ROOT_OF_UNITY = 19103219067921713944291392827692070036145651957329286315305642004821462161904  # 2^28-th
TWO_ADICITY = 28
COSET = 5
G1_GEN = (1, 2)
G2_GEN = ((10857046999023057135944570762232829481370756359578518086990519993285655852781,
           11559732032986387107991004021392285783925812861821192530917403151452391805634),
          (8495653923123431417604973247489272438418190587263600148770280649306958101930,
           4082367875863433681332203403145435568316851327593401208105741076214120093531))


# ------------------------------------------------------------- mt19937_64
class MT64:
    def __init__(self, seed: int):
        self.mt = [0] * 312
        self.mt[0] = seed & 0xFFFFFFFFFFFFFFFF
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & 0xFFFFFFFFFFFFFFFF
        self.idx = 312

    def next(self) -> int:
        if self.idx >= 312:
            mt = self.mt
            for i in range(312):
                y = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
                v = mt[(i + 156) % 312] ^ (y >> 1)
                if y & 1:
                    v ^= 0xB5026F5AA96619E9
                mt[i] = v
            self.idx = 0
        x = self.mt[self.idx]
        self.idx += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & 0xFFFFFFFFFFFFFFFF


def fr_random_mont(rng: MT64) -> int:
    """libff rule: 4 words, clear bits 254-255, reject >= r; the value IS the Montgomery residue."""
    while True:
        w = [rng.next() for _ in range(4)]
        w[3] &= 0x3FFFFFFFFFFFFFFF
        x = sum(v << (64 * i) for i, v in enumerate(w))
        if x < R:
            return x


def mont_to_int(x: int, p: int) -> int:
    return x * pow(MONT, -1, p) % p


def int_to_mont(x: int, p: int) -> int:
    return x * MONT % p


# ------------------------------------------------------------------ Fq2
def f2_add(a, b): return ((a[0] + b[0]) % Q, (a[1] + b[1]) % Q)
def f2_sub(a, b): return ((a[0] - b[0]) % Q, (a[1] - b[1]) % Q)
def f2_mul(a, b): return ((a[0] * b[0] - a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q)
def f2_neg(a): return ((-a[0]) % Q, (-a[1]) % Q)


def f2_inv(a):
    d = pow((a[0] * a[0] + a[1] * a[1]) % Q, -1, Q)
    return (a[0] * d % Q, (-a[1]) * d % Q)


# ----------------------------------------------------- affine group law
class G1:
    b = 3
    zero_f = 0
    one_f = 1

    @staticmethod
    def add(P, Qp):
        if P is None:
            return Qp
        if Qp is None:
            return P
        (x1, y1), (x2, y2) = P, Qp
        if x1 == x2:
            if (y1 + y2) % Q == 0:
                return None
            lam = 3 * x1 * x1 * pow(2 * y1, -1, Q) % Q
        else:
            lam = (y2 - y1) * pow(x2 - x1, -1, Q) % Q
        x3 = (lam * lam - x1 - x2) % Q
        return (x3, (lam * (x1 - x3) - y1) % Q)

    @staticmethod
    def neg(P):
        return None if P is None else (P[0], (-P[1]) % Q)

    @staticmethod
    def on_curve(P):
        return P is None or (P[1] * P[1] - P[0] ** 3 - 3) % Q == 0


class G2:
    b = f2_mul((3, 0), f2_inv((9, 1)))

    @staticmethod
    def add(P, Qp):
        if P is None:
            return Qp
        if Qp is None:
            return P
        (x1, y1), (x2, y2) = P, Qp
        if x1 == x2:
            if f2_add(y1, y2) == (0, 0):
                return None
            lam = f2_mul(f2_mul((3, 0), f2_mul(x1, x1)), f2_inv(f2_add(y1, y1)))
        else:
            lam = f2_mul(f2_sub(y2, y1), f2_inv(f2_sub(x2, x1)))
        x3 = f2_sub(f2_sub(f2_mul(lam, lam), x1), x2)
        return (x3, f2_sub(f2_mul(lam, f2_sub(x1, x3)), y1))

    @staticmethod
    def neg(P):
        return None if P is None else (P[0], f2_neg(P[1]))

    @staticmethod
    def on_curve(P):
        if P is None:
            return True
        x, y = P
        return f2_sub(f2_mul(y, y), f2_add(f2_mul(f2_mul(x, x), x), G2.b)) == (0, 0)


def smul(G, P, k: int):
    k %= R
    acc = None
    while k:
        if k & 1:
            acc = G.add(acc, P)
        P = G.add(P, P)
        k >>= 1
    return acc


# ---------------------------------------------------------- encodings
def hx(x: int) -> str:
    return "%064x" % x


def fe_mont_hex(x_int: int, p: int) -> str:
    """canonical integer -> hex of its Montgomery residue (the in-memory value)."""
    return hx(int_to_mont(x_int % p, p))


def g1_json(P):
    """affine G1 -> libff to_affine layout (Montgomery X, Y, Z=1; infinity (0, 1, 0))."""
    if P is None:
        return [hx(0), fe_mont_hex(1, Q), hx(0)]
    return [fe_mont_hex(P[0], Q), fe_mont_hex(P[1], Q), fe_mont_hex(1, Q)]


def g2_json(P):
    if P is None:
        return [hx(0), hx(0), fe_mont_hex(1, Q), hx(0), hx(0), hx(0)]
    (x0, x1), (y0, y1) = P
    return [fe_mont_hex(x0, Q), fe_mont_hex(x1, Q), fe_mont_hex(y0, Q), fe_mont_hex(y1, Q),
            fe_mont_hex(1, Q), hx(0)]


# ------------------------------------------------------- transforms
def omega_for(n: int) -> int:
    logn = n.bit_length() - 1
    w = ROOT_OF_UNITY
    for _ in range(TWO_ADICITY - logn):
        w = w * w % R
    return w


def dft(a, w):
    n = len(a)
    return [sum(a[j] * pow(w, i * j, R) for j in range(n)) % R for i in range(n)]


def t_fft(a):
    return dft(a, omega_for(len(a)))


def t_ifft(a):
    n = len(a)
    ninv = pow(n, -1, R)
    return [v * ninv % R for v in dft(a, pow(omega_for(n), -1, R))]


def t_coset_fft(a):
    return t_fft([v * pow(COSET, i, R) % R for i, v in enumerate(a)])


def t_icoset_fft(a):
    gi = pow(COSET, -1, R)
    return [v * pow(gi, i, R) % R for i, v in enumerate(t_ifft(a))]


def quotient_math(az, bz, cz):
    """(A*B - C) mod (x^N - g^N), divided by g^N - 1: what the coset chain computes."""
    n = len(az)
    A, B, C = t_ifft(az), t_ifft(bz), t_ifft(cz)
    P = [0] * (2 * n)
    for i, x in enumerate(A):
        if x:
            for j, y in enumerate(B):
                P[i + j] = (P[i + j] + x * y) % R
    for i, c in enumerate(C):
        P[i] = (P[i] - c) % R
    gn = pow(COSET, n, R)
    H = [(P[i] + P[i + n] * gn) % R for i in range(n)]
    zi = pow((gn - 1) % R, -1, R)
    return [h * zi % R for h in H]


# ------------------------------------------------------ circuits
def seeded_mask(count: int, nonzero: int, seed: int):
    """Density mask: `nonzero` ones then zeros, Fisher-Yates shuffled by mt19937_64(seed)."""
    mask = [1] * nonzero + [0] * (count - nonzero)
    rng = MT64(seed)
    for i in range(count - 1, 0, -1):
        j = rng.next() % (i + 1)
        mask[i], mask[j] = mask[j], mask[i]
    return mask


def matmul_circuit(s: int, seed: int, density: str = "full", wide: bool = False):
    """Matmul R1CS with the reference's constraint structure (protocols/groth16/app/gen.cpp:42-84,
    common/include/r1cs/r1cs.hpp:64-99,128-137): P[i][j][l] = A[i][l]*B[l][j] (one nnz per A/B/C row),
    then C[i][j] = sum_l P[i][j][l] (A row: explicit 0 at column 0 plus s ones; B row: 1 at column 0).
    Columns: 1, A (s^2), B (s^2), C (s^2) public, then P (s^3) private; rows and columns padded to 2^k."""
    ss = s * s
    nz = {"full": ss, "hybrid": ss - ss // 2, "sparse": ss // 10 + (ss % 10 != 0)}[density]
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
                v = mont_to_int(fr_random_mont(rng), R) if wide else rng.next() & 0xFFFFFFFF
                if v:
                    break
            out.append(v)
        vals.append(out)
    av, bv = vals
    colA = lambda k: 1 + k
    colB = lambda k: 1 + ss + k
    colC = lambda k: 1 + 2 * ss + k
    colP = lambda k: 1 + 3 * ss + k
    rowsA, rowsB, rowsC = [], [], []
    for i in range(s):
        for j in range(s):
            for l in range(s):
                rowsA.append([(colA(i * s + l), 1)])
                rowsB.append([(colB(l * s + j), 1)])
                rowsC.append([(colP(i * ss + j * s + l), 1)])
    for i in range(s):
        for j in range(s):
            rowsA.append([(0, 0)] + [(colP(i * ss + j * s + l), 1) for l in range(s)])
            rowsB.append([(0, 1)])
            rowsC.append([(colC(i * s + j), 1)])
    live_rows, live_cols = len(rowsA), 1 + 3 * ss + s * ss
    rows = 1 << (live_rows - 1).bit_length()
    cols = 1 << (live_cols - 1).bit_length()
    z = [0] * cols
    z[0] = 1
    for i in range(s):
        for j in range(s):
            acc = 0
            for l in range(s):
                p = av[i * s + l] * bv[l * s + j] % R
                z[colP(i * ss + j * s + l)] = p
                acc += p
            z[colC(i * s + j)] = acc % R
    for k in range(ss):
        z[colA(k)] = av[k]
        z[colB(k)] = bv[k]

    def csr(rowlist):
        rp, ci, va = [0], [], []
        for r in rowlist:
            for c, v in r:
                ci.append(c)
                va.append(v)
            rp.append(len(ci))
        rp += [len(ci)] * (rows - len(rowlist))
        return rp, ci, va

    return dict(rows=rows, cols=cols, public_inputs=3 * ss, a=csr(rowsA), b=csr(rowsB), c=csr(rowsC), z=z)


def spmv_int(m, rows, z):
    rp, ci, va = m
    return [sum(va[j] * z[ci[j]] for j in range(rp[r], rp[r + 1])) % R for r in range(rows)]


def lagrange_at(n: int, t: int):
    w = omega_for(n)
    zt = (pow(t, n, R) - 1) % R
    ninv = pow(n, -1, R)
    out = []
    for i in range(n):
        wi = pow(w, i, R)
        out.append(zt * ninv % R * wi % R * pow((t - wi) % R, -1, R) % R)
    return out


def setup_scalars(circ, tr):
    """Query scalars of the CPU setup (backend.cpp:41-154) for trapdoors tr = (t, alpha, beta, gamma, delta)."""
    t, alpha, beta, gamma, delta = tr
    n, m = circ["rows"], circ["cols"]
    lag = lagrange_at(n, t)
    acc = []
    for key in ("a", "b", "c"):
        rp, ci, va = circ[key]
        v = [0] * m
        for r in range(n):
            for j in range(rp[r], rp[r + 1]):
                v[ci[j]] = (v[ci[j]] + va[j] * lag[r]) % R
        acc.append(v)
    at, bt, ct = acc
    di, gi = pow(delta, -1, R), pow(gamma, -1, R)
    pub = circ["public_inputs"]
    l = [((beta * at[i] + alpha * bt[i] + ct[i]) * di % R) if i > pub else 0 for i in range(m)]
    ic = [(beta * at[i] + alpha * bt[i] + ct[i]) * gi % R for i in range(pub + 1)]
    zt = (pow(t, n, R) - 1) % R
    h = [(pow(t, i, R) * zt % R * di % R) if i + 1 < n else 0 for i in range(n)]
    return dict(at=at, bt=bt, ct=ct, l=l, ic=ic, h=h)


# -------------------------------------------------------------- main
def main() -> None:
    out = {}
    rng = MT64(0x90D)

    # mt19937_64 KAT: C++ [rand.predef] fixes the 10000th output of the default seed (5489).
    m = MT64(5489)
    first = [m.next() for _ in range(10)]
    for _ in range(9989):
        m.next()
    out["mt19937_64"] = {"seed": 5489, "first": [str(v) for v in first], "n10000": str(m.next())}
    assert out["mt19937_64"]["n10000"] == "9981545732273789042"
    out["fr_random"] = {"seed": 0x5EED, "mont": [hx(v) for v in (lambda g: [fr_random_mont(g) for _ in range(8)])(MT64(0x5EED))]}

    # field products: inputs are canonical integers; the fixture holds Montgomery residues.
    def fe_cases(p, count):
        xs = [rng.next() % p for _ in range(count - 4)] + [0, 1, p - 1, p - 2]
        ys = [rng.next() % p for _ in range(count - 4)] + [p - 1, p - 1, p - 1, 2]
        # widen to full 254-bit values
        xs = [(x * (rng.next() | 1)) % p for x in xs[:-4]] + xs[-4:]
        ys = [(y * (rng.next() | 1)) % p for y in ys[:-4]] + ys[-4:]
        return {
            "a": [fe_mont_hex(x, p) for x in xs], "b": [fe_mont_hex(y, p) for y in ys],
            "mul": [fe_mont_hex(x * y, p) for x, y in zip(xs, ys)],
            "add": [fe_mont_hex(x + y, p) for x, y in zip(xs, ys)],
            "sub": [fe_mont_hex(x - y, p) for x, y in zip(xs, ys)],
            "inv": [fe_mont_hex(pow(x, -1, p) if x else 0, p) for x in xs],
        }
    out["fr"] = fe_cases(R, 64)
    out["fq"] = fe_cases(Q, 64)
    f2a = [((rng.next() * rng.next()) % Q, (rng.next() * rng.next()) % Q) for _ in range(16)]
    f2b = [((rng.next() * rng.next()) % Q, (rng.next() * rng.next()) % Q) for _ in range(16)]
    out["fq2"] = {
        "a": [[fe_mont_hex(x[0], Q), fe_mont_hex(x[1], Q)] for x in f2a],
        "b": [[fe_mont_hex(x[0], Q), fe_mont_hex(x[1], Q)] for x in f2b],
        "mul": [[fe_mont_hex(c[0], Q), fe_mont_hex(c[1], Q)] for c in (f2_mul(x, y) for x, y in zip(f2a, f2b))],
        "sqr": [[fe_mont_hex(c[0], Q), fe_mont_hex(c[1], Q)] for c in (f2_mul(x, x) for x in f2a)],
    }

    # groups
    assert G1.on_curve(G1_GEN) and G2.on_curve(G2_GEN)
    assert smul(G1, G1_GEN, R) is None and smul(G2, G2_GEN, R) is None
    ks = [1, 2, 3, 5, R - 1, (rng.next() << 190) % R, (rng.next() << 128 | rng.next()) % R]
    out["g1"] = {"gen": g1_json(G1_GEN), "k": [str(k) for k in ks],
                 "kG": [g1_json(smul(G1, G1_GEN, k)) for k in ks]}
    out["g2"] = {"gen": g2_json(G2_GEN), "k": [str(k) for k in ks],
                 "kG": [g2_json(smul(G2, G2_GEN, k)) for k in ks]}

    # SpMV: ragged rows (including empty rows and repeated columns)
    rows, cols = 40, 24
    rp, ci, va = [0], [], []
    for r in range(rows):
        cnt = rng.next() % 6 if r % 7 else 0
        for _ in range(cnt):
            ci.append(rng.next() % cols)
            va.append(rng.next() % R * (rng.next() | 1) % R)
        rp.append(len(ci))
    zv = [rng.next() * rng.next() % R for _ in range(cols)]
    out["spmv"] = {"rows": rows, "cols": cols, "row_ptr": rp, "col_idx": ci,
                   "values": [fe_mont_hex(v, R) for v in va], "z": [fe_mont_hex(v, R) for v in zv],
                   "y": [fe_mont_hex(v, R) for v in spmv_int((rp, ci, va), rows, zv)]}

    # transforms, N = 2 .. 1024
    ffts = []
    for logn in range(1, 11):
        n = 1 << logn
        a = [rng.next() * rng.next() % R for _ in range(n)]
        ffts.append({"n": n, "a": [fe_mont_hex(v, R) for v in a],
                     "fft": [fe_mont_hex(v, R) for v in t_fft(a)],
                     "ifft": [fe_mont_hex(v, R) for v in t_ifft(a)],
                     "coset_fft": [fe_mont_hex(v, R) for v in t_coset_fft(a)],
                     "icoset_fft": [fe_mont_hex(v, R) for v in t_icoset_fft(a)]})
    out["fft"] = ffts

    quots = []
    for logn in (2, 4, 6, 8):
        n = 1 << logn
        az, bz, cz = ([rng.next() * rng.next() % R for _ in range(n)] for _ in range(3))
        quots.append({"n": n, "az": [fe_mont_hex(v, R) for v in az], "bz": [fe_mont_hex(v, R) for v in bz],
                      "cz": [fe_mont_hex(v, R) for v in cz],
                      "h": [fe_mont_hex(v, R) for v in quotient_math(az, bz, cz)]})
    out["quotient"] = quots

    # MSMs: bases k_i G, result (sum k_i s_i) G; edge cases included.
    msms = []
    for n, flavour in ((1, "uniform"), (2, "dup"), (7, "uniform"), (33, "mix"), (64, "edge"), (100, "uniform")):
        k = [rng.next() * rng.next() % R for _ in range(n)]
        s = [rng.next() * rng.next() * rng.next() % R for _ in range(n)]
        if flavour == "dup":
            k[1] = k[0]
            s[1] = s[0]
        if flavour == "mix":
            s = [(rng.next() & 1) if rng.next() & 1 else v for v in s]
        if flavour == "edge":
            k[1] = k[0]                 # duplicate base
            k[2] = (R - k[0]) % R       # its negation
            k[3] = 0                    # infinity base
            s[4] = 0
            s[5] = 1
            s[6] = R - 1
            k[7] = k[8] = k[9] = k[10]  # four equal bases with equal scalars
            s[7] = s[8] = s[9] = s[10]
        tot = sum(a * b for a, b in zip(k, s)) % R
        for gname, G, gen, enc in (("g1", G1, G1_GEN, g1_json), ("g2", G2, G2_GEN, g2_json)):
            if gname == "g2" and n > 40:
                continue
            msms.append({"group": gname, "n": n, "flavour": flavour, "k": [str(v) for v in k],
                         "scalars": [fe_mont_hex(v, R) for v in s],
                         "bases": [enc(smul(G, gen, v)) for v in k],
                         "result": enc(smul(G, gen, tot))})
    out["msm"] = msms

    json.dump(out, open(os.path.join(HERE, "bn254_golden.json"), "w"), indent=0)

    # ---- tiny end-to-end Groth16 (matmul 2x2, domain 16) -------------------
    circ = matmul_circuit(2, seed=0x4E4B52)
    n, m = circ["rows"], circ["cols"]
    z = circ["z"]
    for key in ("a", "b", "c"):
        pass
    az, bz, cz = (spmv_int(circ[k], n, z) for k in ("a", "b", "c"))
    assert all((x * y - w) % R == 0 for x, y, w in zip(az, bz, cz)), "witness does not satisfy"
    h = quotient_math(az, bz, cz)
    trg = MT64(0x4E4B52 ^ 0x5E7)
    tr = [mont_to_int(fr_random_mont(trg), R) for _ in range(5)]
    sc = setup_scalars(circ, tr)
    t, alpha, beta, gamma, delta = tr
    pk = {
        "a_query": [g1_json(smul(G1, G1_GEN, v)) for v in sc["at"]],
        "b_g1_query": [g1_json(smul(G1, G1_GEN, v)) for v in sc["bt"]],
        "b_g2_query": [g2_json(smul(G2, G2_GEN, v)) for v in sc["bt"]],
        "l_query": [g1_json(smul(G1, G1_GEN, v)) for v in sc["l"]],
        "h_query": [g1_json(smul(G1, G1_GEN, v)) for v in sc["h"]],
        "alpha_g1": g1_json(smul(G1, G1_GEN, alpha)), "beta_g1": g1_json(smul(G1, G1_GEN, beta)),
        "beta_g2": g2_json(smul(G2, G2_GEN, beta)), "delta_g1": g1_json(smul(G1, G1_GEN, delta)),
        "delta_g2": g2_json(smul(G2, G2_GEN, delta)),
    }
    dot = lambda u, v: sum(a * b for a, b in zip(u, v)) % R
    ar, bs, zk, qz = dot(z, sc["at"]), dot(z, sc["bt"]), dot(z, sc["l"]), dot(h, sc["h"])
    rsr = MT64(0x4E4B52 ^ 0x72)
    r_m, s_m = fr_random_mont(rsr), fr_random_mont(rsr)
    r_, s_ = mont_to_int(r_m, R), mont_to_int(s_m, R)
    A = (alpha + ar + r_ * delta) % R
    B = (beta + bs + s_ * delta) % R
    B1 = B
    C = (qz + zk + s_ * A + r_ * B1 - r_ * s_ * delta) % R
    pub = circ["public_inputs"]
    # exponent form of the verifier equation e(A,B) = e(alpha,beta) e(IC,gamma) e(C,delta)
    ic_acc = (sc["ic"][0] + sum(z[i + 1] * sc["ic"][i + 1] for i in range(pub))) % R
    assert A * B % R == (alpha * beta + ic_acc * gamma + C * delta) % R
    e2e = {
        "seed": 0x4E4B52, "rows": n, "cols": m, "public_inputs": pub,
        "a": [circ["a"][0], circ["a"][1], [fe_mont_hex(v, R) for v in circ["a"][2]]],
        "b": [circ["b"][0], circ["b"][1], [fe_mont_hex(v, R) for v in circ["b"][2]]],
        "c": [circ["c"][0], circ["c"][1], [fe_mont_hex(v, R) for v in circ["c"][2]]],
        "witness": [fe_mont_hex(v, R) for v in z],
        "trapdoors": [str(v) for v in tr],
        "pk": pk,
        "az": [fe_mont_hex(v, R) for v in az], "bz": [fe_mont_hex(v, R) for v in bz],
        "cz": [fe_mont_hex(v, R) for v in cz], "h": [fe_mont_hex(v, R) for v in h],
        "raw": {"ar": g1_json(smul(G1, G1_GEN, ar)), "bs1": g1_json(smul(G1, G1_GEN, bs)),
                "zk": g1_json(smul(G1, G1_GEN, zk)), "qz": g1_json(smul(G1, G1_GEN, qz)),
                "bs2": g2_json(smul(G2, G2_GEN, bs))},
        "r": hx(r_m), "s": hx(s_m),
        "proof": {"a": g1_json(smul(G1, G1_GEN, A)), "b": g2_json(smul(G2, G2_GEN, B)),
                  "c": g1_json(smul(G1, G1_GEN, C))},
        "vk_ic": [g1_json(smul(G1, G1_GEN, v)) for v in sc["ic"]],
    }
    json.dump(e2e, open(os.path.join(HERE, "groth16_matmul2.json"), "w"), indent=0)
    print("wrote", os.listdir(HERE), file=sys.stderr)


if __name__ == "__main__":
    main()

"""Golden vectors for the BN254 reduced pairing and the Groth16 verifier.

An independent model: Fq12 is the flat ring Fq[w]/(w^12 - 18 w^6 + 82)
(w^6 = 9 + u, u^2 = -1), not the library's Fq2/Fq6/Fq12 tower; points are
affine on E(Fq12): y^2 = x^3 + 3; the Frobenius map is computed as x -> x^q
by exponentiation; the final exponent is the integer m (q^12 - 1)/r.  What is
shared with the library is only the definition being modelled:

  * G2 (on y^2 = x^3 + 3/(9+u) over Fq2) enters E(Fq12) as (x w^2, y w^3), the
    embedding under which the reference's ate lines (libff
    alt_bn128_pairing.cpp:241-330, mul_by_024 with ell_0 at 1, ell_VW * yP at
    w^3, ell_VV * xP at w^4) are tangents and chords;
  * the Miller function of 6x + 2 followed by the chords through pi(Q) and
    -pi^2(Q) (alt_bn128_pairing.cpp:368-420);
  * libff's final exponentiation (alt_bn128_pairing.cpp:110-237) raises to
    (q^6 - 1)(q^2 + 1) times q^3 l3 + q^2 l2 + q l1 + l0, which equals
    m (q^12 - 1)/r with m = 2x(6x^2 + 3x + 1) as integers (asserted below).

Output: tests/golden/bn254_pairing.json with GT values in the in-memory layout
(c0.c0.c0, c0.c0.c1, c0.c1.c0, ..., c1.c2.c1; Montgomery hex), pairing inputs
as points in the libff affine layout, and the verification key of the
groth16_matmul2 fixture (alpha_g1_beta_g2 from its trapdoors).
Run: python3 tests/golden/gen_pairing.py  (about a minute).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import gen_golden as GG  # noqa: E402

Q, R = GG.Q, GG.R
X = 4965661367192848881  # BN parameter of alt_bn128
ATE = 6 * X + 2


# ------------------------------------------------- flat Fq12 = Fq[w]/(w^12 - 18 w^6 + 82)
def k_mul(a, b):
    t = [0] * 23
    for i, ai in enumerate(a):
        if ai:
            for j, bj in enumerate(b):
                t[i + j] += ai * bj
    for k in range(22, 11, -1):  # w^k = 18 w^(k-6) - 82 w^(k-12)
        c = t[k]
        if c:
            t[k - 6] += 18 * c
            t[k - 12] -= 82 * c
    return [v % Q for v in t[:12]]


def k_add(a, b): return [(x + y) % Q for x, y in zip(a, b)]
def k_sub(a, b): return [(x - y) % Q for x, y in zip(a, b)]
def k_const(c): return [c % Q] + [0] * 11


K_ONE = k_const(1)
MODPOLY = [82, 0, 0, 0, 0, 0, Q - 18, 0, 0, 0, 0, 0, 1]  # w^12 - 18 w^6 + 82, low first


def poly_trim(p):
    while p and p[-1] == 0:
        p = p[:-1]
    return p


def poly_divmod(a, b):
    a, b = poly_trim(list(a)), poly_trim(list(b))
    inv_lead = pow(b[-1], -1, Q)
    quo = [0] * max(len(a) - len(b) + 1, 1)
    while len(a) >= len(b) and a:
        c = a[-1] * inv_lead % Q
        d = len(a) - len(b)
        quo[d] = c
        for i, bi in enumerate(b):
            a[d + i] = (a[d + i] - c * bi) % Q
        a = poly_trim(a)
    return quo, a


def poly_mul(a, b):
    if not a or not b:
        return []
    t = [0] * (len(a) + len(b) - 1)
    for i, ai in enumerate(a):
        for j, bj in enumerate(b):
            t[i + j] = (t[i + j] + ai * bj) % Q
    return t


def poly_sub(a, b):
    n = max(len(a), len(b))
    return poly_trim([((a[i] if i < len(a) else 0) - (b[i] if i < len(b) else 0)) % Q for i in range(n)])


def k_inv(a):
    """Extended Euclid in Fq[w] against the modulus polynomial."""
    r0, r1 = MODPOLY[:], poly_trim(list(a))
    s0, s1 = [], [1]
    while len(r1) > 1:
        qt, rem = poly_divmod(r0, r1)
        r0, r1 = r1, rem
        s0, s1 = s1, poly_sub(s0, poly_mul(qt, s1))
    assert len(r1) == 1 and r1[0] != 0, "not invertible"
    c = pow(r1[0], -1, Q)
    out = [v * c % Q for v in s1] + [0] * 12
    return out[:12]


def k_pow(a, e):
    acc, base = K_ONE, a
    while e:
        if e & 1:
            acc = k_mul(acc, base)
        base = k_mul(base, base)
        e >>= 1
    return acc


def from_fq2(c):
    """a0 + a1 u -> flat, u = w^6 - 9."""
    out = [0] * 12
    out[0] = (c[0] - 9 * c[1]) % Q
    out[6] = c[1] % Q
    return out


def w_pow(k):
    out = [0] * 12
    out[k] = 1
    return out


def to_tower(a):
    """flat -> [(c0.c0), (c0.c1), (c0.c2), (c1.c0), (c1.c1), (c1.c2)] Fq2 pairs:
    the coefficient of w^i is (b_i + 9 b_(i+6)) + b_(i+6) u; c0 holds even i, c1 odd i."""
    coef = [((a[i] + 9 * a[i + 6]) % Q, a[i + 6] % Q) for i in range(6)]
    return [coef[0], coef[2], coef[4], coef[1], coef[3], coef[5]]


def from_tower(t):
    order = [0, 2, 4, 1, 3, 5]  # tower slot -> power of w
    out = [0] * 12
    for slot, (c0, c1) in enumerate(t):
        f = k_mul(from_fq2((c0, c1)), w_pow(order[slot]))
        out = k_add(out, f)
    return out


def gt_hex(a):
    return [GG.fe_mont_hex(c, Q) for pair in to_tower(a) for c in pair]


def gt_from_hex(h):
    vals = [GG.mont_to_int(int(x, 16), Q) for x in h]
    return from_tower([(vals[2 * i], vals[2 * i + 1]) for i in range(6)])


# ------------------------------------------------------ E(Fq12), affine
def e_add(P, S):
    if P is None:
        return S
    if S is None:
        return P
    (x1, y1), (x2, y2) = P, S
    if x1 == x2:
        if k_add(y1, y2) == [0] * 12:
            return None
        lam = k_mul(k_mul(k_const(3), k_mul(x1, x1)), k_inv(k_add(y1, y1)))
    else:
        lam = k_mul(k_sub(y2, y1), k_inv(k_sub(x2, x1)))
    x3 = k_sub(k_sub(k_mul(lam, lam), x1), x2)
    return (x3, k_sub(k_mul(lam, k_sub(x1, x3)), y1))


def line(T, S, P):
    """Value at P of the line through T and S (tangent when T == S)."""
    (x1, y1), (x2, y2) = T, S
    if x1 == x2 and y1 == y2:
        lam = k_mul(k_mul(k_const(3), k_mul(x1, x1)), k_inv(k_add(y1, y1)))
    else:
        lam = k_mul(k_sub(y2, y1), k_inv(k_sub(x2, x1)))
    xp, yp = P
    return k_sub(k_sub(yp, y1), k_mul(lam, k_sub(xp, x1)))


def untwist(Q2):
    (x, y) = Q2
    return (k_mul(from_fq2(x), w_pow(2)), k_mul(from_fq2(y), w_pow(3)))


def on_e12(P):
    x, y = P
    return k_sub(k_mul(y, y), k_add(k_mul(k_mul(x, x), x), k_const(3))) == [0] * 12


def miller(P1, Q2):
    if P1 is None or Q2 is None:
        return K_ONE
    P = (k_const(P1[0]), k_const(P1[1]))
    Qe = untwist(Q2)
    assert on_e12(Qe)
    T, f = Qe, K_ONE
    for i in range(ATE.bit_length() - 2, -1, -1):
        f = k_mul(k_mul(f, f), line(T, T, P))
        T = e_add(T, T)
        if (ATE >> i) & 1:
            f = k_mul(f, line(T, Qe, P))
            T = e_add(T, Qe)
    Q1 = (k_pow(Qe[0], Q), k_pow(Qe[1], Q))
    Q2n = (k_pow(Q1[0], Q), [(-v) % Q for v in k_pow(Q1[1], Q)])
    assert on_e12(Q1) and on_e12(Q2n)
    f = k_mul(f, line(T, Q1, P))
    T = e_add(T, Q1)
    f = k_mul(f, line(T, Q2n, P))
    return f


M_EXP = 2 * X * (6 * X * X + 3 * X + 1)
FINAL_EXP = M_EXP * ((Q ** 12 - 1) // R)


def final_exp(f):
    return k_pow(f, FINAL_EXP)


def pairing(P1, Q2):
    return final_exp(miller(P1, Q2))


def check_libff_exponent():
    z = X
    lam = (Q ** 3 * (12 * z ** 3 + 6 * z ** 2 + 4 * z - 1) + Q ** 2 * (12 * z ** 3 + 6 * z ** 2 + 6 * z)
           + Q * (12 * z ** 3 + 6 * z ** 2 + 4 * z) + (12 * z ** 3 + 12 * z ** 2 + 6 * z + 1))
    phi = Q ** 4 - Q ** 2 + 1
    assert (Q ** 12 - 1) % R == 0 and phi % R == 0
    assert lam == M_EXP * phi // R
    assert (Q ** 6 - 1) * (Q ** 2 + 1) * lam == FINAL_EXP


def main():
    check_libff_exponent()
    rng = GG.MT64(0x4E4B52 ^ 0xE12)
    rand = lambda: GG.mont_to_int(GG.fr_random_mont(rng), R)
    g1, g2 = GG.G1_GEN, GG.G2_GEN
    e11 = pairing(g1, g2)
    assert e11 != K_ONE and k_pow(e11, R) == K_ONE
    cases = []
    for _ in range(2):
        a, b = rand(), rand()
        P, S = GG.smul(GG.G1, g1, a), GG.smul(GG.G2, g2, b)
        e = pairing(P, S)
        assert e == k_pow(e11, a * b % R), "bilinearity"
        cases.append({"a": str(a), "b": str(b), "g1": GG.g1_json(P), "g2": GG.g2_json(S), "gt": gt_hex(e)})
    e2e = json.load(open(os.path.join(HERE, "groth16_matmul2.json")))
    t, alpha, beta, gamma, delta = (int(v) for v in e2e["trapdoors"])
    ab = k_pow(e11, alpha * beta % R)
    vk = {"alpha_g1_beta_g2": gt_hex(ab), "gamma_g2": GG.g2_json(GG.smul(GG.G2, g2, gamma)),
          "delta_g2": GG.g2_json(GG.smul(GG.G2, g2, delta))}
    out = {"e_g1_g2": gt_hex(e11), "miller_g1_g2": gt_hex(miller(g1, g2)), "cases": cases,
           "matmul2_vk": vk}
    json.dump(out, open(os.path.join(HERE, "bn254_pairing.json"), "w"), indent=0)
    print("wrote bn254_pairing.json", file=sys.stderr)


if __name__ == "__main__":
    main()

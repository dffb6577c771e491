// BN254 optimal ate pairing and the Groth16 verifier, on the host.
//
// Verifier: protocols/groth16/common/src/protocol.cpp:96-108.  Pairing: the
// optimal ate pairing of BN curves with x = 4965661367192848881: the Miller
// function of 6x + 2 at Q, then the chords through pi(Q) and -pi^2(Q)
// (alt_bn128_ate_miller_loop, alt_bn128_pairing.cpp:368-420, computes the
// same function in projective coordinates), then the final exponentiation,
// whose value is fixed by alt_bn128_pairing.cpp:152-237.  G2 lives on the D-type twist
// y^2 = x^3 + 3/(9+u) and enters E(Fq12) as (x, y) -> (x w^2, y w^3), the
// embedding implied by the reference's line layout (mul_by_024 with ell_0 at
// 1, ell_VW * yP at w^3, ell_VV * xP at w^4).  Lines are evaluated affinely;
// factors in proper subfields vanish in the final exponentiation.
#include "pairing.h"

#include <cstring>

namespace nkr {

namespace {

using F2 = fq2_t;
using F6 = fq6_t;
using F12 = fq12_t;

// --- small unsigned integers as little-endian u64 words (constant setup) ---

using Words = std::vector<uint64_t>;

Words words_of(const uint64_t *w, int n) { return Words(w, w + n); }

Words words_mul(const Words &a, const Words &b)
{
    Words r(a.size() + b.size(), 0);
    for (size_t i = 0; i < a.size(); ++i) {
        unsigned __int128 carry = 0;
        for (size_t j = 0; j < b.size(); ++j) {
            const unsigned __int128 t = (unsigned __int128)a[i] * b[j] + r[i + j] + carry;
            r[i + j] = (uint64_t)t;
            carry = t >> 64;
        }
        r[i + b.size()] = (uint64_t)carry;
    }
    return r;
}

Words words_sub1(Words a)  // a - 1, a > 0
{
    for (auto &w : a)
        if (w-- != 0) break;
    return a;
}

Words words_div_small(const Words &a, uint64_t d)  // floor division
{
    Words r(a.size());
    unsigned __int128 rem = 0;
    for (size_t i = a.size(); i-- > 0;) {
        const unsigned __int128 cur = (rem << 64) | a[i];
        r[i] = (uint64_t)(cur / d);
        rem = cur % d;
    }
    return r;
}

// --- Fq2, Fq6, Fq12 ---

F2 f2_mul_xi(const F2 &a)  // (9 + u) a = (9 a0 - a1) + (9 a1 + a0) u
{
    fq_t e0 = fp_dbl(fp_dbl(fp_dbl(a.c0))), e1 = fp_dbl(fp_dbl(fp_dbl(a.c1)));
    e0 = fp_add(e0, a.c0);
    e1 = fp_add(e1, a.c1);
    return {fp_sub(e0, a.c1), fp_add(e1, a.c0)};
}

F2 f2_conj(const F2 &a) { return {a.c0, fp_neg(a.c1)}; }

F2 f2_mul_fq(const F2 &a, const fq_t &s) { return {fp_mul(a.c0, s), fp_mul(a.c1, s)}; }

F2 f2_pow(const F2 &a, const Words &e)
{
    F2 acc = F2::one();
    for (size_t w = e.size(); w-- > 0;)
        for (int b = 63; b >= 0; --b) {
            acc = fq2_sqr(acc);
            if ((e[w] >> b) & 1) acc = fq2_mul(acc, a);
        }
    return acc;
}

F6 f6_zero() { return {F2::zero(), F2::zero(), F2::zero()}; }
F6 f6_one() { return {F2::one(), F2::zero(), F2::zero()}; }
F6 f6_add(const F6 &a, const F6 &b) { return {fq2_add(a.c0, b.c0), fq2_add(a.c1, b.c1), fq2_add(a.c2, b.c2)}; }
F6 f6_sub(const F6 &a, const F6 &b) { return {fq2_sub(a.c0, b.c0), fq2_sub(a.c1, b.c1), fq2_sub(a.c2, b.c2)}; }
F6 f6_neg(const F6 &a) { return {fq2_neg(a.c0), fq2_neg(a.c1), fq2_neg(a.c2)}; }
F6 f6_mul_v(const F6 &a) { return {f2_mul_xi(a.c2), a.c0, a.c1}; }
F6 f6_mul_fq(const F6 &a, const fq_t &s) { return {f2_mul_fq(a.c0, s), f2_mul_fq(a.c1, s), f2_mul_fq(a.c2, s)}; }

// Karatsuba over the three coefficients: 6 Fq2 products.
F6 f6_mul(const F6 &a, const F6 &b)
{
    const F2 t0 = fq2_mul(a.c0, b.c0), t1 = fq2_mul(a.c1, b.c1), t2 = fq2_mul(a.c2, b.c2);
    const F2 m12 = fq2_sub(fq2_sub(fq2_mul(fq2_add(a.c1, a.c2), fq2_add(b.c1, b.c2)), t1), t2);
    const F2 m01 = fq2_sub(fq2_sub(fq2_mul(fq2_add(a.c0, a.c1), fq2_add(b.c0, b.c1)), t0), t1);
    const F2 m02 = fq2_sub(fq2_sub(fq2_mul(fq2_add(a.c0, a.c2), fq2_add(b.c0, b.c2)), t0), t2);
    return {fq2_add(t0, f2_mul_xi(m12)), fq2_add(m01, f2_mul_xi(t2)), fq2_add(m02, t1)};
}

// a * (b + c v): the w-part of a line.
F6 f6_mul_01(const F6 &a, const F2 &b, const F2 &c)
{
    return {fq2_add(fq2_mul(a.c0, b), f2_mul_xi(fq2_mul(a.c2, c))), fq2_add(fq2_mul(a.c0, c), fq2_mul(a.c1, b)),
            fq2_add(fq2_mul(a.c1, c), fq2_mul(a.c2, b))};
}

F6 f6_inv(const F6 &a)
{
    const F2 A = fq2_sub(fq2_sqr(a.c0), f2_mul_xi(fq2_mul(a.c1, a.c2)));
    const F2 B = fq2_sub(f2_mul_xi(fq2_sqr(a.c2)), fq2_mul(a.c0, a.c1));
    const F2 C = fq2_sub(fq2_sqr(a.c1), fq2_mul(a.c0, a.c2));
    const F2 n = fq2_add(fq2_mul(a.c0, A), f2_mul_xi(fq2_add(fq2_mul(a.c2, B), fq2_mul(a.c1, C))));
    const F2 ni = fq2_inv(n);
    return {fq2_mul(A, ni), fq2_mul(B, ni), fq2_mul(C, ni)};
}

F12 f12_mul(const F12 &a, const F12 &b)
{
    const F6 t0 = f6_mul(a.c0, b.c0), t1 = f6_mul(a.c1, b.c1);
    return {f6_add(t0, f6_mul_v(t1)), f6_sub(f6_sub(f6_mul(f6_add(a.c0, a.c1), f6_add(b.c0, b.c1)), t0), t1)};
}

// (a0 + a1 w)^2 = (a0 + a1)(a0 + v a1) - t - v t + 2 t w, t = a0 a1.
F12 f12_sqr(const F12 &a)
{
    const F6 t = f6_mul(a.c0, a.c1);
    const F6 s = f6_mul(f6_add(a.c0, a.c1), f6_add(a.c0, f6_mul_v(a.c1)));
    return {f6_sub(f6_sub(s, t), f6_mul_v(t)), f6_add(t, t)};
}

F12 f12_conj(const F12 &a) { return {a.c0, f6_neg(a.c1)}; }  // a^(q^6)

F12 f12_inv(const F12 &a)  // (a0 - a1 w) / (a0^2 - v a1^2)
{
    const F6 d = f6_inv(f6_sub(f6_mul(a.c0, a.c0), f6_mul_v(f6_mul(a.c1, a.c1))));
    return {f6_mul(a.c0, d), f6_neg(f6_mul(a.c1, d))};
}

// f * (yP + (b + c v) w): multiplication by a line value.
F12 f12_mul_line(const F12 &f, const fq_t &yp, const F2 &b, const F2 &c)
{
    const F6 f0l0 = f6_mul_fq(f.c0, yp), f1l0 = f6_mul_fq(f.c1, yp);
    const F6 f0l1 = f6_mul_01(f.c0, b, c), f1l1 = f6_mul_01(f.c1, b, c);
    return {f6_add(f0l0, f6_mul_v(f1l1)), f6_add(f0l1, f1l0)};
}

fq_t fq_from_u32(uint32_t v) { return fp_to_mont(fq_t{{v, 0, 0, 0, 0, 0, 0, 0}}); }

struct Constants {
    F2 frob_x, frob_y;  // xi^((q-1)/3), xi^((q-1)/2): q-power Frobenius on the twist
    F2 frob2[6];        // gamma^i, gamma = xi^((q^2-1)/6): q^2-power Frobenius on w^i
    F2 twist_b;         // b' = 3 / (9 + u)
    Constants()
    {
        const Words q = words_of(FqParams::P64, 4);
        const Words qm1 = words_sub1(q);
        const Words q2m1 = words_sub1(words_mul(q, q));
        F2 xi;
        xi.c0 = fq_from_u32(9);
        xi.c1 = fq_t::one();
        frob_x = f2_pow(xi, words_div_small(qm1, 3));
        frob_y = f2_pow(xi, words_div_small(qm1, 2));
        const F2 gamma = f2_pow(xi, words_div_small(q2m1, 6));
        frob2[0] = F2::one();
        for (int i = 1; i < 6; ++i) frob2[i] = fq2_mul(frob2[i - 1], gamma);
        twist_b = fq2_mul(F2{fq_from_u32(3), fq_t::zero()}, fq2_inv(xi));
    }
};

const Constants &constants()
{
    static const Constants c;
    return c;
}

// a^(q^2): the w^i coefficients (c0.c0, c1.c0, c0.c1, c1.c1, c0.c2, c1.c2
// hold i = 0..5) are scaled by gamma^i; Fq2 is fixed by the q^2-power map.
F12 f12_frob2(const F12 &a)
{
    const F2 *g = constants().frob2;
    return {{a.c0.c0, fq2_mul(a.c0.c1, g[2]), fq2_mul(a.c0.c2, g[4])},
            {fq2_mul(a.c1.c0, g[1]), fq2_mul(a.c1.c1, g[3]), fq2_mul(a.c1.c2, g[5])}};
}

// m (q^4 - q^2 + 1) / r, m = 2x(6x^2 + 3x + 1): the hard part of the final
// exponent, little-endian (951 bits).
const uint64_t kHardExp[15] = {
    0xcaa4152366144ab4ull, 0x114dc0ec2cab7ffdull, 0x0cf0888c7a0ff6cfull, 0x65c5644e949b6a90ull,
    0x1be2458885117085ull, 0x5b35eb719e58db4bull, 0x2566c550aeb7e0e2ull, 0x0c974024a316619full,
    0xd147cb7d3a5203dcull, 0x621d9bfed77c2ad0ull, 0x26473fbcd1c3ec1eull, 0xe86518527b5e4036ull,
    0x29259e9712ca7b71ull, 0x1891045f68d15763ull, 0x00679dd974c68787ull};

// 6x + 2 (65 bits).
constexpr uint64_t kAteLoop[2] = {0x9d797039be763ba8ull, 0x1ull};

// a^e with a fixed 4-bit window.
F12 f12_pow(const F12 &a, const uint64_t *e, int words)
{
    F12 table[16];
    table[0] = F12{f6_one(), f6_zero()};
    for (int i = 1; i < 16; ++i) table[i] = f12_mul(table[i - 1], a);
    F12 acc = table[0];
    bool started = false;
    for (int nib = words * 16 - 1; nib >= 0; --nib) {
        if (started)
            for (int d = 0; d < 4; ++d) acc = f12_sqr(acc);
        const unsigned v = (unsigned)(e[nib >> 4] >> ((nib & 15) * 4)) & 15u;
        if (v) {
            acc = started ? f12_mul(acc, table[v]) : table[v];
            started = true;
        }
    }
    return acc;
}

struct TwistPoint {
    F2 x, y;
};

// f * (value at P of the line of slope lam through T):
// yP - lam xP w + (lam xT - yT) w^3.
F12 mul_by_line(const F12 &f, const F2 &lam, const TwistPoint &T, const fq_t &xp, const fq_t &yp)
{
    const F2 b = {fp_neg(fp_mul(lam.c0, xp)), fp_neg(fp_mul(lam.c1, xp))};
    const F2 c = fq2_sub(fq2_mul(lam, T.x), T.y);
    return f12_mul_line(f, yp, b, c);
}

// f <- f * tangent(T)(P); T <- 2T.
void dbl_step(F12 &f, TwistPoint &T, const fq_t &xp, const fq_t &yp)
{
    const F2 x2 = fq2_sqr(T.x);
    const F2 lam = fq2_mul(fq2_add(fq2_dbl(x2), x2), fq2_inv(fq2_dbl(T.y)));
    f = mul_by_line(f, lam, T, xp, yp);
    const F2 x3 = fq2_sub(fq2_sqr(lam), fq2_dbl(T.x));
    T.y = fq2_sub(fq2_mul(lam, fq2_sub(T.x, x3)), T.y);
    T.x = x3;
}

// f <- f * chord(T, R)(P); T <- T + R.  T != +-R for points of order r
// (T is then a multiple of Q below the group order).
void add_step(F12 &f, TwistPoint &T, const TwistPoint &R, const fq_t &xp, const fq_t &yp)
{
    const F2 lam = fq2_mul(fq2_sub(R.y, T.y), fq2_inv(fq2_sub(R.x, T.x)));
    f = mul_by_line(f, lam, T, xp, yp);
    const F2 x3 = fq2_sub(fq2_sub(fq2_sqr(lam), T.x), R.x);
    T.y = fq2_sub(fq2_mul(lam, fq2_sub(T.x, x3)), T.y);
    T.x = x3;
}

// Jacobian -> affine for many points with one inversion (Montgomery's
// simultaneous inversion); infinity maps to the flagged affine infinity.
std::vector<g1_affine> normalize_batch(const g1_jac *p, size_t n)
{
    std::vector<g1_affine> out(n);
    std::vector<fq_t> prefix(n);
    fq_t acc = fq_t::one();
    for (size_t i = 0; i < n; ++i) {
        prefix[i] = acc;
        if (!p[i].Z.is_zero()) acc = fp_mul(acc, p[i].Z);
    }
    fq_t inv = fp_inv(acc);
    for (size_t i = n; i-- > 0;) {
        if (p[i].Z.is_zero()) {
            out[i] = affine_infinity<fq_t>();
            continue;
        }
        const fq_t zi = fp_mul(inv, prefix[i]);
        inv = fp_mul(inv, p[i].Z);
        const fq_t zi2 = fp_sqr(zi);
        out[i].x = fp_mul(p[i].X, zi2);
        out[i].y = fp_mul(p[i].Y, fp_mul(zi2, zi));
    }
    return out;
}

}  // namespace

fq12_t gt_one() { return F12{f6_one(), f6_zero()}; }
fq12_t gt_mul(const fq12_t &a, const fq12_t &b) { return f12_mul(a, b); }
fq12_t gt_inv(const fq12_t &a) { return f12_inv(a); }
bool gt_eq(const fq12_t &a, const fq12_t &b) { return std::memcmp(&a, &b, sizeof(F12)) == 0; }

fq12_t bn254_miller_loop(const g1_jac &Pj, const g2_jac &Qj)
{
    g1_affine P;
    g2_affine Qa;
    if (!jacobian_to_affine(Pj, P) || !jacobian_to_affine(Qj, Qa)) return gt_one();
    const TwistPoint Q{Qa.x, Qa.y};
    TwistPoint T = Q;
    F12 f = gt_one();
    for (int i = 63; i >= 0; --i) {  // the bits of 6x + 2 below its top bit (bit 64)
        f = f12_sqr(f);
        dbl_step(f, T, P.x, P.y);
        if ((kAteLoop[i >> 6] >> (i & 63)) & 1) add_step(f, T, Q, P.x, P.y);
    }
    const Constants &k = constants();
    const TwistPoint Q1{fq2_mul(f2_conj(Q.x), k.frob_x), fq2_mul(f2_conj(Q.y), k.frob_y)};
    const TwistPoint Q2{fq2_mul(f2_conj(Q1.x), k.frob_x), fq2_neg(fq2_mul(f2_conj(Q1.y), k.frob_y))};
    add_step(f, T, Q1, P.x, P.y);
    add_step(f, T, Q2, P.x, P.y);
    return f;
}

fq12_t bn254_final_exponentiation(const fq12_t &f)
{
    F12 t = f12_mul(f12_conj(f), f12_inv(f));  // f^(q^6 - 1)
    t = f12_mul(f12_frob2(t), t);              // ^(q^2 + 1)
    return f12_pow(t, kHardExp, 15);
}

fq12_t bn254_pairing(const g1_jac &P, const g2_jac &Q) { return bn254_final_exponentiation(bn254_miller_loop(P, Q)); }

fq12_t bn254_pairing_product(const g1_jac *P, const g2_jac *Q, size_t k)
{
    F12 f = gt_one();
    for (size_t i = 0; i < k; ++i) f = f12_mul(f, bn254_miller_loop(P[i], Q[i]));
    return bn254_final_exponentiation(f);
}

bool g1_on_curve(const g1_jac &p)
{
    if (p.Z.is_zero()) return true;
    // Y^2 = X^3 + 3 Z^6
    const fq_t z2 = fp_sqr(p.Z), z6 = fp_mul(fp_sqr(z2), z2);
    const fq_t rhs = fp_add(fp_mul(fp_sqr(p.X), p.X), fp_mul(fq_from_u32(3), z6));
    return fp_sqr(p.Y) == rhs;
}

bool g2_on_curve(const g2_jac &p)
{
    if (p.Z.is_zero()) return true;
    const F2 z2 = fq2_sqr(p.Z), z6 = fq2_mul(fq2_sqr(z2), z2);
    const F2 rhs = fq2_add(fq2_mul(fq2_sqr(p.X), p.X), fq2_mul(constants().twist_b, z6));
    return fq2_sqr(p.Y) == rhs;
}

bool g2_in_subgroup(const g2_jac &p) { return xyzz_mul_u256(jacobian_to_xyzz(p), FrParams::P).is_inf(); }

g1_jac g1_msm_host(const g1_jac *bases, const fr_t *scalars, size_t n)
{
    const std::vector<g1_affine> pts = normalize_batch(bases, n);
    std::vector<fr_t> k(n);
    for (size_t i = 0; i < n; ++i) k[i] = fp_from_mont(scalars[i]);
    int c = 1;
    while (c < 16 && (size_t(1) << (c + 2)) < n) ++c;
    const int windows = (254 + c - 1) / c;
    std::vector<g1_xyzz> buckets(size_t(1) << c);
    g1_xyzz acc = g1_xyzz::infinity();
    for (int w = windows - 1; w >= 0; --w) {
        for (int d = 0; d < c; ++d) acc = xyzz_dbl(acc);
        for (auto &b : buckets) b = g1_xyzz::infinity();
        const int bit = w * c, limb = bit >> 5, sh = bit & 31;
        for (size_t i = 0; i < n; ++i) {
            if (affine_is_inf(pts[i])) continue;
            // the c-bit digit at bit w*c spans at most two limbs (c <= 16)
            uint64_t word = k[i].l[limb];
            if (limb + 1 < 8) word |= (uint64_t)k[i].l[limb + 1] << 32;
            const uint32_t digit = (uint32_t)(word >> sh) & ((1u << c) - 1);
            if (digit) buckets[digit] = xyzz_add_affine(buckets[digit], pts[i].x, pts[i].y);
        }
        g1_xyzz running = g1_xyzz::infinity(), sum = g1_xyzz::infinity();
        for (size_t d = buckets.size() - 1; d >= 1; --d) {
            running = xyzz_add(running, buckets[d]);
            sum = xyzz_add(sum, running);
        }
        acc = xyzz_add(acc, sum);
    }
    return xyzz_to_jacobian_affine(acc);
}

bool groth16_verify(const VerifyingKeyView &vk, const fr_t *public_input, size_t n_public, const g1_jac &A,
                    const g2_jac &B, const g1_jac &C)
{
    if (vk.ic_len != n_public + 1) return false;
    if (!g1_on_curve(A) || !g1_on_curve(C) || !g2_on_curve(B) || !g2_in_subgroup(B)) return false;
    g1_jac acc = vk.ic[0];
    if (n_public) {
        const g1_jac sum = g1_msm_host(vk.ic + 1, public_input, n_public);
        acc = xyzz_to_jacobian_affine(xyzz_add(jacobian_to_xyzz(vk.ic[0]), jacobian_to_xyzz(sum)));
    }
    // e(A, B) e(-IC, gamma) e(-C, delta) == alpha_g1_beta_g2
    g1_jac neg_acc = acc, neg_c = C;
    neg_acc.Y = fp_neg(neg_acc.Y);
    neg_c.Y = fp_neg(neg_c.Y);
    const g1_jac ps[3] = {A, neg_acc, neg_c};
    const g2_jac qs[3] = {B, *vk.gamma_g2, *vk.delta_g2};
    const F12 lhs = bn254_pairing_product(ps, qs, 3);
    // the caller's GT value in canonical form before the byte comparison
    const F12 want = f12_mul(*vk.alpha_g1_beta_g2, gt_one());
    return gt_eq(lhs, want);
}

}  // namespace nkr

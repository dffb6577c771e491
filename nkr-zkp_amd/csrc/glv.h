// Host scalar multiplication on BN254 G1 with the curve's endomorphism
// (Gallant, Lambert, Vanstone, "Faster point multiplication on elliptic
// curves with efficient endomorphisms", CRYPTO 2001).
//
// phi(x, y) = (beta x, y) equals lambda (x, y) on G1, with beta a cube root of
// unity mod q and lambda one mod r (the pair that matches on the generator).
// A scalar k splits as k = k1 + k2 lambda (mod r) with |k1|, |k2| < 2^128:
// c1 = floor(k g1 / 2^256), c2 = floor(k g2 / 2^256) with g1 = floor(b2 2^256 / r),
// g2 = floor(b1 2^256 / r), then k1 = k - c1 a1 - c2 a2 and k2 = c1 b1 - c2 b2,
// where (a1, -b1) and (a2, b2) are the short basis of the lattice
// {(u, v) : u + v lambda = 0 mod r} that the extended Euclidean algorithm on
// (r, lambda) yields.  Any integers c1, c2 keep k1 + k2 lambda = k (mod r); the
// rounding only bounds the halves.  k P is then k1 P + k2 phi(P): two 128-bit
// scalars sharing 128 doublings.
#pragma once

#include "curve.h"

namespace nkr {

struct GlvSplit {
    uint64_t k1[2], k2[2];  // |k1|, |k2|
    bool neg1, neg2;
};

namespace glv_detail {

typedef unsigned __int128 u128;

constexpr uint64_t kA1 = 0x89d3256894d213e3ull;                               // a1 = 9931322734385697763
constexpr uint64_t kB1[2] = {0x8211bbeb7d4f1128ull, 0x6f4d8248eeb859fcull};   // b1
constexpr uint64_t kA2[2] = {0x0be4e1541221250bull, 0x6f4d8248eeb859fdull};   // a2
constexpr uint64_t kB2 = 0x89d3256894d213e3ull;                               // b2 = a1
constexpr uint64_t kG1[2] = {0xd91d232ec7e0b3d7ull, 0x2ull};
constexpr uint64_t kG2[3] = {0x7a7bd9d4391eb18dull, 0x4ccef014a773d2cfull, 0x2ull};

// out[0..na+nb) = a * b
inline void mul_words(const uint64_t *a, int na, const uint64_t *b, int nb, uint64_t *out)
{
    for (int i = 0; i < na + nb; ++i) out[i] = 0;
    for (int i = 0; i < na; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < nb; ++j) {
            const u128 t = (u128)a[i] * b[j] + out[i + j] + carry;
            out[i + j] = (uint64_t)t;
            carry = (uint64_t)(t >> 64);
        }
        out[i + nb] = carry;
    }
}

// x -= y over 4 words (mod 2^256)
inline void sub4(uint64_t *x, const uint64_t *y)
{
    uint64_t borrow = 0;
    for (int i = 0; i < 4; ++i) {
        const u128 t = (u128)x[i] - y[i] - borrow;
        x[i] = (uint64_t)t;
        borrow = (uint64_t)(t >> 64) & 1;
    }
}

// |x| and its sign for a 4-word two's-complement value below 2^128 in size.
inline bool abs128(const uint64_t *x, uint64_t *out)
{
    const bool neg = (x[3] >> 63) != 0;
    if (!neg) {
        out[0] = x[0];
        out[1] = x[1];
        return false;
    }
    uint64_t z[4] = {0, 0, 0, 0};
    sub4(z, x);
    out[0] = z[0];
    out[1] = z[1];
    return true;
}

}  // namespace glv_detail

// k (Montgomery form) = k1 + k2 lambda (mod r).
inline GlvSplit glv_split(const fr_t &k_mont)
{
    using namespace glv_detail;
    const fr_t kc = fp_from_mont(k_mont);
    uint64_t k[4];
    for (int i = 0; i < 4; ++i) k[i] = (uint64_t)kc.l[2 * i] | ((uint64_t)kc.l[2 * i + 1] << 32);
    uint64_t p1[6], p2[7];
    mul_words(k, 4, kG1, 2, p1);
    mul_words(k, 4, kG2, 3, p2);
    const uint64_t c1 = p1[4];  // k g1 < 2^320
    const uint64_t c2[2] = {p2[4], p2[5]};  // k g2 < 2^384
    // k1 = k - c1 a1 - c2 a2
    uint64_t t[4] = {k[0], k[1], k[2], k[3]};
    uint64_t m[4];
    const uint64_t a1[1] = {kA1};
    uint64_t m3[3];
    mul_words(&c1, 1, a1, 1, m3);
    m[0] = m3[0];
    m[1] = m3[1];
    m[2] = m[3] = 0;
    sub4(t, m);
    mul_words(c2, 2, kA2, 2, m);
    sub4(t, m);
    GlvSplit g;
    g.neg1 = abs128(t, g.k1);
    // k2 = c1 b1 - c2 b2
    uint64_t u[4];
    mul_words(&c1, 1, kB1, 2, m3);
    u[0] = m3[0];
    u[1] = m3[1];
    u[2] = m3[2];
    u[3] = 0;
    const uint64_t b2[1] = {kB2};
    mul_words(c2, 2, b2, 1, m3);
    m[0] = m3[0];
    m[1] = m3[1];
    m[2] = m3[2];
    m[3] = 0;
    sub4(u, m);
    g.neg2 = abs128(u, g.k2);
    return g;
}

// phi(P) in XYZZ coordinates: x = X / ZZ, so X picks up the factor beta.
inline g1_xyzz g1_endo(const g1_xyzz &p)
{
    // beta = 0x59e26bcea0d48bacd4f263f1acdb5c4f5763473177fffffe, Montgomery form
    fq_t beta;
    const uint32_t bm[8] = {0xd782e155u, 0x71930c11u, 0xffbe3323u, 0xa6bb947cu,
                            0xd4741444u, 0xaa303344u, 0x26594943u, 0x2c3b3f0du};
    for (int i = 0; i < 8; ++i) beta.l[i] = bm[i];
    g1_xyzz r = p;
    r.X = fp_mul(p.X, beta);
    return r;
}

// k P on G1 for a Montgomery-form scalar k: the GLV halves with 4-bit joint
// windows (124 doublings, at most 64 additions, 14 more for the table).
inline g1_xyzz g1_mul_glv(const g1_xyzz &P, const fr_t &k_mont)
{
    const GlvSplit s = glv_split(k_mont);
    g1_xyzz t1[16], t2[16];
    t1[0] = t2[0] = g1_xyzz::infinity();
    t1[1] = s.neg1 ? xyzz_neg(P) : P;
    for (int i = 2; i < 16; ++i) t1[i] = xyzz_add(t1[i - 1], t1[1]);
    for (int i = 1; i < 16; ++i) {
        const g1_xyzz e = g1_endo(t1[i]);
        t2[i] = (s.neg1 != s.neg2) ? xyzz_neg(e) : e;
    }
    g1_xyzz acc = g1_xyzz::infinity();
    for (int nib = 31; nib >= 0; --nib) {
        if (!acc.is_inf())
            for (int d = 0; d < 4; ++d) acc = xyzz_dbl(acc);
        const uint32_t sh = (nib & 15) * 4;
        const uint32_t v1 = (uint32_t)(s.k1[nib >> 4] >> sh) & 15u, v2 = (uint32_t)(s.k2[nib >> 4] >> sh) & 15u;
        if (v1) acc = xyzz_add(acc, t1[v1]);
        if (v2) acc = xyzz_add(acc, t2[v2]);
    }
    return acc;
}

// k P + l Q on G1: both scalars' GLV halves (four 128-bit scalars) with
// 4-bit joint windows, sharing one chain of 124 doublings (two g1_mul_glv
// calls take 248).
inline g1_xyzz g1_mul2_glv(const g1_xyzz &P, const fr_t &k_mont, const g1_xyzz &Q, const fr_t &l_mont)
{
    const GlvSplit s[2] = {glv_split(k_mont), glv_split(l_mont)};
    const g1_xyzz *base[2] = {&P, &Q};
    g1_xyzz t[2][2][16];
    for (int b = 0; b < 2; ++b) {
        t[b][0][0] = t[b][1][0] = g1_xyzz::infinity();
        t[b][0][1] = s[b].neg1 ? xyzz_neg(*base[b]) : *base[b];
        for (int i = 2; i < 16; ++i) t[b][0][i] = xyzz_add(t[b][0][i - 1], t[b][0][1]);
        for (int i = 1; i < 16; ++i) {
            const g1_xyzz e = g1_endo(t[b][0][i]);
            t[b][1][i] = (s[b].neg1 != s[b].neg2) ? xyzz_neg(e) : e;
        }
    }
    g1_xyzz acc = g1_xyzz::infinity();
    for (int nib = 31; nib >= 0; --nib) {
        if (!acc.is_inf())
            for (int d = 0; d < 4; ++d) acc = xyzz_dbl(acc);
        const uint32_t sh = (nib & 15) * 4;
        for (int b = 0; b < 2; ++b) {
            const uint32_t v1 = (uint32_t)(s[b].k1[nib >> 4] >> sh) & 15u;
            const uint32_t v2 = (uint32_t)(s[b].k2[nib >> 4] >> sh) & 15u;
            if (v1) acc = xyzz_add(acc, t[b][0][v1]);
            if (v2) acc = xyzz_add(acc, t[b][1][v2]);
        }
    }
    return acc;
}

}  // namespace nkr

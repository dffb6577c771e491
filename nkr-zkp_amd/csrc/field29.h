// Reduced-radix BN254 field arithmetic for the hot loops (gfx950).
//
// An element is 9 limbs of 29 bits (value = sum l_k 2^(29k), room for 261
// bits) holding a Montgomery residue with R = 2^261, lazily reduced: values
// may exceed p by a small multiple, as each operation below states.
//
// Why: with 32-bit limbs every 32x32+64 multiply-add of a Montgomery product
// can overflow its 64-bit accumulator, so each v_mad_u64_u32 needs a carry
// add behind it (and the carry travels through an SGPR, which costs wait
// states).  With 29-bit limbs a product term is below 2^58 and a column of
// at most 18 terms plus the carry-in stays below 2^63: one v_mad_u64_u32 per
// term and no carries inside a column -- 162 multiply-adds and ~90 other
// instructions per product instead of ~450 instructions.  Additions become
// carry-normalised limb sums without reduction; the bounds below keep every
// product input inside what the reduction tolerates.
//
// Conversions: a canonical 32-bit-limb element x*2^256 (the reference's
// layout, field.h) becomes x*2^261 by a 5-bit shift (a value below 32p, no
// reduction); back, v*2^-5 mod p is one 5-bit Montgomery step and one
// conditional subtraction.  Results are exact field elements, so kernels that
// use this form produce bit-identical canonical outputs.
#pragma once

#include "field.h"

namespace nkr {

constexpr uint32_t kMask29 = (1u << 29) - 1;

// Limbs of k * p in the 29-bit radix, computed at compile time (k * p < 2^261).
template <class P>
struct P29 {
    struct Limbs {
        uint32_t l[9];
    };
    static constexpr uint32_t bits(int k)
    {
        const int lo = 29 * k, wi = lo / 32, sh = lo % 32;
        uint64_t v = wi < 8 ? ((uint64_t)P::P[wi] >> sh) : 0;
        if (wi + 1 < 8 && sh > 3) v |= (uint64_t)P::P[wi + 1] << (32 - sh);
        return (uint32_t)v & kMask29;
    }
    static constexpr Limbs times(uint32_t k)
    {
        Limbs r{};
        uint64_t carry = 0;
        for (int i = 0; i < 9; ++i) {
            const uint64_t t = (uint64_t)k * bits(i) + carry;
            r.l[i] = (uint32_t)t & kMask29;
            carry = t >> 29;
        }
        return r;
    }
    // -p^-1 mod 2^29
    static constexpr uint32_t inv() { return P::INV & kMask29; }
    // 2^261 mod p, the Montgomery form of 1: (2^256 mod p) * 32 reduced.
    static constexpr Limbs one()
    {
        Limbs v{};
        // ONE (< p, 8 x 32 bits) shifted left by 5 into 29-bit limbs
        for (int k = 0; k < 9; ++k) {
            const int lo = 29 * k - 5;
            uint64_t x = 0;
            if (lo < 0) {
                x = (uint64_t)P::ONE[0] << (-lo);
            } else {
                const int wi = lo / 32, sh = lo % 32;
                x = wi < 8 ? ((uint64_t)P::ONE[wi] >> sh) : 0;
                if (wi + 1 < 8 && sh > 0) x |= (uint64_t)P::ONE[wi + 1] << (32 - sh);
            }
            v.l[k] = (uint32_t)x & kMask29;
        }
        for (uint32_t m = 16; m >= 1; m >>= 1) {  // subtract 16p, 8p, 4p, 2p, p where possible
            const Limbs mp = times(m);
            Limbs d{};
            int64_t borrow = 0;
            for (int i = 0; i < 9; ++i) {
                const int64_t t = (int64_t)v.l[i] - mp.l[i] - borrow;
                d.l[i] = (uint32_t)(t & kMask29);
                borrow = t < 0 ? 1 : 0;
            }
            if (!borrow) v = d;
        }
        return v;
    }
};

template <class P>
struct Fp29 {
    uint32_t l[9];

    NKR_HD static Fp29 zero()
    {
        Fp29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i) r.l[i] = 0;
        return r;
    }
    NKR_HD static Fp29 one()
    {
        constexpr typename P29<P>::Limbs o = P29<P>::one();
        Fp29 r;
#pragma unroll
        for (int i = 0; i < 9; ++i) r.l[i] = o.l[i];
        return r;
    }
    // all limbs zero (the exact zero this code writes for the point at infinity)
    NKR_HD bool is_exact_zero() const
    {
        uint32_t acc = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i) acc |= l[i];
        return acc == 0;
    }
};

// x*2^256 (canonical, 8 x 32 bits) -> x*2^261 as 29-bit limbs: a 5-bit shift,
// no reduction; the value is below 32p.
template <class P>
NKR_HD Fp29<P> fp29_from_mont(const Fp<P> &x)
{
    Fp29<P> r;
    r.l[0] = (x.l[0] << 5) & kMask29;
#pragma unroll
    for (int k = 1; k < 9; ++k) {
        const int lo = 29 * k - 5, wi = lo / 32, sh = lo % 32;
        uint32_t v = x.l[wi] >> sh;
        if (wi + 1 < 8 && sh > 3) v |= x.l[wi + 1] << (32 - sh);
        r.l[k] = v & kMask29;
    }
    return r;
}

// Montgomery product a*b*2^-261 mod p by product scanning.  Inputs: limbs
// below 2^29 and a*b < 2^261 * p (e.g. both below 10p, or a < 32p and
// b < 2p).  Output: limbs below 2^29, value below p + a*b/2^261 < 2p.
template <class P>
NKR_HD Fp29<P> fp29_mul(const Fp29<P> &a, const Fp29<P> &b)
{
    constexpr uint32_t INV = P29<P>::inv();
    uint32_t m[9];
    Fp29<P> r;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
#pragma unroll
        for (int i = 0; i <= k; ++i) acc += (uint64_t)a.l[i] * b.l[k - i];
#pragma unroll
        for (int i = 0; i < k; ++i) acc += (uint64_t)m[i] * P29<P>::bits(k - i);
        m[k] = ((uint32_t)acc * INV) & kMask29;
        acc += (uint64_t)m[k] * P29<P>::bits(0);
        acc >>= 29;
    }
#pragma unroll
    for (int k = 9; k < 17; ++k) {
#pragma unroll
        for (int i = k - 8; i < 9; ++i) acc += (uint64_t)a.l[i] * b.l[k - i];
#pragma unroll
        for (int i = k - 8; i < 9; ++i) acc += (uint64_t)m[i] * P29<P>::bits(k - i);
        r.l[k - 9] = (uint32_t)acc & kMask29;
        acc >>= 29;
    }
    r.l[8] = (uint32_t)acc;
    return r;
}

// Two independent products r0 = a0*b0 and r1 = a1*b1 (fp29_mul's bounds for
// each) with their multiply-adds interleaved: each product's column sum is
// one dependent chain through its accumulator, so a wave that alternates two
// chains issues while the other's v_mad_u64_u32 is still in flight.  The same
// results as two fp29_mul calls.  Outputs may alias inputs.
template <class P>
NKR_HD void fp29_mul2(const Fp29<P> &a0, const Fp29<P> &b0, const Fp29<P> &a1, const Fp29<P> &b1, Fp29<P> &r0,
                      Fp29<P> &r1)
{
    constexpr uint32_t INV = P29<P>::inv();
    uint32_t m0[9], m1[9], t0[9], t1[9];
    uint64_t acc0 = 0, acc1 = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
#pragma unroll
        for (int i = 0; i <= k; ++i) {
            acc0 += (uint64_t)a0.l[i] * b0.l[k - i];
            acc1 += (uint64_t)a1.l[i] * b1.l[k - i];
        }
#pragma unroll
        for (int i = 0; i < k; ++i) {
            acc0 += (uint64_t)m0[i] * P29<P>::bits(k - i);
            acc1 += (uint64_t)m1[i] * P29<P>::bits(k - i);
        }
        m0[k] = ((uint32_t)acc0 * INV) & kMask29;
        m1[k] = ((uint32_t)acc1 * INV) & kMask29;
        acc0 += (uint64_t)m0[k] * P29<P>::bits(0);
        acc1 += (uint64_t)m1[k] * P29<P>::bits(0);
        acc0 >>= 29;
        acc1 >>= 29;
    }
#pragma unroll
    for (int k = 9; k < 17; ++k) {
#pragma unroll
        for (int i = k - 8; i < 9; ++i) {
            acc0 += (uint64_t)a0.l[i] * b0.l[k - i];
            acc1 += (uint64_t)a1.l[i] * b1.l[k - i];
        }
#pragma unroll
        for (int i = k - 8; i < 9; ++i) {
            acc0 += (uint64_t)m0[i] * P29<P>::bits(k - i);
            acc1 += (uint64_t)m1[i] * P29<P>::bits(k - i);
        }
        t0[k - 9] = (uint32_t)acc0 & kMask29;
        t1[k - 9] = (uint32_t)acc1 & kMask29;
        acc0 >>= 29;
        acc1 >>= 29;
    }
    t0[8] = (uint32_t)acc0;
    t1[8] = (uint32_t)acc1;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        r0.l[i] = t0[i];
        r1.l[i] = t1[i];
    }
}

// a + b, limbs carry-normalised, not reduced (the caller keeps a + b < 2^261).
template <class P>
NKR_HD Fp29<P> fp29_add(const Fp29<P> &a, const Fp29<P> &b)
{
    Fp29<P> r;
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const uint32_t t = a.l[i] + b.l[i] + c;
        r.l[i] = t & kMask29;
        c = t >> 29;
    }
    return r;
}

// a + K*p - b for b <= K*p: congruent to a - b, value a + K*p - b (not
// reduced), limbs carry-normalised.
template <class P, uint32_t K>
NKR_HD Fp29<P> fp29_sub(const Fp29<P> &a, const Fp29<P> &b)
{
    constexpr typename P29<P>::Limbs kp = P29<P>::times(K);
    Fp29<P> r;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int32_t t = (int32_t)(a.l[i] + kp.l[i]) - (int32_t)b.l[i] + c;
        r.l[i] = (uint32_t)t & kMask29;
        c = t >> 29;  // arithmetic: -1, 0 or a small positive carry
    }
    return r;
}

// x - K*p if x >= K*p (x < 2^261).
template <class P, uint32_t K>
NKR_HD Fp29<P> fp29_sub_if_ge(const Fp29<P> &x)
{
    constexpr typename P29<P>::Limbs kp = P29<P>::times(K);
    Fp29<P> d;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int32_t t = (int32_t)x.l[i] - (int32_t)kp.l[i] + c;
        d.l[i] = (uint32_t)t & kMask29;
        c = t >> 29;
    }
    const bool keep = c < 0;  // x < K*p
    Fp29<P> r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.l[i] = keep ? x.l[i] : d.l[i];
    return r;
}

// x below 32p -> the same residue below 2p.
template <class P>
NKR_HD Fp29<P> fp29_reduce_to_2p(const Fp29<P> &x)
{
    return fp29_sub_if_ge<P, 2>(fp29_sub_if_ge<P, 4>(fp29_sub_if_ge<P, 8>(fp29_sub_if_ge<P, 16>(x))));
}

// x below 32p -> the same residue below 8p.
template <class P>
NKR_HD Fp29<P> fp29_reduce_to_8p(const Fp29<P> &x)
{
    return fp29_sub_if_ge<P, 8>(fp29_sub_if_ge<P, 16>(x));
}

// v*2^261 (v < 32p) -> canonical v*2^256 in 32-bit limbs: v*2^-5 mod p as
// (v + k*p) / 32 with k = -v*p^-1 mod 32, below 2p, then one subtraction.
template <class P>
NKR_HD Fp<P> fp29_to_mont(const Fp29<P> &v)
{
    const uint32_t k = (v.l[0] * P29<P>::inv()) & 31u;
    uint32_t s[9];
    uint64_t carry = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const uint64_t t = (uint64_t)k * P29<P>::bits(i) + v.l[i] + carry;
        s[i] = (uint32_t)t & kMask29;
        carry = t >> 29;
    }
    // s = v + k p < 64p < 2^261, divisible by 32: word i of s / 32 is bits
    // [32i + 5, 32i + 37) of s.
    Fp<P> r;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int lo = 32 * i + 5, k0 = lo / 29, sh = lo % 29;
        uint64_t w = (uint64_t)s[k0] >> sh;
        if (k0 + 1 < 9) w |= (uint64_t)s[k0 + 1] << (29 - sh);
        if (k0 + 2 < 9 && 58 - sh < 32) w |= (uint64_t)s[k0 + 2] << (58 - sh);
        r.l[i] = (uint32_t)w;
    }
    return fp_reduce_once(r);
}

// Could x (below k_max * p, k_max <= 16) be 0 mod p?  A necessary condition
// on the low limb only: x = j*p for some j < k_max.  Exact answers come from
// fp29_to_mont(x).is_zero().
template <class P>
struct P29LowMultiples {
    uint32_t v[16];
};

template <class P>
constexpr P29LowMultiples<P> p29_low_multiples()
{
    P29LowMultiples<P> r{};
    for (uint32_t j = 0; j < 16; ++j) r.v[j] = P29<P>::times(j).l[0];
    return r;
}

template <class P, uint32_t KMax>
NKR_HD bool fp29_may_be_zero(const Fp29<P> &x)
{
    static_assert(KMax <= 16, "at most 16 multiples of p");
    constexpr P29LowMultiples<P> lm = p29_low_multiples<P>();
    bool any = false;
#pragma unroll
    for (uint32_t j = 0; j < KMax; ++j) any = any || x.l[0] == lm.v[j];
    return any;
}

}  // namespace nkr

// BN254 prime-field arithmetic for gfx950 (CDNA4), device side.
//
// Elements are 8 little-endian 32-bit limbs holding the Montgomery residue
// x*2^256 mod p, fully reduced (< p).  This is byte-for-byte the layout of the
// reference's host scalar/base-field type (libff Fp_model<4,..>, 4 x u64
// mont_repr, R = 2^256: third_party/lightzk/depends/libsnark/depends/libff/
// libff/algebra/fields/fp.hpp:42), so buffers move between host and device
// without conversion.
//
// Multiplication is Montgomery with the operands regrouped into 29-bit limbs
// (fp_mul_r29 below): each 29x29+64 step is one v_mad_u64_u32 with no carry
// to propagate; additions and subtractions stay on the 32-bit limbs, whose
// carries the hardware chains.  The integer product is not a dense
// contraction, so nothing here uses MFMA.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define NKR_DEV __device__ __forceinline__
#define NKR_HD __host__ __device__ __forceinline__

namespace nkr {

// Scalar field Fr: r = 0x30644e72e131a029b85045b68181585d2833e84879b9709143e1f593f0000001
struct FrParams {
    static constexpr uint32_t P[8] = {0xf0000001u, 0x43e1f593u, 0x79b97091u, 0x2833e848u,
                                      0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
    static constexpr uint32_t ONE[8] = {0x4ffffffbu, 0xac96341cu, 0x9f60cd29u, 0x36fc7695u,
                                        0x7879462eu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
    static constexpr uint32_t R2[8] = {0xae216da7u, 0x1bb8e645u, 0xe35c59e3u, 0x53fe3ab1u,
                                       0x53bb8085u, 0x8c49833du, 0x7f4e44a5u, 0x0216d0b1u};
    static constexpr uint32_t INV = 0xefffffffu;  // -r^-1 mod 2^32
    static constexpr uint64_t P64[4] = {0x43e1f593f0000001ull, 0x2833e84879b97091ull, 0xb85045b68181585dull,
                                        0x30644e72e131a029ull};
    static constexpr uint64_t INV64 = 0xc2e1f593efffffffull;  // -r^-1 mod 2^64
};

// Base field Fq: q = 0x30644e72e131a029b85045b68181585d97816a916871ca8d3c208c16d87cfd47
struct FqParams {
    static constexpr uint32_t P[8] = {0xd87cfd47u, 0x3c208c16u, 0x6871ca8du, 0x97816a91u,
                                      0x8181585du, 0xb85045b6u, 0xe131a029u, 0x30644e72u};
    static constexpr uint32_t ONE[8] = {0xc58f0d9du, 0xd35d438du, 0xf5c70b3du, 0x0a78eb28u,
                                        0x7879462cu, 0x666ea36fu, 0x9a07df2fu, 0x0e0a77c1u};
    static constexpr uint32_t R2[8] = {0x538afa89u, 0xf32cfc5bu, 0xd44501fbu, 0xb5e71911u,
                                       0x0a417ff6u, 0x47ab1effu, 0xcab8351fu, 0x06d89f71u};
    static constexpr uint32_t INV = 0xe4866389u;  // -q^-1 mod 2^32
    static constexpr uint64_t P64[4] = {0x3c208c16d87cfd47ull, 0x97816a916871ca8dull, 0xb85045b68181585dull,
                                        0x30644e72e131a029ull};
    static constexpr uint64_t INV64 = 0x87d20782e4866389ull;  // -q^-1 mod 2^64
};

template <class Params>
struct Fp {
    uint32_t l[8];

    NKR_HD static Fp zero()
    {
        Fp r;
#pragma unroll
        for (int i = 0; i < 8; ++i) r.l[i] = 0;
        return r;
    }
    NKR_HD static Fp one()
    {
        Fp r;
#pragma unroll
        for (int i = 0; i < 8; ++i) r.l[i] = Params::ONE[i];
        return r;
    }
    NKR_HD static Fp modulus()
    {
        Fp r;
#pragma unroll
        for (int i = 0; i < 8; ++i) r.l[i] = Params::P[i];
        return r;
    }

    NKR_HD bool is_zero() const
    {
        uint32_t acc = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) acc |= l[i];
        return acc == 0;
    }
    NKR_HD bool operator==(const Fp &o) const
    {
        uint32_t acc = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) acc |= l[i] ^ o.l[i];
        return acc == 0;
    }
    NKR_HD bool operator!=(const Fp &o) const { return !(*this == o); }
};

// 32-bit add and subtract with carry / borrow (0 or 1) in and out.  On the
// device these are the compiler's carry builtins, one v_add_co / v_addc_co
// (v_sub_co / v_subb_co) per limb with the carry in VCC; the same chain
// written with 64-bit intermediates compiles for gfx950 into 64-bit adds
// plus the moves that build their register pairs, about three times the
// instructions (fp_add: 100 VALU instructions against 35).  Host code keeps
// the portable form.
NKR_HD uint32_t addc32(uint32_t a, uint32_t b, uint32_t cin, uint32_t &cout)
{
#if defined(__HIP_DEVICE_COMPILE__)
    unsigned int c;
    const uint32_t r = __builtin_addc(a, b, cin, &c);
    cout = c;
    return r;
#else
    const uint64_t t = (uint64_t)a + b + cin;
    cout = (uint32_t)(t >> 32);
    return (uint32_t)t;
#endif
}

NKR_HD uint32_t subc32(uint32_t a, uint32_t b, uint32_t bin, uint32_t &bout)
{
#if defined(__HIP_DEVICE_COMPILE__)
    unsigned int c;
    const uint32_t r = __builtin_subc(a, b, bin, &c);
    bout = c;
    return r;
#else
    const uint64_t t = (uint64_t)a - b - bin;
    bout = (uint32_t)(t >> 63);
    return (uint32_t)t;
#endif
}

// x - p if x >= p (x < 2p).  Borrow chain, then a select.
template <class P>
NKR_HD Fp<P> fp_reduce_once(const Fp<P> &x)
{
    Fp<P> d;
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) d.l[i] = subc32(x.l[i], P::P[i], borrow, borrow);
    Fp<P> r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.l[i] = borrow ? x.l[i] : d.l[i];
    return r;
}

template <class P>
NKR_HD Fp<P> fp_add(const Fp<P> &a, const Fp<P> &b)
{
    Fp<P> s;
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) s.l[i] = addc32(a.l[i], b.l[i], carry, carry);
    // a + b < 2p < 2^255: no carry out of the top word.
    return fp_reduce_once(s);
}

template <class P>
NKR_HD Fp<P> fp_sub(const Fp<P> &a, const Fp<P> &b)
{
    Fp<P> d;
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) d.l[i] = subc32(a.l[i], b.l[i], borrow, borrow);
    // On borrow add p back; mask instead of branch keeps the wave converged.
    const uint32_t mask = 0u - borrow;
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) d.l[i] = addc32(d.l[i], P::P[i] & mask, carry, carry);
    return d;
}

template <class P>
NKR_HD Fp<P> fp_dbl(const Fp<P> &a)
{
    return fp_add(a, a);
}

template <class P>
NKR_HD Fp<P> fp_neg(const Fp<P> &a)
{
    // p - a, and 0 stays 0.
    Fp<P> d;
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) d.l[i] = subc32(P::P[i], a.l[i], borrow, borrow);
    const bool z = a.is_zero();
#pragma unroll
    for (int i = 0; i < 8; ++i) d.l[i] = z ? 0u : d.l[i];
    return d;
}

// ---------------------------------------------------------------------------
// Reduced-radix products.  An element of the 29-bit form is 9 limbs of 29 bits
// (value = sum l_k 2^(29k), room for 261 bits) holding a Montgomery residue
// with R = 2^261.  With 32-bit limbs every 32x32+64 multiply-add of a
// Montgomery product can overflow its 64-bit accumulator, so each
// v_mad_u64_u32 needs a carry add behind it, and the carry travels through
// an SGPR, which costs wait states.  With 29-bit limbs a product term is
// below 2^58 and a column of at most 18 terms plus the carry-in stays below
// 2^63: one v_mad_u64_u32 per term and no carries inside a column (162
// multiply-adds and ~90 other instructions per product).  The canonical
// product below (fp_mul) goes through it: x*y*2^-256 = (32x)*y*2^-261, the
// shift by 5 being free in the limb conversion.  Measured on MI355X in
// chains of dependent products: 126.5 G canonical products/s against 106.6 G
// for the 32-bit product-scanning form (field29.h builds lazily reduced
// arithmetic on the same form for the MSM accumulation).

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

// The bits of x (any 256-bit value) as 29-bit limbs, value unchanged.
template <class P>
NKR_HD Fp29<P> fp29_from_bits(const Fp<P> &x)
{
    Fp29<P> r;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int lo = 29 * k, wi = lo / 32, sh = lo % 32;
        uint32_t v = x.l[wi] >> sh;
        if (wi + 1 < 8 && sh > 3) v |= x.l[wi + 1] << (32 - sh);
        r.l[k] = v & kMask29;
    }
    return r;
}

// The value of x (below 2^256) as 8 x 32-bit limbs.
template <class P>
NKR_HD Fp<P> fp29_to_bits(const Fp29<P> &x)
{
    Fp<P> r;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int lo = 32 * i, k = lo / 29, sh = lo % 29;
        uint32_t w = x.l[k] >> sh;
        if (k + 1 < 9) w |= x.l[k + 1] << (29 - sh);
        if (k + 2 < 9 && 58 - sh < 32) w |= x.l[k + 2] << (58 - sh);
        r.l[i] = w;
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

// a*a*2^-261 mod p, the same value and bounds as fp29_mul(a, a): each
// symmetric pair of partial products is formed once from a doubled limb
// (2 a_i < 2^30, so a pair term stays below 2^59 and a column below 2^63),
// 45 multiply-adds for the product half instead of 81.
template <class P>
NKR_HD Fp29<P> fp29_sqr(const Fp29<P> &a)
{
    constexpr uint32_t INV = P29<P>::inv();
    uint32_t d[9], m[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) d[i] = a.l[i] << 1;
    Fp29<P> r;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
#pragma unroll
        for (int i = 0; 2 * i < k; ++i) acc += (uint64_t)d[i] * a.l[k - i];
        if (k % 2 == 0) acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
#pragma unroll
        for (int i = 0; i < k; ++i) acc += (uint64_t)m[i] * P29<P>::bits(k - i);
        m[k] = ((uint32_t)acc * INV) & kMask29;
        acc += (uint64_t)m[k] * P29<P>::bits(0);
        acc >>= 29;
    }
#pragma unroll
    for (int k = 9; k < 17; ++k) {
#pragma unroll
        for (int i = k - 8; 2 * i < k; ++i) acc += (uint64_t)d[i] * a.l[k - i];
        if (k % 2 == 0) acc += (uint64_t)a.l[k / 2] * a.l[k / 2];
#pragma unroll
        for (int i = k - 8; i < 9; ++i) acc += (uint64_t)m[i] * P29<P>::bits(k - i);
        r.l[k - 9] = (uint32_t)acc & kMask29;
        acc >>= 29;
    }
    r.l[8] = (uint32_t)acc;
    return r;
}

// (a*b + c*d)*2^-261 mod p with one Montgomery reduction: both products'
// partial sums go into the same columns (27 terms below 2^58 each, below
// 2^63).  Inputs: limbs below 2^29, a*b + c*d < 2^261 * p.  Output: limbs
// below 2^29, value below p + (a*b + c*d)/2^261 < 2p.
template <class P>
NKR_HD Fp29<P> fp29_mul_add(const Fp29<P> &a, const Fp29<P> &b, const Fp29<P> &c, const Fp29<P> &d)
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
        for (int i = 0; i <= k; ++i) acc += (uint64_t)c.l[i] * d.l[k - i];
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
        for (int i = k - 8; i < 9; ++i) acc += (uint64_t)c.l[i] * d.l[k - i];
#pragma unroll
        for (int i = k - 8; i < 9; ++i) acc += (uint64_t)m[i] * P29<P>::bits(k - i);
        r.l[k - 9] = (uint32_t)acc & kMask29;
        acc >>= 29;
    }
    r.l[8] = (uint32_t)acc;
    return r;
}

// (a*b + c*d + e*f + g*h)*2^-261 mod p with one Montgomery reduction (36
// product terms and 9 reduction terms per column, below 45 * 2^58 < 2^64).
// Inputs: limbs below 2^29, the sum of the four products below 2^261 * p.
// Output: limbs below 2^29, value below p + sum/2^261 < 2p.
template <class P>
NKR_HD Fp29<P> fp29_mul_add4(const Fp29<P> &a, const Fp29<P> &b, const Fp29<P> &c, const Fp29<P> &d,
                             const Fp29<P> &e, const Fp29<P> &f, const Fp29<P> &g, const Fp29<P> &h)
{
    constexpr uint32_t INV = P29<P>::inv();
    uint32_t m[9];
    Fp29<P> r;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
#pragma unroll
        for (int i = 0; i <= k; ++i)
            acc += (uint64_t)a.l[i] * b.l[k - i] + (uint64_t)c.l[i] * d.l[k - i] + (uint64_t)e.l[i] * f.l[k - i] +
                   (uint64_t)g.l[i] * h.l[k - i];
#pragma unroll
        for (int i = 0; i < k; ++i) acc += (uint64_t)m[i] * P29<P>::bits(k - i);
        m[k] = ((uint32_t)acc * INV) & kMask29;
        acc += (uint64_t)m[k] * P29<P>::bits(0);
        acc >>= 29;
    }
#pragma unroll
    for (int k = 9; k < 17; ++k) {
#pragma unroll
        for (int i = k - 8; i < 9; ++i)
            acc += (uint64_t)a.l[i] * b.l[k - i] + (uint64_t)c.l[i] * d.l[k - i] + (uint64_t)e.l[i] * f.l[k - i] +
                   (uint64_t)g.l[i] * h.l[k - i];
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

// x*w*2^-256 mod p (the canonical form's Montgomery product) through the
// 29-bit product: (32x)*w*2^-261.  x: any 256-bit value; w: below p.  The
// result is below p + x*w/2^256 < 2p, not reduced, in 32-bit limbs.
template <class P>
NKR_HD Fp<P> fp_mul_r29(const Fp<P> &x, const Fp<P> &w)
{
    return fp29_to_bits(fp29_mul(fp29_from_mont(x), fp29_from_bits(w)));
}



#if !defined(__HIP_DEVICE_COMPILE__)
// Host build: the same no-carry CIOS over 4 x 64-bit words with 128-bit
// products (the top word of either modulus is below 2^62).
template <class P>
inline Fp<P> fp_mul_host64(const Fp<P> &a, const Fp<P> &b)
{
    typedef unsigned __int128 u128;
    uint64_t x[4], y[4], t[4] = {0, 0, 0, 0};
    __builtin_memcpy(x, a.l, 32);
    __builtin_memcpy(y, b.l, 32);
    for (int i = 0; i < 4; ++i) {
        u128 s = (u128)x[0] * y[i] + t[0];
        uint64_t A = (uint64_t)(s >> 64);
        const uint64_t t0 = (uint64_t)s;
        const uint64_t m = t0 * P::INV64;
        u128 u = (u128)m * P::P64[0] + t0;
        uint64_t C = (uint64_t)(u >> 64);
        for (int j = 1; j < 4; ++j) {
            s = (u128)x[j] * y[i] + t[j] + A;
            A = (uint64_t)(s >> 64);
            u = (u128)m * P::P64[j] + (uint64_t)s + C;
            C = (uint64_t)(u >> 64);
            t[j - 1] = (uint64_t)u;
        }
        t[3] = C + A;
    }
    Fp<P> r;
    __builtin_memcpy(r.l, t, 32);
    return fp_reduce_once(r);
}
#endif

// Montgomery product a*b*2^-256 mod p, inputs and output below p.  On gfx950
// through the 29-bit product (fp_mul_r29) and one conditional subtraction;
// the host build uses 64-bit words.
template <class P>
NKR_HD Fp<P> fp_mul(const Fp<P> &a, const Fp<P> &b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return fp_reduce_once(fp_mul_r29(a, b));
#else
    return fp_mul_host64(a, b);
#endif
}

// r0 = a0*b0 and r1 = a1*b1 (independent products; outputs may alias inputs).
template <class P>
NKR_HD void fp_mul2(const Fp<P> &a0, const Fp<P> &b0, const Fp<P> &a1, const Fp<P> &b1, Fp<P> &r0, Fp<P> &r1)
{
#if defined(__HIP_DEVICE_COMPILE__)
    Fp29<P> x, y;
    fp29_mul2(fp29_from_mont(a0), fp29_from_bits(b0), fp29_from_mont(a1), fp29_from_bits(b1), x, y);
    r0 = fp_reduce_once(fp29_to_bits(x));
    r1 = fp_reduce_once(fp29_to_bits(y));
#else
    const Fp<P> x = fp_mul_host64(a0, b0), y = fp_mul_host64(a1, b1);
    r0 = x;
    r1 = y;
#endif
}

template <class P>
NKR_HD Fp<P> fp_sqr(const Fp<P> &a)
{
    return fp_mul(a, a);
}

// Integer x (< p) to Montgomery form, and back.
template <class P>
NKR_HD Fp<P> fp_to_mont(const Fp<P> &x)
{
    Fp<P> r2;
#pragma unroll
    for (int i = 0; i < 8; ++i) r2.l[i] = P::R2[i];
    return fp_mul(x, r2);
}

template <class P>
NKR_HD Fp<P> fp_from_mont(const Fp<P> &x)
{
    Fp<P> one_int = Fp<P>::zero();
    one_int.l[0] = 1;
    return fp_mul(x, one_int);
}

// x^e for a 256-bit exponent given as 8 words (left-to-right square and multiply).
template <class P>
NKR_HD Fp<P> fp_pow(const Fp<P> &x, const uint32_t e[8])
{
    Fp<P> acc = Fp<P>::one();
    for (int w = 7; w >= 0; --w) {
        for (int b = 31; b >= 0; --b) {
            acc = fp_sqr(acc);
            if ((e[w] >> b) & 1u) acc = fp_mul(acc, x);
        }
    }
    return acc;
}

// Multiplicative inverse by Fermat (x^(p-2)); 0 maps to 0.
template <class P>
NKR_HD Fp<P> fp_inv(const Fp<P> &x)
{
    uint32_t e[8];
    uint64_t borrow = 2;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint64_t t = (uint64_t)P::P[i] - borrow;
        e[i] = (uint32_t)t;
        borrow = (t >> 63) & 1;
    }
    return fp_pow(x, e);
}

using fr_t = Fp<FrParams>;
using fq_t = Fp<FqParams>;

// Fq2 = Fq[u]/(u^2 + 1)  (non-residue -1: alt_bn128_init.cpp:95 of the reference's libff).
struct fq2_t {
    fq_t c0, c1;
    NKR_HD static fq2_t zero() { return {fq_t::zero(), fq_t::zero()}; }
    NKR_HD static fq2_t one() { return {fq_t::one(), fq_t::zero()}; }
    NKR_HD bool is_zero() const { return c0.is_zero() && c1.is_zero(); }
    NKR_HD bool operator==(const fq2_t &o) const { return c0 == o.c0 && c1 == o.c1; }
    NKR_HD bool operator!=(const fq2_t &o) const { return !(*this == o); }
};

NKR_HD fq2_t fq2_add(const fq2_t &a, const fq2_t &b) { return {fp_add(a.c0, b.c0), fp_add(a.c1, b.c1)}; }
NKR_HD fq2_t fq2_sub(const fq2_t &a, const fq2_t &b) { return {fp_sub(a.c0, b.c0), fp_sub(a.c1, b.c1)}; }
NKR_HD fq2_t fq2_dbl(const fq2_t &a) { return {fp_dbl(a.c0), fp_dbl(a.c1)}; }
NKR_HD fq2_t fq2_neg(const fq2_t &a) { return {fp_neg(a.c0), fp_neg(a.c1)}; }

// Karatsuba: 3 base-field products.
NKR_HD fq2_t fq2_mul(const fq2_t &a, const fq2_t &b)
{
    fq_t aa, bb;
    fp_mul2(a.c0, b.c0, a.c1, b.c1, aa, bb);
    const fq_t s = fp_mul(fp_add(a.c0, a.c1), fp_add(b.c0, b.c1));
    return {fp_sub(aa, bb), fp_sub(fp_sub(s, aa), bb)};
}

// (a0 + a1 u)^2 = (a0+a1)(a0-a1) + 2 a0 a1 u: 2 base-field products.
NKR_HD fq2_t fq2_sqr(const fq2_t &a)
{
    fq_t p, m;
    fp_mul2(fp_add(a.c0, a.c1), fp_sub(a.c0, a.c1), a.c0, a.c1, p, m);
    return {p, fp_dbl(m)};
}

NKR_HD fq2_t fq2_inv(const fq2_t &a)
{
    // 1/(a0 + a1 u) = (a0 - a1 u) / (a0^2 + a1^2)
    const fq_t n = fp_add(fp_sqr(a.c0), fp_sqr(a.c1));
    const fq_t ni = fp_inv(n);
    return {fp_mul(a.c0, ni), fp_neg(fp_mul(a.c1, ni))};
}

// Field traits so curve code is written once for G1 (Fq) and G2 (Fq2).
template <class F>
struct FieldOps;

template <>
struct FieldOps<fq_t> {
    NKR_HD static fq_t add(const fq_t &a, const fq_t &b) { return fp_add(a, b); }
    NKR_HD static fq_t sub(const fq_t &a, const fq_t &b) { return fp_sub(a, b); }
    NKR_HD static fq_t dbl(const fq_t &a) { return fp_dbl(a); }
    NKR_HD static fq_t neg(const fq_t &a) { return fp_neg(a); }
    NKR_HD static fq_t mul(const fq_t &a, const fq_t &b) { return fp_mul(a, b); }
    NKR_HD static void mul2(const fq_t &a0, const fq_t &b0, const fq_t &a1, const fq_t &b1, fq_t &r0, fq_t &r1)
    {
        fp_mul2(a0, b0, a1, b1, r0, r1);
    }
    // ra = a^2, rb = b^2
    NKR_HD static void sqr2(const fq_t &a, const fq_t &b, fq_t &ra, fq_t &rb) { fp_mul2(a, a, b, b, ra, rb); }
    // ra = a^2, rbc = b*c
    NKR_HD static void sqr_mul(const fq_t &a, const fq_t &b, const fq_t &c, fq_t &ra, fq_t &rbc)
    {
        fp_mul2(a, a, b, c, ra, rbc);
    }
    NKR_HD static fq_t sqr(const fq_t &a) { return fp_sqr(a); }
    NKR_HD static fq_t inv(const fq_t &a) { return fp_inv(a); }
    NKR_HD static fq_t zero() { return fq_t::zero(); }
    NKR_HD static fq_t one() { return fq_t::one(); }
};

template <>
struct FieldOps<fq2_t> {
    NKR_HD static fq2_t add(const fq2_t &a, const fq2_t &b) { return fq2_add(a, b); }
    NKR_HD static fq2_t sub(const fq2_t &a, const fq2_t &b) { return fq2_sub(a, b); }
    NKR_HD static fq2_t dbl(const fq2_t &a) { return fq2_dbl(a); }
    NKR_HD static fq2_t neg(const fq2_t &a) { return fq2_neg(a); }
    NKR_HD static fq2_t mul(const fq2_t &a, const fq2_t &b) { return fq2_mul(a, b); }
    // G2 formulas already hold many live values: the pairs are formed inside
    // each Fq2 product (fq2_mul, fq2_sqr) instead of across two of them.
    NKR_HD static void mul2(const fq2_t &a0, const fq2_t &b0, const fq2_t &a1, const fq2_t &b1, fq2_t &r0,
                            fq2_t &r1)
    {
        const fq2_t x = fq2_mul(a0, b0), y = fq2_mul(a1, b1);
        r0 = x;
        r1 = y;
    }
    NKR_HD static void sqr2(const fq2_t &a, const fq2_t &b, fq2_t &ra, fq2_t &rb)
    {
        const fq2_t x = fq2_sqr(a), y = fq2_sqr(b);
        ra = x;
        rb = y;
    }
    NKR_HD static void sqr_mul(const fq2_t &a, const fq2_t &b, const fq2_t &c, fq2_t &ra, fq2_t &rbc)
    {
        const fq2_t x = fq2_sqr(a), y = fq2_mul(b, c);
        ra = x;
        rbc = y;
    }
    NKR_HD static fq2_t sqr(const fq2_t &a) { return fq2_sqr(a); }
    NKR_HD static fq2_t inv(const fq2_t &a) { return fq2_inv(a); }
    NKR_HD static fq2_t zero() { return fq2_t::zero(); }
    NKR_HD static fq2_t one() { return fq2_t::one(); }
};

}  // namespace nkr

// BN254 prime-field arithmetic for gfx950 (CDNA4), device side.
//
// Elements are 8 little-endian 32-bit limbs holding the Montgomery residue
// x*2^256 mod p, fully reduced (< p).  This is byte-for-byte the layout of the
// reference's host scalar/base-field type (libff Fp_model<4,..>, 4 x u64
// mont_repr, R = 2^256: third_party/lightzk/depends/libsnark/depends/libff/
// libff/algebra/fields/fp.hpp:42), so buffers move between host and device
// without conversion.
//
// Multiplication is Montgomery over 32-bit words.  Both BN254 moduli have a
// top word below 2^30, so partial results stay below 2p < 2^255 and never
// spill into a ninth word.  Each 32x32+64 step maps onto one v_mad_u64_u32
// (full rate on gfx950); the integer product is not a dense contraction, so
// nothing here uses MFMA.
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

// x - p if x >= p (x < 2p).  Borrow chain, then a select.
template <class P>
NKR_HD Fp<P> fp_reduce_once(const Fp<P> &x)
{
    Fp<P> d;
    uint64_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint64_t t = (uint64_t)x.l[i] - P::P[i] - borrow;
        d.l[i] = (uint32_t)t;
        borrow = (t >> 63) & 1;
    }
    Fp<P> r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.l[i] = borrow ? x.l[i] : d.l[i];
    return r;
}

template <class P>
NKR_HD Fp<P> fp_add(const Fp<P> &a, const Fp<P> &b)
{
    Fp<P> s;
    uint64_t carry = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint64_t t = (uint64_t)a.l[i] + b.l[i] + carry;
        s.l[i] = (uint32_t)t;
        carry = t >> 32;
    }
    // a + b < 2p < 2^255: no carry out of the top word.
    return fp_reduce_once(s);
}

template <class P>
NKR_HD Fp<P> fp_sub(const Fp<P> &a, const Fp<P> &b)
{
    Fp<P> d;
    uint64_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint64_t t = (uint64_t)a.l[i] - b.l[i] - borrow;
        d.l[i] = (uint32_t)t;
        borrow = (t >> 63) & 1;
    }
    // On borrow add p back; mask instead of branch keeps the wave converged.
    const uint32_t mask = 0u - (uint32_t)borrow;
    uint64_t carry = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint64_t t = (uint64_t)d.l[i] + (P::P[i] & mask) + carry;
        d.l[i] = (uint32_t)t;
        carry = t >> 32;
    }
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
    uint64_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint64_t t = (uint64_t)P::P[i] - a.l[i] - borrow;
        d.l[i] = (uint32_t)t;
        borrow = (t >> 63) & 1;
    }
    const bool z = a.is_zero();
#pragma unroll
    for (int i = 0; i < 8; ++i) d.l[i] = z ? 0u : d.l[i];
    return d;
}

// Montgomery product a*b*2^-256 mod p, inputs and output < p (word-serial
// CIOS in plain C; the host build of this header uses it).
template <class P>
NKR_HD Fp<P> fp_mul_cios(const Fp<P> &a, const Fp<P> &b)
{
    uint32_t t[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) t[j] = 0;

#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t bi = b.l[i];
        uint64_t s = (uint64_t)a.l[0] * bi + t[0];
        uint32_t A = (uint32_t)(s >> 32);
        const uint32_t t0 = (uint32_t)s;
        const uint32_t m = t0 * P::INV;
        uint64_t u = (uint64_t)m * P::P[0] + t0;
        uint32_t C = (uint32_t)(u >> 32);
#pragma unroll
        for (int j = 1; j < 8; ++j) {
            s = (uint64_t)a.l[j] * bi + t[j] + A;
            A = (uint32_t)(s >> 32);
            u = (uint64_t)m * P::P[j] + (uint32_t)s + C;
            C = (uint32_t)(u >> 32);
            t[j - 1] = (uint32_t)u;
        }
        t[7] = C + A;
    }
    Fp<P> r;
#pragma unroll
    for (int j = 0; j < 8; ++j) r.l[j] = t[j];
    return fp_reduce_once(r);
}

#if defined(__HIP_DEVICE_COMPILE__)
// acc(96 bit) += x*y.  One v_mad_u64_u32 into the low 64 bits; its carry-out
// (an SGPR lane mask) feeds v_addc_co_u32 on the top word.
NKR_DEV void mac96(uint64_t &acc, uint32_t &top, uint32_t x, uint32_t y)
{
    uint64_t cc;
    asm("v_mad_u64_u32 %0, %2, %3, %4, %0\n\t"
        "v_addc_co_u32 %1, %2, %1, 0, %2"
        : "+v"(acc), "+v"(top), "=&s"(cc)
        : "v"(x), "v"(y));
}

// Montgomery product by product scanning (column-wise, "FIPS" order): column
// k accumulates a_j*b_{k-j} and m_j*p_{k-j} into a 96-bit accumulator, so each
// partial product costs a v_mad_u64_u32 plus one carry add and no moves.
//
// With kReduce = false the final subtraction is skipped: for a*b < 4p^2
// (say a < 4p and b < p, or both below 2p) and 4p < 2^256 the result is
// below 2p, which lazily reduced transforms keep as their invariant.
template <class P, bool kReduce = true>
NKR_DEV Fp<P> fp_mul_fips(const Fp<P> &a, const Fp<P> &b)
{
    uint32_t m[8];
    uint32_t t[8];
    uint64_t acc = 0;
    uint32_t top = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
#pragma unroll
        for (int j = 0; j < i; ++j) {
            mac96(acc, top, a.l[j], b.l[i - j]);
            mac96(acc, top, m[j], P::P[i - j]);
        }
        mac96(acc, top, a.l[i], b.l[0]);
        m[i] = (uint32_t)acc * P::INV;
        mac96(acc, top, m[i], P::P[0]);
        acc = (acc >> 32) | ((uint64_t)top << 32);
        top = 0;
    }
#pragma unroll
    for (int i = 8; i < 15; ++i) {
#pragma unroll
        for (int j = i - 7; j < 8; ++j) {
            mac96(acc, top, a.l[j], b.l[i - j]);
            mac96(acc, top, m[j], P::P[i - j]);
        }
        t[i - 8] = (uint32_t)acc;
        acc = (acc >> 32) | ((uint64_t)top << 32);
        top = 0;
    }
    // (a*b + m*p) / 2^256 < 2p < 2^255, so the last column fits one word.
    t[7] = (uint32_t)acc;
    Fp<P> r;
#pragma unroll
    for (int j = 0; j < 8; ++j) r.l[j] = t[j];
    if constexpr (kReduce)
        return fp_reduce_once(r);
    else
        return r;
}

// Two independent mac96 steps in one block: each v_mad_u64_u32 has the
// other chain's instruction beside it while its result is pending, so one
// wave keeps issuing where a single chain would wait on the multiplier.
NKR_DEV void mac96x2(uint64_t &acc0, uint32_t &top0, uint32_t x0, uint32_t y0, uint64_t &acc1, uint32_t &top1,
                     uint32_t x1, uint32_t y1)
{
    uint64_t c0, c1;
    asm("v_mad_u64_u32 %0, %4, %6, %7, %0\n\t"
        "v_mad_u64_u32 %2, %5, %8, %9, %2\n\t"
        "v_addc_co_u32 %1, %4, %1, 0, %4\n\t"
        "v_addc_co_u32 %3, %5, %3, 0, %5"
        : "+v"(acc0), "+v"(top0), "+v"(acc1), "+v"(top1), "=&s"(c0), "=&s"(c1)
        : "v"(x0), "v"(y0), "v"(x1), "v"(y1));
}

// Two independent Montgomery products (r0 = a0*b0, r1 = a1*b1) with their
// column steps interleaved: the same instructions as two fp_mul_fips calls,
// in half the dependent-chain latency.  Outputs may alias inputs.
template <class P>
NKR_DEV void fp_mul_fips2(const Fp<P> &a0, const Fp<P> &b0, const Fp<P> &a1, const Fp<P> &b1, Fp<P> &r0,
                          Fp<P> &r1)
{
    uint32_t m0[8], m1[8], t0[8], t1[8];
    uint64_t acc0 = 0, acc1 = 0;
    uint32_t top0 = 0, top1 = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
#pragma unroll
        for (int j = 0; j < i; ++j) {
            mac96x2(acc0, top0, a0.l[j], b0.l[i - j], acc1, top1, a1.l[j], b1.l[i - j]);
            mac96x2(acc0, top0, m0[j], P::P[i - j], acc1, top1, m1[j], P::P[i - j]);
        }
        mac96x2(acc0, top0, a0.l[i], b0.l[0], acc1, top1, a1.l[i], b1.l[0]);
        m0[i] = (uint32_t)acc0 * P::INV;
        m1[i] = (uint32_t)acc1 * P::INV;
        mac96x2(acc0, top0, m0[i], P::P[0], acc1, top1, m1[i], P::P[0]);
        acc0 = (acc0 >> 32) | ((uint64_t)top0 << 32);
        acc1 = (acc1 >> 32) | ((uint64_t)top1 << 32);
        top0 = top1 = 0;
    }
#pragma unroll
    for (int i = 8; i < 15; ++i) {
#pragma unroll
        for (int j = i - 7; j < 8; ++j) {
            mac96x2(acc0, top0, a0.l[j], b0.l[i - j], acc1, top1, a1.l[j], b1.l[i - j]);
            mac96x2(acc0, top0, m0[j], P::P[i - j], acc1, top1, m1[j], P::P[i - j]);
        }
        t0[i - 8] = (uint32_t)acc0;
        t1[i - 8] = (uint32_t)acc1;
        acc0 = (acc0 >> 32) | ((uint64_t)top0 << 32);
        acc1 = (acc1 >> 32) | ((uint64_t)top1 << 32);
        top0 = top1 = 0;
    }
    t0[7] = (uint32_t)acc0;
    t1[7] = (uint32_t)acc1;
    Fp<P> x0, x1;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        x0.l[j] = t0[j];
        x1.l[j] = t1[j];
    }
    r0 = fp_reduce_once(x0);
    r1 = fp_reduce_once(x1);
}
#endif

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

// Montgomery product a*b*2^-256 mod p.  On gfx950 the product-scanning form
// issues ~300 VALU instructions against ~880 for the C form (measured 129 vs
// 100 G products/s on MI355X); the host build uses 64-bit words.
template <class P>
NKR_HD Fp<P> fp_mul(const Fp<P> &a, const Fp<P> &b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return fp_mul_fips(a, b);
#else
    return fp_mul_host64(a, b);
#endif
}

// r0 = a0*b0 and r1 = a1*b1 (independent products; outputs may alias inputs).
template <class P>
NKR_HD void fp_mul2(const Fp<P> &a0, const Fp<P> &b0, const Fp<P> &a1, const Fp<P> &b1, Fp<P> &r0, Fp<P> &r1)
{
#if defined(__HIP_DEVICE_COMPILE__)
    fp_mul_fips2(a0, b0, a1, b1, r0, r1);
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

// Lazily reduced arithmetic on the reduced-radix form of field.h (Fp29: 9
// limbs of 29 bits, Montgomery R = 2^261), for loops that keep values in that
// form across many operations (the MSM bucket accumulation).  Sums and
// differences are carry-normalised limb sums without reduction; values may
// exceed p by a small multiple, as each operation states, and the callers'
// bounds keep every product input inside what fp29_mul tolerates.
//
// Conversions: a canonical element x*2^256 becomes x*2^261 by a 5-bit shift
// (fp29_from_mont: a value below 32p, no reduction); back, v*2^-5 mod p is
// one 5-bit Montgomery step and one conditional subtraction
// (fp29_to_mont).  Results are exact field elements, so kernels that use this
// form produce bit-identical canonical outputs.
#pragma once

#include "field.h"

namespace nkr {

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

// Hides x's provenance from the optimiser without changing it (an empty
// asm statement per limb, no instruction).  Where a product's operand comes
// from code the compiler also sees feeding 64-bit arithmetic (the exact zero
// test of the addition's exceptional case), it may carry the limbs as 64-bit
// values and square or multiply them as such: on gfx950 that cost the G1
// bucket addition 135 extra multiply-adds and 260 register moves.
template <class P>
NKR_HD void fp29_opaque(Fp29<P> &x)
{
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
    for (int i = 0; i < 9; ++i) asm volatile("" : "+v"(x.l[i]));
#else
    (void)x;
#endif
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

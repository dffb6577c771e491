// Mixed XYZZ addition over the reduced-radix field (field29.h), for the MSM
// bucket accumulation: acc + (x2, y2) with the accumulator kept in 29-bit
// limbs across a whole piece of a bucket.
//
// Bounds (p the modulus; every limb below 2^29):
//   accumulator  X < 8p, Y < 8p, ZZ < 2p, ZZZ < 2p; infinity is ZZ = 0 exactly
//   base         x2, y2 < 32p (fp29_from_mont of canonical coordinates)
// madd-2008-s (a = 0), as xyzz_add_affine in curve.h:
//   U2 = x2 ZZ, S2 = y2 ZZZ                   (32p x 2p inputs)    < 2p
//   P = U2 + 8p - X, R = S2 + 8p - Y                               < 10p
//   PP = P^2, RR = R^2, PPP = P PP, Q = X PP  (at most 10p x 10p)  < 2p
//   X3 = (RR + 2p - PPP) + 4p - 2Q                                 < 8p
//   Y3 = R (Q + 8p - X3) + Y (2p - PPP), one Montgomery reduction
//        for both products (10p x 10p + 8p x 2p = 116 p^2)         < 2p
//   ZZ3 = ZZ PP, ZZZ3 = ZZZ PPP                                    < 2p
// so the invariant holds without a single reduction.  PP and RR are
// squarings (fp29_sqr: the symmetric partial products once).  A product's inputs stay
// below a*b < 2^261 p (fp29_mul).  The exceptional cases (P = 0: the base
// equals the accumulator, or its negation) take the canonical 32-bit path.
// Every result is the same field element as xyzz_add_affine's, so the
// canonical outputs are bit-identical.
#pragma once

#include "curve.h"
#include "field29.h"

namespace nkr {

template <class P>
struct Xyzz29 {
    Fp29<P> X, Y, ZZ, ZZZ;

    NKR_HD static Xyzz29 infinity()
    {
        Xyzz29 r;
        r.X = r.Y = r.ZZ = r.ZZZ = Fp29<P>::zero();
        return r;
    }
    NKR_HD bool is_inf() const { return ZZ.is_exact_zero(); }
};

// Canonical XYZZ (32-bit limbs) -> 29-bit form within the invariant.
template <class P>
NKR_HD Xyzz29<P> xyzz29_from(const Xyzz<Fp<P>> &p)
{
    if (p.ZZ.is_zero()) return Xyzz29<P>::infinity();
    Xyzz29<P> r;
    r.X = fp29_reduce_to_8p(fp29_from_mont(p.X));
    r.Y = fp29_reduce_to_8p(fp29_from_mont(p.Y));
    r.ZZ = fp29_reduce_to_2p(fp29_from_mont(p.ZZ));
    r.ZZZ = fp29_reduce_to_2p(fp29_from_mont(p.ZZZ));
    return r;
}

// 29-bit form -> canonical XYZZ (infinity: all coordinates zero).
template <class P>
NKR_HD Xyzz<Fp<P>> xyzz29_to(const Xyzz29<P> &p)
{
    if (p.is_inf()) return Xyzz<Fp<P>>::infinity();
    Xyzz<Fp<P>> r;
    r.X = fp29_to_mont(p.X);
    r.Y = fp29_to_mont(p.Y);
    r.ZZ = fp29_to_mont(p.ZZ);
    r.ZZZ = fp29_to_mont(p.ZZZ);
    return r;
}

// acc + (x2, y2), complete.  x2, y2: the affine base (29-bit, below 32p).
template <class P>
NKR_HD Xyzz29<P> xyzz29_add_affine(const Xyzz29<P> &acc, const Fp29<P> &x2, const Fp29<P> &y2)
{
    if (acc.is_inf()) {
        Xyzz29<P> r;
        r.X = fp29_reduce_to_8p(x2);
        r.Y = fp29_reduce_to_8p(y2);
        r.ZZ = Fp29<P>::one();
        r.ZZZ = Fp29<P>::one();
        return r;
    }
    const Fp29<P> U2 = fp29_mul(x2, acc.ZZ);
    const Fp29<P> S2 = fp29_mul(y2, acc.ZZZ);
    Fp29<P> Pd = fp29_sub<P, 8>(U2, acc.X);
    Fp29<P> R = fp29_sub<P, 8>(S2, acc.Y);
    if (fp29_may_be_zero<P, 10>(Pd) && fp29_to_mont(Pd).is_zero()) {
        // the base is +-(the accumulator): the canonical formulas decide
        // (doubling or infinity)
        if (fp29_to_mont(R).is_zero())
            return xyzz29_from<P>(xyzz_dbl_affine(fp29_to_mont(x2), fp29_to_mont(y2)));
        return Xyzz29<P>::infinity();
    }
    // One product at a time, each operand's last use as early as possible:
    // the accumulation runs at three waves per SIMD only while the working
    // set stays near 150 registers (pairing the products, or holding the next
    // base in registers, measured slower on MI355X).
    fp29_opaque(Pd);
    fp29_opaque(R);
    const Fp29<P> PP = fp29_sqr(Pd);
    const Fp29<P> PPP = fp29_mul(Pd, PP);
    const Fp29<P> RR = fp29_sqr(R);
    const Fp29<P> Q = fp29_mul(acc.X, PP);
    Xyzz29<P> r;
    r.ZZ = fp29_mul(acc.ZZ, PP);
    r.X = fp29_sub<P, 4>(fp29_sub<P, 2>(RR, PPP), fp29_add(Q, Q));
    const Fp29<P> T = fp29_sub<P, 8>(Q, r.X);
    const Fp29<P> N = fp29_sub<P, 2>(Fp29<P>::zero(), PPP);
    r.ZZZ = fp29_mul(acc.ZZZ, PPP);
    r.Y = fp29_mul_add(R, T, acc.Y, N);
    return r;
}

// a + b, complete (add-2008-s), both in the accumulator invariant (X, Y
// below 8p; ZZ, ZZZ below 2p; infinity ZZ = 0 exactly), for the bucket
// reduction (fold, completion, bit-plane tree), whose partial sums stay in
// this form from the accumulation to the bit planes:
//   U1 = X1 ZZ2, U2 = X2 ZZ1, S1 = Y1 ZZZ2, S2 = Y2 ZZZ1  (8p x 2p)  < 2p
//   P = U2 + 2p - U1, R = S2 + 2p - S1                           < 4p
//   PP = P^2, RR = R^2, PPP = P PP, Q = U1 PP                     < 2p
//   X3 = (RR + 2p - PPP) + 4p - 2Q                                < 8p
//   Y3 = R (Q + 8p - X3) + S1 (2p - PPP), one reduction
//        (4p x 10p + 2p x 2p = 44 p^2)                            < 2p
//   ZZ3 = (ZZ1 ZZ2) PP, ZZZ3 = (ZZZ1 ZZZ2) PPP                    < 2p
// The exceptional case P = 0 (b = +-a) takes the canonical doubling or
// gives infinity.
template <class P>
NKR_HD Xyzz29<P> xyzz29_add(const Xyzz29<P> &a, const Xyzz29<P> &b)
{
    if (a.is_inf()) return b;
    if (b.is_inf()) return a;
    const Fp29<P> U1 = fp29_mul(a.X, b.ZZ);
    const Fp29<P> U2 = fp29_mul(b.X, a.ZZ);
    const Fp29<P> S1 = fp29_mul(a.Y, b.ZZZ);
    const Fp29<P> S2 = fp29_mul(b.Y, a.ZZZ);
    Fp29<P> Pd = fp29_sub<P, 2>(U2, U1);
    Fp29<P> R = fp29_sub<P, 2>(S2, S1);
    if (fp29_may_be_zero<P, 4>(Pd) && fp29_to_mont(Pd).is_zero()) {
        if (fp29_to_mont(R).is_zero()) return xyzz29_from<P>(xyzz_dbl(xyzz29_to(a)));
        return Xyzz29<P>::infinity();
    }
    fp29_opaque(Pd);
    fp29_opaque(R);
    const Fp29<P> PP = fp29_sqr(Pd);
    const Fp29<P> PPP = fp29_mul(Pd, PP);
    const Fp29<P> Q = fp29_mul(U1, PP);
    const Fp29<P> RR = fp29_sqr(R);
    Xyzz29<P> r;
    r.ZZ = fp29_mul(fp29_mul(a.ZZ, b.ZZ), PP);
    r.ZZZ = fp29_mul(fp29_mul(a.ZZZ, b.ZZZ), PPP);
    r.X = fp29_sub<P, 4>(fp29_sub<P, 2>(RR, PPP), fp29_add(Q, Q));
    const Fp29<P> T = fp29_sub<P, 8>(Q, r.X);
    const Fp29<P> N = fp29_sub<P, 2>(Fp29<P>::zero(), PPP);
    r.Y = fp29_mul_add(R, T, S1, N);
    return r;
}

// ---------------------------------------------------------------- G2 (Fq2)
// Fq2 = Fq[u]/(u^2 + 1) over the 29-bit form.  Products are schoolbook with
// each output component one Montgomery reduction of a sum of two (or, for
// Y3, four) partial products (fp29_mul_add): (a b)_0 = a0 b0 + a1 (Kp - b1),
// (a b)_1 = a0 b1 + a1 b0 -- the multiply-adds of Karatsuba's three products
// without its five additions of sums, and with input bounds up to A*B = 84
// (components below A p and B p; 2AB + 2AK below 2^261/p^2 = 169.28 p^2,
// K the multiple of p that negates b1), outputs below 2p.  Squares use
// (a0 + a1)(a0 - a1) and a0 (2 a1), inputs below 6.5p.
//   accumulator  X below 8p, Y, ZZ, ZZZ below 2p; infinity is ZZ = 0 exactly
//   base         x2, y2 below 32p (fp29_from_mont, no reduction)
// madd-2008-s, the sequence of xyzz29_add_affine:
//   U2 = x2 ZZ, S2 = y2 ZZZ (32p x 2p: 2 * 64 = 128)         below 2p
//   P = U2 + 8p - X, R = S2 + 2p - Y, each reduced by 4p      below 6p
//   PP = P^2, RR = R^2 ((12p)^2 = 144: c0 < 1.86p; c1 = P0 (2 P1)
//        72: < 1.43p)                                         below 2p
//   PPP = P PP, Q = X PP, ZZ3 = ZZ PP (at most 8p x 2p)       below 2p
//   X3 = (RR + 2p - PPP) + 4p - 2Q                            below 8p
//   Y3 = R (Q + 8p - X3) + Y (2p - PPP): per component one
//        reduction of four partial products (6p x 10p twice +
//        2p x 2p twice = 128)                                 below 2p
//   ZZZ3 = ZZZ PPP                                            below 2p
template <class P>
struct Fq2_29 {
    Fp29<P> c0, c1;
    NKR_HD bool is_exact_zero() const { return c0.is_exact_zero() && c1.is_exact_zero(); }
};

// canonical -> 29-bit components below 32p (no reduction)
template <class P>
NKR_HD Fq2_29<P> fq2_29_from(const fq2_t &a)
{
    return {fp29_from_mont(a.c0), fp29_from_mont(a.c1)};
}

template <class P>
NKR_HD fq2_t fq2_29_to(const Fq2_29<P> &a)
{
    return {fp29_to_mont(a.c0), fp29_to_mont(a.c1)};
}

// a * b; components below A p and B p with A*B <= 84 and b1 <= KB p,
// A*(B + KB) <= 169: components below 2p.
template <class P, uint32_t KB>
NKR_HD Fq2_29<P> fq2_29_mul(const Fq2_29<P> &a, const Fq2_29<P> &b)
{
    const Fp29<P> nb1 = fp29_sub<P, KB>(Fp29<P>::zero(), b.c1);
    return {fp29_mul_add(a.c0, b.c0, a.c1, nb1), fp29_mul_add(a.c0, b.c1, a.c1, b.c0)};
}

// a^2, components below 6.5p: c0 below 1.86p, c1 below 1.43p.
template <class P, uint32_t KA>
NKR_HD Fq2_29<P> fq2_29_sqr(const Fq2_29<P> &a)
{
    return {fp29_mul(fp29_add(a.c0, a.c1), fp29_sub<P, KA>(a.c0, a.c1)), fp29_mul(a.c0, fp29_add(a.c1, a.c1))};
}

template <class P, uint32_t K>
NKR_HD Fq2_29<P> fq2_29_sub(const Fq2_29<P> &a, const Fq2_29<P> &b)
{
    return {fp29_sub<P, K>(a.c0, b.c0), fp29_sub<P, K>(a.c1, b.c1)};
}

template <class P, uint32_t K>
NKR_HD Fq2_29<P> fq2_29_sub_if_ge(const Fq2_29<P> &a)
{
    return {fp29_sub_if_ge<P, K>(a.c0), fp29_sub_if_ge<P, K>(a.c1)};
}

// a below 32p -> below 2p, component-wise
template <class P>
NKR_HD Fq2_29<P> fq2_29_reduce_to_2p(const Fq2_29<P> &a)
{
    return {fp29_reduce_to_2p(a.c0), fp29_reduce_to_2p(a.c1)};
}

template <class P>
NKR_HD bool fq2_29_is_zero_below_6p(const Fq2_29<P> &a)
{
    if (!(fp29_may_be_zero<P, 6>(a.c0) && fp29_may_be_zero<P, 6>(a.c1))) return false;
    return fp29_to_mont(a.c0).is_zero() && fp29_to_mont(a.c1).is_zero();
}

template <class P>
struct Xyzz2_29 {
    Fq2_29<P> X, Y, ZZ, ZZZ;
    NKR_HD static Xyzz2_29 infinity()
    {
        Xyzz2_29 r;
        r.X.c0 = r.X.c1 = r.Y.c0 = r.Y.c1 = Fp29<P>::zero();
        r.ZZ.c0 = r.ZZ.c1 = r.ZZZ.c0 = r.ZZZ.c1 = Fp29<P>::zero();
        return r;
    }
    NKR_HD bool is_inf() const { return ZZ.is_exact_zero(); }
};

// Canonical XYZZ over Fq2 -> the accumulator's form.
template <class P>
NKR_HD Xyzz2_29<P> xyzz2_29_from(const Xyzz<fq2_t> &p)
{
    if (p.ZZ.is_zero()) return Xyzz2_29<P>::infinity();
    Xyzz2_29<P> r;
    r.X = fq2_29_reduce_to_2p<P>(fq2_29_from<P>(p.X));
    r.Y = fq2_29_reduce_to_2p<P>(fq2_29_from<P>(p.Y));
    r.ZZ = fq2_29_reduce_to_2p<P>(fq2_29_from<P>(p.ZZ));
    r.ZZZ = fq2_29_reduce_to_2p<P>(fq2_29_from<P>(p.ZZZ));
    return r;
}

template <class P>
NKR_HD Xyzz<fq2_t> xyzz2_29_to(const Xyzz2_29<P> &p)
{
    if (p.is_inf()) return Xyzz<fq2_t>::infinity();
    Xyzz<fq2_t> r;
    r.X = fq2_29_to(p.X);
    r.Y = fq2_29_to(p.Y);
    r.ZZ = fq2_29_to(p.ZZ);
    r.ZZZ = fq2_29_to(p.ZZZ);
    return r;
}

// a + b over G2, complete (add-2008-s), both in the accumulator invariant,
// for the bucket reduction (the G2 partial sums stay in this form):
//   U1 = X1 ZZ2, U2 = X2 ZZ1 (8p x 2p), S1 = Y1 ZZZ2, S2 = Y2 ZZZ1  below 2p
//   P = U2 + 2p - U1, R = S2 + 2p - S1                          below 4p
//   PP = P^2, RR = R^2, PPP = P PP, Q = U1 PP                    below 2p
//   X3 = (RR + 2p - PPP) + 4p - 2Q                               below 8p
//   Y3 = R (Q + 8p - X3) + S1 (2p - PPP): per component one
//        reduction of four partial products (4p x 10p twice +
//        2p x 2p twice = 88)                                     below 2p
//   ZZ3 = (ZZ1 ZZ2) PP, ZZZ3 = (ZZZ1 ZZZ2) PPP                   below 2p
// P = 0 (b = +-a) takes the canonical doubling or gives infinity.
template <class P>
NKR_HD Xyzz2_29<P> xyzz2_29_add(const Xyzz2_29<P> &a, const Xyzz2_29<P> &b)
{
    if (a.is_inf()) return b;
    if (b.is_inf()) return a;
    const Fq2_29<P> U1 = fq2_29_mul<P, 2>(a.X, b.ZZ);
    const Fq2_29<P> U2 = fq2_29_mul<P, 2>(b.X, a.ZZ);
    const Fq2_29<P> S1 = fq2_29_mul<P, 2>(a.Y, b.ZZZ);
    const Fq2_29<P> S2 = fq2_29_mul<P, 2>(b.Y, a.ZZZ);
    Fq2_29<P> Pd = fq2_29_sub<P, 2>(U2, U1);
    Fq2_29<P> R = fq2_29_sub<P, 2>(S2, S1);
    if (fq2_29_is_zero_below_6p(Pd)) {
        if (fq2_29_is_zero_below_6p(R)) return xyzz2_29_from<P>(xyzz_dbl(xyzz2_29_to(a)));
        return Xyzz2_29<P>::infinity();
    }
    fp29_opaque(Pd.c0);
    fp29_opaque(Pd.c1);
    fp29_opaque(R.c0);
    fp29_opaque(R.c1);
    const Fq2_29<P> PP = fq2_29_sqr<P, 4>(Pd);
    const Fq2_29<P> PPP = fq2_29_mul<P, 2>(Pd, PP);
    const Fq2_29<P> Q = fq2_29_mul<P, 2>(U1, PP);
    const Fq2_29<P> RR = fq2_29_sqr<P, 4>(R);
    Xyzz2_29<P> r;
    r.ZZ = fq2_29_mul<P, 2>(fq2_29_mul<P, 2>(a.ZZ, b.ZZ), PP);
    r.ZZZ = fq2_29_mul<P, 2>(fq2_29_mul<P, 2>(a.ZZZ, b.ZZZ), PPP);
    const Fq2_29<P> Q2 = {fp29_add(Q.c0, Q.c0), fp29_add(Q.c1, Q.c1)};
    r.X = fq2_29_sub<P, 4>(fq2_29_sub<P, 2>(RR, PPP), Q2);
    const Fq2_29<P> T = fq2_29_sub<P, 8>(Q, r.X);
    const Fq2_29<P> N = fq2_29_sub<P, 2>({Fp29<P>::zero(), Fp29<P>::zero()}, PPP);
    const Fp29<P> nT1 = fp29_sub<P, 10>(Fp29<P>::zero(), T.c1);
    const Fp29<P> nN1 = fp29_sub<P, 2>(Fp29<P>::zero(), N.c1);
    r.Y.c0 = fp29_mul_add4(R.c0, T.c0, R.c1, nT1, S1.c0, N.c0, S1.c1, nN1);
    r.Y.c1 = fp29_mul_add4(R.c0, T.c1, R.c1, T.c0, S1.c0, N.c1, S1.c1, N.c0);
    return r;
}

// The G2 accumulator as the mixed addition below reads and writes it: X(),
// Y(), ZZ(), ZZZ() read a coordinate, set() replaces all four at once.  The
// formula is written once over this interface, so a caller may keep some
// coordinates outside the registers (the accumulation kernel holds Y and
// ZZZ in LDS between their uses) with the same arithmetic.
template <class P>
struct Xyzz2_29Regs {
    Xyzz2_29<P> p;
    NKR_HD bool is_inf() const { return p.is_inf(); }
    NKR_HD Fq2_29<P> X() const { return p.X; }
    NKR_HD Fq2_29<P> Y() const { return p.Y; }
    NKR_HD Fq2_29<P> ZZ() const { return p.ZZ; }
    NKR_HD Fq2_29<P> ZZZ() const { return p.ZZZ; }
    NKR_HD void set(const Fq2_29<P> &X, const Fq2_29<P> &Y, const Fq2_29<P> &ZZ, const Fq2_29<P> &ZZZ)
    {
        p.X = X;
        p.Y = Y;
        p.ZZ = ZZ;
        p.ZZZ = ZZZ;
    }
    NKR_HD void set(const Xyzz2_29<P> &q) { p = q; }
};

// acc += (x2, y2) over G2, complete.  x2, y2: the affine base from
// fq2_29_from (below 32p).  Each coordinate of acc is read where the
// formula uses it (Y and ZZZ twice each) and all four are replaced at the end.
template <class P, class Acc>
NKR_HD void xyzz2_29_madd(Acc &acc, const Fq2_29<P> &x2, const Fq2_29<P> &y2)
{
    if (acc.is_inf()) {
        const Fq2_29<P> one = {Fp29<P>::one(), Fp29<P>::zero()};
        acc.set(fq2_29_reduce_to_2p(x2), fq2_29_reduce_to_2p(y2), one, one);
        return;
    }
    const Fq2_29<P> U2 = fq2_29_mul<P, 2>(x2, acc.ZZ());
    const Fq2_29<P> S2 = fq2_29_mul<P, 2>(y2, acc.ZZZ());
    Fq2_29<P> Pd = fq2_29_sub_if_ge<P, 4>(fq2_29_sub<P, 8>(U2, acc.X()));
    Fq2_29<P> R = fq2_29_sub_if_ge<P, 4>(fq2_29_sub<P, 2>(S2, acc.Y()));
    if (fq2_29_is_zero_below_6p(Pd)) {
        // the base is +-(the accumulator): the canonical formulas decide
        if (fq2_29_is_zero_below_6p(R))
            acc.set(xyzz2_29_from<P>(xyzz_dbl_affine(fq2_29_to(x2), fq2_29_to(y2))));
        else
            acc.set(Xyzz2_29<P>::infinity());
        return;
    }
    fp29_opaque(Pd.c0);
    fp29_opaque(Pd.c1);
    fp29_opaque(R.c0);
    fp29_opaque(R.c1);
    const Fq2_29<P> PP = fq2_29_sqr<P, 6>(Pd);
    const Fq2_29<P> PPP = fq2_29_mul<P, 2>(Pd, PP);
    const Fq2_29<P> RR = fq2_29_sqr<P, 6>(R);
    const Fq2_29<P> Q = fq2_29_mul<P, 2>(acc.X(), PP);
    const Fq2_29<P> ZZ3 = fq2_29_mul<P, 2>(acc.ZZ(), PP);
    const Fq2_29<P> Q2 = {fp29_add(Q.c0, Q.c0), fp29_add(Q.c1, Q.c1)};
    const Fq2_29<P> X3 = fq2_29_sub<P, 4>(fq2_29_sub<P, 2>(RR, PPP), Q2);
    const Fq2_29<P> T = fq2_29_sub<P, 8>(Q, X3);
    const Fq2_29<P> N = fq2_29_sub<P, 2>({Fp29<P>::zero(), Fp29<P>::zero()}, PPP);
    const Fq2_29<P> ZZZ3 = fq2_29_mul<P, 2>(acc.ZZZ(), PPP);
    // Y3 = R T + Y N, per component one reduction of four partial products
    const Fp29<P> nT1 = fp29_sub<P, 10>(Fp29<P>::zero(), T.c1);
    const Fp29<P> nN1 = fp29_sub<P, 2>(Fp29<P>::zero(), N.c1);
    const Fq2_29<P> Y1 = acc.Y();
    Fq2_29<P> Y3;
    Y3.c0 = fp29_mul_add4(R.c0, T.c0, R.c1, nT1, Y1.c0, N.c0, Y1.c1, nN1);
    Y3.c1 = fp29_mul_add4(R.c0, T.c1, R.c1, T.c0, Y1.c0, N.c1, Y1.c1, N.c0);
    acc.set(X3, Y3, ZZ3, ZZZ3);
}

// acc + (x2, y2) with the accumulator in registers.
template <class P>
NKR_HD Xyzz2_29<P> xyzz2_29_add_affine(const Xyzz2_29<P> &acc, const Fq2_29<P> &x2, const Fq2_29<P> &y2)
{
    Xyzz2_29Regs<P> a{acc};
    xyzz2_29_madd<P>(a, x2, y2);
    return a.p;
}

}  // namespace nkr

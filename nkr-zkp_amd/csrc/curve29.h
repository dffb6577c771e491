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
//   Y3 = R (Q + 8p - X3) + 2p - Y PPP                              < 4p
//   ZZ3 = ZZ PP, ZZZ3 = ZZZ PPP                                    < 2p
// so the invariant holds without a single reduction.  A product's inputs stay
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
    const Fp29<P> Pd = fp29_sub<P, 8>(U2, acc.X);
    const Fp29<P> R = fp29_sub<P, 8>(S2, acc.Y);
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
    const Fp29<P> PP = fp29_mul(Pd, Pd);
    const Fp29<P> PPP = fp29_mul(Pd, PP);
    const Fp29<P> RR = fp29_mul(R, R);
    const Fp29<P> Q = fp29_mul(acc.X, PP);
    Xyzz29<P> r;
    r.ZZ = fp29_mul(acc.ZZ, PP);
    r.X = fp29_sub<P, 4>(fp29_sub<P, 2>(RR, PPP), fp29_add(Q, Q));
    const Fp29<P> t1 = fp29_mul(R, fp29_sub<P, 8>(Q, r.X));
    const Fp29<P> t2 = fp29_mul(acc.Y, PPP);
    r.ZZZ = fp29_mul(acc.ZZZ, PPP);
    r.Y = fp29_sub<P, 2>(t1, t2);
    return r;
}

}  // namespace nkr

// BN254 G1 (over Fq) and G2 (over Fq2) point arithmetic for gfx950.
//
// Buckets and partial sums use extended Jacobian "XYZZ" coordinates
// (x = X/ZZ, y = Y/ZZZ, ZZ^3 = ZZZ^2); the point at infinity is ZZ = 0.
// Bases are affine (x, y) with an explicit infinity flag kept beside them.
//
// Every addition here is complete for the inputs a Pippenger bucket can see:
// P + P falls through to a doubling and P + (-P) to infinity.  (The
// reference's HIP bucket kernel uses an addition that returns infinity for
// P + P, operators/hip/include/curves/xyzz_t.hip:436-474; duplicate proving
// key bases make that reachable.)
//
// Formulas: madd-2008-s, add-2008-s, dbl-2008-s-1 and mdbl-2008-s-1 from the
// Explicit-Formulas Database, specialised to a = 0.
#pragma once

#include "field.h"

namespace nkr {

template <class F>
struct Affine {
    F x, y;
};

// The word that carries the infinity flag of an affine base: bit 31 of the
// top limb of y (field elements are < 2^254, so the bit is otherwise clear).
NKR_HD uint32_t &inf_word(fq_t &y) { return y.l[7]; }
NKR_HD uint32_t inf_word(const fq_t &y) { return y.l[7]; }
NKR_HD uint32_t &inf_word(fq2_t &y) { return y.c1.l[7]; }
NKR_HD uint32_t inf_word(const fq2_t &y) { return y.c1.l[7]; }

template <class F>
NKR_HD Affine<F> affine_infinity()
{
    Affine<F> a;
    a.x = FieldOps<F>::zero();
    a.y = FieldOps<F>::zero();
    inf_word(a.y) = 0x80000000u;
    return a;
}

template <class F>
NKR_HD bool affine_is_inf(const Affine<F> &a)
{
    return (inf_word(a.y) >> 31) != 0;
}

template <class F>
struct Xyzz {
    F X, Y, ZZ, ZZZ;

    NKR_HD static Xyzz infinity()
    {
        Xyzz r;
        r.X = FieldOps<F>::zero();
        r.Y = FieldOps<F>::zero();
        r.ZZ = FieldOps<F>::zero();
        r.ZZZ = FieldOps<F>::zero();
        return r;
    }
    NKR_HD bool is_inf() const { return ZZ.is_zero(); }
    NKR_HD static Xyzz from_affine(const Affine<F> &p)
    {
        Xyzz r;
        r.X = p.x;
        r.Y = p.y;
        r.ZZ = FieldOps<F>::one();
        r.ZZZ = FieldOps<F>::one();
        return r;
    }
};

// The formulas below take their products two at a time (FieldOps::mul2 /
// sqr2 / sqr_mul): independent products issued together halve the length
// of the dependent chain a single wave waits on; the arithmetic, and so
// every result, is that of the one-at-a-time formulas.

// 2*(x, y) for an affine point (mdbl-2008-s-1).
template <class F>
NKR_HD Xyzz<F> xyzz_dbl_affine(const F &x, const F &y)
{
    using O = FieldOps<F>;
    const F U = O::dbl(y);
    F V, x2, W, S, MM;
    O::sqr2(U, x, V, x2);
    O::mul2(U, V, x, V, W, S);
    const F M = O::add(O::dbl(x2), x2);
    MM = O::sqr(M);
    Xyzz<F> r;
    r.X = O::sub(MM, O::dbl(S));
    F t1, t2;
    O::mul2(M, O::sub(S, r.X), W, y, t1, t2);
    r.Y = O::sub(t1, t2);
    r.ZZ = V;
    r.ZZZ = W;
    return r;
}

// 2*P (dbl-2008-s-1); infinity stays infinity.
template <class F>
NKR_HD Xyzz<F> xyzz_dbl(const Xyzz<F> &p)
{
    using O = FieldOps<F>;
    const F U = O::dbl(p.Y);
    F V, x2, W, S, MM, ZZ3;
    O::sqr2(U, p.X, V, x2);
    O::mul2(U, V, p.X, V, W, S);
    const F M = O::add(O::dbl(x2), x2);
    O::sqr_mul(M, V, p.ZZ, MM, ZZ3);
    Xyzz<F> r;
    r.X = O::sub(MM, O::dbl(S));
    F t1, t2;
    O::mul2(M, O::sub(S, r.X), W, p.Y, t1, t2);
    r.Y = O::sub(t1, t2);
    r.ZZ = ZZ3;
    r.ZZZ = O::mul(W, p.ZZZ);
    return r;  // p = infinity gives ZZ = 0 again
}

// acc + (x2, y2), complete (madd-2008-s with the equal / opposite cases).
template <class F>
NKR_HD Xyzz<F> xyzz_add_affine(const Xyzz<F> &acc, const F &x2, const F &y2)
{
    using O = FieldOps<F>;
    if (acc.is_inf()) {
        Xyzz<F> r;
        r.X = x2;
        r.Y = y2;
        r.ZZ = O::one();
        r.ZZZ = O::one();
        return r;
    }
    F U2, S2;
    O::mul2(x2, acc.ZZ, y2, acc.ZZZ, U2, S2);
    const F P = O::sub(U2, acc.X);
    const F R = O::sub(S2, acc.Y);
    if (P.is_zero()) {
        if (R.is_zero()) return xyzz_dbl_affine(x2, y2);
        return Xyzz<F>::infinity();
    }
    F PP, RR, PPP, Qv;
    O::sqr2(P, R, PP, RR);
    O::mul2(P, PP, acc.X, PP, PPP, Qv);
    Xyzz<F> r;
    r.X = O::sub(O::sub(RR, PPP), O::dbl(Qv));
    F t1, t2;
    O::mul2(R, O::sub(Qv, r.X), acc.Y, PPP, t1, t2);
    r.Y = O::sub(t1, t2);
    O::mul2(acc.ZZ, PP, acc.ZZZ, PPP, r.ZZ, r.ZZZ);
    return r;
}

// p + q, complete (add-2008-s with the equal / opposite / infinity cases).
template <class F>
NKR_HD Xyzz<F> xyzz_add(const Xyzz<F> &p, const Xyzz<F> &q)
{
    using O = FieldOps<F>;
    if (p.is_inf()) return q;
    if (q.is_inf()) return p;
    F U1, U2, S1, S2;
    O::mul2(p.X, q.ZZ, q.X, p.ZZ, U1, U2);
    O::mul2(p.Y, q.ZZZ, q.Y, p.ZZZ, S1, S2);
    const F P = O::sub(U2, U1);
    const F R = O::sub(S2, S1);
    if (P.is_zero()) {
        if (R.is_zero()) return xyzz_dbl(p);
        return Xyzz<F>::infinity();
    }
    F PP, ZZ12, PPP, Qv, RR, ZZZ12;
    O::sqr_mul(P, p.ZZ, q.ZZ, PP, ZZ12);
    O::mul2(P, PP, U1, PP, PPP, Qv);
    O::sqr_mul(R, p.ZZZ, q.ZZZ, RR, ZZZ12);
    Xyzz<F> r;
    r.X = O::sub(O::sub(RR, PPP), O::dbl(Qv));
    F t1, t2;
    O::mul2(R, O::sub(Qv, r.X), S1, PPP, t1, t2);
    r.Y = O::sub(t1, t2);
    O::mul2(ZZ12, PP, ZZZ12, PPP, r.ZZ, r.ZZZ);
    return r;
}

template <class F>
NKR_HD Xyzz<F> xyzz_neg(const Xyzz<F> &p)
{
    Xyzz<F> r = p;
    r.Y = FieldOps<F>::neg(p.Y);
    return r;
}

// Jacobian (X, Y, Z) with x = X/Z^2, y = Y/Z^3: the reference's host layout
// (libff alt_bn128_G1 / alt_bn128_G2: 3 coordinates, infinity is Z = 0).
template <class F>
struct Jacobian {
    F X, Y, Z;
};

// Affine image in the reference's to_affine_coordinates form: (x, y, 1), or
// (0, 1, 0) for infinity (alt_bn128_g1.cpp:64-80).
template <class F>
NKR_HD Jacobian<F> xyzz_to_jacobian_affine(const Xyzz<F> &p)
{
    using O = FieldOps<F>;
    Jacobian<F> r;
    if (p.is_inf()) {
        r.X = O::zero();
        r.Y = O::one();
        r.Z = O::zero();
        return r;
    }
    // x = X/ZZ, y = Y/ZZZ with one inversion: t = 1/(ZZ*ZZZ).
    const F t = O::inv(O::mul(p.ZZ, p.ZZZ));
    const F zz_inv = O::mul(t, p.ZZZ);
    const F zzz_inv = O::mul(t, p.ZZ);
    r.X = O::mul(p.X, zz_inv);
    r.Y = O::mul(p.Y, zzz_inv);
    r.Z = O::one();
    return r;
}

// xyzz_to_jacobian_affine of n <= 8 points with one inversion shared by all
// (Montgomery's simultaneous inversion of the ZZ * ZZZ products; infinity
// takes no part): the same points, so the same canonical coordinates.
template <class F>
NKR_HD void xyzz_to_jacobian_affine_batch(const Xyzz<F> *in, Jacobian<F> *out, int n)
{
    using O = FieldOps<F>;
    F d[8], prefix[8];
    F acc = O::one();
    for (int i = 0; i < n; ++i) {
        d[i] = in[i].is_inf() ? O::one() : O::mul(in[i].ZZ, in[i].ZZZ);
        prefix[i] = acc;
        acc = O::mul(acc, d[i]);
    }
    F inv = O::inv(acc);  // 1 / (d_0 ... d_{n-1})
    for (int i = n - 1; i >= 0; --i) {
        const F t = O::mul(inv, prefix[i]);  // 1 / d_i
        inv = O::mul(inv, d[i]);
        Jacobian<F> &r = out[i];
        if (in[i].is_inf()) {
            r.X = O::zero();
            r.Y = O::one();
            r.Z = O::zero();
            continue;
        }
        r.X = O::mul(in[i].X, O::mul(t, in[i].ZZZ));
        r.Y = O::mul(in[i].Y, O::mul(t, in[i].ZZ));
        r.Z = O::one();
    }
}

template <class F>
NKR_HD Affine<F> xyzz_to_affine(const Xyzz<F> &p)
{
    using O = FieldOps<F>;
    if (p.is_inf()) return affine_infinity<F>();
    const F t = O::inv(O::mul(p.ZZ, p.ZZZ));
    Affine<F> r;
    r.x = O::mul(p.X, O::mul(t, p.ZZZ));
    r.y = O::mul(p.Y, O::mul(t, p.ZZ));
    return r;
}

// Jacobian -> affine; returns false for infinity.
template <class F>
NKR_HD bool jacobian_to_affine(const Jacobian<F> &p, Affine<F> &out)
{
    using O = FieldOps<F>;
    if (p.Z.is_zero()) {
        out.x = O::zero();
        out.y = O::zero();
        return false;
    }
    const F zi = O::inv(p.Z);
    const F zi2 = O::sqr(zi);
    out.x = O::mul(p.X, zi2);
    out.y = O::mul(p.Y, O::mul(zi2, zi));
    return true;
}

template <class F>
NKR_HD Xyzz<F> jacobian_to_xyzz(const Jacobian<F> &p)
{
    using O = FieldOps<F>;
    if (p.Z.is_zero()) return Xyzz<F>::infinity();
    Xyzz<F> r;
    r.X = p.X;
    r.Y = p.Y;
    r.ZZ = O::sqr(p.Z);
    r.ZZZ = O::mul(r.ZZ, p.Z);
    return r;
}

// k*P for a 256-bit little-endian integer k, left-to-right double and add.
template <class F>
NKR_HD Xyzz<F> xyzz_mul_u256(const Xyzz<F> &p, const uint32_t k[8])
{
    Xyzz<F> acc = Xyzz<F>::infinity();
    for (int w = 7; w >= 0; --w)
        for (int b = 31; b >= 0; --b) {
            acc = xyzz_dbl(acc);
            if ((k[w] >> b) & 1u) acc = xyzz_add(acc, p);
        }
    return acc;
}

// k*P with a fixed 4-bit window: 14 additions of precomputation, then 4
// doublings and at most one addition per nibble (host-side blinding).
template <class F>
NKR_HD Xyzz<F> xyzz_mul_u256_w4(const Xyzz<F> &p, const uint32_t k[8])
{
    Xyzz<F> table[16];
    table[0] = Xyzz<F>::infinity();
    table[1] = p;
    for (int i = 2; i < 16; ++i) table[i] = xyzz_add(table[i - 1], p);
    Xyzz<F> acc = Xyzz<F>::infinity();
    bool started = false;
    for (int nib = 63; nib >= 0; --nib) {
        if (started)
            for (int d = 0; d < 4; ++d) acc = xyzz_dbl(acc);
        const uint32_t v = (k[nib >> 3] >> ((nib & 7) * 4)) & 15u;
        if (v) {
            acc = xyzz_add(acc, table[v]);
            started = true;
        }
    }
    return acc;
}

using g1_affine = Affine<fq_t>;
using g2_affine = Affine<fq2_t>;
using g1_xyzz = Xyzz<fq_t>;
using g2_xyzz = Xyzz<fq2_t>;
using g1_jac = Jacobian<fq_t>;
using g2_jac = Jacobian<fq2_t>;

static_assert(sizeof(g1_jac) == 96, "G1 Jacobian must match the reference's 96-byte layout");
static_assert(sizeof(g2_jac) == 192, "G2 Jacobian must match the reference's 192-byte layout");

}  // namespace nkr

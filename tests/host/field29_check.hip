// CPU check of the reduced-radix field and its mixed addition
// (nkr-zkp_amd/csrc/field29.h, curve29.h) against the canonical 32-bit-limb
// code (field.h, curve.h), compiled as host code.
//   * conversions: canonical -> 29-bit -> canonical is the identity; the
//     29-bit Montgomery one converts to the canonical one;
//   * products: every input bound the formulas use (below 2p, 10p, and 32p
//     against 2p), built as x + k*p, gives fp_mul's canonical result;
//     squarings give fp29_mul(a, a)'s limbs, and a sum of two products with
//     one reduction the canonical sum;
//   * lazy sums and differences: a + b and a + K*p - b are congruent to the
//     canonical ones;
//   * mixed addition: chains of xyzz29_add_affine over random bases, repeated
//     bases (doubling), negated ones (infinity) and infinity restarts give,
//     after every step, the canonical coordinates of xyzz_add_affine;
//   * full addition: chains of xyzz29_add over partial sums (ZZ != 1),
//     doublings, cancellations and infinity operands give xyzz_add's
//     coordinates after every step;
//   * the same over Fq2 (G2): products, squares and the four-product sum at
//     their input bounds, and chains of xyzz2_29_add_affine against
//     xyzz_add_affine.
//     The formulas are field identities, so random coordinates (not on the
//     curve) exercise them as well as curve points do.
// Prints "ok <checks>" or the first mismatches; exit status 1 on a mismatch.
#include <cstdio>
#include <vector>

#include "curve29.h"

using namespace nkr;

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static uint32_t rnd()
{
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return (uint32_t)(g_state >> 11);
}

template <class P>
static Fp<P> rand_fp()
{
    Fp<P> x;
    for (int i = 0; i < 8; ++i) x.l[i] = rnd();
    x.l[7] %= P::P[7] + 1;  // below 2p
    return fp_reduce_once(x);
}

static int g_bad = 0;
static long g_doublings = 0, g_cancels = 0;
static long g_checks = 0;
static void expect(bool ok, const char *what, int i)
{
    ++g_checks;
    if (!ok && g_bad++ < 10) std::printf("mismatch: %s at %d\n", what, i);
}

// x + K*p in the 29-bit form, x canonical (value below (K + 2) p)
template <class P, uint32_t K>
static Fp29<P> widen(const Fp<P> &x)
{
    return fp29_sub<P, K>(fp29_reduce_to_2p(fp29_from_mont(x)), Fp29<P>::zero());
}

template <class P>
static void check_field(int n)
{
    for (int i = 0; i < n; ++i) {
        Fp<P> x = rand_fp<P>(), y = rand_fp<P>();
        if (i == 0) x = y = Fp<P>::zero();
        if (i == 1) x = y = Fp<P>::one();
        if (i == 2) {
            x = Fp<P>::modulus();
            x.l[0] -= 1;  // p - 1
            y = x;
        }
        const Fp<P> want = fp_mul(x, y);
        expect(fp29_to_mont(fp29_from_mont(x)) == x, "round trip", i);
        expect(fp29_to_mont(fp29_mul(fp29_from_mont(x), fp29_reduce_to_2p(fp29_from_mont(y)))) == want,
               "mul 32p x 2p", i);
        expect(fp29_to_mont(fp29_mul(widen<P, 8>(x), widen<P, 8>(y))) == want, "mul 10p x 10p", i);
        expect(fp29_to_mont(fp29_mul(widen<P, 0>(x), widen<P, 0>(y))) == want, "mul 2p x 2p", i);
        expect(fp29_to_mont(fp29_sqr(widen<P, 8>(x))) == fp_mul(x, x), "sqr 10p", i);
        {
            const Fp29<P> wx = widen<P, 8>(x), s0 = fp29_sqr(wx), s1 = fp29_mul(wx, wx);
            bool same = true;
            for (int k = 0; k < 9; ++k) same = same && s0.l[k] == s1.l[k];
            expect(same, "sqr limbs equal mul(a, a)", i);
        }
        {
            // the G1 Y3 shape: 10p x 10p + 8p x 2p in one reduction
            const Fp<P> u = rand_fp<P>(), v = rand_fp<P>();
            const Fp29<P> s = fp29_mul_add(widen<P, 8>(x), widen<P, 8>(y), widen<P, 6>(u), widen<P, 0>(v));
            bool limbs = true;
            for (int k = 0; k < 9; ++k) limbs = limbs && (s.l[k] >> 29) == 0;
            expect(limbs && fp29_to_mont(s) == fp_add(fp_mul(x, y), fp_mul(u, v)), "mul_add", i);
        }
        const Fp29<P> m = fp29_mul(widen<P, 8>(x), widen<P, 8>(y));
        bool limbs = true;
        for (int k = 0; k < 9; ++k) limbs = limbs && (m.l[k] >> 29) == 0;
        expect(limbs, "product limbs below 2^29", i);
        expect(fp29_to_mont(fp29_add(widen<P, 0>(x), widen<P, 4>(y))) == fp_add(x, y), "add", i);
        expect(fp29_to_mont(fp29_sub<P, 8>(widen<P, 0>(x), widen<P, 4>(y))) == fp_sub(x, y), "sub", i);
        expect(fp29_to_mont(fp29_reduce_to_8p(fp29_from_mont(x))) == x, "reduce to 8p", i);
        const bool z = fp_sub(x, y).is_zero();
        const Fp29<P> d = fp29_sub<P, 8>(widen<P, 0>(x), widen<P, 0>(y));
        expect(!z || fp29_may_be_zero<P, 10>(d), "zero filter", i);
    }
    expect(fp29_to_mont(Fp29<P>::one()) == Fp<P>::one(), "one", 0);
    // the zero filter passes every multiple j*p below 10p
    const Fp29<P> z = Fp29<P>::zero();
    expect(fp29_may_be_zero<P, 10>(z) && fp29_may_be_zero<P, 10>(fp29_sub<P, 1>(z, z)) &&
               fp29_may_be_zero<P, 10>(fp29_sub<P, 5>(z, z)) && fp29_may_be_zero<P, 10>(fp29_sub<P, 9>(z, z)),
           "zero filter on multiples of p", 0);
}

static Affine<fq_t> random_base(const Affine<fq_t> &g)
{
    uint32_t k[8];
    for (int i = 0; i < 8; ++i) k[i] = rnd();
    k[7] &= 0x0fffffffu;
    return xyzz_to_affine(xyzz_mul_u256(Xyzz<fq_t>::from_affine(g), k));
}

static void check_madd(int chains, int len)
{
    Affine<fq_t> g;
    g.x = fp_to_mont(fq_t::zero());
    g.x.l[0] = 1;
    g.x = fp_to_mont(g.x);
    g.y = fq_t::zero();
    g.y.l[0] = 2;
    g.y = fp_to_mont(g.y);
    std::vector<Affine<fq_t>> pool;
    for (int i = 0; i < 24; ++i) pool.push_back(random_base(g));
    for (int c = 0; c < chains; ++c) {
        Xyzz<fq_t> a32 = Xyzz<fq_t>::infinity();
        Xyzz29<FqParams> a29 = Xyzz29<FqParams>::infinity();
        for (int e = 0; e < len; ++e) {
            const uint32_t r = rnd();
            Affine<fq_t> b = pool[r % pool.size()];
            if ((r >> 8) % 5 == 0 && !a32.is_inf()) {  // doubling
                b = xyzz_to_affine(a32);
                ++g_doublings;
            }
            if ((r >> 12) % 7 == 0) b.y = fp_neg(b.y);
            if ((r >> 16) % 11 == 0 && !a32.is_inf()) {  // back to infinity
                b = xyzz_to_affine(a32);
                b.y = fp_neg(b.y);
                ++g_cancels;
            }
            a32 = xyzz_add_affine(a32, b.x, b.y);
            a29 = xyzz29_add_affine(a29, fp29_from_mont(b.x), fp29_from_mont(b.y));
            const Xyzz<fq_t> got = xyzz29_to(a29);
            expect(got.X == a32.X && got.Y == a32.Y && got.ZZ == a32.ZZ && got.ZZZ == a32.ZZZ, "madd chain",
                   c * len + e);
        }
    }
}

// Full additions: sums of random partial sums (each a chain of mixed
// additions, so ZZ != 1), including a + a (doubling), a + (-a) (infinity)
// and infinity operands, against xyzz_add after every step.
static void check_add(int chains, int len)
{
    Affine<fq_t> g;
    g.x = fq_t::zero();
    g.x.l[0] = 1;
    g.x = fp_to_mont(g.x);
    g.y = fq_t::zero();
    g.y.l[0] = 2;
    g.y = fp_to_mont(g.y);
    std::vector<Affine<fq_t>> pool;
    for (int i = 0; i < 16; ++i) pool.push_back(random_base(g));
    auto partial = [&](int n, Xyzz<fq_t> &c32, Xyzz29<FqParams> &c29) {
        c32 = Xyzz<fq_t>::infinity();
        c29 = Xyzz29<FqParams>::infinity();
        for (int e = 0; e < n; ++e) {
            const Affine<fq_t> b = pool[rnd() % pool.size()];
            c32 = xyzz_add_affine(c32, b.x, b.y);
            c29 = xyzz29_add_affine(c29, fp29_from_mont(b.x), fp29_from_mont(b.y));
        }
    };
    for (int c = 0; c < chains; ++c) {
        Xyzz<fq_t> a32;
        Xyzz29<FqParams> a29;
        partial(1 + rnd() % 3, a32, a29);
        for (int e = 0; e < len; ++e) {
            const uint32_t r = rnd();
            Xyzz<fq_t> b32;
            Xyzz29<FqParams> b29;
            partial((r >> 4) % 4, b32, b29);  // 0: infinity
            if ((r >> 8) % 6 == 0) {          // a + a
                b32 = a32;
                b29 = a29;
                ++g_doublings;
            } else if ((r >> 12) % 7 == 0 && !a32.is_inf()) {  // a + (-a)
                b32 = a32;
                b32.Y = fp_neg(b32.Y);
                b29 = a29;
                b29.Y = fp29_sub<FqParams, 8>(Fp29<FqParams>::zero(), a29.Y);
                ++g_cancels;
            }
            a32 = xyzz_add(a32, b32);
            a29 = xyzz29_add(a29, b29);
            const Xyzz<fq_t> got = xyzz29_to(a29);
            const Affine<fq_t> w = xyzz_to_affine(a32), h = xyzz_to_affine(got);
            expect(got.is_inf() == a32.is_inf() && (a32.is_inf() || (w.x == h.x && w.y == h.y)), "add chain",
                   c * len + e);
            expect(got.X == a32.X && got.Y == a32.Y && got.ZZ == a32.ZZ && got.ZZZ == a32.ZZZ, "add chain coords",
                   c * len + e);
            if (a32.is_inf()) partial(1, a32, a29);
        }
    }
}

static fq2_t rand_fq2() { return {rand_fp<FqParams>(), rand_fp<FqParams>()}; }

// a + K*p per component (values below (K + 2) p)
template <uint32_t K>
static Fq2_29<FqParams> widen2(const fq2_t &a)
{
    return {widen<FqParams, K>(a.c0), widen<FqParams, K>(a.c1)};
}

static void check_fq2(int n)
{
    for (int i = 0; i < n; ++i) {
        const fq2_t a = rand_fq2(), b = rand_fq2();
        const fq2_t want = fq2_mul(a, b);
        // the products' input bounds in the G2 addition: 32p x 2p (U2, S2),
        // 8p x 2p (Q), 6p x 2p (PPP), 2p x 2p (ZZ3, ZZZ3)
        expect(fq2_29_to(fq2_29_mul<FqParams, 2>(widen2<30>(a), widen2<0>(b))) == want, "fq2 mul 32p x 2p", i);
        expect(fq2_29_to(fq2_29_mul<FqParams, 2>(widen2<6>(a), widen2<0>(b))) == want, "fq2 mul 8p x 2p", i);
        expect(fq2_29_to(fq2_29_mul<FqParams, 2>(widen2<0>(a), widen2<0>(b))) == want, "fq2 mul 2p x 2p", i);
        expect(fq2_29_to(fq2_29_sqr<FqParams, 6>(widen2<4>(a))) == fq2_sqr(a), "fq2 sqr 6p", i);
        {
            // Y3's shape: R (6p) T (9.2p, as 10p) + Y (2p) N (2p), per component
            const fq2_t c = rand_fq2(), d = rand_fq2();
            const Fq2_29<FqParams> R = widen2<4>(a), T = widen2<8>(b), Y = widen2<0>(c), N = widen2<0>(d);
            const Fp29<FqParams> nT1 = fp29_sub<FqParams, 10>(Fp29<FqParams>::zero(), T.c1);
            const Fp29<FqParams> nN1 = fp29_sub<FqParams, 2>(Fp29<FqParams>::zero(), N.c1);
            const Fq2_29<FqParams> y = {fp29_mul_add4(R.c0, T.c0, R.c1, nT1, Y.c0, N.c0, Y.c1, nN1),
                                        fp29_mul_add4(R.c0, T.c1, R.c1, T.c0, Y.c0, N.c1, Y.c1, N.c0)};
            expect(fq2_29_to(y) == fq2_add(fq2_mul(a, b), fq2_mul(c, d)), "fq2 four-product sum", i);
            bool limbs = true;
            for (int k = 0; k < 9; ++k) limbs = limbs && (y.c0.l[k] >> 29) == 0 && (y.c1.l[k] >> 29) == 0;
            expect(limbs, "fq2 four-product limbs", i);
        }
        const Fq2_29<FqParams> m = fq2_29_mul<FqParams, 2>(widen2<30>(a), widen2<0>(b));
        bool limbs = true;
        for (int k = 0; k < 9; ++k) limbs = limbs && (m.c0.l[k] >> 29) == 0 && (m.c1.l[k] >> 29) == 0;
        expect(limbs, "fq2 product limbs", i);
        expect(fq2_29_to(fq2_29_from<FqParams>(a)) == a, "fq2 round trip", i);
    }
}

static void check_madd_g2(int chains, int len)
{
    std::vector<Affine<fq2_t>> pool;
    for (int i = 0; i < 24; ++i) pool.push_back({rand_fq2(), rand_fq2()});
    for (int c = 0; c < chains; ++c) {
        Xyzz<fq2_t> a32 = Xyzz<fq2_t>::infinity();
        Xyzz2_29<FqParams> a29 = Xyzz2_29<FqParams>::infinity();
        for (int e = 0; e < len; ++e) {
            const uint32_t r = rnd();
            Affine<fq2_t> b = pool[r % pool.size()];
            if ((r >> 8) % 5 == 0 && !a32.is_inf()) {
                b = xyzz_to_affine(a32);
                ++g_doublings;
            }
            if ((r >> 12) % 7 == 0) b.y = fq2_neg(b.y);
            if ((r >> 16) % 11 == 0 && !a32.is_inf()) {
                b = xyzz_to_affine(a32);
                b.y = fq2_neg(b.y);
                ++g_cancels;
            }
            a32 = xyzz_add_affine(a32, b.x, b.y);
            a29 = xyzz2_29_add_affine(a29, fq2_29_from<FqParams>(b.x), fq2_29_from<FqParams>(b.y));
            const Xyzz<fq2_t> got = xyzz2_29_to(a29);
            expect(got.X == a32.X && got.Y == a32.Y && got.ZZ == a32.ZZ && got.ZZZ == a32.ZZZ, "g2 madd chain",
                   c * len + e);
        }
    }
}

// Full G2 additions over partial sums, as check_add.
static void check_add_g2(int chains, int len)
{
    std::vector<Affine<fq2_t>> pool;
    for (int i = 0; i < 16; ++i) pool.push_back({rand_fq2(), rand_fq2()});
    auto partial = [&](int n, Xyzz<fq2_t> &c32, Xyzz2_29<FqParams> &c29) {
        c32 = Xyzz<fq2_t>::infinity();
        c29 = Xyzz2_29<FqParams>::infinity();
        for (int e = 0; e < n; ++e) {
            const Affine<fq2_t> b = pool[rnd() % pool.size()];
            c32 = xyzz_add_affine(c32, b.x, b.y);
            c29 = xyzz2_29_add_affine(c29, fq2_29_from<FqParams>(b.x), fq2_29_from<FqParams>(b.y));
        }
    };
    for (int c = 0; c < chains; ++c) {
        Xyzz<fq2_t> a32;
        Xyzz2_29<FqParams> a29;
        partial(1 + rnd() % 3, a32, a29);
        for (int e = 0; e < len; ++e) {
            const uint32_t r = rnd();
            Xyzz<fq2_t> b32;
            Xyzz2_29<FqParams> b29;
            partial((r >> 4) % 4, b32, b29);
            if ((r >> 8) % 6 == 0) {
                b32 = a32;
                b29 = a29;
                ++g_doublings;
            } else if ((r >> 12) % 7 == 0 && !a32.is_inf()) {
                b32 = a32;
                b32.Y = fq2_neg(b32.Y);
                b29 = a29;
                b29.Y = fq2_29_sub<FqParams, 2>({Fp29<FqParams>::zero(), Fp29<FqParams>::zero()}, a29.Y);
                ++g_cancels;
            }
            a32 = xyzz_add(a32, b32);
            a29 = xyzz2_29_add(a29, b29);
            const Xyzz<fq2_t> got = xyzz2_29_to(a29);
            expect(got.X == a32.X && got.Y == a32.Y && got.ZZ == a32.ZZ && got.ZZZ == a32.ZZZ, "g2 add chain",
                   c * len + e);
            if (a32.is_inf()) partial(1, a32, a29);
        }
    }
}

int main()
{
    check_field<FqParams>(20000);
    check_field<FrParams>(20000);
    check_madd(200, 80);
    check_add(150, 60);
    check_fq2(5000);
    check_madd_g2(60, 60);
    check_add_g2(60, 40);
    if (g_bad) {
        std::printf("FAILED %d of %ld checks\n", g_bad, g_checks);
        return 1;
    }
    std::printf("ok %ld checks (%ld doublings, %ld cancellations)\n", g_checks, g_doublings, g_cancels);
    return 0;
}

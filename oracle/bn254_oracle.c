/*
 * CPU oracle for the Groth16 prove path on BN254 -- TEST INFRASTRUCTURE ONLY
 * (the checker and the CPU baseline; never part of the product library).
 *
 * Restates, in plain C, the algorithms of the reference CPU backend and of the
 * vendored libff / libfqfft code it runs.  Paths below are relative to the
 * reference tree; LIBFF = third_party/lightzk/depends/libsnark/depends/libff/libff,
 * LIBFQFFT = third_party/lightzk/depends/libsnark/depends/libfqfft/libfqfft.
 *
 *   field arithmetic   LIBFF/algebra/fields/fp.tcc:23 (mul_reduce, Montgomery
 *                      CIOS on 4 x u64), :529-536 (operator*=), :712 (inverse;
 *                      GMP mpn_gcdext there, Fermat here -- the inverse is unique)
 *   random element     LIBFF/algebra/fields/fp.tcc:719-744, bigint.tcc:171-183
 *   Fq2                LIBFF/algebra/fields/fp2.tcc:73-118 (Karatsuba product,
 *                      complex squaring, non-residue -1)
 *   G1 / G2            LIBFF/algebra/curves/alt_bn128/alt_bn128_g1.cpp:64-362,
 *                      alt_bn128_g2.cpp:74-372 (Jacobian add-2007-bl with a
 *                      doubling check, dbl-2009-l, to_affine_coordinates)
 *   scalar_mul         LIBFF/algebra/curves/curve_utils.tcc:14-37
 *   multi_exp BDLO12   LIBFF/algebra/scalar_multiplication/multiexp.tcc:165-282,402-433
 *   radix-2 FFT        LIBFQFFT/evaluation_domain/domains/basic_radix2_domain.tcc:26-112,
 *                      basic_radix2_domain_aux.tcc:45-76,172-180;
 *                      LIBFF/algebra/fields/field_utils.tcc:36-52 (root of unity)
 *   SpMV, prove        protocols/groth16/backends/cpu/src/backend.cpp:25-35,156-201
 *   blind_proof        protocols/groth16/common/src/protocol.cpp:85-94
 *   constants          LIBFF/algebra/curves/alt_bn128/alt_bn128_init.cpp:38-62,134-150,207-215
 *
 * The pseudo-random source is the C++ standard mt19937_64 ([rand.predef]),
 * the engine libff's bigint::randomize draws from.
 */
#include "bn254_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef struct {
    uint64_t v[4];
} fe;
typedef struct {
    fe c0, c1;
} fe2;

/* ---------------------------------------------------------------- constants */
static const uint64_t Q_P[4] = {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL,
                                0x30644e72e131a029ULL};
static const uint64_t Q_INV = 0x87d20782e4866389ULL;
static const uint64_t Q_R2[4] = {0xf32cfc5b538afa89ULL, 0xb5e71911d44501fbULL, 0x47ab1eff0a417ff6ULL,
                                 0x06d89f71cab8351fULL};
static const uint64_t R_P[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL,
                                0x30644e72e131a029ULL};
static const uint64_t R_INV = 0xc2e1f593efffffffULL;
static const uint64_t R_R2[4] = {0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL,
                                 0x0216d0b17f4e44a5ULL};
#define FR_S 28

/* ------------------------------------------------------------------- fields */
#define DEFINE_FIELD(F, P, INV, R2)                                                          \
    static inline int F##_is_zero(const fe *a)                                               \
    {                                                                                        \
        return (a->v[0] | a->v[1] | a->v[2] | a->v[3]) == 0;                                 \
    }                                                                                        \
    static inline int F##_eq(const fe *a, const fe *b)                                       \
    {                                                                                        \
        return a->v[0] == b->v[0] && a->v[1] == b->v[1] && a->v[2] == b->v[2] &&             \
               a->v[3] == b->v[3];                                                           \
    }                                                                                        \
    /* t (< 2^256) minus p when t >= p */                                                    \
    static inline void F##_sub_p_if_ge(fe *r, const uint64_t t[4], uint64_t hi)              \
    {                                                                                        \
        uint64_t d[4], borrow = 0;                                                           \
        for (int i = 0; i < 4; ++i) {                                                        \
            const u128 s = (u128)t[i] - P[i] - borrow;                                       \
            d[i] = (uint64_t)s;                                                              \
            borrow = (uint64_t)(s >> 64) & 1;                                                \
        }                                                                                    \
        const int ge = hi != 0 || borrow == 0;                                               \
        memcpy(r->v, ge ? d : t, 32);                                                        \
    }                                                                                        \
    static inline void F##_add(fe *r, const fe *a, const fe *b)                              \
    {                                                                                        \
        uint64_t t[4], carry = 0;                                                            \
        for (int i = 0; i < 4; ++i) {                                                        \
            const u128 s = (u128)a->v[i] + b->v[i] + carry;                                  \
            t[i] = (uint64_t)s;                                                              \
            carry = (uint64_t)(s >> 64);                                                     \
        }                                                                                    \
        F##_sub_p_if_ge(r, t, carry);                                                        \
    }                                                                                        \
    static inline void F##_sub(fe *r, const fe *a, const fe *b)                              \
    {                                                                                        \
        uint64_t t[4], borrow = 0;                                                           \
        for (int i = 0; i < 4; ++i) {                                                        \
            const u128 s = (u128)a->v[i] - b->v[i] - borrow;                                 \
            t[i] = (uint64_t)s;                                                              \
            borrow = (uint64_t)(s >> 64) & 1;                                                \
        }                                                                                    \
        if (borrow) {                                                                        \
            uint64_t carry = 0;                                                              \
            for (int i = 0; i < 4; ++i) {                                                    \
                const u128 s = (u128)t[i] + P[i] + carry;                                    \
                t[i] = (uint64_t)s;                                                          \
                carry = (uint64_t)(s >> 64);                                                 \
            }                                                                                \
        }                                                                                    \
        memcpy(r->v, t, 32);                                                                 \
    }                                                                                        \
    static inline void F##_neg(fe *r, const fe *a)                                           \
    {                                                                                        \
        if (F##_is_zero(a)) {                                                                \
            memset(r, 0, 32);                                                                \
            return;                                                                          \
        }                                                                                    \
        uint64_t borrow = 0;                                                                 \
        fe t;                                                                                \
        for (int i = 0; i < 4; ++i) {                                                        \
            const u128 s = (u128)P[i] - a->v[i] - borrow;                                    \
            t.v[i] = (uint64_t)s;                                                            \
            borrow = (uint64_t)(s >> 64) & 1;                                                \
        }                                                                                    \
        *r = t;                                                                              \
    }                                                                                        \
    /* Montgomery CIOS: a*b/2^256 mod p (fp.tcc:23 mul_reduce) */                            \
    static inline void F##_mul(fe *r, const fe *a, const fe *b)                              \
    {                                                                                        \
        uint64_t t[6] = {0, 0, 0, 0, 0, 0};                                                  \
        for (int i = 0; i < 4; ++i) {                                                        \
            uint64_t C = 0;                                                                  \
            for (int j = 0; j < 4; ++j) {                                                    \
                const u128 s = (u128)a->v[j] * b->v[i] + t[j] + C;                           \
                t[j] = (uint64_t)s;                                                          \
                C = (uint64_t)(s >> 64);                                                     \
            }                                                                                \
            u128 s = (u128)t[4] + C;                                                         \
            t[4] = (uint64_t)s;                                                              \
            t[5] = (uint64_t)(s >> 64);                                                      \
            const uint64_t m = t[0] * INV;                                                   \
            s = (u128)m * P[0] + t[0];                                                       \
            C = (uint64_t)(s >> 64);                                                         \
            for (int j = 1; j < 4; ++j) {                                                    \
                s = (u128)m * P[j] + t[j] + C;                                               \
                t[j - 1] = (uint64_t)s;                                                      \
                C = (uint64_t)(s >> 64);                                                     \
            }                                                                                \
            s = (u128)t[4] + C;                                                              \
            t[3] = (uint64_t)s;                                                              \
            t[4] = t[5] + (uint64_t)(s >> 64);                                               \
        }                                                                                    \
        F##_sub_p_if_ge(r, t, t[4]);                                                         \
    }                                                                                        \
    static inline void F##_sqr(fe *r, const fe *a) { F##_mul(r, a, a); }                     \
    static inline void F##_set_u64(fe *r, uint64_t x) /* Fp_model(long): to Montgomery */    \
    {                                                                                        \
        fe t = {{x, 0, 0, 0}}, r2;                                                           \
        memcpy(r2.v, R2, 32);                                                                \
        F##_mul(r, &t, &r2);                                                                 \
    }                                                                                        \
    static inline void F##_one(fe *r) { F##_set_u64(r, 1); }                                 \
    static inline void F##_to_mont(fe *r, const fe *x)                                       \
    {                                                                                        \
        fe r2;                                                                               \
        memcpy(r2.v, R2, 32);                                                                \
        F##_mul(r, x, &r2);                                                                  \
    }                                                                                        \
    static inline void F##_from_mont(fe *r, const fe *x) /* as_bigint */                     \
    {                                                                                        \
        fe one = {{1, 0, 0, 0}};                                                             \
        F##_mul(r, x, &one);                                                                 \
    }                                                                                        \
    /* x^e for a little-endian multi-word exponent (square and multiply, MSB first) */      \
    static inline void F##_pow(fe *r, const fe *x, const uint64_t *e, int words)             \
    {                                                                                        \
        fe acc;                                                                              \
        F##_one(&acc);                                                                       \
        int started = 0;                                                                     \
        for (int w = words - 1; w >= 0; --w)                                                 \
            for (int b = 63; b >= 0; --b) {                                                  \
                if (started) F##_sqr(&acc, &acc);                                            \
                if ((e[w] >> b) & 1) {                                                       \
                    F##_mul(&acc, &acc, x);                                                  \
                    started = 1;                                                             \
                }                                                                            \
            }                                                                                \
        *r = acc;                                                                            \
    }                                                                                        \
    static inline void F##_inv(fe *r, const fe *x)                                           \
    {                                                                                        \
        uint64_t e[4];                                                                       \
        memcpy(e, P, 32);                                                                    \
        e[0] -= 2; /* p - 2; p is odd and > 2, no borrow */                                  \
        F##_pow(r, x, e, 4);                                                                 \
    }

DEFINE_FIELD(fq, Q_P, Q_INV, Q_R2)
DEFINE_FIELD(fr, R_P, R_INV, R_R2)

/* ------------------------------------------------------------------------ Fq2 */
static fe FQ_NONRES; /* -1 in Montgomery form (alt_bn128_init.cpp:95) */

static inline int fq2_is_zero(const fe2 *a) { return fq_is_zero(&a->c0) && fq_is_zero(&a->c1); }
static inline int fq2_eq(const fe2 *a, const fe2 *b) { return fq_eq(&a->c0, &b->c0) && fq_eq(&a->c1, &b->c1); }
static inline void fq2_add(fe2 *r, const fe2 *a, const fe2 *b)
{
    fq_add(&r->c0, &a->c0, &b->c0);
    fq_add(&r->c1, &a->c1, &b->c1);
}
static inline void fq2_sub(fe2 *r, const fe2 *a, const fe2 *b)
{
    fq_sub(&r->c0, &a->c0, &b->c0);
    fq_sub(&r->c1, &a->c1, &b->c1);
}
static inline void fq2_neg(fe2 *r, const fe2 *a)
{
    fq_neg(&r->c0, &a->c0);
    fq_neg(&r->c1, &a->c1);
}
/* fp2.tcc:73-84: (aA + nr*bB, (a+b)(A+B) - aA - bB) */
static inline void fq2_mul(fe2 *r, const fe2 *x, const fe2 *y)
{
    fe aA, bB, t, s0, s1, c0, c1;
    fq_mul(&aA, &x->c0, &y->c0);
    fq_mul(&bB, &x->c1, &y->c1);
    fq_mul(&t, &FQ_NONRES, &bB);
    fq_add(&c0, &aA, &t);
    fq_add(&s0, &x->c0, &x->c1);
    fq_add(&s1, &y->c0, &y->c1);
    fq_mul(&c1, &s0, &s1);
    fq_sub(&c1, &c1, &aA);
    fq_sub(&c1, &c1, &bB);
    r->c0 = c0;
    r->c1 = c1;
}
/* fp2.tcc:109-118 complex squaring: ((a+b)(a+nr*b) - ab - nr*ab, 2ab) */
static inline void fq2_sqr(fe2 *r, const fe2 *x)
{
    fe ab, nb, s, t, nab, c0, c1;
    fq_mul(&ab, &x->c0, &x->c1);
    fq_mul(&nb, &FQ_NONRES, &x->c1);
    fq_add(&s, &x->c0, &x->c1);
    fq_add(&t, &x->c0, &nb);
    fq_mul(&c0, &s, &t);
    fq_sub(&c0, &c0, &ab);
    fq_mul(&nab, &FQ_NONRES, &ab);
    fq_sub(&c0, &c0, &nab);
    fq_add(&c1, &ab, &ab);
    r->c0 = c0;
    r->c1 = c1;
}
/* 1/(a + bu) = (a - bu) / (a^2 - nr*b^2) */
static inline void fq2_inv(fe2 *r, const fe2 *x)
{
    fe t0, t1, t, ti;
    fq_sqr(&t0, &x->c0);
    fq_sqr(&t1, &x->c1);
    fq_mul(&t1, &FQ_NONRES, &t1);
    fq_sub(&t, &t0, &t1);
    fq_inv(&ti, &t);
    fq_mul(&r->c0, &x->c0, &ti);
    fq_mul(&t, &x->c1, &ti);
    fq_neg(&r->c1, &t);
}
static inline void fq2_zero(fe2 *r) { memset(r, 0, sizeof(*r)); }
static inline void fq2_one(fe2 *r)
{
    fq_one(&r->c0);
    memset(&r->c1, 0, 32);
}

/* ------------------------------------------------------------------- groups */
/* One body for G1 (F = fe, ops fq_*) and G2 (F = fe2, ops fq2_*). */
#define DEFINE_GROUP(G, F, OP)                                                                \
    typedef struct {                                                                          \
        F X, Y, Z;                                                                            \
    } G##_t;                                                                                  \
    static inline int G##_is_zero(const G##_t *p) { return OP##_is_zero(&p->Z); }             \
    static inline void G##_zero(G##_t *p)                                                     \
    {                                                                                         \
        memset(p, 0, sizeof(*p));                                                             \
        OP##_one(&p->Y);                                                                      \
    }                                                                                         \
    /* dbl-2009-l (alt_bn128_g1.cpp:329-362) */                                               \
    static inline void G##_dbl(G##_t *r, const G##_t *p)                                      \
    {                                                                                         \
        if (G##_is_zero(p)) {                                                                 \
            *r = *p;                                                                          \
            return;                                                                           \
        }                                                                                     \
        F A, B, C, D, E, Fv, X3, Y3, Z3, t, eightC;                                           \
        OP##_sqr(&A, &p->X);                                                                  \
        OP##_sqr(&B, &p->Y);                                                                  \
        OP##_sqr(&C, &B);                                                                     \
        OP##_add(&t, &p->X, &B);                                                              \
        OP##_sqr(&D, &t);                                                                     \
        OP##_sub(&D, &D, &A);                                                                 \
        OP##_sub(&D, &D, &C);                                                                 \
        OP##_add(&D, &D, &D);                                                                 \
        OP##_add(&E, &A, &A);                                                                 \
        OP##_add(&E, &E, &A);                                                                 \
        OP##_sqr(&Fv, &E);                                                                    \
        OP##_add(&t, &D, &D);                                                                 \
        OP##_sub(&X3, &Fv, &t);                                                               \
        OP##_add(&eightC, &C, &C);                                                            \
        OP##_add(&eightC, &eightC, &eightC);                                                  \
        OP##_add(&eightC, &eightC, &eightC);                                                  \
        OP##_sub(&t, &D, &X3);                                                                \
        OP##_mul(&Y3, &E, &t);                                                                \
        OP##_sub(&Y3, &Y3, &eightC);                                                          \
        OP##_mul(&t, &p->Y, &p->Z);                                                           \
        OP##_add(&Z3, &t, &t);                                                                \
        r->X = X3;                                                                            \
        r->Y = Y3;                                                                            \
        r->Z = Z3;                                                                            \
    }                                                                                         \
    /* operator+ (alt_bn128_g1.cpp:143-210): add-2007-bl, doubling detected */                \
    static inline void G##_add(G##_t *r, const G##_t *p, const G##_t *q)                      \
    {                                                                                         \
        if (G##_is_zero(p)) {                                                                 \
            *r = *q;                                                                          \
            return;                                                                           \
        }                                                                                     \
        if (G##_is_zero(q)) {                                                                 \
            *r = *p;                                                                          \
            return;                                                                           \
        }                                                                                     \
        F Z1Z1, Z2Z2, U1, U2, Z1c, Z2c, S1, S2, H, S2mS1, I, J, rr, V, X3, Y3, Z3, t, S1J;    \
        OP##_sqr(&Z1Z1, &p->Z);                                                               \
        OP##_sqr(&Z2Z2, &q->Z);                                                               \
        OP##_mul(&U1, &p->X, &Z2Z2);                                                          \
        OP##_mul(&U2, &q->X, &Z1Z1);                                                          \
        OP##_mul(&Z1c, &p->Z, &Z1Z1);                                                         \
        OP##_mul(&Z2c, &q->Z, &Z2Z2);                                                         \
        OP##_mul(&S1, &p->Y, &Z2c);                                                           \
        OP##_mul(&S2, &q->Y, &Z1c);                                                           \
        if (OP##_eq(&U1, &U2) && OP##_eq(&S1, &S2)) {                                         \
            G##_dbl(r, p);                                                                    \
            return;                                                                           \
        }                                                                                     \
        OP##_sub(&H, &U2, &U1);                                                               \
        OP##_sub(&S2mS1, &S2, &S1);                                                           \
        OP##_add(&t, &H, &H);                                                                 \
        OP##_sqr(&I, &t);                                                                     \
        OP##_mul(&J, &H, &I);                                                                 \
        OP##_add(&rr, &S2mS1, &S2mS1);                                                        \
        OP##_mul(&V, &U1, &I);                                                                \
        OP##_sqr(&X3, &rr);                                                                   \
        OP##_sub(&X3, &X3, &J);                                                               \
        OP##_add(&t, &V, &V);                                                                 \
        OP##_sub(&X3, &X3, &t);                                                               \
        OP##_mul(&S1J, &S1, &J);                                                              \
        OP##_sub(&t, &V, &X3);                                                                \
        OP##_mul(&Y3, &rr, &t);                                                               \
        OP##_add(&t, &S1J, &S1J);                                                             \
        OP##_sub(&Y3, &Y3, &t);                                                               \
        OP##_add(&t, &p->Z, &q->Z);                                                           \
        OP##_sqr(&Z3, &t);                                                                    \
        OP##_sub(&Z3, &Z3, &Z1Z1);                                                            \
        OP##_sub(&Z3, &Z3, &Z2Z2);                                                            \
        OP##_mul(&Z3, &Z3, &H);                                                               \
        r->X = X3;                                                                            \
        r->Y = Y3;                                                                            \
        r->Z = Z3;                                                                            \
    }                                                                                         \
    static inline void G##_neg(G##_t *r, const G##_t *p)                                      \
    {                                                                                         \
        r->X = p->X;                                                                          \
        OP##_neg(&r->Y, &p->Y);                                                               \
        r->Z = p->Z;                                                                          \
    }                                                                                         \
    /* to_affine_coordinates (alt_bn128_g1.cpp:64-80) */                                      \
    static inline void G##_to_affine(G##_t *r, const G##_t *p)                                \
    {                                                                                         \
        if (G##_is_zero(p)) {                                                                 \
            G##_zero(r);                                                                      \
            return;                                                                           \
        }                                                                                     \
        F zi, z2, z3;                                                                         \
        OP##_inv(&zi, &p->Z);                                                                 \
        OP##_sqr(&z2, &zi);                                                                   \
        OP##_mul(&z3, &z2, &zi);                                                              \
        OP##_mul(&r->X, &p->X, &z2);                                                          \
        OP##_mul(&r->Y, &p->Y, &z3);                                                          \
        OP##_one(&r->Z);                                                                      \
    }                                                                                         \
    /* scalar_mul (curve_utils.tcc:14-37) over the 256 bits of bigint<4> */                   \
    static inline void G##_mul_bigint(G##_t *r, const G##_t *base, const uint64_t k[4])       \
    {                                                                                         \
        G##_t acc;                                                                            \
        G##_zero(&acc);                                                                       \
        int found = 0;                                                                        \
        for (int i = 255; i >= 0; --i) {                                                      \
            if (found) G##_dbl(&acc, &acc);                                                   \
            if ((k[i / 64] >> (i % 64)) & 1) {                                                \
                found = 1;                                                                    \
                G##_add(&acc, &acc, base);                                                    \
            }                                                                                 \
        }                                                                                     \
        *r = acc;                                                                             \
    }                                                                                         \
    /* multi_exp_inner<BDLO12> (multiexp.tcc:165-282) */                                      \
    static void G##_msm_bdlo12(G##_t *out, const G##_t *bases, const fe *scalars, size_t n)   \
    {                                                                                         \
        G##_t result;                                                                         \
        G##_zero(&result);                                                                    \
        if (n == 0) {                                                                         \
            *out = result;                                                                    \
            return;                                                                           \
        }                                                                                     \
        const size_t log2n = or_log2(n);                                                      \
        const size_t c = log2n - (log2n / 3 - 2);                                             \
        fe *bn = (fe *)malloc(n * sizeof(fe));                                                \
        size_t num_bits = 0;                                                                  \
        for (size_t i = 0; i < n; ++i) {                                                      \
            fr_from_mont(&bn[i], &scalars[i]);                                                \
            const size_t b = bigint_num_bits(bn[i].v);                                        \
            if (b > num_bits) num_bits = b;                                                   \
        }                                                                                     \
        const size_t num_groups = (num_bits + c - 1) / c;                                     \
        const size_t nb = (size_t)1 << c;                                                     \
        G##_t *buckets = (G##_t *)malloc(nb * sizeof(G##_t));                                 \
        unsigned char *nonzero = (unsigned char *)malloc(nb);                                 \
        int result_nonzero = 0;                                                               \
        for (size_t k = num_groups - 1; k <= num_groups; k--) {                               \
            if (result_nonzero)                                                               \
                for (size_t i = 0; i < c; ++i) G##_dbl(&result, &result);                     \
            memset(nonzero, 0, nb);                                                           \
            for (size_t i = 0; i < n; ++i) {                                                  \
                size_t id = 0;                                                                \
                for (size_t j = 0; j < c; ++j)                                                \
                    if (bigint_test_bit(bn[i].v, k * c + j)) id |= (size_t)1 << j;            \
                if (id == 0) continue;                                                        \
                if (nonzero[id])                                                              \
                    G##_add(&buckets[id], &buckets[id], &bases[i]);                           \
                else {                                                                        \
                    buckets[id] = bases[i];                                                   \
                    nonzero[id] = 1;                                                          \
                }                                                                             \
            }                                                                                 \
            G##_t running;                                                                    \
            int running_nonzero = 0;                                                          \
            for (size_t i = nb - 1; i > 0; i--) {                                             \
                if (nonzero[i]) {                                                             \
                    if (running_nonzero)                                                      \
                        G##_add(&running, &running, &buckets[i]);                             \
                    else {                                                                    \
                        running = buckets[i];                                                 \
                        running_nonzero = 1;                                                  \
                    }                                                                         \
                }                                                                             \
                if (running_nonzero) {                                                        \
                    if (result_nonzero)                                                       \
                        G##_add(&result, &result, &running);                                  \
                    else {                                                                    \
                        result = running;                                                     \
                        result_nonzero = 1;                                                   \
                    }                                                                         \
                }                                                                             \
            }                                                                                 \
        }                                                                                     \
        free(bn);                                                                             \
        free(buckets);                                                                        \
        free(nonzero);                                                                        \
        *out = result;                                                                        \
    }

/* libff::log2: ceil(log2(n)) (common/utils.cpp:32-45) */
static size_t or_log2(size_t n)
{
    size_t r = ((n & (n - 1)) == 0 ? 0 : 1);
    while (n > 1) {
        n >>= 1;
        r++;
    }
    return r;
}
static size_t bigint_num_bits(const uint64_t v[4])
{
    for (int w = 3; w >= 0; --w)
        if (v[w]) return (size_t)w * 64 + 64 - (size_t)__builtin_clzll(v[w]);
    return 0;
}
static int bigint_test_bit(const uint64_t v[4], size_t bit)
{
    if (bit >= 256) return 0;
    return (int)((v[bit / 64] >> (bit % 64)) & 1);
}

DEFINE_GROUP(g1, fe, fq)
DEFINE_GROUP(g2, fe2, fq2)

/* ------------------------------------------------------------ initialisation */
static pthread_once_t g_init_once = PTHREAD_ONCE_INIT;
static void or_init_once(void)
{
    fe one;
    fq_one(&one);
    fq_neg(&FQ_NONRES, &one);
}
static void or_init(void) { pthread_once(&g_init_once, or_init_once); }

/* ------------------------------------------------------------------ mt19937_64 */
typedef struct {
    uint64_t mt[312];
    int idx;
} mt64;
static void mt64_seed(mt64 *s, uint64_t seed)
{
    s->mt[0] = seed;
    for (int i = 1; i < 312; ++i)
        s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
    s->idx = 312;
}
static uint64_t mt64_next(mt64 *s)
{
    if (s->idx >= 312) {
        for (int i = 0; i < 312; ++i) {
            const uint64_t y = (s->mt[i] & 0xFFFFFFFF80000000ULL) | (s->mt[(i + 1) % 312] & 0x7FFFFFFFULL);
            uint64_t v = s->mt[(i + 156) % 312] ^ (y >> 1);
            if (y & 1) v ^= 0xB5026F5AA96619E9ULL;
            s->mt[i] = v;
        }
        s->idx = 0;
    }
    uint64_t x = s->mt[s->idx++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= x >> 43;
    return x;
}

void or_mt19937_64(uint64_t seed, uint64_t *out, size_t n)
{
    mt64 s;
    mt64_seed(&s, seed);
    for (size_t i = 0; i < n; ++i) out[i] = mt64_next(&s);
}

/* fp.tcc:719-744: draw 4 words, clear bits above the modulus MSB (254, 255),
 * reject if >= r; the words are taken as the Montgomery residue. */
void or_fr_random(uint64_t seed, void *out, size_t n)
{
    mt64 s;
    mt64_seed(&s, seed);
    fe *o = (fe *)out;
    for (size_t i = 0; i < n; ++i) {
        for (;;) {
            fe x;
            for (int w = 0; w < 4; ++w) x.v[w] = mt64_next(&s);
            x.v[3] &= 0x3FFFFFFFFFFFFFFFULL;
            int lt = 0;
            for (int w = 3; w >= 0; --w)
                if (x.v[w] != R_P[w]) {
                    lt = x.v[w] < R_P[w];
                    break;
                }
            if (lt) {
                o[i] = x;
                break;
            }
        }
    }
}

/* ---------------------------------------------------------- element batches */
#define BATCH2(NAME, T, FN)                                                   \
    void NAME(const void *a, const void *b, void *out, size_t n)              \
    {                                                                         \
        or_init();                                                            \
        const T *x = (const T *)a, *y = (const T *)b;                         \
        T *o = (T *)out;                                                      \
        for (size_t i = 0; i < n; ++i) FN(&o[i], &x[i], &y[i]);               \
    }
#define BATCH1(NAME, T, FN)                                                   \
    void NAME(const void *a, void *out, size_t n)                             \
    {                                                                         \
        or_init();                                                            \
        const T *x = (const T *)a;                                            \
        T *o = (T *)out;                                                      \
        for (size_t i = 0; i < n; ++i) FN(&o[i], &x[i]);                      \
    }
BATCH2(or_fr_mul, fe, fr_mul)
BATCH2(or_fr_add, fe, fr_add)
BATCH2(or_fr_sub, fe, fr_sub)
BATCH1(or_fr_inv, fe, fr_inv)
BATCH1(or_fr_to_mont, fe, fr_to_mont)
BATCH1(or_fr_from_mont, fe, fr_from_mont)
BATCH2(or_fq_mul, fe, fq_mul)
BATCH2(or_fq_add, fe, fq_add)
BATCH2(or_fq_sub, fe, fq_sub)
BATCH1(or_fq_to_mont, fe, fq_to_mont)
BATCH1(or_fq_from_mont, fe, fq_from_mont)
BATCH2(or_fq2_mul, fe2, fq2_mul)
BATCH1(or_fq2_sqr, fe2, fq2_sqr)

/* ------------------------------------------------------------------ groups API */
void or_g1_generator(void *out96)
{
    or_init();
    g1_t g;
    fq_set_u64(&g.X, 1);
    fq_set_u64(&g.Y, 2);
    fq_one(&g.Z);
    memcpy(out96, &g, sizeof(g));
}

/* alt_bn128_init.cpp:207-215, coordinates as integers (little-endian words) */
void or_g2_generator(void *out192)
{
    or_init();
    static const uint64_t x0[4] = {0x46debd5cd992f6edULL, 0x674322d4f75edaddULL, 0x426a00665e5c4479ULL,
                                   0x1800deef121f1e76ULL};
    static const uint64_t x1[4] = {0x97e485b7aef312c2ULL, 0xf1aa493335a9e712ULL, 0x7260bfb731fb5d25ULL,
                                   0x198e9393920d483aULL};
    static const uint64_t y0[4] = {0x4ce6cc0166fa7daaULL, 0xe3d1e7690c43d37bULL, 0x4aab71808dcb408fULL,
                                   0x12c85ea5db8c6debULL};
    static const uint64_t y1[4] = {0x55acdadcd122975bULL, 0xbc4b313370b38ef3ULL, 0xec9e99ad690c3395ULL,
                                   0x090689d0585ff075ULL};
    g2_t g;
    fe t;
    memcpy(t.v, x0, 32);
    fq_to_mont(&g.X.c0, &t);
    memcpy(t.v, x1, 32);
    fq_to_mont(&g.X.c1, &t);
    memcpy(t.v, y0, 32);
    fq_to_mont(&g.Y.c0, &t);
    memcpy(t.v, y1, 32);
    fq_to_mont(&g.Y.c1, &t);
    fq2_one(&g.Z);
    memcpy(out192, &g, sizeof(g));
}

void or_g1_to_affine(const void *in96, void *out96, size_t n)
{
    or_init();
    const g1_t *p = (const g1_t *)in96;
    g1_t *o = (g1_t *)out96;
    for (size_t i = 0; i < n; ++i) g1_to_affine(&o[i], &p[i]);
}
void or_g2_to_affine(const void *in192, void *out192, size_t n)
{
    or_init();
    const g2_t *p = (const g2_t *)in192;
    g2_t *o = (g2_t *)out192;
    for (size_t i = 0; i < n; ++i) g2_to_affine(&o[i], &p[i]);
}
void or_g1_add(const void *a96, const void *b96, void *out96)
{
    or_init();
    g1_add((g1_t *)out96, (const g1_t *)a96, (const g1_t *)b96);
}
void or_g2_add(const void *a192, const void *b192, void *out192)
{
    or_init();
    g2_add((g2_t *)out192, (const g2_t *)a192, (const g2_t *)b192);
}
void or_g1_dbl(const void *a96, void *out96)
{
    or_init();
    g1_dbl((g1_t *)out96, (const g1_t *)a96);
}
void or_g1_neg(const void *a96, void *out96)
{
    or_init();
    g1_neg((g1_t *)out96, (const g1_t *)a96);
}
void or_g1_mul(const void *base96, const void *scalar, void *out96, size_t n)
{
    or_init();
    const g1_t *b = (const g1_t *)base96;
    const fe *s = (const fe *)scalar;
    g1_t *o = (g1_t *)out96;
    for (size_t i = 0; i < n; ++i) {
        fe k;
        fr_from_mont(&k, &s[i]);
        g1_mul_bigint(&o[i], &b[i], k.v);
    }
}
void or_g2_mul(const void *base192, const void *scalar, void *out192, size_t n)
{
    or_init();
    const g2_t *b = (const g2_t *)base192;
    const fe *s = (const fe *)scalar;
    g2_t *o = (g2_t *)out192;
    for (size_t i = 0; i < n; ++i) {
        fe k;
        fr_from_mont(&k, &s[i]);
        g2_mul_bigint(&o[i], &b[i], k.v);
    }
}

/* y^2 Z^0 = x^3 + b on the affine image (alt_bn128_init.cpp:134-136 twist b' = 3/(9+u)) */
int or_g1_on_curve(const void *p96)
{
    or_init();
    g1_t a;
    g1_to_affine(&a, (const g1_t *)p96);
    if (g1_is_zero(&a)) return 1;
    fe y2, x3, b;
    fq_sqr(&y2, &a.Y);
    fq_sqr(&x3, &a.X);
    fq_mul(&x3, &x3, &a.X);
    fq_set_u64(&b, 3);
    fq_add(&x3, &x3, &b);
    return fq_eq(&y2, &x3);
}
int or_g2_on_curve(const void *p192)
{
    or_init();
    g2_t a;
    g2_to_affine(&a, (const g2_t *)p192);
    if (g2_is_zero(&a)) return 1;
    fe2 y2, x3, tw, twi, b;
    fq2_sqr(&y2, &a.Y);
    fq2_sqr(&x3, &a.X);
    fq2_mul(&x3, &x3, &a.X);
    fq_set_u64(&tw.c0, 9);
    fq_set_u64(&tw.c1, 1);
    fq2_inv(&twi, &tw);
    fe2 three;
    fq_set_u64(&three.c0, 3);
    memset(&three.c1, 0, 32);
    fq2_mul(&b, &three, &twi);
    fq2_add(&x3, &x3, &b);
    return fq2_eq(&y2, &x3);
}

/* -------------------------------------------------------------------- SpMV */
/* backend.cpp:25-35: row sums in the order of the row's entries, from zero. */
void or_spmv(const or_csr *m, const void *z, void *out)
{
    or_init();
    const fe *vals = (const fe *)m->values;
    const fe *x = (const fe *)z;
    fe *y = (fe *)out;
    for (size_t row = 0; row < m->rows; ++row) {
        fe acc = {{0, 0, 0, 0}};
        for (uint32_t j = m->row_ptr[row]; j < m->row_ptr[row + 1]; ++j) {
            fe t;
            fr_mul(&t, &vals[j], &x[m->col_idx[j]]);
            fr_add(&acc, &acc, &t);
        }
        y[row] = acc;
    }
}

/* --------------------------------------------------------------------- FFT */
static size_t bitreverse(size_t n, size_t l)
{
    size_t r = 0;
    for (size_t k = 0; k < l; ++k) {
        r = (r << 1) | (n & 1);
        n >>= 1;
    }
    return r;
}

/* field_utils.tcc:36-52: root_of_unity squared (s - logn) times. */
static void fr_root_of_unity(fe *omega, size_t logn)
{
    /* 19103219067921713944291392827692070036145651957329286315305642004821462161904 */
    static const uint64_t ROOT[4] = {0x9bd61b6e725b19f0ULL, 0x402d111e41112ed4ULL, 0x00e0a7eb8ef62abcULL,
                                     0x2a3c09f0a58a7e85ULL};
    fe t;
    memcpy(t.v, ROOT, 32);
    fr_to_mont(omega, &t);
    for (size_t i = FR_S; i > logn; --i) fr_sqr(omega, omega);
}

/* Fp_model::operator^(unsigned long) */
static void fr_pow_u64(fe *r, const fe *x, uint64_t e)
{
    fr_pow(r, x, &e, 1);
}

/* _basic_serial_radix2_FFT (basic_radix2_domain_aux.tcc:45-76) */
static void serial_radix2_fft(fe *a, size_t n, const fe *omega)
{
    const size_t logn = or_log2(n);
    for (size_t k = 0; k < n; ++k) {
        const size_t rk = bitreverse(k, logn);
        if (k < rk) {
            fe t = a[k];
            a[k] = a[rk];
            a[rk] = t;
        }
    }
    size_t m = 1;
    for (size_t s = 1; s <= logn; ++s) {
        fe w_m;
        fr_pow_u64(&w_m, omega, n / (2 * m));
        for (size_t k = 0; k < n; k += 2 * m) {
            fe w;
            fr_one(&w);
            for (size_t j = 0; j < m; ++j) {
                fe t;
                fr_mul(&t, &w, &a[k + j + m]);
                fr_sub(&a[k + j + m], &a[k + j], &t);
                fr_add(&a[k + j], &a[k + j], &t);
                fr_mul(&w, &w, &w_m);
            }
        }
        m *= 2;
    }
}

/* _multiply_by_coset (basic_radix2_domain_aux.tcc:172-180) */
static void multiply_by_coset(fe *a, size_t n, const fe *g)
{
    fe u = *g;
    for (size_t i = 1; i < n; ++i) {
        fr_mul(&a[i], &a[i], &u);
        fr_mul(&u, &u, g);
    }
}

static void domain_fft(fe *a, size_t n)
{
    fe omega;
    fr_root_of_unity(&omega, or_log2(n));
    serial_radix2_fft(a, n, &omega);
}
static void domain_ifft(fe *a, size_t n)
{
    fe omega, oi, sconst, nn;
    fr_root_of_unity(&omega, or_log2(n));
    fr_inv(&oi, &omega);
    serial_radix2_fft(a, n, &oi);
    fr_set_u64(&nn, (uint64_t)n);
    fr_inv(&sconst, &nn);
    for (size_t i = 0; i < n; ++i) fr_mul(&a[i], &a[i], &sconst);
}
static void coset_gen(fe *g) { fr_set_u64(g, 5); } /* multiplicative_generator */

int or_fft(void *data, size_t n, int kind)
{
    or_init();
    if (n < 2 || (n & (n - 1)) != 0 || or_log2(n) > FR_S) return -1;
    fe *a = (fe *)data;
    fe g, gi;
    coset_gen(&g);
    switch (kind) {
    case 0: domain_fft(a, n); break;
    case 1: domain_ifft(a, n); break;
    case 2:
        multiply_by_coset(a, n, &g);
        domain_fft(a, n);
        break;
    case 3:
        domain_ifft(a, n);
        fr_inv(&gi, &g);
        multiply_by_coset(a, n, &gi);
        break;
    default: return -1;
    }
    return 0;
}

/* backend.cpp:171-179 */
int or_quotient(const void *az_in, const void *bz_in, const void *cz_in, size_t n, void *h)
{
    or_init();
    if (n < 2 || (n & (n - 1)) != 0 || or_log2(n) > FR_S) return -1;
    fe *a = (fe *)h;
    fe *b = (fe *)malloc(n * sizeof(fe));
    fe *c = (fe *)malloc(n * sizeof(fe));
    memcpy(a, az_in, n * sizeof(fe));
    memcpy(b, bz_in, n * sizeof(fe));
    memcpy(c, cz_in, n * sizeof(fe));
    domain_ifft(a, n);
    domain_ifft(b, n);
    domain_ifft(c, n);
    or_fft(a, n, 2);
    or_fft(b, n, 2);
    or_fft(c, n, 2);
    for (size_t i = 0; i < n; ++i) {
        fr_mul(&a[i], &a[i], &b[i]);
        fr_sub(&a[i], &a[i], &c[i]);
    }
    /* divide_by_Z_on_coset (basic_radix2_domain.tcc:104-112): Z(g) = g^n - 1 */
    fe g, gn, one, zi;
    coset_gen(&g);
    fr_pow_u64(&gn, &g, (uint64_t)n);
    fr_one(&one);
    fr_sub(&gn, &gn, &one);
    fr_inv(&zi, &gn);
    for (size_t i = 0; i < n; ++i) fr_mul(&a[i], &a[i], &zi);
    or_fft(a, n, 3);
    free(b);
    free(c);
    return 0;
}

/* -------------------------------------------------------------------- MSM */
int or_msm_g1(const void *bases96, const void *scalars, size_t n, void *out96)
{
    or_init();
    g1_msm_bdlo12((g1_t *)out96, (const g1_t *)bases96, (const fe *)scalars, n);
    return 0;
}
int or_msm_g2(const void *bases192, const void *scalars, size_t n, void *out192)
{
    or_init();
    g2_msm_bdlo12((g2_t *)out192, (const g2_t *)bases192, (const fe *)scalars, n);
    return 0;
}

/* ------------------------------------------------------------------ Groth16 */
/* CpuBackend::prove (backend.cpp:156-201) */
int or_prove(const or_circuit *c, const or_pk *pk, const void *witness, void *raw_out, void *az_out,
             void *bz_out, void *cz_out, void *h_out)
{
    or_init();
    const size_t n = c->a.rows, m = c->a.cols;
    if (n < 2 || (n & (n - 1)) != 0 || or_log2(n) > FR_S) return -1;
    fe *az = (fe *)malloc(n * sizeof(fe));
    fe *bz = (fe *)malloc(n * sizeof(fe));
    fe *cz = (fe *)malloc(n * sizeof(fe));
    fe *h = (fe *)malloc(n * sizeof(fe));
    or_spmv(&c->a, witness, az);
    or_spmv(&c->b, witness, bz);
    or_spmv(&c->c, witness, cz);
    if (az_out) memcpy(az_out, az, n * sizeof(fe));
    if (bz_out) memcpy(bz_out, bz, n * sizeof(fe));
    if (cz_out) memcpy(cz_out, cz, n * sizeof(fe));
    or_quotient(az, bz, cz, n, h);
    if (h_out) memcpy(h_out, h, n * sizeof(fe));
    unsigned char *o = (unsigned char *)raw_out;
    g1_msm_bdlo12((g1_t *)(o + 0), (const g1_t *)pk->a_query, (const fe *)witness, m);
    g1_msm_bdlo12((g1_t *)(o + 96), (const g1_t *)pk->b_g1_query, (const fe *)witness, m);
    g2_msm_bdlo12((g2_t *)(o + 384), (const g2_t *)pk->b_g2_query, (const fe *)witness, m);
    g1_msm_bdlo12((g1_t *)(o + 192), (const g1_t *)pk->l_query, (const fe *)witness, m);
    g1_msm_bdlo12((g1_t *)(o + 288), (const g1_t *)pk->h_query, h, n);
    free(az);
    free(bz);
    free(cz);
    free(h);
    return 0;
}

/* blind_proof (protocol.cpp:85-94) with r, s given instead of sampled. */
void or_blind(const or_pk *pk, const void *raw, const void *r, const void *s, void *proof_out)
{
    or_init();
    const unsigned char *rw = (const unsigned char *)raw;
    const g1_t *ar = (const g1_t *)(rw + 0), *bs1 = (const g1_t *)(rw + 96), *zk = (const g1_t *)(rw + 192),
               *qz = (const g1_t *)(rw + 288);
    const g2_t *bs2 = (const g2_t *)(rw + 384);
    const g1_t *alpha = (const g1_t *)pk->alpha_g1, *beta1 = (const g1_t *)pk->beta_g1,
               *delta1 = (const g1_t *)pk->delta_g1;
    const g2_t *beta2 = (const g2_t *)pk->beta_g2, *delta2 = (const g2_t *)pk->delta_g2;
    fe rb, sb, rs, rsb;
    fr_from_mont(&rb, (const fe *)r);
    fr_from_mont(&sb, (const fe *)s);
    fr_mul(&rs, (const fe *)r, (const fe *)s);
    fr_from_mont(&rsb, &rs);

    g1_t t, a, b1, cc;
    /* a = alpha + ar + r*delta1 */
    g1_add(&a, alpha, ar);
    g1_mul_bigint(&t, delta1, rb.v);
    g1_add(&a, &a, &t);
    /* b_g1 = beta1 + bs1 + s*delta1 */
    g1_add(&b1, beta1, bs1);
    g1_mul_bigint(&t, delta1, sb.v);
    g1_add(&b1, &b1, &t);
    /* b = beta2 + bs2 + s*delta2 */
    g2_t b, t2;
    g2_add(&b, beta2, bs2);
    g2_mul_bigint(&t2, delta2, sb.v);
    g2_add(&b, &b, &t2);
    /* c = qz + zk + s*a + r*b_g1 - (r*s)*delta1 */
    g1_add(&cc, qz, zk);
    g1_mul_bigint(&t, &a, sb.v);
    g1_add(&cc, &cc, &t);
    g1_mul_bigint(&t, &b1, rb.v);
    g1_add(&cc, &cc, &t);
    g1_mul_bigint(&t, delta1, rsb.v);
    g1_neg(&t, &t);
    g1_add(&cc, &cc, &t);
    unsigned char *o = (unsigned char *)proof_out;
    memcpy(o, &a, 96);
    memcpy(o + 96, &b, 192);
    memcpy(o + 288, &cc, 96);
}

/* ------------------------------------------------------------ all host cores */
/* The reference's parallel MSM (operators/include/nkr/operators/
 * variable_base_msm.hpp:17-49): workers = min(n, threads), chunk =
 * ceil(n / workers); each worker runs single-threaded BDLO12 on its slice and
 * the partial sums are added in worker order.  Here the chunks of several
 * MSMs share one pool of `threads` POSIX threads taking tasks in order. */
typedef struct {
    int g2;           /* 0: G1 (96 B points), 1: G2 (192 B) */
    const void *bases;
    const fe *scalars;
    size_t n;
    void *out;        /* final sum */
    size_t workers, chunk;
    void *partial;    /* workers points */
} msm_job;

typedef struct {
    msm_job *jobs;
    size_t njobs, ntasks;
    size_t next;
    pthread_mutex_t lock;
} msm_pool;

static void msm_job_init(msm_job *j, int g2, const void *bases, const fe *scalars, size_t n, void *out,
                         size_t threads)
{
    j->g2 = g2;
    j->bases = bases;
    j->scalars = scalars;
    j->n = n;
    j->out = out;
    j->workers = n == 0 ? 0 : (threads < n ? (threads ? threads : 1) : n);
    j->chunk = j->workers ? (n + j->workers - 1) / j->workers : 0;
    j->partial = calloc(j->workers ? j->workers : 1, g2 ? sizeof(g2_t) : sizeof(g1_t));
}

static void *msm_pool_worker(void *arg)
{
    msm_pool *p = (msm_pool *)arg;
    for (;;) {
        pthread_mutex_lock(&p->lock);
        const size_t t = p->next++;
        pthread_mutex_unlock(&p->lock);
        if (t >= p->ntasks) return NULL;
        size_t k = t, j = 0;
        while (k >= p->jobs[j].workers) k -= p->jobs[j++].workers;
        msm_job *jb = &p->jobs[j];
        const size_t begin = k * jb->chunk;
        const size_t end = begin + jb->chunk < jb->n ? begin + jb->chunk : jb->n;
        if (jb->g2) {
            g2_t *part = (g2_t *)jb->partial + k;
            if (begin >= end)
                g2_zero(part);
            else
                g2_msm_bdlo12(part, (const g2_t *)jb->bases + begin, jb->scalars + begin, end - begin);
        } else {
            g1_t *part = (g1_t *)jb->partial + k;
            if (begin >= end)
                g1_zero(part);
            else
                g1_msm_bdlo12(part, (const g1_t *)jb->bases + begin, jb->scalars + begin, end - begin);
        }
    }
}

static void msm_run_jobs(msm_job *jobs, size_t njobs, size_t threads)
{
    msm_pool p;
    p.jobs = jobs;
    p.njobs = njobs;
    p.ntasks = 0;
    p.next = 0;
    for (size_t j = 0; j < njobs; ++j) p.ntasks += jobs[j].workers;
    pthread_mutex_init(&p.lock, NULL);
    size_t nt = threads ? threads : 1;
    if (nt > p.ntasks) nt = p.ntasks ? p.ntasks : 1;
    pthread_t *th = (pthread_t *)malloc(nt * sizeof(pthread_t));
    size_t started = 0;
    for (; started + 1 < nt; ++started)
        if (pthread_create(&th[started], NULL, msm_pool_worker, &p) != 0) break;
    msm_pool_worker(&p); /* the calling thread works too */
    for (size_t i = 0; i < started; ++i) pthread_join(th[i], NULL);
    free(th);
    pthread_mutex_destroy(&p.lock);
    for (size_t j = 0; j < njobs; ++j) {
        msm_job *jb = &jobs[j];
        if (jb->g2) {
            g2_t acc;
            g2_zero(&acc);
            for (size_t k = 0; k < jb->workers; ++k) g2_add(&acc, &acc, (const g2_t *)jb->partial + k);
            memcpy(jb->out, &acc, sizeof(acc));
        } else {
            g1_t acc;
            g1_zero(&acc);
            for (size_t k = 0; k < jb->workers; ++k) g1_add(&acc, &acc, (const g1_t *)jb->partial + k);
            memcpy(jb->out, &acc, sizeof(acc));
        }
        free(jb->partial);
        jb->partial = NULL;
    }
}

int or_msm_g1_mt(const void *bases96, const void *scalars, size_t n, void *out96, size_t threads)
{
    or_init();
    msm_job j;
    msm_job_init(&j, 0, bases96, (const fe *)scalars, n, out96, threads);
    msm_run_jobs(&j, 1, threads);
    return 0;
}

int or_msm_g2_mt(const void *bases192, const void *scalars, size_t n, void *out192, size_t threads)
{
    or_init();
    msm_job j;
    msm_job_init(&j, 1, bases192, (const fe *)scalars, n, out192, threads);
    msm_run_jobs(&j, 1, threads);
    return 0;
}

typedef struct {
    const or_csr *m;
    const void *z;
    fe *out;
    int transform; /* also iFFT then cosetFFT the product in place */
} poly_task;

static void *poly_worker(void *arg)
{
    poly_task *t = (poly_task *)arg;
    or_spmv(t->m, t->z, t->out);
    return NULL;
}

static void *coset_worker(void *arg)
{
    poly_task *t = (poly_task *)arg;
    const size_t n = t->m->rows;
    domain_ifft(t->out, n);
    or_fft(t->out, n, 2);
    return NULL;
}

/* or_prove with the work spread over `threads` host threads: the three
 * SpMVs and the three iFFT + cosetFFT pairs one thread per matrix, the
 * pointwise quotient and icosetFFT on the calling thread, then the five
 * MSMs chunked as variable_base_msm does.  Same results as or_prove. */
int or_prove_mt(const or_circuit *c, const or_pk *pk, const void *witness, void *raw_out, void *az_out,
                void *bz_out, void *cz_out, void *h_out, size_t threads)
{
    or_init();
    const size_t n = c->a.rows, m = c->a.cols;
    if (n < 2 || (n & (n - 1)) != 0 || or_log2(n) > FR_S) return -1;
    if (threads == 0) threads = 1;
    fe *poly[3];
    for (int k = 0; k < 3; ++k) poly[k] = (fe *)malloc(n * sizeof(fe));
    const or_csr *mats[3] = {&c->a, &c->b, &c->c};
    poly_task tasks[3];
    pthread_t th[3];
    int started[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        tasks[k].m = mats[k];
        tasks[k].z = witness;
        tasks[k].out = poly[k];
        started[k] = threads > 1 && pthread_create(&th[k], NULL, poly_worker, &tasks[k]) == 0;
        if (!started[k]) poly_worker(&tasks[k]);
    }
    for (int k = 0; k < 3; ++k)
        if (started[k]) pthread_join(th[k], NULL);
    void *outs[3] = {az_out, bz_out, cz_out};
    for (int k = 0; k < 3; ++k)
        if (outs[k]) memcpy(outs[k], poly[k], n * sizeof(fe));
    for (int k = 0; k < 3; ++k) {
        started[k] = threads > 1 && pthread_create(&th[k], NULL, coset_worker, &tasks[k]) == 0;
        if (!started[k]) coset_worker(&tasks[k]);
    }
    for (int k = 0; k < 3; ++k)
        if (started[k]) pthread_join(th[k], NULL);
    fe *a = poly[0], *b = poly[1], *cc = poly[2];
    for (size_t i = 0; i < n; ++i) {
        fr_mul(&a[i], &a[i], &b[i]);
        fr_sub(&a[i], &a[i], &cc[i]);
    }
    fe g, gn, one, zi;
    coset_gen(&g);
    fr_pow_u64(&gn, &g, (uint64_t)n);
    fr_one(&one);
    fr_sub(&gn, &gn, &one);
    fr_inv(&zi, &gn);
    for (size_t i = 0; i < n; ++i) fr_mul(&a[i], &a[i], &zi);
    or_fft(a, n, 3);
    if (h_out) memcpy(h_out, a, n * sizeof(fe));
    unsigned char *o = (unsigned char *)raw_out;
    const fe *w = (const fe *)witness;
    msm_job jobs[5];
    msm_job_init(&jobs[0], 1, pk->b_g2_query, w, m, o + 384, threads);
    msm_job_init(&jobs[1], 0, pk->a_query, w, m, o + 0, threads);
    msm_job_init(&jobs[2], 0, pk->b_g1_query, w, m, o + 96, threads);
    msm_job_init(&jobs[3], 0, pk->l_query, w, m, o + 192, threads);
    msm_job_init(&jobs[4], 0, pk->h_query, a, n, o + 288, threads);
    msm_run_jobs(jobs, 5, threads);
    for (int k = 0; k < 3; ++k) free(poly[k]);
    return 0;
}

// CPU check of the host GLV scalar multiplication (nkr-zkp_amd/csrc/glv.h).
// stdin: lines of 16 words (decimal u64): a G1 point as libff Jacobian
// Montgomery words X[4] Y[4] Z[4], then a Montgomery Fr scalar k[4].
// stdout: per line, k*P as 12 words in the same layout, normalised to Z = 1
// (or all zero for the point at infinity).
// With the argument "pair": lines of 32 words, P, k, Q, l; per line three
// points, k*P + l*Q (g1_mul2_glv), k*P and l*Q (g1_mul_glv), normalised
// together by xyzz_to_jacobian_affine_batch (36 words).
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "glv.h"

using namespace nkr;

static void get(fq_t &f, const uint64_t *w)
{
    for (int i = 0; i < 4; ++i) {
        f.l[2 * i] = (uint32_t)w[i];
        f.l[2 * i + 1] = (uint32_t)(w[i] >> 32);
    }
}

static void put(const fq_t &f, uint64_t *w)
{
    for (int i = 0; i < 4; ++i) w[i] = (uint64_t)f.l[2 * i] | ((uint64_t)f.l[2 * i + 1] << 32);
}

static void put_point(const g1_jac &a, bool inf, uint64_t *out)
{
    std::memset(out, 0, 12 * sizeof(uint64_t));
    if (inf) return;
    put(a.X, out);
    put(a.Y, out + 4);
    put(a.Z, out + 8);
}

static void get_scalar(fr_t &k, const uint64_t *w)
{
    for (int i = 0; i < 4; ++i) {
        k.l[2 * i] = (uint32_t)w[i];
        k.l[2 * i + 1] = (uint32_t)(w[i] >> 32);
    }
}

static int pair_mode()
{
    uint64_t in[32];
    for (;;) {
        for (int i = 0; i < 32; ++i)
            if (std::scanf("%lu", (unsigned long *)&in[i]) != 1) return 0;
        g1_jac P, Q;
        get(P.X, in);
        get(P.Y, in + 4);
        get(P.Z, in + 8);
        get(Q.X, in + 16);
        get(Q.Y, in + 20);
        get(Q.Z, in + 24);
        fr_t k, l;
        get_scalar(k, in + 12);
        get_scalar(l, in + 28);
        const g1_xyzz xp = jacobian_to_xyzz(P), xq = jacobian_to_xyzz(Q);
        const g1_xyzz pts[3] = {g1_mul2_glv(xp, k, xq, l), g1_mul_glv(xp, k), g1_mul_glv(xq, l)};
        g1_jac aff[3];
        xyzz_to_jacobian_affine_batch(pts, aff, 3);
        uint64_t out[36];
        for (int i = 0; i < 3; ++i) put_point(aff[i], pts[i].is_inf(), out + 12 * i);
        for (int i = 0; i < 36; ++i) std::printf("%lu%c", (unsigned long)out[i], i == 35 ? '\n' : ' ');
    }
}

int main(int argc, char **argv)
{
    if (argc > 1 && std::strcmp(argv[1], "pair") == 0) return pair_mode();
    uint64_t in[16];
    for (;;) {
        for (int i = 0; i < 16; ++i)
            if (std::scanf("%lu", (unsigned long *)&in[i]) != 1) return 0;
        g1_jac P;
        get(P.X, in);
        get(P.Y, in + 4);
        get(P.Z, in + 8);
        fr_t k;
        for (int i = 0; i < 4; ++i) {
            k.l[2 * i] = (uint32_t)in[12 + i];
            k.l[2 * i + 1] = (uint32_t)(in[12 + i] >> 32);
        }
        const g1_xyzz r = g1_mul_glv(jacobian_to_xyzz(P), k);
        uint64_t out[12];
        std::memset(out, 0, sizeof(out));
        if (!r.is_inf()) {
            const g1_jac a = xyzz_to_jacobian_affine(r);
            put(a.X, out);
            put(a.Y, out + 4);
            put(a.Z, out + 8);
        }
        for (int i = 0; i < 12; ++i) std::printf("%lu%c", (unsigned long)out[i], i == 11 ? '\n' : ' ');
    }
}

// CPU check of the host GLV scalar multiplication (nkr-zkp_amd/csrc/glv.h).
// stdin: lines of 16 words (decimal u64): a G1 point as libff Jacobian
// Montgomery words X[4] Y[4] Z[4], then a Montgomery Fr scalar k[4].
// stdout: per line, k*P as 12 words in the same layout, normalised to Z = 1
// (or all zero for the point at infinity).
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

int main()
{
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

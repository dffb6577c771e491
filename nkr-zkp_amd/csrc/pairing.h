// BN254 optimal ate pairing, GT arithmetic and the Groth16 verifier (host).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "curve.h"

namespace nkr {

// Fq6 = Fq2[v]/(v^3 - (9+u)), Fq12 = Fq6[w]/(w^2 - v): the reference's tower
// (alt_bn128_init.cpp:95-134).  An Fq12 is stored as c0.c0 .. c0.c2, c1.c0 ..
// c1.c2, each an Fq2 (c0, c1) of Montgomery Fq elements: 384 bytes, the
// in-memory layout of alt_bn128_GT.
struct fq6_t {
    fq2_t c0, c1, c2;
};

struct fq12_t {
    fq6_t c0, c1;
};

static_assert(sizeof(fq12_t) == 384, "GT must match the reference's 384-byte Fq12 layout");

fq12_t gt_one();
fq12_t gt_mul(const fq12_t &a, const fq12_t &b);
fq12_t gt_inv(const fq12_t &a);
bool gt_eq(const fq12_t &a, const fq12_t &b);

// Miller function of the optimal ate pairing; 1 if either point is infinity.
fq12_t bn254_miller_loop(const g1_jac &P, const g2_jac &Q);

// f^(m (q^12 - 1) / r) with m = 2x(6x^2 + 3x + 1): the value the reference's
// final exponentiation produces (alt_bn128_pairing.cpp:152-237; its
// Fuentes-Castaneda chain equals this exponent exactly).
fq12_t bn254_final_exponentiation(const fq12_t &f);

// alt_bn128_reduced_pairing(P, Q) (alt_bn128_pairing.cpp:542-547).
fq12_t bn254_pairing(const g1_jac &P, const g2_jac &Q);

// prod_i reduced_pairing(P_i, Q_i) with one final exponentiation.
fq12_t bn254_pairing_product(const g1_jac *P, const g2_jac *Q, size_t k);

// Curve equations y^2 = x^3 + 3 and y^2 = x^3 + 3/(9+u) (infinity passes).
bool g1_on_curve(const g1_jac &p);
bool g2_on_curve(const g2_jac &p);
// r * Q == infinity for a G2 point on the twist.
bool g2_in_subgroup(const g2_jac &p);

// sum_i s_i * P_i on the host (Pippenger over batch-normalized bases), in the
// reference's affine Jacobian form.  For the verifier's input accumulation.
g1_jac g1_msm_host(const g1_jac *bases, const fr_t *scalars, size_t n);

struct VerifyingKeyView {
    const fq12_t *alpha_g1_beta_g2;
    const g2_jac *gamma_g2;
    const g2_jac *delta_g2;
    const g1_jac *ic;
    size_t ic_len;
};

// protocol.cpp:96-108: false when ic_len != n_public + 1, otherwise
// e(A, B) == alpha_g1_beta_g2 * e(IC, gamma) * e(C, delta) with
// IC = ic[0] + sum_i x_i ic[i+1].  Proof points that are off their curves
// (or B outside the order-r subgroup) are rejected.
bool groth16_verify(const VerifyingKeyView &vk, const fr_t *public_input, size_t n_public, const g1_jac &A,
                    const g2_jac &B, const g1_jac &C);

}  // namespace nkr

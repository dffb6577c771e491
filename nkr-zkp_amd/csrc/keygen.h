// Groth16 key generation on one GPU: the counterpart of Backend::setup
// (protocols/groth16/common/include/groth16/backend.hpp:20) as the CPU
// backend computes it (protocols/groth16/backends/cpu/src/backend.cpp:41-154)
// and the HIP backend accelerates it (backends/hip/src/backend.hip:111-268).
#pragma once

#include "../../include/nkr_g16.h"
#include "curve.h"

namespace nkr {

// t, alpha, beta, gamma, delta in the order the reference draws them
// (backend.cpp:48-52), Montgomery form.
struct Trapdoors {
    fr_t t, alpha, beta, gamma, delta;
};

// Lagrange basis at t over the 2^k-th roots of unity, the QAP column
// evaluations A^T L, B^T L, C^T L (transposed SpMV), the query scalars, then
// fixed-base multiples of the generators for every query; the key's
// alpha/beta/delta points; the verification key with alpha_g1_beta_g2 from
// the host pairing.  Outputs are host memory in the reference layouts,
// normalised (Z = 1, or (0, 1, 0)).
void groth16_keygen(const nkr_g16_circuit &circuit, const Trapdoors &td, int device, const nkr_g16_pk_out &pk,
                    const nkr_g16_vk_out &vk);

// The alt_bn128 generators (alt_bn128_init.cpp:148-150, 211-215), Montgomery.
g1_jac g1_generator();
g2_jac g2_generator();

}  // namespace nkr

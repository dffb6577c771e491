// Seeded R1CS generators: the circuit families of the reference's gen driver
// (protocols/groth16/app/gen.cpp:42-197) -- matmul, Fibonacci and Poseidon
// hashes -- with every value drawn from std::mt19937_64 streams of one seed
// instead of std::random_device, so a workload is reproducible.  The
// circuits keep the reference's constraint structure (R1CSManager ordering,
// protocols/groth16/common/include/r1cs/r1cs.hpp:26-160; explicit zero
// constant terms; rows and columns padded to powers of two) and equal,
// array for array, those of the Python generator nkr-zkp_amd/circuits.py.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "groth16.hpp"

namespace nkr::groth16::gen {

// The reference's input densities (common/include/density.hpp:55-64): the
// share of input entries that are non-zero, ceil(count * density) of them,
// at positions chosen by a seeded shuffle.
enum class Density { Full, Hybrid, Sparse };
Density parse_density(const std::string &name);
const char *density_name(Density d);
size_t density_nonzero_count(size_t count, Density d);

struct Instance {
    Circuit circuit;
    std::vector<Scalar> witness;  // z, z[0] = 1, Montgomery form
};

// C = A x B for s x s matrices: s^3 product constraints and s^2 sums.
// uniform_fr: A and B entries uniform in Fr (libff random_element rule, as
// the reference's gen draws them); otherwise uniform non-zero 32-bit values.
Instance matmul(size_t s, uint64_t seed, Density density, bool uniform_fr);
// F_steps from public F0, F1 (uniform Fr), one constraint per step.
Instance fibonacci(size_t steps, uint64_t seed);
// count independent Poseidon-style hashes (width 3, x^5, 8 full + 57 partial
// rounds, seeded constants), public digest, private preimage.  folded keeps
// the MDS rows as linear combinations (243 + 1 constraints per hash instead
// of the reference circuit's 439).
Instance poseidon(size_t count, uint64_t seed, Density density, bool folded);

}  // namespace nkr::groth16::gen

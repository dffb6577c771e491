// nkr-g16-gen: the reference's groth16-gen driver (protocols/groth16/app/
// gen.cpp:197-300) with seeded values.  Writes <stem>.bin (R1CS: public
// input count, columns, A, B, C) and <stem>.witness.bin in the reference's
// BinaryArchive format under --path (default app/data), with the
// reference's stems: matmul_<s>, fibonacci_<steps>, poseidon_<count>, and
// _<density> appended for hybrid and sparse inputs.  Additions: --seed
// (default 0x4E4B52; the reference draws from std::random_device), --u32
// (matmul entries uniform non-zero 32-bit values instead of uniform Fr;
// stem suffix _u32) and --folded (Poseidon MDS rows folded into the S-box
// constraints, ~244 constraints per hash; stem suffix _folded).  The
// sql-aggregate case is outside this build's scope.
#include <cstdlib>
#include <iostream>
#include <limits>
#include <string>

#include "cli.hpp"
#include "gen.hpp"

using namespace nkr::groth16;
using nkr::cli::Scope;

namespace {

size_t positive(const std::string &s, const char *what)
{
    size_t pos = 0;
    unsigned long long v = 0;
    try {
        v = std::stoull(s, &pos, 0);
    } catch (const std::exception &) {
        pos = 0;
    }
    if (pos != s.size() || v == 0) throw std::invalid_argument(std::string(what) + " must be a positive integer");
    return (size_t)v;
}

}  // namespace

int main(int argc, char **argv)
{
    const std::string usage =
        "[--profile] [--path <dir>] [--seed <u64>] [--density full|hybrid|sparse] [--u32] [--folded] "
        "<case> [params...]\n"
        "  cases:\n"
        "    matmul <size>      matrix-multiplication R1CS of dimension <size>\n"
        "    fibonacci <steps>  compute F_steps from public F0 and F1; steps >= 2\n"
        "    poseidon [count]   <count> independent private-input Poseidon-style hashes\n"
        "  input density (matmul and poseidon; default full): full 100%, hybrid 50%, sparse 10% non-zero\n"
        "  Outputs: <stem>.bin (R1CS) and <stem>.witness.bin (witness z)";
    std::string density_arg = "full";
    bool u32 = false, folded = false;
    const nkr::cli::Args args = nkr::cli::parse(
        argc, argv, nkr::cli::kPath | nkr::cli::kSeed | nkr::cli::kProfile, usage,
        [&](const std::string &flag, const char *value) -> int {
            if (flag == "--density" && value) {
                density_arg = value;
                return 2;
            }
            if (flag.rfind("--density=", 0) == 0) {
                density_arg = flag.substr(10);
                return 1;
            }
            if (flag == "--u32") {
                u32 = true;
                return 1;
            }
            if (flag == "--folded") {
                folded = true;
                return 1;
            }
            return 0;
        });
    if (args.positional.empty()) {
        std::cerr << "Usage: " << argv[0] << ' ' << usage << '\n';
        return 2;
    }
    const uint64_t seed = args.seeded ? args.seed : 0x4E4B52;
    try {
        const gen::Density density = gen::parse_density(density_arg);
        const std::string name = args.positional[0];
        if (density != gen::Density::Full && name != "matmul" && name != "poseidon")
            throw std::invalid_argument("--density is supported only for matmul and poseidon");
        Scope total("groth16.gen.total");
        gen::Instance inst;
        std::string stem;
        if (name == "matmul") {
            if (args.positional.size() < 2) throw std::invalid_argument("matmul needs <size>");
            const size_t s = positive(args.positional[1], "matmul size");
            {
                Scope stage("groth16.gen.build_matmul");
                inst = gen::matmul(s, seed, density, !u32);
            }
            stem = "matmul_" + std::to_string(s);
            if (density != gen::Density::Full) stem += std::string("_") + gen::density_name(density);
            if (u32) stem += "_u32";
            const size_t nz = gen::density_nonzero_count(s * s, density);
            std::cout << "Density: " << gen::density_name(density) << " (A: " << nz << "/" << s * s
                      << " nonzero, B: " << nz << "/" << s * s << " nonzero)\n";
        } else if (name == "fibonacci") {
            if (args.positional.size() < 2) throw std::invalid_argument("fibonacci needs <steps>");
            const size_t steps = positive(args.positional[1], "fibonacci steps");
            {
                Scope stage("groth16.gen.build_fibonacci");
                inst = gen::fibonacci(steps, seed);
            }
            stem = "fibonacci_" + std::to_string(steps);
        } else if (name == "poseidon") {
            const size_t count = args.positional.size() < 2 ? 1 : positive(args.positional[1], "Poseidon count");
            if (count > std::numeric_limits<size_t>::max() / 2) throw std::invalid_argument("Poseidon count is too large");
            {
                Scope stage("groth16.gen.build_poseidon");
                inst = gen::poseidon(count, seed, density, folded);
            }
            stem = "poseidon_" + std::to_string(count);
            if (density != gen::Density::Full) stem += std::string("_") + gen::density_name(density);
            if (folded) stem += "_folded";
            const size_t nz = gen::density_nonzero_count(2 * count, density);
            std::cout << "Density: " << gen::density_name(density) << " (inputs: " << nz << "/" << 2 * count
                      << " nonzero)\n";
        } else {
            throw std::invalid_argument("unknown case '" + name + "' (matmul, fibonacci or poseidon)");
        }
        std::string error;
        if (!validate_circuit(inst.circuit, error)) throw std::runtime_error("generated circuit is invalid: " + error);
        const auto paths = artifact_paths(args.directory, stem + ".bin");
        {
            Scope stage("groth16.gen.write_outputs");
            std::filesystem::create_directories(args.directory);
            write_circuit(paths.r1cs, inst.circuit);
            write_witness(paths.witness, inst.witness);
        }
        std::cout << "Constraints: " << inst.circuit.a.num_rows << ", variables: " << inst.circuit.a.num_cols
                  << ", public inputs: " << inst.circuit.public_inputs << '\n';
        std::cout << "Wrote " << paths.r1cs << " and " << paths.witness << '\n';
        return 0;
    } catch (const std::exception &e) {
        std::cerr << "gen: " << e.what() << '\n';
        return 1;
    }
}

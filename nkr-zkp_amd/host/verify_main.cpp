// nkr-g16-verify: the reference's groth16-verify driver (protocols/groth16/
// app/verify.cpp:26-77) over this library's host pairing.  Same flags, same
// input files (r1cs, witness for the public inputs, verification key, proof),
// same messages and exit codes: 0 when the proof verifies, 1 when it does
// not or on an error ("verify: <what>"), 2 on a usage error.
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <string>

#include "groth16.hpp"

using namespace nkr::groth16;

namespace {

bool g_profile = false;

class Scope {
public:
    explicit Scope(const char *label) : label_(label), t0_(std::chrono::steady_clock::now()) {}
    ~Scope()
    {
        if (!g_profile) return;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
        std::cout << "[profile] " << label_ << " = " << ms << " ms\n";
    }

private:
    const char *label_;
    std::chrono::steady_clock::time_point t0_;
};

[[noreturn]] void usage(const char *program)
{
    std::cerr << "Usage: " << program << " [--path <dir>] [--proof <proof-file>] [--profile] <r1cs-file>\n";
    std::exit(2);
}

}  // namespace

int main(int argc, char **argv)
{
    std::filesystem::path directory = "app/data";
    std::filesystem::path proof_input;
    std::string filename;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        if (arg == "--path" && i + 1 < argc)
            directory = argv[++i];
        else if (arg == "--proof" && i + 1 < argc)
            proof_input = argv[++i];
        else if (arg == "--profile")
            g_profile = true;
        else if (arg == "--help" || arg == "-h" || (!arg.empty() && arg.front() == '-'))
            usage(argv[0]);
        else
            filename = arg;
    }
    if (filename.empty()) usage(argv[0]);

    try {
        Scope total("groth16.verify.total");
        auto paths = artifact_paths(directory, filename);
        if (!proof_input.empty()) paths.proof = proof_input;
        Circuit circuit;
        std::vector<Scalar> witness;
        VerificationKey key;
        Proof proof;
        {
            Scope s("groth16.verify.read_inputs");
            circuit = read_circuit(paths.r1cs);
            witness = read_witness(paths.witness);
            key = read_verification_key(paths.verification_key);
            proof = read_proof(paths.proof);
        }
        if (witness.size() < circuit.public_inputs + 1) {
            std::cerr << "Witness is too short for the public inputs.\n";
            return 1;
        }
        const std::vector<Scalar> public_input(witness.begin() + 1, witness.begin() + 1 + circuit.public_inputs);
        bool ok;
        {
            Scope s("groth16.verify.pairing_check");
            ok = verify(key, public_input, proof);
        }
        std::cout << (ok ? "Proof verification succeeded." : "Proof verification FAILED.") << '\n';
        return ok ? 0 : 1;
    } catch (const std::exception &e) {
        std::cerr << "verify: " << e.what() << '\n';
        return 1;
    }
}

// nkr-g16-verify: the reference's groth16-verify driver (protocols/groth16/
// app/verify.cpp:26-77) over this library's host pairing.  Same flags, same
// input files (r1cs, witness for the public inputs, verification key, proof),
// same messages and exit codes: 0 when the proof verifies, 1 when it does
// not or on an error ("verify: <what>"), 2 on a usage error.
#include <iostream>
#include <string>

#include "cli.hpp"
#include "groth16.hpp"

using namespace nkr::groth16;
using nkr::cli::Scope;

int main(int argc, char **argv)
{
    std::filesystem::path proof_input;
    const std::string usage = "[--path <dir>] [--proof <proof-file>] [--profile] <r1cs-file>";
    const nkr::cli::Args args = nkr::cli::parse(argc, argv, nkr::cli::kPath | nkr::cli::kProfile, usage,
                                                [&](const std::string &flag, const char *value) -> int {
                                                    if (flag == "--proof" && value) {
                                                        proof_input = value;
                                                        return 2;
                                                    }
                                                    return 0;
                                                });
    if (args.positional.empty()) {
        std::cerr << "Usage: " << argv[0] << ' ' << usage << '\n';
        return 2;
    }
    const std::filesystem::path directory = args.directory;
    const std::string filename = args.positional.back();

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

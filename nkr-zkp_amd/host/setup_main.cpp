// nkr-g16-setup: the reference's groth16-setup driver (protocols/groth16/
// app/setup.cpp:26-76) over the MI355X backend: reads <r1cs-file>, validates
// it, runs Backend::setup on the GPU and writes the proving and verification
// key files.  --seed fixes the trapdoors (the reference draws them from
// std::random_device).
#include <iostream>
#include <string>

#include "cli.hpp"
#include "groth16.hpp"

using namespace nkr::groth16;
using nkr::cli::Scope;

int main(int argc, char **argv)
{
    const std::string usage = "[--backend hip] [--path <dir>] [--seed <u64>] [--profile] <r1cs-file>";
    const nkr::cli::Args args = nkr::cli::parse(
        argc, argv, nkr::cli::kPath | nkr::cli::kBackend | nkr::cli::kSeed | nkr::cli::kProfile, usage);
    if (args.positional.empty()) {
        std::cerr << "Usage: " << argv[0] << ' ' << usage << '\n';
        return 2;
    }
    const std::filesystem::path directory = args.directory;
    const std::string backend_name = args.backend, filename = args.positional.back();
    const bool seeded = args.seeded;
    const uint64_t seed = args.seed;

    try {
        Scope total("groth16.setup.total");
        const auto paths = artifact_paths(directory, filename);
        Circuit circuit;
        {
            Scope s("groth16.setup.read_circuit");
            circuit = read_circuit(paths.r1cs);
        }
        {
            Scope s("groth16.setup.validate_circuit");
            std::string error;
            if (!validate_circuit(circuit, error)) throw std::runtime_error("invalid circuit: " + error);
        }
        auto backend = make_backend(backend_name);
        std::cout << "Groth16 setup backend: " << backend->name() << '\n';
        ProvingKey pk;
        VerificationKey vk;
        {
            Scope s("groth16.setup.backend");
            if (seeded)
                setup_with_trapdoors(circuit, sample_trapdoors(seed), pk, vk);
            else
                backend->setup(circuit, pk, vk);
        }
        {
            Scope s("groth16.setup.write_keys");
            write_proving_key(paths.proving_key, pk);
            write_verification_key(paths.verification_key, vk);
        }
        std::cout << "Wrote " << paths.proving_key << " and " << paths.verification_key << '\n';
        return 0;
    } catch (const std::exception &e) {
        std::cerr << "setup: " << e.what() << '\n';
        return 1;
    }
}

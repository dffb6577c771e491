// nkr-g16-setup: the reference's groth16-setup driver (protocols/groth16/
// app/setup.cpp:26-76) over the MI355X backend: reads <r1cs-file>, validates
// it, runs Backend::setup on the GPU and writes the proving and verification
// key files.  --seed fixes the trapdoors (the reference draws them from
// std::random_device).
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
    std::cerr << "Usage: " << program
              << " [--backend hip] [--path <dir>] [--seed <u64>] [--profile] <r1cs-file>\n";
    std::exit(2);
}

}  // namespace

int main(int argc, char **argv)
{
    std::filesystem::path directory = "app/data";
    std::string backend_name = "hip", filename;
    bool seeded = false;
    uint64_t seed = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        if (arg == "--backend" && i + 1 < argc)
            backend_name = argv[++i];
        else if (arg == "--path" && i + 1 < argc)
            directory = argv[++i];
        else if (arg == "--seed" && i + 1 < argc) {
            seed = std::stoull(argv[++i], nullptr, 0);
            seeded = true;
        } else if (arg == "--profile")
            g_profile = true;
        else if (arg == "--help" || arg == "-h" || (!arg.empty() && arg.front() == '-'))
            usage(argv[0]);
        else
            filename = arg;
    }
    if (filename.empty()) usage(argv[0]);

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

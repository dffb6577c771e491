// nkr-g16-prove: the reference's groth16-prove driver (protocols/groth16/app/
// prove.cpp:27-99) over the MI355X backend.  Same flags, same input and
// output files, same timed scopes and "[profile] <label> = <ms> ms" lines.
// --seed fixes r and s (the reference draws them from std::random_device).
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <random>
#include <string>

#include "../../include/nkr_g16.h"
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
              << " [--backend hip] [--path <dir>] [--output <proof-file>] [--seed <u64>]"
                 " [--profile] <r1cs-file>\n";
    std::exit(2);
}

}  // namespace

int main(int argc, char **argv)
{
    std::filesystem::path directory = "app/data";
    std::filesystem::path proof_output;
    std::string backend_name = "hip", filename;
    bool seeded = false;
    uint64_t seed = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        if (arg == "--backend" && i + 1 < argc)
            backend_name = argv[++i];
        else if (arg == "--path" && i + 1 < argc)
            directory = argv[++i];
        else if (arg == "--output" && i + 1 < argc)
            proof_output = argv[++i];
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
        Scope program_total("groth16.prove.program_total");
        auto paths = artifact_paths(directory, filename);
        if (!proof_output.empty()) paths.proof = proof_output;
        Circuit circuit;
        std::vector<Scalar> witness;
        ProvingKey key;
        {
            Scope s("groth16.prove.read_inputs");
            circuit = read_circuit(paths.r1cs);
            witness = read_witness(paths.witness);
            key = read_proving_key(paths.proving_key);
        }
        {
            Scope s("groth16.prove.validate_inputs");
            std::string error;
            if (!validate_circuit(circuit, error)) throw std::runtime_error("invalid circuit: " + error);
            if (!validate_proving_key(circuit, key, error)) throw std::runtime_error("invalid proving key: " + error);
            if (witness.size() != circuit.a.num_cols) throw std::runtime_error("witness does not satisfy the R1CS");
            nkr_g16_circuit cv{};
            cv.rows = circuit.a.num_rows;
            cv.cols = circuit.a.num_cols;
            cv.public_inputs = circuit.public_inputs;
            const CsrMatrix<Scalar> *m[3] = {&circuit.a, &circuit.b, &circuit.c};
            const uint32_t **rp[3] = {&cv.a_row_ptr, &cv.b_row_ptr, &cv.c_row_ptr};
            const uint32_t **ci[3] = {&cv.a_col_idx, &cv.b_col_idx, &cv.c_col_idx};
            const void **va[3] = {&cv.a_values, &cv.b_values, &cv.c_values};
            uint64_t *nz[3] = {&cv.a_nnz, &cv.b_nnz, &cv.c_nnz};
            for (int k = 0; k < 3; ++k) {
                *rp[k] = m[k]->row_ptr.data();
                *ci[k] = m[k]->col_idx.data();
                *va[k] = m[k]->values.data();
                *nz[k] = m[k]->values.size();
            }
            int ok = 0;
            if (nkr_g16_check_witness(&cv, witness.data(), &ok) != 0) throw std::runtime_error(nkr_last_error());
            if (!ok) throw std::runtime_error("witness does not satisfy the R1CS");
        }
        auto backend = make_backend(backend_name);
        std::cout << "Groth16 prove backend: " << backend->name() << '\n';
        {
            Scope s("groth16.prove.prepare_backend");
            backend->prepare(circuit, key);
        }
        const auto rs = sample_blinding(seeded ? seed : std::random_device{}());
        RawProof raw;
        Proof proof;
        {
            Scope total("groth16.prove.total");
            {
                Scope s("groth16.prove.backend");
                backend->prove(circuit, key, witness, raw);
            }
            {
                Scope s("groth16.prove.blind_proof");
                proof = blind_proof(key, raw, rs[0], rs[1]);
            }
        }
        {
            Scope s("groth16.prove.write_proof");
            write_proof(paths.proof, proof);
        }
        std::cout << "Wrote " << paths.proof << '\n';
        return 0;
    } catch (const std::exception &e) {
        std::cerr << "prove: " << e.what() << '\n';
        return 1;
    }
}

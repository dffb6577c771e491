// nkr-g16-prove: the reference's groth16-prove driver (protocols/groth16/app/
// prove.cpp:27-99) over the MI355X backend.  Same flags, same input and
// output files, same timed scopes and "[profile] <label> = <ms> ms" lines.
// --seed fixes r and s (the reference draws them from std::random_device).
#include <iostream>
#include <random>
#include <string>

#include "../../include/nkr_g16.h"
#include "cli.hpp"
#include "groth16.hpp"

using namespace nkr::groth16;
using nkr::cli::Scope;

int main(int argc, char **argv)
{
    std::filesystem::path proof_output;
    const std::string usage = "[--backend hip] [--path <dir>] [--output <proof-file>] [--seed <u64>] [--profile] <r1cs-file>";
    const nkr::cli::Args args = nkr::cli::parse(
        argc, argv, nkr::cli::kPath | nkr::cli::kBackend | nkr::cli::kSeed | nkr::cli::kProfile, usage,
        [&](const std::string &flag, const char *value) -> int {
            if (flag == "--output" && value) {
                proof_output = value;
                return 2;
            }
            return 0;
        });
    if (args.positional.empty()) {
        std::cerr << "Usage: " << argv[0] << ' ' << usage << '\n';
        return 2;
    }
    const std::filesystem::path directory = args.directory;
    const std::string backend_name = args.backend, filename = args.positional.back();
    const bool seeded = args.seeded;
    const uint64_t seed = args.seed;

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

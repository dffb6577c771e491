// Host-side mirror of the reference's Groth16 backend interface, with the
// MI355X backend behind it.
//
//   types     protocols/groth16/common/include/groth16/types.hpp:11-59
//   Backend   protocols/groth16/common/include/groth16/backend.hpp:11-39
//   I/O       protocols/groth16/common/src/io.cpp:5-96 and the BinaryArchive
//             format of common/include/binary_archive.hpp:17-97
//   checks    protocols/groth16/common/src/protocol.cpp:9-83
//
// The element types here are plain byte-layout twins of libff's (32-byte
// Montgomery field elements, 96-byte G1 and 192-byte G2 Jacobian points), so
// files written by the reference are read unchanged and vice versa.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace nkr::groth16 {

struct Scalar {
    uint64_t mont[4];
};
struct G1 {
    uint64_t words[12];
};
struct G2 {
    uint64_t words[24];
};
struct GT {
    uint64_t words[48];  // Fq12 = 12 Fq elements (libff alt_bn128_GT)
};
static_assert(sizeof(Scalar) == 32 && sizeof(G1) == 96 && sizeof(G2) == 192 && sizeof(GT) == 384,
              "layouts must match libff's alt_bn128 types");

template <typename S>
struct CsrMatrix {
    size_t num_rows = 0, num_cols = 0;
    std::vector<S> values;
    std::vector<uint32_t> row_ptr, col_idx;
    size_t nnz() const { return values.size(); }
};

struct Circuit {
    size_t public_inputs = 0;
    CsrMatrix<Scalar> a, b, c;
    size_t constraints() const { return a.num_rows; }
    size_t variables() const { return a.num_cols > 0 ? a.num_cols - 1 : 0; }
};

struct ProvingKey {
    std::vector<G1> a_query, b_g1_query;
    std::vector<G2> b_g2_query;
    std::vector<G1> l_query, h_query;
    G1 alpha_g1, beta_g1;
    G2 beta_g2;
    G1 delta_g1;
    G2 delta_g2;
};

struct VerificationKey {
    GT alpha_g1_beta_g2;
    G2 gamma_g2, delta_g2;
    std::vector<G1> ic;
};

struct RawProof {
    G1 ar, bs1, zk, qz;
    G2 bs2;
};

struct Proof {
    G1 a;
    G2 b;
    G1 c;
};

class Backend {
public:
    virtual ~Backend() = default;
    virtual const char *name() const = 0;
    virtual void setup(const Circuit &circuit, ProvingKey &proving_key, VerificationKey &verification_key) = 0;
    virtual void prepare(const Circuit &, const ProvingKey &) {}
    virtual void prove(const Circuit &circuit, const ProvingKey &proving_key, const std::vector<Scalar> &witness,
                       RawProof &proof) = 0;
};

// "hip" (and its alias "dcu"): the gfx950 backend.
std::unique_ptr<Backend> make_mi355x_backend(int device = 0);
std::unique_ptr<Backend> make_backend(const std::string &name);
std::vector<std::string> available_backends();

// blind_proof (protocol.cpp:85-94) with caller-supplied r, s.
Proof blind_proof(const ProvingKey &key, const RawProof &raw, const Scalar &r, const Scalar &s);
// Seeded r, s by libff's random_element rule on mt19937_64(seed).
std::array<Scalar, 2> sample_blinding(uint64_t seed);
// Seeded t, alpha, beta, gamma, delta (the order backend.cpp:48-52 draws them).
std::array<Scalar, 5> sample_trapdoors(uint64_t seed);
// Backend::setup with given trapdoors, on the GPU (nkr_g16_setup).
void setup_with_trapdoors(const Circuit &circuit, const std::array<Scalar, 5> &trapdoors, ProvingKey &pk,
                          VerificationKey &vk, int device = 0);
// verify (protocol.cpp:96-108): the pairing check on the host.
bool verify(const VerificationKey &key, const std::vector<Scalar> &public_input, const Proof &proof);

// ------------------------------------------------------------- validation
bool validate_circuit(const Circuit &circuit, std::string &error);
bool validate_proving_key(const Circuit &circuit, const ProvingKey &key, std::string &error);

// ---------------------------------------------------------------- archive
class BinaryArchive {
public:
    static constexpr char MAGIC[4] = {'L', 'K', 'Z', '1'};
    static constexpr uint32_t VERSION = 1;
    explicit BinaryArchive(uint64_t max_bytes = uint64_t{1} << 30) : max_vector_bytes_(max_bytes) {}
    void open_for_write(const std::string &path);
    void open_for_read(const std::string &path);
    template <typename T>
    void write(const T &v);
    template <typename T>
    void read(T &v);
    template <typename T>
    void write(const std::vector<T> &v);
    template <typename T>
    void read(std::vector<T> &v);

private:
    std::fstream fs_;
    uint64_t max_vector_bytes_;
};

struct ArtifactPaths {
    std::filesystem::path r1cs, witness, proving_key, verification_key, proof;
};
ArtifactPaths artifact_paths(const std::filesystem::path &directory, std::string filename);
Circuit read_circuit(const std::filesystem::path &path);
void write_circuit(const std::filesystem::path &path, const Circuit &circuit);
std::vector<Scalar> read_witness(const std::filesystem::path &path);
void write_witness(const std::filesystem::path &path, const std::vector<Scalar> &witness);
ProvingKey read_proving_key(const std::filesystem::path &path);
void write_proving_key(const std::filesystem::path &path, const ProvingKey &key);
VerificationKey read_verification_key(const std::filesystem::path &path);
void write_verification_key(const std::filesystem::path &path, const VerificationKey &key);
Proof read_proof(const std::filesystem::path &path);
void write_proof(const std::filesystem::path &path, const Proof &proof);

}  // namespace nkr::groth16

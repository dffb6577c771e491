// Host side of the MI355X Groth16 backend: the reference's Backend interface
// (protocols/groth16/common/include/groth16/backend.hpp:11-39) implemented
// over the C ABI, plus artifact I/O in the reference's BinaryArchive format
// (common/include/binary_archive.hpp:17-97, protocols/groth16/common/src/io.cpp).
#include "groth16.hpp"

#include <cstring>
#include <random>

#include "../../include/nkr_g16.h"

namespace nkr::groth16 {

namespace {

void check(int rc, const char *what)
{
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + nkr_last_error());
}

nkr_g16_circuit circuit_view(const Circuit &c)
{
    nkr_g16_circuit v{};
    v.rows = c.a.num_rows;
    v.cols = c.a.num_cols;
    v.public_inputs = c.public_inputs;
    v.a_row_ptr = c.a.row_ptr.data();
    v.a_col_idx = c.a.col_idx.data();
    v.a_values = c.a.values.data();
    v.a_nnz = c.a.values.size();
    v.b_row_ptr = c.b.row_ptr.data();
    v.b_col_idx = c.b.col_idx.data();
    v.b_values = c.b.values.data();
    v.b_nnz = c.b.values.size();
    v.c_row_ptr = c.c.row_ptr.data();
    v.c_col_idx = c.c.col_idx.data();
    v.c_values = c.c.values.data();
    v.c_nnz = c.c.values.size();
    return v;
}

nkr_g16_pk key_view(const ProvingKey &k)
{
    nkr_g16_pk v{};
    v.a_query = k.a_query.data();
    v.b_g1_query = k.b_g1_query.data();
    v.b_g2_query = k.b_g2_query.data();
    v.l_query = k.l_query.data();
    v.h_query = k.h_query.data();
    v.alpha_g1 = &k.alpha_g1;
    v.beta_g1 = &k.beta_g1;
    v.beta_g2 = &k.beta_g2;
    v.delta_g1 = &k.delta_g1;
    v.delta_g2 = &k.delta_g2;
    return v;
}

// HipBackend's prepare/prove split (protocols/groth16/backends/hip/src/backend.hip:270-335):
// prepare builds the device context once; prove prepares lazily if needed.
class Mi355xBackend final : public Backend {
public:
    explicit Mi355xBackend(int device) : device_(device) {}
    ~Mi355xBackend() override
    {
        if (ctx_) nkr_g16_ctx_destroy(ctx_);
    }
    const char *name() const override { return "hip"; }

    // Trapdoors from std::random_device, as the CPU backend draws them
    // (backend.cpp:48-52); setup_with_trapdoors takes them from the caller.
    void setup(const Circuit &circuit, ProvingKey &pk, VerificationKey &vk) override
    {
        std::random_device rd;
        const uint64_t seed = (uint64_t(rd()) << 32) ^ rd();
        setup_with_trapdoors(circuit, sample_trapdoors(seed), pk, vk, device_);
    }

    void prepare(const Circuit &circuit, const ProvingKey &key) override
    {
        if (ctx_ && prepared_for_ == std::make_pair(&circuit, &key)) return;
        if (ctx_) {
            nkr_g16_ctx_destroy(ctx_);
            ctx_ = nullptr;
        }
        const nkr_g16_circuit cv = circuit_view(circuit);
        const nkr_g16_pk kv = key_view(key);
        check(nkr_g16_ctx_create(&cv, &kv, device_, &ctx_), "prepare");
        prepared_for_ = {&circuit, &key};
    }

    void prove(const Circuit &circuit, const ProvingKey &key, const std::vector<Scalar> &witness,
               RawProof &proof) override
    {
        if (witness.size() != circuit.a.num_cols) throw std::invalid_argument("witness length mismatch");
        prepare(circuit, key);
        check(nkr_g16_prove(ctx_, witness.data(), &proof), "prove");
    }

private:
    int device_;
    nkr_g16_ctx *ctx_ = nullptr;
    std::pair<const Circuit *, const ProvingKey *> prepared_for_{nullptr, nullptr};
};

}  // namespace

std::unique_ptr<Backend> make_mi355x_backend(int device) { return std::make_unique<Mi355xBackend>(device); }

std::vector<std::string> available_backends() { return {"hip"}; }

std::unique_ptr<Backend> make_backend(const std::string &name)
{
    if (name == "hip" || name == "dcu") return make_mi355x_backend(0);
    throw std::invalid_argument("Groth16 backend is not available in this build: " + name);
}

Proof blind_proof(const ProvingKey &key, const RawProof &raw, const Scalar &r, const Scalar &s)
{
    Proof p;
    const nkr_g16_pk kv = key_view(key);
    check(nkr_g16_blind(&kv, &raw, &r, &s, &p), "blind_proof");
    return p;
}

bool verify(const VerificationKey &key, const std::vector<Scalar> &public_input, const Proof &proof)
{
    const nkr_g16_vk vk{&key.alpha_g1_beta_g2, &key.gamma_g2, &key.delta_g2, key.ic.data(), key.ic.size()};
    int accepted = 0;
    check(nkr_g16_verify(&vk, public_input.data(), public_input.size(), &proof, &accepted), "verify");
    return accepted != 0;
}

template <size_t N>
std::array<Scalar, N> sample_scalars(uint64_t seed);

std::array<Scalar, 2> sample_blinding(uint64_t seed) { return sample_scalars<2>(seed); }

std::array<Scalar, 5> sample_trapdoors(uint64_t seed) { return sample_scalars<5>(seed); }

void setup_with_trapdoors(const Circuit &circuit, const std::array<Scalar, 5> &trapdoors, ProvingKey &pk,
                          VerificationKey &vk, int device)
{
    const size_t rows = circuit.a.num_rows, cols = circuit.a.num_cols;
    pk.a_query.resize(cols);
    pk.b_g1_query.resize(cols);
    pk.b_g2_query.resize(cols);
    pk.l_query.resize(cols);
    pk.h_query.resize(rows);
    vk.ic.resize(circuit.public_inputs + 1);
    const nkr_g16_pk_out po{pk.a_query.data(), pk.b_g1_query.data(), pk.b_g2_query.data(), pk.l_query.data(),
                            pk.h_query.data(),  &pk.alpha_g1,          &pk.beta_g1,           &pk.beta_g2,
                            &pk.delta_g1,       &pk.delta_g2};
    const nkr_g16_vk_out vo{&vk.alpha_g1_beta_g2, &vk.gamma_g2, &vk.delta_g2, vk.ic.data()};
    const nkr_g16_circuit cv = circuit_view(circuit);
    check(nkr_g16_setup(&cv, trapdoors.data(), device, &po, &vo), "setup");
}

template <size_t N>
std::array<Scalar, N> sample_scalars(uint64_t seed)
{
    // Fp_model::random_element (libff fp.tcc:719-744): four words, bits above
    // the modulus MSB cleared, rejected while >= r.
    static const uint64_t R[4] = {0x43e1f593f0000001ull, 0x2833e84879b97091ull, 0xb85045b68181585dull,
                                  0x30644e72e131a029ull};
    std::mt19937_64 rng(seed);
    std::array<Scalar, N> out{};
    for (Scalar &x : out) {
        for (;;) {
            for (uint64_t &w : x.mont) w = rng();
            x.mont[3] &= 0x3FFFFFFFFFFFFFFFull;
            int lt = 0;
            for (int w = 3; w >= 0; --w)
                if (x.mont[w] != R[w]) {
                    lt = x.mont[w] < R[w];
                    break;
                }
            if (lt) break;
        }
    }
    return out;
}

bool validate_circuit(const Circuit &circuit, std::string &error)
{
    const auto valid_matrix = [&](const CsrMatrix<Scalar> &m, const char *name) {
        if (m.num_rows == 0 || m.num_cols == 0) {
            error = std::string(name) + " has an empty shape";
            return false;
        }
        if (m.row_ptr.size() != m.num_rows + 1 || m.row_ptr.front() != 0 || m.row_ptr.back() != m.values.size() ||
            m.col_idx.size() != m.values.size()) {
            error = std::string(name) + " has invalid CSR offsets";
            return false;
        }
        for (size_t r = 1; r < m.row_ptr.size(); ++r)
            if (m.row_ptr[r] < m.row_ptr[r - 1]) {
                error = std::string(name) + " has non-monotonic CSR offsets";
                return false;
            }
        for (const auto col : m.col_idx)
            if (col >= m.num_cols) {
                error = std::string(name) + " contains an out-of-range column";
                return false;
            }
        return true;
    };
    if (!valid_matrix(circuit.a, "A") || !valid_matrix(circuit.b, "B") || !valid_matrix(circuit.c, "C"))
        return false;
    if (circuit.a.num_rows != circuit.b.num_rows || circuit.a.num_rows != circuit.c.num_rows ||
        circuit.a.num_cols != circuit.b.num_cols || circuit.a.num_cols != circuit.c.num_cols) {
        error = "R1CS matrices have different shapes";
        return false;
    }
    if (circuit.public_inputs + 1 > circuit.a.num_cols) {
        error = "public input count exceeds the variable count";
        return false;
    }
    if ((circuit.constraints() & (circuit.constraints() - 1)) != 0) {
        error = "constraint count must be a power of two";
        return false;
    }
    return true;
}

bool validate_proving_key(const Circuit &circuit, const ProvingKey &key, std::string &error)
{
    const size_t columns = circuit.a.num_cols;
    if (key.a_query.size() != columns || key.b_g1_query.size() != columns || key.b_g2_query.size() != columns ||
        key.l_query.size() != columns) {
        error = "query length does not match the circuit variable count";
        return false;
    }
    if (key.h_query.size() != circuit.constraints()) {
        error = "H query length does not match the constraint count";
        return false;
    }
    return true;
}

// ------------------------------------------------------------ BinaryArchive
constexpr char BinaryArchive::MAGIC[4];

void BinaryArchive::open_for_write(const std::string &path)
{
    fs_.open(path, std::ios::binary | std::ios::out | std::ios::trunc);
    if (!fs_) throw std::runtime_error("BinaryArchive: cannot open file for writing: " + path);
    fs_.write(MAGIC, sizeof(MAGIC));
    fs_.write(reinterpret_cast<const char *>(&VERSION), sizeof(VERSION));
}

void BinaryArchive::open_for_read(const std::string &path)
{
    fs_.open(path, std::ios::binary | std::ios::in);
    if (!fs_) throw std::runtime_error("BinaryArchive: cannot open file for reading: " + path);
    char magic[4] = {0, 0, 0, 0};
    uint32_t version = 0;
    fs_.read(magic, sizeof(magic));
    fs_.read(reinterpret_cast<char *>(&version), sizeof(version));
    if (!fs_ || std::memcmp(magic, MAGIC, 4) != 0 || version != VERSION)
        throw std::runtime_error("BinaryArchive: incompatible or truncated file: " + path);
}

template <typename T>
void BinaryArchive::write(const T &value)
{
    fs_.write(reinterpret_cast<const char *>(&value), sizeof(T));
    if (!fs_) throw std::runtime_error("BinaryArchive: write failed");
}

template <typename T>
void BinaryArchive::read(T &value)
{
    fs_.read(reinterpret_cast<char *>(&value), sizeof(T));
    if (!fs_) throw std::runtime_error("BinaryArchive: unexpected end of file");
}

template <typename T>
void BinaryArchive::write(const std::vector<T> &v)
{
    const uint64_t bytes = v.size() * sizeof(T);
    if (bytes > max_vector_bytes_) throw std::runtime_error("BinaryArchive: vector exceeds configured size limit");
    fs_.write(reinterpret_cast<const char *>(&bytes), sizeof(bytes));
    if (bytes) fs_.write(reinterpret_cast<const char *>(v.data()), static_cast<std::streamsize>(bytes));
    if (!fs_) throw std::runtime_error("BinaryArchive: vector write failed");
}

template <typename T>
void BinaryArchive::read(std::vector<T> &v)
{
    uint64_t bytes = 0;
    fs_.read(reinterpret_cast<char *>(&bytes), sizeof(bytes));
    if (!fs_) throw std::runtime_error("BinaryArchive: unexpected end of file");
    if (bytes > max_vector_bytes_) throw std::runtime_error("BinaryArchive: vector exceeds configured size limit");
    if (bytes % sizeof(T) != 0) throw std::runtime_error("BinaryArchive: corrupt vector length");
    v.resize(bytes / sizeof(T));
    if (bytes) {
        fs_.read(reinterpret_cast<char *>(v.data()), static_cast<std::streamsize>(bytes));
        if (!fs_) throw std::runtime_error("BinaryArchive: truncated vector data");
    }
}

ArtifactPaths artifact_paths(const std::filesystem::path &directory, std::string filename)
{
    if (filename.size() >= 4 && filename.compare(filename.size() - 4, 4, ".bin") == 0)
        filename.resize(filename.size() - 4);
    return {directory / (filename + ".bin"), directory / (filename + ".witness.bin"),
            directory / (filename + ".pk.bin"), directory / (filename + ".vk.bin"),
            directory / (filename + ".proof.bin")};
}

Circuit read_circuit(const std::filesystem::path &path)
{
    BinaryArchive ar;
    ar.open_for_read(path.string());
    Circuit c;
    size_t columns = 0;
    ar.read(c.public_inputs);
    ar.read(columns);
    for (CsrMatrix<Scalar> *m : {&c.a, &c.b, &c.c}) {
        ar.read(m->row_ptr);
        ar.read(m->col_idx);
        ar.read(m->values);
        if (m->row_ptr.empty()) throw std::runtime_error("read_circuit: empty row offsets");
        m->num_rows = m->row_ptr.size() - 1;
        m->num_cols = columns;
    }
    return c;
}

void write_circuit(const std::filesystem::path &path, const Circuit &c)
{
    BinaryArchive ar;
    ar.open_for_write(path.string());
    ar.write(c.public_inputs);
    ar.write(c.a.num_cols);
    for (const CsrMatrix<Scalar> *m : {&c.a, &c.b, &c.c}) {
        ar.write(m->row_ptr);
        ar.write(m->col_idx);
        ar.write(m->values);
    }
}

std::vector<Scalar> read_witness(const std::filesystem::path &path)
{
    BinaryArchive ar;
    ar.open_for_read(path.string());
    std::vector<Scalar> w;
    ar.read(w);
    return w;
}

void write_witness(const std::filesystem::path &path, const std::vector<Scalar> &w)
{
    BinaryArchive ar;
    ar.open_for_write(path.string());
    ar.write(w);
}

ProvingKey read_proving_key(const std::filesystem::path &path)
{
    BinaryArchive ar;
    ar.open_for_read(path.string());
    ProvingKey k;
    ar.read(k.a_query);
    ar.read(k.b_g1_query);
    ar.read(k.b_g2_query);
    ar.read(k.l_query);
    ar.read(k.h_query);
    ar.read(k.alpha_g1);
    ar.read(k.beta_g1);
    ar.read(k.beta_g2);
    ar.read(k.delta_g1);
    ar.read(k.delta_g2);
    return k;
}

void write_proving_key(const std::filesystem::path &path, const ProvingKey &k)
{
    BinaryArchive ar;
    ar.open_for_write(path.string());
    ar.write(k.a_query);
    ar.write(k.b_g1_query);
    ar.write(k.b_g2_query);
    ar.write(k.l_query);
    ar.write(k.h_query);
    ar.write(k.alpha_g1);
    ar.write(k.beta_g1);
    ar.write(k.beta_g2);
    ar.write(k.delta_g1);
    ar.write(k.delta_g2);
}

VerificationKey read_verification_key(const std::filesystem::path &path)
{
    BinaryArchive ar;
    ar.open_for_read(path.string());
    VerificationKey k;
    ar.read(k.alpha_g1_beta_g2);
    ar.read(k.gamma_g2);
    ar.read(k.delta_g2);
    ar.read(k.ic);
    return k;
}

void write_verification_key(const std::filesystem::path &path, const VerificationKey &k)
{
    BinaryArchive ar;
    ar.open_for_write(path.string());
    ar.write(k.alpha_g1_beta_g2);
    ar.write(k.gamma_g2);
    ar.write(k.delta_g2);
    ar.write(k.ic);
}

Proof read_proof(const std::filesystem::path &path)
{
    BinaryArchive ar;
    ar.open_for_read(path.string());
    Proof p;
    ar.read(p.a);
    ar.read(p.b);
    ar.read(p.c);
    return p;
}

void write_proof(const std::filesystem::path &path, const Proof &p)
{
    BinaryArchive ar;
    ar.open_for_write(path.string());
    ar.write(p.a);
    ar.write(p.b);
    ar.write(p.c);
}

}  // namespace nkr::groth16

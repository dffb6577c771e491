// Seeded R1CS generators (see gen.hpp).
#include "gen.hpp"

#include <algorithm>
#include <random>
#include <stdexcept>
#include <utility>

namespace nkr::groth16::gen {

namespace {

using u128 = unsigned __int128;

// BN254 scalar field Fr in Montgomery form (R = 2^256), 4 x 64-bit words.
constexpr uint64_t kR[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL,
                            0x30644e72e131a029ULL};
constexpr uint64_t kRInv = 0xc2e1f593efffffffULL;  // -r^-1 mod 2^64
constexpr uint64_t kR2[4] = {0x1bb8e645ae216da7ULL, 0x53fe3ab1e35c59e3ULL, 0x8c49833d53bb8085ULL,
                             0x0216d0b17f4e44a5ULL};  // 2^512 mod r

struct Fr {
    uint64_t v[4] = {0, 0, 0, 0};
    bool is_zero() const { return (v[0] | v[1] | v[2] | v[3]) == 0; }
};

bool geq_modulus(const uint64_t t[4])
{
    for (int i = 3; i >= 0; --i) {
        if (t[i] != kR[i]) return t[i] > kR[i];
    }
    return true;
}

void sub_modulus(uint64_t t[4])
{
    uint64_t borrow = 0;
    for (int i = 0; i < 4; ++i) {
        const u128 d = (u128)t[i] - kR[i] - borrow;
        t[i] = (uint64_t)d;
        borrow = (uint64_t)(d >> 64) & 1;
    }
}

Fr add(const Fr &a, const Fr &b)
{
    Fr r;
    uint64_t carry = 0;
    for (int i = 0; i < 4; ++i) {
        const u128 s = (u128)a.v[i] + b.v[i] + carry;
        r.v[i] = (uint64_t)s;
        carry = (uint64_t)(s >> 64);
    }
    if (carry || geq_modulus(r.v)) sub_modulus(r.v);
    return r;
}

// Montgomery product a * b / 2^256 mod r (CIOS).
Fr mul(const Fr &a, const Fr &b)
{
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < 4; ++j) {
            const u128 p = (u128)a.v[j] * b.v[i] + t[j] + carry;
            t[j] = (uint64_t)p;
            carry = (uint64_t)(p >> 64);
        }
        u128 s = (u128)t[4] + carry;
        t[4] = (uint64_t)s;
        t[5] = (uint64_t)(s >> 64);
        const uint64_t m = t[0] * kRInv;
        u128 p = (u128)m * kR[0] + t[0];
        carry = (uint64_t)(p >> 64);
        for (int j = 1; j < 4; ++j) {
            p = (u128)m * kR[j] + t[j] + carry;
            t[j - 1] = (uint64_t)p;
            carry = (uint64_t)(p >> 64);
        }
        s = (u128)t[4] + carry;
        t[3] = (uint64_t)s;
        t[4] = t[5] + (uint64_t)(s >> 64);
    }
    Fr r;
    for (int i = 0; i < 4; ++i) r.v[i] = t[i];
    if (t[4] || geq_modulus(r.v)) sub_modulus(r.v);
    return r;
}

Fr from_u64(uint64_t x)
{
    Fr a;
    a.v[0] = x;
    Fr r2;
    for (int i = 0; i < 4; ++i) r2.v[i] = kR2[i];
    return mul(a, r2);
}

Fr one() { return from_u64(1); }

Fr pow5(const Fr &x)
{
    const Fr x2 = mul(x, x);
    return mul(mul(x2, x2), x);
}

Scalar to_scalar(const Fr &a)
{
    Scalar s;
    for (int i = 0; i < 4; ++i) s.mont[i] = a.v[i];
    return s;
}

// libff Fp_model::random_element (fp.tcc:719-744): four words, the bits
// above the modulus' top bit (254, 255) cleared, rejected while >= r; the
// words are the Montgomery residue.
Fr random_element(std::mt19937_64 &rng)
{
    for (;;) {
        Fr x;
        for (int i = 0; i < 4; ++i) x.v[i] = rng();
        x.v[3] &= 0x3FFFFFFFFFFFFFFFULL;
        if (!geq_modulus(x.v)) return x;
    }
}

Fr random_nonzero(std::mt19937_64 &rng)
{
    for (;;) {
        const Fr x = random_element(rng);
        if (!x.is_zero()) return x;
    }
}

// The non-zero positions of `count` entries: `nonzero` ones then zeros,
// Fisher-Yates shuffled by mt19937_64(seed).
std::vector<uint8_t> seeded_mask(size_t count, size_t nonzero, uint64_t seed)
{
    std::vector<uint8_t> mask(count, 0);
    std::fill(mask.begin(), mask.begin() + (std::ptrdiff_t)nonzero, uint8_t(1));
    if (nonzero == count) return mask;
    std::mt19937_64 rng(seed);
    for (size_t i = count - 1; i > 0; --i) {
        const size_t j = (size_t)(rng() % (uint64_t)(i + 1));
        std::swap(mask[i], mask[j]);
    }
    return mask;
}

size_t pow2_at_least(size_t n)
{
    size_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

void pad_rows(std::vector<uint32_t> &row_ptr, size_t rows)
{
    const uint32_t last = row_ptr.back();
    row_ptr.resize(rows + 1, last);
}

// ------------------------------------------------------------- R1CS builder
// A linear combination: (id, coefficient) sorted by id; id 0 is the constant.
using Lc = std::vector<std::pair<uint64_t, Fr>>;

Lc term(uint64_t id, const Fr &c) { return Lc{{id, c}}; }
Lc constant(const Fr &c) { return Lc{{0, c}}; }

Lc lc_add(const Lc &a, const Lc &b)
{
    Lc r;
    r.reserve(a.size() + b.size());
    size_t i = 0, j = 0;
    while (i < a.size() || j < b.size()) {
        if (j == b.size() || (i < a.size() && a[i].first < b[j].first))
            r.push_back(a[i++]);
        else if (i == a.size() || b[j].first < a[i].first)
            r.push_back(b[j++]);
        else {
            r.push_back({a[i].first, add(a[i].second, b[j].second)});  // zero sums are kept
            ++i;
            ++j;
        }
    }
    return r;
}

Lc lc_scale(const Lc &a, const Fr &s)
{
    Lc r = a;
    for (auto &t : r) t.second = mul(t.second, s);
    return r;
}

// R1CSManager semantics (r1cs.hpp:26-99): public ids count up from 1,
// private ids from 2^63; columns are 0 (the constant), the public ids, then
// the private ones in allocation order; add_constraint(out, a, b) is the row
// a * b = out (b = the constant 1 when absent).
class Builder {
public:
    static constexpr uint64_t kHalf = uint64_t(1) << 63;
    uint64_t add_public(const Fr &value)
    {
        values_.push_back({next_public_, value});
        return next_public_++;
    }
    uint64_t add_private(const Fr &value)
    {
        values_.push_back({next_private_, value});
        return next_private_++;
    }
    void set_value(uint64_t id, const Fr &value) { values_.push_back({id, value}); }
    void constrain(const Lc &out, const Lc &a, const Lc &b) { rows_.push_back({a, b, out}); }
    void constrain(const Lc &out, const Lc &a) { rows_.push_back({a, constant(one()), out}); }

    Instance finish() const
    {
        const size_t live_rows = rows_.size();
        const size_t n_public = next_public_;  // the constant column plus the public ids
        const size_t live_cols = n_public + (size_t)(next_private_ - kHalf);
        const size_t rows = pow2_at_least(live_rows), cols = pow2_at_least(live_cols);
        if (cols > (size_t(1) << 32)) throw std::invalid_argument("circuit too large for u32 columns");
        auto dense = [&](uint64_t id) -> uint32_t {
            return (uint32_t)(id < kHalf ? id : id - kHalf + n_public);
        };
        Instance inst;
        inst.circuit.public_inputs = n_public - 1;
        CsrMatrix<Scalar> *mats[3] = {&inst.circuit.a, &inst.circuit.b, &inst.circuit.c};
        for (int k = 0; k < 3; ++k) {
            CsrMatrix<Scalar> &m = *mats[k];
            m.num_rows = rows;
            m.num_cols = cols;
            m.row_ptr.reserve(rows + 1);
            m.row_ptr.push_back(0);
            for (const Row &r : rows_) {
                const Lc &lc = k == 0 ? r.a : (k == 1 ? r.b : r.out);
                for (const auto &t : lc) {
                    m.col_idx.push_back(dense(t.first));
                    m.values.push_back(to_scalar(t.second));
                }
                if (m.values.size() >= (size_t(1) << 32)) throw std::invalid_argument("matrix too large");
                m.row_ptr.push_back((uint32_t)m.values.size());
            }
            pad_rows(m.row_ptr, rows);
        }
        inst.witness.assign(cols, Scalar{});
        inst.witness[0] = to_scalar(one());
        for (const auto &v : values_) inst.witness[dense(v.first)] = to_scalar(v.second);
        return inst;
    }

private:
    struct Row {
        Lc a, b, out;
    };
    uint64_t next_public_ = 1;
    uint64_t next_private_ = kHalf;
    std::vector<std::pair<uint64_t, Fr>> values_;
    std::vector<Row> rows_;
};

// ------------------------------------------------------------------ Poseidon
struct PoseidonParams {
    static constexpr int kFull = 8, kPartial = 57;
    std::vector<Fr> rc;  // 3 per round
    Fr mds[3][3];
    explicit PoseidonParams(uint64_t seed = 0x9053)
    {
        std::mt19937_64 rng(seed);
        for (int i = 0; i < 3 * (kFull + kPartial); ++i) rc.push_back(random_element(rng));
        for (auto &row : mds)
            for (Fr &x : row) x = random_element(rng);
    }
    static bool full_round(int rnd) { return rnd < kFull / 2 || rnd >= kFull / 2 + kPartial; }
    Fr hash2(const Fr &left, const Fr &right) const
    {
        Fr s[3] = {Fr{}, left, right};
        size_t k = 0;
        for (int rnd = 0; rnd < kFull + kPartial; ++rnd) {
            for (Fr &x : s) x = add(x, rc[k++]);
            for (int i = 0; i < 3; ++i)
                if (full_round(rnd) || i == 0) s[i] = pow5(s[i]);
            Fr t[3];
            for (int r = 0; r < 3; ++r) {
                t[r] = Fr{};
                for (int c = 0; c < 3; ++c) t[r] = add(t[r], mul(mds[r][c], s[c]));
            }
            std::copy(t, t + 3, s);
        }
        return s[0];
    }
};

// add_hash2_constraints (protocols/groth16/common/include/r1cs/poseidon.hpp:
// 33-105): per round the state takes its constants, each S-box is three
// multiplications (squared = s*s, fourth = squared^2, fifth = fourth*s),
// then each MDS row is one private output with one linear constraint -- or,
// folded, stays a linear combination of the round's S-box outputs.
void hash_constraints(Builder &b, const PoseidonParams &prm, uint64_t left, const Fr &lv, uint64_t right,
                      const Fr &rv, uint64_t digest, bool folded)
{
    const Fr unit = one();
    Lc state[3] = {constant(Fr{}), term(left, unit), term(right, unit)};
    Fr vals[3] = {Fr{}, lv, rv};
    size_t k = 0;
    for (int rnd = 0; rnd < PoseidonParams::kFull + PoseidonParams::kPartial; ++rnd) {
        Lc sboxed[3];
        Fr sv[3];
        for (int i = 0; i < 3; ++i) {
            state[i] = lc_add(state[i], constant(prm.rc[k]));
            vals[i] = add(vals[i], prm.rc[k]);
            ++k;
        }
        const int nsb = PoseidonParams::full_round(rnd) ? 3 : 1;
        for (int i = 0; i < nsb; ++i) {
            const Fr x = vals[i], x2 = mul(x, x), x4 = mul(x2, x2), x5 = mul(x4, x);
            const uint64_t sq = b.add_private(x2), fo = b.add_private(x4), fi = b.add_private(x5);
            b.constrain(term(sq, unit), state[i], state[i]);
            b.constrain(term(fo, unit), term(sq, unit), term(sq, unit));
            b.constrain(term(fi, unit), term(fo, unit), state[i]);
            sboxed[i] = term(fi, unit);
            sv[i] = x5;
        }
        for (int i = nsb; i < 3; ++i) {
            sboxed[i] = state[i];
            sv[i] = vals[i];
        }
        Lc next[3];
        Fr nv[3];
        for (int r = 0; r < 3; ++r) {
            Lc mixed = constant(Fr{});
            Fr v{};
            for (int c = 0; c < 3; ++c) {
                mixed = lc_add(mixed, lc_scale(sboxed[c], prm.mds[r][c]));
                v = add(v, mul(prm.mds[r][c], sv[c]));
            }
            if (folded) {
                next[r] = std::move(mixed);
            } else {
                const uint64_t out = b.add_private(v);
                b.constrain(term(out, unit), mixed);
                next[r] = term(out, unit);
            }
            nv[r] = v;
        }
        for (int r = 0; r < 3; ++r) {
            state[r] = std::move(next[r]);
            vals[r] = nv[r];
        }
    }
    b.constrain(term(digest, unit), state[0]);
}

}  // namespace

Density parse_density(const std::string &name)
{
    if (name == "full") return Density::Full;
    if (name == "hybrid") return Density::Hybrid;
    if (name == "sparse") return Density::Sparse;
    throw std::invalid_argument("unknown density '" + name + "' (full, hybrid or sparse)");
}

const char *density_name(Density d)
{
    return d == Density::Full ? "full" : (d == Density::Hybrid ? "hybrid" : "sparse");
}

size_t density_nonzero_count(size_t count, Density d)
{
    switch (d) {
    case Density::Full:
        return count;
    case Density::Hybrid:
        return count - count / 2;
    default:
        return count / 10 + (count % 10 != 0);
    }
}

// build_matmul (gen.cpp:42-70): A entries from mt19937_64(seed ^ 0xA11CE),
// B from mt19937_64(seed ^ 0xB0B); zero positions from seeded shuffles
// (seed ^ 0x3A5C0 for A, seed ^ 0x3A5C1 for B).  Columns: the constant, A,
// B, C (public, s^2 each), then the products P (private, s^3).
Instance matmul(size_t s, uint64_t seed, Density density, bool uniform_fr)
{
    if (s == 0) throw std::invalid_argument("matmul size must be positive");
    const size_t ss = s * s, nz = density_nonzero_count(ss, density);
    std::vector<Fr> vals[2];
    const uint64_t tags[2] = {0xA11CE, 0xB0B}, mask_tags[2] = {0x3A5C0, 0x3A5C1};
    for (int m = 0; m < 2; ++m) {
        const std::vector<uint8_t> mask = seeded_mask(ss, nz, seed ^ mask_tags[m]);
        std::mt19937_64 rng(seed ^ tags[m]);
        vals[m].resize(ss);
        for (size_t i = 0; i < ss; ++i) {
            if (!mask[i]) continue;
            if (uniform_fr) {
                vals[m][i] = random_nonzero(rng);
            } else {
                uint64_t v = 0;
                while (v == 0) v = rng() & 0xFFFFFFFFULL;
                vals[m][i] = from_u64(v);
            }
        }
    }
    const std::vector<Fr> &av = vals[0], &bv = vals[1];
    const size_t prod_rows = s * ss, live_rows = prod_rows + ss;
    const size_t live_cols = 1 + 3 * ss + prod_rows;
    const size_t rows = pow2_at_least(live_rows), cols = pow2_at_least(live_cols);
    if (cols > (size_t(1) << 31)) throw std::invalid_argument("matmul size too large");
    const uint32_t colA = 1, colB = (uint32_t)(1 + ss), colC = (uint32_t)(1 + 2 * ss),
                   colP = (uint32_t)(1 + 3 * ss);
    const Scalar unit = to_scalar(one()), zero{};

    Instance inst;
    Circuit &c = inst.circuit;
    c.public_inputs = 3 * ss;
    for (CsrMatrix<Scalar> *m : {&c.a, &c.b, &c.c}) {
        m->num_rows = rows;
        m->num_cols = cols;
        m->row_ptr.push_back(0);
    }
    // product rows P_ijl = A_il * B_lj, in (i, j, l) order
    for (size_t i = 0; i < s; ++i)
        for (size_t j = 0; j < s; ++j)
            for (size_t l = 0; l < s; ++l) {
                const size_t p = i * ss + j * s + l;
                c.a.col_idx.push_back((uint32_t)(colA + i * s + l));
                c.a.values.push_back(unit);
                c.b.col_idx.push_back((uint32_t)(colB + l * s + j));
                c.b.values.push_back(unit);
                c.c.col_idx.push_back((uint32_t)(colP + p));
                c.c.values.push_back(unit);
                for (CsrMatrix<Scalar> *m : {&c.a, &c.b, &c.c}) m->row_ptr.push_back((uint32_t)m->values.size());
            }
    // sum rows C_ij = sum_l P_ijl: A = {0: 0, P_ij0 .. P_ij(s-1): 1}, B = {0: 1}
    for (size_t k = 0; k < ss; ++k) {
        c.a.col_idx.push_back(0);
        c.a.values.push_back(zero);
        for (size_t l = 0; l < s; ++l) {
            c.a.col_idx.push_back((uint32_t)(colP + (k / s) * ss + (k % s) * s + l));
            c.a.values.push_back(unit);
        }
        c.b.col_idx.push_back(0);
        c.b.values.push_back(unit);
        c.c.col_idx.push_back((uint32_t)(colC + k));
        c.c.values.push_back(unit);
        for (CsrMatrix<Scalar> *m : {&c.a, &c.b, &c.c}) m->row_ptr.push_back((uint32_t)m->values.size());
    }
    for (CsrMatrix<Scalar> *m : {&c.a, &c.b, &c.c}) pad_rows(m->row_ptr, rows);

    std::vector<Fr> z(cols);
    z[0] = one();
    for (size_t i = 0; i < s; ++i)
        for (size_t j = 0; j < s; ++j) {
            Fr acc{};
            for (size_t l = 0; l < s; ++l) {
                const Fr p = mul(av[i * s + l], bv[l * s + j]);
                z[colP + i * ss + j * s + l] = p;
                acc = add(acc, p);
            }
            z[colC + i * s + j] = acc;
        }
    for (size_t k = 0; k < ss; ++k) {
        z[colA + k] = av[k];
        z[colB + k] = bv[k];
    }
    inst.witness.resize(cols);
    for (size_t k = 0; k < cols; ++k) inst.witness[k] = to_scalar(z[k]);
    return inst;
}

// build_fibonacci (gen.cpp:87-120): F0, F1 from mt19937_64(seed ^ 0xF1B).
Instance fibonacci(size_t steps, uint64_t seed)
{
    if (steps < 2) throw std::invalid_argument("fibonacci steps must be at least 2");
    std::mt19937_64 rng(seed ^ 0xF1B);
    const Fr f0 = random_element(rng), f1 = random_element(rng);
    Builder b;
    const Fr unit = one();
    uint64_t prev = b.add_public(f0), cur = b.add_public(f1);
    const uint64_t expected = b.add_public(Fr{});
    Fr pv = f0, cv = f1;
    for (size_t k = 2; k <= steps; ++k) {
        const Fr nv = add(pv, cv);
        const uint64_t nxt = b.add_private(nv);
        b.constrain(term(nxt, unit), lc_add(term(prev, unit), term(cur, unit)));
        pv = cv;
        cv = nv;
        prev = cur;
        cur = nxt;
    }
    b.set_value(expected, cv);
    b.constrain(term(expected, unit), term(cur, unit));
    return b.finish();
}

// build_poseidon (gen.cpp:136-162): per hash a public digest then the
// private preimage (left, right), inputs uniform non-zero Fr from
// mt19937_64(seed ^ 0x9E5) with the density's zeros at seeded_mask(seed ^
// 0x3A5C2) positions; round constants and MDS from mt19937_64(0x9053).
Instance poseidon(size_t count, uint64_t seed, Density density, bool folded)
{
    if (count == 0) throw std::invalid_argument("Poseidon count must be positive");
    const PoseidonParams prm;
    const size_t n_in = 2 * count, nz = density_nonzero_count(n_in, density);
    const std::vector<uint8_t> mask = seeded_mask(n_in, nz, seed ^ 0x3A5C2);
    std::mt19937_64 rng(seed ^ 0x9E5);
    std::vector<Fr> inputs(n_in);
    for (size_t i = 0; i < n_in; ++i)
        if (mask[i]) inputs[i] = random_nonzero(rng);
    Builder b;
    for (size_t h = 0; h < count; ++h) {
        const Fr &lv = inputs[2 * h], &rv = inputs[2 * h + 1];
        const uint64_t digest = b.add_public(prm.hash2(lv, rv));
        const uint64_t left = b.add_private(lv), right = b.add_private(rv);
        hash_constraints(b, prm, left, lv, right, rv, digest, folded);
    }
    return b.finish();
}

}  // namespace nkr::groth16::gen

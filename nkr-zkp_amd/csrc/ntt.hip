// Fr NTT kernels for gfx950: LDS-resident sub-transforms with four-step
// inter-pass twiddles.  Algorithmic reference: libfqfft basic_radix2_domain
// (third_party/lightzk/depends/libsnark/depends/libfqfft/libfqfft/
// evaluation_domain/domains/basic_radix2_domain.tcc:41-112 and
// basic_radix2_domain_aux.tcc:45-76,172-180): the same omega, coset g = 5 and
// 1/N scaling, so outputs equal the reference's element for element.
#include "ntt.h"

namespace nkr {

namespace {

constexpr unsigned kMaxThreads = 512;

// Workgroup shape of the pass kernels: 2^lds_log elements of LDS per
// workgroup and 2^radix_log elements per thread per register round
// (threads = 2^(lds_log - radix_log)).  Chosen by measurement on MI355X
// inside whole proofs, where the MSM side stream shares the CUs: radix-4
// rounds need about half the registers of radix-8 ones, which leaves room
// for the other stream's waves.
unsigned g_lds_log = 11;
unsigned g_radix_log = 2;
// Pass split for transforms larger than one LDS tile: 0 balances the bits
// between passes, 1 gives the unit-stride last pass a full tile and keeps
// the strided passes narrow enough to gather several adjacent columns.
unsigned g_split_policy = 1;

struct PassArgs {
    fr_t *data;
    fr_t *const *batch;  // when set, workgroup row blockIdx.y transforms batch[blockIdx.y]
    const fr_t *tw_lo;
    const fr_t *tw_hi;
    const fr_t *tw_full;  // the pass's whole twiddle table, or null (two-level tables)
    const fr_t *roots;
    const fr_t *pre;
    const fr_t *post;
    fr_t post_const;
    uint32_t has_post_const;
    uint32_t log_n, S, np, log_L, log_cols, log_roots, m_lo;
    uint32_t twiddle;
};

__device__ __forceinline__ uint32_t brev_bits(uint32_t x, uint32_t bits)
{
    return bits == 0 ? 0u : (__brev(x) >> (32 - bits));
}

// omega^e from the two-level table.
__device__ __forceinline__ fr_t tw_pow(const PassArgs &a, uint32_t e)
{
    const uint32_t lo = e & ((1u << a.m_lo) - 1u);
    const uint32_t hi = e >> a.m_lo;
    return fp_mul(a.tw_hi[hi], a.tw_lo[lo]);
}

// The inter-pass twiddle of element j of the sub-transform at column lo:
// omega^e, e = (lo * bitrev(j)) << S, from the whole table when the domain has
// one.
__device__ __forceinline__ fr_t tw_at(const PassArgs &a, uint32_t j, uint32_t lo, uint32_t e)
{
    return a.tw_full ? a.tw_full[((uint64_t)j << a.log_L) + lo] : tw_pow(a, e);
}

// 2r as eight 32-bit words (r < 2^254, so 2r < 2^255).
constexpr Fp<FrParams> twice_modulus()
{
    Fp<FrParams> t{};
    uint64_t c = 0;
    for (int i = 0; i < 8; ++i) {
        const uint64_t v = 2ull * FrParams::P[i] + c;
        t.l[i] = (uint32_t)v;
        c = v >> 32;
    }
    return t;
}

// Lazy reduction: inside a lazily reduced pass every value is kept below 2r
// (r < 2^254, so 4r < 2^256 and sums of two such values fit the eight words);
// products skip their final subtraction (below 2r again when the other factor
// is a canonical twiddle), and the pass stores canonical values.

// x - 2r if x >= 2r, for x < 4r.
__device__ __forceinline__ fr_t sub_twice_r_if(const fr_t &x)
{
    constexpr Fp<FrParams> two_r = twice_modulus();
    fr_t d;
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) d.l[i] = subc32(x.l[i], two_r.l[i], borrow, borrow);
    fr_t r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.l[i] = borrow ? x.l[i] : d.l[i];
    return r;
}

// x + y for x, y < 2r, reduced below 2r.
__device__ __forceinline__ fr_t add_lazy(const fr_t &x, const fr_t &y)
{
    fr_t s;
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) s.l[i] = addc32(x.l[i], y.l[i], carry, carry);
    return sub_twice_r_if(s);
}

// x - y + 2r for x, y < 2r: in (0, 4r), not reduced.
__device__ __forceinline__ fr_t sub_lazy(const fr_t &x, const fr_t &y)
{
    constexpr Fp<FrParams> two_r = twice_modulus();
    fr_t s;
    uint32_t carry = 0, borrow = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) s.l[i] = addc32(x.l[i], two_r.l[i], carry, carry);
#pragma unroll
    for (int i = 0; i < 8; ++i) s.l[i] = subc32(s.l[i], y.l[i], borrow, borrow);
    return s;
}

// Decimation-in-frequency passes run lazily (their butterfly's difference
// goes straight into the product); decimation-in-time passes stay canonical,
// since their lazy butterfly needs a second conditional subtraction that
// costs what the product saves (measured on MI355X).
template <bool DIT>
constexpr bool kLazyNtt = !DIT;

// x * w * 2^-256 for x < 4r and a canonical w: below r + x*w/2^256 < 2r
// (the 29-bit product without its final subtraction).
__device__ __forceinline__ fr_t mul_lazy(const fr_t &x, const fr_t &w)
{
    return fp_mul_r29(x, w);
}

// LDS tile of n field elements: two planes of 16-byte halves (element i's
// low half in plane 0, high half in plane 1) so that every access is one
// 16-byte LDS instruction per half, and within each aligned run of 64
// elements the 16-element rows are XOR-permuted (row r by 5 * (r mod 4)),
// which spreads the strided accesses of every radix-4 round over distinct
// bank groups.
__device__ __forceinline__ uint32_t lds_slot(uint32_t i) { return i ^ (5u * ((i >> 4) & 3u)); }

__device__ __forceinline__ fr_t lds_get(const uint4 *tile, uint32_t n, uint32_t i)
{
    const uint32_t k = lds_slot(i);
    const uint4 a = tile[k], b = tile[n + k];
    fr_t r;
    r.l[0] = a.x;
    r.l[1] = a.y;
    r.l[2] = a.z;
    r.l[3] = a.w;
    r.l[4] = b.x;
    r.l[5] = b.y;
    r.l[6] = b.z;
    r.l[7] = b.w;
    return r;
}

__device__ __forceinline__ void lds_put(uint4 *tile, uint32_t n, uint32_t i, const fr_t &v)
{
    const uint32_t k = lds_slot(i);
    tile[k] = make_uint4(v.l[0], v.l[1], v.l[2], v.l[3]);
    tile[n + k] = make_uint4(v.l[4], v.l[5], v.l[6], v.l[7]);
}

// One round of R consecutive radix-2 stages (t .. t+R-1) on groups of 2^R
// elements held in registers.  A group fixes every index bit except the R
// bits [np-t-R, np-t) that the round's butterflies pair.  The twiddle of the
// pair (u, u + 2^(R-1-k)) at stage s = t + k is root^((j mod 2^(np-1-s)) * 2^s)
// for its lower index j, as in the radix-2 DIF recursion.
template <bool DIT, int R, bool LOW0>
__device__ __forceinline__ void ntt_round(uint4 *smem, uint32_t n, const PassArgs &a, uint32_t t, uint32_t cols,
                                          uint32_t root_shift)
{
    constexpr uint32_t G = 1u << R;
    const uint32_t low = LOW0 ? 0u : a.np - t - R;
    const uint32_t groups = (1u << (a.np - R)) << a.log_cols;
    for (uint32_t gi = threadIdx.x; gi < groups; gi += blockDim.x) {
        const uint32_t c = gi & (cols - 1);
        const uint32_t q = gi >> a.log_cols;
        const uint32_t lo_part = q & ((1u << low) - 1u);
        const uint32_t j0 = ((q >> low) << (a.np - t)) | lo_part;
        fr_t x[G];
#pragma unroll
        for (uint32_t u = 0; u < G; ++u) x[u] = lds_get(smem, n, (j0 + (u << low)) * cols + c);
#pragma unroll
        for (int kk = 0; kk < R; ++kk) {
            const int k = DIT ? (R - 1 - kk) : kk;
            const uint32_t s = t + (uint32_t)k;
            const uint32_t hs = 1u << (R - 1 - k);
#pragma unroll
            for (uint32_t u = 0; u < G; ++u) {
                if (u & hs) continue;
                const uint32_t e = (lo_part + ((u & (hs - 1u)) << low)) << s;
                const fr_t w = a.roots[e << root_shift];
                const fr_t X = x[u], Y = x[u + hs];
                // twiddle 1 needs no product.  In the round whose pairs are
                // adjacent (low = 0, so lo_part = 0) e depends on u alone and
                // the test folds away; other rounds multiply every pair.
                const bool one = LOW0 && (u & (hs - 1u)) == 0;
                if (!DIT) {
                    if (kLazyNtt<DIT>) {
                        x[u] = add_lazy(X, Y);
                        const fr_t D = sub_lazy(X, Y);
                        x[u + hs] = one ? sub_twice_r_if(D) : mul_lazy(D, w);
                    } else {
                        x[u] = fp_add(X, Y);
                        const fr_t D = fp_sub(X, Y);
                        x[u + hs] = one ? D : fp_mul(D, w);
                    }
                } else {
                    if (kLazyNtt<DIT>) {
                        const fr_t wy = one ? Y : mul_lazy(Y, w);
                        x[u] = add_lazy(X, wy);
                        x[u + hs] = sub_twice_r_if(sub_lazy(X, wy));
                    } else {
                        const fr_t wy = one ? Y : fp_mul(Y, w);
                        x[u] = fp_add(X, wy);
                        x[u + hs] = fp_sub(X, wy);
                    }
                }
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < G; ++u) lds_put(smem, n, (j0 + (u << low)) * cols + c, x[u]);
    }
}

// Radix-4 passes are held to 128 registers, four waves per SIMD: as many as
// the LDS tiles allow (two 64 KiB tiles per CU), at the price of a few spills.
template <bool DIT, int RMAX>
__global__ __launch_bounds__(kMaxThreads, RMAX == 2 ? 4 : 1) void k_ntt_pass(PassArgs a)
{
    extern __shared__ uint4 smem[];
    if (a.batch) a.data = a.batch[blockIdx.y];
    const uint32_t G = 1u << a.np;
    const uint32_t cols = 1u << a.log_cols;
    const uint32_t total = G << a.log_cols;
    const uint32_t L = 1u << a.log_L;
    const bool col_fast = L >= cols;  // consecutive groups are consecutive addresses
    const uint64_t g0 = (uint64_t)blockIdx.x << a.log_cols;

    for (uint32_t idx = threadIdx.x; idx < total; idx += blockDim.x) {
        uint32_t j, c;
        if (col_fast) {
            c = idx & (cols - 1);
            j = idx >> a.log_cols;
        } else {
            j = idx & (G - 1);
            c = idx >> a.np;
        }
        const uint64_t g = g0 + c;
        const uint64_t hi = g >> a.log_L;
        const uint32_t lo = (uint32_t)(g & (L - 1));
        const uint64_t pos = (hi << (a.log_n - a.S)) + ((uint64_t)j << a.log_L) + lo;
        fr_t v = a.data[pos];
        if (a.pre) v = kLazyNtt<DIT> ? mul_lazy(v, a.pre[pos]) : fp_mul(v, a.pre[pos]);
        if (DIT && a.twiddle) {
            const uint32_t e = (lo * brev_bits(j, a.np)) << a.S;
            if (e) v = kLazyNtt<DIT> ? mul_lazy(v, tw_at(a, j, lo, e)) : fp_mul(v, tw_at(a, j, lo, e));
        }
        lds_put(smem, total, j * cols + c, v);
    }
    __syncthreads();

    // Sub-transform in rounds of up to RMAX radix-2 stages: each thread keeps
    // a group of 2^RMAX elements in registers for the whole round, so LDS is
    // visited once per RMAX stages.  DIF rounds run stages t..t+r-1, DIT
    // rounds the same stages in reverse (its transpose).
    const uint32_t root_shift = a.log_roots - a.np;
    for (uint32_t done = 0; done < a.np;) {
        const uint32_t r = min((uint32_t)RMAX, a.np - done);
        const uint32_t t = DIT ? (a.np - done - r) : done;
        const bool low0 = (a.np - t - r == 0);
        if (r == 3) {
            if (low0)
                ntt_round<DIT, 3, true>(smem, total, a, t, cols, root_shift);
            else
                ntt_round<DIT, 3, false>(smem, total, a, t, cols, root_shift);
        } else if (r == 2) {
            if (low0)
                ntt_round<DIT, 2, true>(smem, total, a, t, cols, root_shift);
            else
                ntt_round<DIT, 2, false>(smem, total, a, t, cols, root_shift);
        } else {
            if (low0)
                ntt_round<DIT, 1, true>(smem, total, a, t, cols, root_shift);
            else
                ntt_round<DIT, 1, false>(smem, total, a, t, cols, root_shift);
        }
        __syncthreads();
        done += r;
    }

    for (uint32_t idx = threadIdx.x; idx < total; idx += blockDim.x) {
        uint32_t j, c;
        if (col_fast) {
            c = idx & (cols - 1);
            j = idx >> a.log_cols;
        } else {
            j = idx & (G - 1);
            c = idx >> a.np;
        }
        const uint64_t g = g0 + c;
        const uint64_t hi = g >> a.log_L;
        const uint32_t lo = (uint32_t)(g & (L - 1));
        const uint64_t pos = (hi << (a.log_n - a.S)) + ((uint64_t)j << a.log_L) + lo;
        fr_t v = lds_get(smem, total, j * cols + c);
        if (!DIT && a.twiddle) {
            const uint32_t e = (lo * brev_bits(j, a.np)) << a.S;
            if (e) v = kLazyNtt<DIT> ? mul_lazy(v, tw_at(a, j, lo, e)) : fp_mul(v, tw_at(a, j, lo, e));
        }
        if (a.post) v = kLazyNtt<DIT> ? mul_lazy(v, a.post[pos]) : fp_mul(v, a.post[pos]);
        if (a.has_post_const) v = kLazyNtt<DIT> ? mul_lazy(v, a.post_const) : fp_mul(v, a.post_const);
        a.data[pos] = kLazyNtt<DIT> ? fp_reduce_once(v) : v;
    }
}

// out[e] = base^e for e < count.
__global__ void k_pow_table(fr_t *out, fr_t base, uint32_t count)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    fr_t acc = fr_t::one();
    fr_t sq = base;
    for (uint32_t x = e; x; x >>= 1) {
        if (x & 1) acc = fp_mul(acc, sq);
        sq = fp_sqr(sq);
    }
    out[e] = acc;
}

// out[j] = base^bitrev(j) * scale for j < 2^log_n.
__global__ void k_pow_rev_table(fr_t *out, fr_t base, fr_t scale, uint32_t log_n)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (1u << log_n)) return;
    fr_t acc = scale;
    fr_t sq = base;
    for (uint32_t x = brev_bits(j, log_n); x; x >>= 1) {
        if (x & 1) acc = fp_mul(acc, sq);
        sq = fp_sqr(sq);
    }
    out[j] = acc;
}

// out[j] = scale * base^e, e = j (or bitrev(j) when rev), from the two-level
// tables lo[x] = base^x (x < 2^m) and hi[y] = base^(2^m y): one product per
// element (two with a scale) instead of an exponentiation.
__global__ void k_pow_two_level(fr_t *out, const fr_t *__restrict__ hi, const fr_t *__restrict__ lo, fr_t scale,
                                uint32_t m, uint32_t log_n, uint32_t rev, uint32_t has_scale)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (1u << log_n)) return;
    const uint32_t e = rev ? brev_bits(j, log_n) : j;
    fr_t v = fp_mul(hi[e >> m], lo[e & ((1u << m) - 1u)]);
    if (has_scale) v = fp_mul(v, scale);
    out[j] = v;
}

// out[(j << log_L) + lo] = omega^((lo * bitrev(j, np)) << S) for j < 2^np,
// lo < 2^log_L, from the two-level tables (one product per element).
__global__ void k_tw_full(fr_t *out, const fr_t *__restrict__ hi, const fr_t *__restrict__ lo_t, uint32_t m,
                          uint32_t np, uint32_t log_L, uint32_t S)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t(1) << (np + log_L))) return;
    const uint32_t j = (uint32_t)(i >> log_L), lo = (uint32_t)(i & ((1u << log_L) - 1u));
    const uint32_t e = (lo * brev_bits(j, np)) << S;
    out[i] = fp_mul(hi[e >> m], lo_t[e & ((1u << m) - 1u)]);
}

__global__ void k_bitrev(const fr_t *__restrict__ in, fr_t *__restrict__ out, uint32_t log_n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1u << log_n)) return;
    out[i] = in[brev_bits(i, log_n)];
}

__global__ void k_quotient(const fr_t *__restrict__ a, const fr_t *__restrict__ b, const fr_t *__restrict__ c,
                           fr_t *__restrict__ h, fr_t zinv, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    h[i] = fp_mul(fp_sub(fp_mul(a[i], b[i]), c[i]), zinv);
}

// k_quotient for job blockIdx.y with pointers abch[4 * job + {0, 1, 2, 3}] = a, b, c, h.
__global__ void k_quotient_batch(fr_t *const *__restrict__ abch, fr_t zinv, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fr_t *const *p = abch + 4 * blockIdx.y;
    p[3][i] = fp_mul(fp_sub(fp_mul(p[0][i], p[1][i]), p[2][i]), zinv);
}

// 2^28-th root of unity of Fr (libff alt_bn128_init.cpp:56, as an integer).
const uint32_t kRootOfUnity28[8] = {0x725b19f0u, 0x9bd61b6eu, 0x41112ed4u, 0x402d111eu,
                                    0x8ef62abcu, 0x00e0a7ebu, 0xa58a7e85u, 0x2a3c09f0u};
constexpr unsigned kTwoAdicity = 28;

}  // namespace

void ntt_set_shape(unsigned lds_log, unsigned radix_log, unsigned split_policy)
{
    NKR_REQUIRE(lds_log == 10 || lds_log == 11, "LDS elements per workgroup must be 2^10 or 2^11");
    NKR_REQUIRE(radix_log == 2 || radix_log == 3, "register radix must be 2^2 or 2^3");
    NKR_REQUIRE(split_policy <= 1, "pass split policy must be 0 (balanced) or 1 (contiguous last)");
    g_lds_log = lds_log;
    g_radix_log = radix_log;
    g_split_policy = split_policy;
}

fr_t fr_from_u64(uint64_t x)
{
    fr_t t = fr_t::zero();
    t.l[0] = (uint32_t)x;
    t.l[1] = (uint32_t)(x >> 32);
    return fp_to_mont(t);
}

fr_t fr_pow_u64(const fr_t &x, uint64_t e)
{
    fr_t acc = fr_t::one(), sq = x;
    for (; e; e >>= 1) {
        if (e & 1) acc = fp_mul(acc, sq);
        sq = fp_sqr(sq);
    }
    return acc;
}

fr_t fr_root_of_unity(unsigned log_n)
{
    NKR_REQUIRE(log_n <= kTwoAdicity, "domain exceeds the 2-adicity of Fr (2^28)");
    fr_t w;
    for (int i = 0; i < 8; ++i) w.l[i] = kRootOfUnity28[i];
    w = fp_to_mont(w);
    for (unsigned i = kTwoAdicity; i > log_n; --i) w = fp_sqr(w);
    return w;
}

void NttDomain::plan(unsigned log_n, const NttShape *shape)
{
    NKR_REQUIRE(log_n >= 1 && log_n <= kTwoAdicity, "NTT size must be 2^1 .. 2^28");
    const size_t n = size_t(1) << log_n;
    omega_[0] = fr_root_of_unity(log_n);
    omega_[1] = fp_inv(omega_[0]);
    n_inv_ = fp_inv(fr_from_u64(n));
    const fr_t g = fr_from_u64(5);  // multiplicative generator (alt_bn128_init.cpp:55)
    zinv_ = fp_inv(fp_sub(fr_pow_u64(g, n), fr_t::one()));

    // pass plan: sub-transforms of at most lds_log bits
    lds_log_ = shape && shape->lds_log ? shape->lds_log : g_lds_log;
    radix_log_ = shape && shape->radix_log ? shape->radix_log : g_radix_log;
    const unsigned split_policy = shape && shape->split_set ? shape->split_policy : g_split_policy;
    NKR_REQUIRE(lds_log_ == 10 || lds_log_ == 11, "LDS elements per workgroup must be 2^10 or 2^11");
    NKR_REQUIRE(radix_log_ == 2 || radix_log_ == 3, "register radix must be 2^2 or 2^3");
    NKR_REQUIRE(split_policy <= 1, "pass split policy must be 0 (balanced) or 1 (contiguous last)");
    std::vector<unsigned> split;
    if (log_n <= lds_log_) {
        split = {log_n};
    } else if (split_policy == 1) {
        // The last pass (unit stride) takes a full LDS tile; the strided passes
        // before it take at most lds_log - 2 bits, so that each workgroup
        // gathers >= 4 adjacent columns (>= 128 contiguous bytes per row).
        const unsigned rest = log_n - lds_log_, cap = lds_log_ - 2;
        const unsigned k = (rest + cap - 1) / cap;
        for (unsigned i = 0; i < k; ++i) split.push_back(rest / k + (i < rest % k ? 1 : 0));
        split.push_back(lds_log_);
    } else if (log_n <= 2 * lds_log_) {
        split = {(log_n + 1) / 2, log_n / 2};
    } else {
        const unsigned a = (log_n + 2) / 3, b = (log_n + 1) / 3, c = log_n / 3;
        NKR_REQUIRE(a <= lds_log_, "NTT size too large for the configured workgroup shape");
        split = {a, b, c};
    }
    unsigned S = 0;
    for (unsigned np : split) {
        Pass p;
        p.S = S;
        p.np = np;
        p.log_L = log_n - S - np;
        const unsigned groups_log = log_n - np;
        unsigned lc = std::min(groups_log, lds_log_ - np);
        if (p.log_L > 0 && p.log_L < lc) lc = p.log_L;
        p.log_cols = lc;
        passes_.push_back(p);
        S += np;
    }

    m_lo_ = (log_n + 1) / 2;
    log_roots_ = std::min(log_n, 12u);
    lo_count_ = 1u << m_lo_;
    hi_count_ = 1u << (log_n - m_lo_);
    root_count_ = std::max(1u, 1u << (log_roots_ - 1));
}

// The inter-pass twiddle tables and sub-transform roots of one direction.
static void build_twiddles(fr_t *lo, fr_t *hi, fr_t *roots, const fr_t &omega, uint32_t lo_count, uint32_t hi_count,
                           uint32_t root_count, unsigned log_n, unsigned log_roots, hipStream_t stream)
{
    k_pow_table<<<div_up(lo_count, 256), 256, 0, stream>>>(lo, omega, lo_count);
    NKR_LAUNCH_CHECK();
    k_pow_table<<<div_up(hi_count, 256), 256, 0, stream>>>(hi, fr_pow_u64(omega, lo_count), hi_count);
    NKR_LAUNCH_CHECK();
    const fr_t root = fr_pow_u64(omega, uint64_t(1) << (log_n - log_roots));
    k_pow_table<<<div_up(root_count, 256), 256, 0, stream>>>(roots, root, root_count);
    NKR_LAUNCH_CHECK();
}

NttDomain::NttDomain(unsigned log_n, hipStream_t stream, const NttShape *shape) : log_n_(log_n)
{
    plan(log_n, shape);
    const size_t n = size_t(1) << log_n;
    const fr_t g = fr_from_u64(5);
    const fr_t g_inv = fp_inv(g);
    for (int d = 0; d < 2; ++d) {
        tw_lo_[d].allocate(lo_count_);
        tw_hi_[d].allocate(hi_count_);
        roots_[d].allocate(root_count_);
        build_twiddles(tw_lo_[d].get(), tw_hi_[d].get(), roots_[d].get(), omega_[d], lo_count_, hi_count_,
                       root_count_, log_n, log_roots_, stream);
        tw_lo_p_[d] = tw_lo_[d].get();
        tw_hi_p_[d] = tw_hi_[d].get();
        roots_p_[d] = roots_[d].get();
        // whole twiddle tables of the passes that have a twiddle (all but the last)
        tw_full_[d].resize(passes_.size());
        for (size_t pi = 0; pi + 1 < passes_.size(); ++pi) {
            const Pass &p = passes_[pi];
            const size_t count = size_t(1) << (p.np + p.log_L);
            tw_full_[d][pi].allocate(count);
            k_tw_full<<<(uint32_t)div_up(count, 256), 256, 0, stream>>>(tw_full_[d][pi].get(), tw_hi_p_[d],
                                                                      tw_lo_p_[d], m_lo_, p.np, p.log_L, p.S);
            NKR_LAUNCH_CHECK();
        }
    }
    coset_nat_.allocate(n);
    coset_rev_.allocate(n);
    icoset_rev_.allocate(n);
    k_pow_table<<<div_up(n, 256), 256, 0, stream>>>(coset_nat_.get(), g, (uint32_t)n);
    NKR_LAUNCH_CHECK();
    k_pow_rev_table<<<div_up(n, 256), 256, 0, stream>>>(coset_rev_.get(), g, n_inv_, log_n);
    NKR_LAUNCH_CHECK();
    k_pow_rev_table<<<div_up(n, 256), 256, 0, stream>>>(icoset_rev_.get(), g_inv, n_inv_, log_n);
    NKR_LAUNCH_CHECK();
    coset_nat_p_ = coset_nat_.get();
    coset_rev_p_ = coset_rev_.get();
    icoset_rev_p_ = icoset_rev_.get();
}

NttDomain::NttDomain(unsigned log_n, int kind, fr_t *arena, hipStream_t stream) : log_n_(log_n)
{
    NKR_REQUIRE(kind >= 0 && kind <= 3, "unknown transform kind (0 FFT, 1 iFFT, 2 cosetFFT, 3 icosetFFT)");
    NKR_REQUIRE(arena, "null table arena");
    plan(log_n, nullptr);
    arena_kind_ = kind;
    const int d = (kind == 1 || kind == 3) ? 1 : 0;
    fr_t *p = arena;
    fr_t *lo = p, *hi = lo + lo_count_, *roots = hi + hi_count_;
    p = roots + root_count_;
    build_twiddles(lo, hi, roots, omega_[d], lo_count_, hi_count_, root_count_, log_n, log_roots_, stream);
    tw_lo_p_[d] = lo;
    tw_hi_p_[d] = hi;
    roots_p_[d] = roots;
    if (kind >= 2) {
        // g^i (cosetFFT, natural order) or g^-rev(j) / N (icosetFFT, applied
        // after the bit-reversed output), from two-level power tables of g^(+-1)
        const size_t n = size_t(1) << log_n;
        const fr_t base = kind == 2 ? fr_from_u64(5) : fp_inv(fr_from_u64(5));
        fr_t *glo = p, *ghi = glo + lo_count_, *table = ghi + hi_count_;
        NKR_REQUIRE(size_t(table + n - arena) <= standalone_arena_elems(log_n), "table arena layout overflow");
        k_pow_table<<<div_up(lo_count_, 256), 256, 0, stream>>>(glo, base, lo_count_);
        NKR_LAUNCH_CHECK();
        k_pow_table<<<div_up(hi_count_, 256), 256, 0, stream>>>(ghi, fr_pow_u64(base, lo_count_), hi_count_);
        NKR_LAUNCH_CHECK();
        k_pow_two_level<<<div_up(n, 256), 256, 0, stream>>>(table, ghi, glo, n_inv_, m_lo_, log_n,
                                                            kind == 3 ? 1u : 0u, kind == 3 ? 1u : 0u);
        NKR_LAUNCH_CHECK();
        (kind == 2 ? coset_nat_p_ : icoset_rev_p_) = table;
    }
}

void NttDomain::run(fr_t *data, fr_t *const *batch, uint32_t count, bool dit, bool inverse, const fr_t *pre,
                    const fr_t *post, const fr_t *post_const, hipStream_t stream) const
{
    NKR_REQUIRE(count >= 1 && count <= 65535, "NTT batch must hold 1 .. 65535 transforms");
    const int d = inverse ? 1 : 0;
    NKR_REQUIRE(tw_lo_p_[d], "this transform direction has no tables in this domain");
    const int P = (int)passes_.size();
    for (int k = 0; k < P; ++k) {
        const int pi = dit ? (P - 1 - k) : k;
        const Pass &p = passes_[pi];
        PassArgs a{};
        a.data = data;
        a.batch = batch;
        a.tw_lo = tw_lo_p_[d];
        a.tw_hi = tw_hi_p_[d];
        a.tw_full = (size_t)pi < tw_full_[d].size() ? tw_full_[d][pi].get() : nullptr;
        a.roots = roots_p_[d];
        a.pre = (k == 0) ? pre : nullptr;
        a.post = (k == P - 1) ? post : nullptr;
        a.has_post_const = (k == P - 1 && post_const) ? 1u : 0u;
        if (a.has_post_const) a.post_const = *post_const;
        a.log_n = log_n_;
        a.S = p.S;
        a.np = p.np;
        a.log_L = p.log_L;
        a.log_cols = p.log_cols;
        a.log_roots = log_roots_;
        a.m_lo = m_lo_;
        a.twiddle = (pi != P - 1) ? 1u : 0u;
        const dim3 blocks((uint32_t)((size_t(1) << (log_n_ - p.np)) >> p.log_cols), count);
        const uint32_t elems = (1u << p.np) << p.log_cols;
        const size_t lds = size_t(sizeof(fr_t)) * elems;
        const uint32_t threads = std::min(kMaxThreads, std::max(64u, elems >> radix_log_));
        if (radix_log_ >= 3) {
            if (dit)
                k_ntt_pass<true, 3><<<blocks, threads, lds, stream>>>(a);
            else
                k_ntt_pass<false, 3><<<blocks, threads, lds, stream>>>(a);
        } else {
            if (dit)
                k_ntt_pass<true, 2><<<blocks, threads, lds, stream>>>(a);
            else
                k_ntt_pass<false, 2><<<blocks, threads, lds, stream>>>(a);
        }
        NKR_LAUNCH_CHECK();
    }
}

void NttDomain::dif(fr_t *data, bool inverse, const fr_t *pre, const fr_t *post, const fr_t *post_const,
                    hipStream_t stream) const
{
    run(data, nullptr, 1, false, inverse, pre, post, post_const, stream);
}

void NttDomain::dit(fr_t *data, bool inverse, const fr_t *pre, const fr_t *post, hipStream_t stream) const
{
    run(data, nullptr, 1, true, inverse, pre, post, nullptr, stream);
}

void NttDomain::dif_batch(fr_t *const *batch_dev, uint32_t count, bool inverse, const fr_t *pre, const fr_t *post,
                          const fr_t *post_const, hipStream_t stream) const
{
    run(nullptr, batch_dev, count, false, inverse, pre, post, post_const, stream);
}

void NttDomain::dit_batch(fr_t *const *batch_dev, uint32_t count, bool inverse, const fr_t *pre, const fr_t *post,
                          hipStream_t stream) const
{
    run(nullptr, batch_dev, count, true, inverse, pre, post, nullptr, stream);
}

void NttDomain::quotient_batch(fr_t *const *abch_dev, uint32_t count, hipStream_t stream) const
{
    NKR_REQUIRE(count >= 1 && count <= 65535, "quotient batch must hold 1 .. 65535 jobs");
    const size_t n = size();
    k_quotient_batch<<<dim3(div_up(n, 256), count), 256, 0, stream>>>(abch_dev, zinv_, (uint32_t)n);
    NKR_LAUNCH_CHECK();
}

void NttDomain::bitrev(const fr_t *in, fr_t *out, hipStream_t stream) const
{
    const size_t n = size();
    k_bitrev<<<div_up(n, 256), 256, 0, stream>>>(in, out, log_n_);
    NKR_LAUNCH_CHECK();
}

void NttDomain::standalone(fr_t *data, fr_t *scratch, int kind, hipStream_t stream) const
{
    NKR_REQUIRE(arena_kind_ < 0 || kind == arena_kind_, "this domain's tables serve another transform kind");
    switch (kind) {
    case 0: dif(data, false, nullptr, nullptr, nullptr, stream); break;
    case 1: dif(data, true, nullptr, nullptr, &n_inv_, stream); break;
    case 2: dif(data, false, coset_nat_p_, nullptr, nullptr, stream); break;
    case 3: dif(data, true, nullptr, icoset_rev_p_, nullptr, stream); break;
    default: throw Error("unknown transform kind (0 FFT, 1 iFFT, 2 cosetFFT, 3 icosetFFT)");
    }
    bitrev(data, scratch, stream);
    NKR_HIP(hipMemcpyAsync(data, scratch, size() * sizeof(fr_t), hipMemcpyDeviceToDevice, stream));
}

void NttDomain::quotient(const fr_t *a, const fr_t *b, const fr_t *c, fr_t *h, hipStream_t stream) const
{
    const size_t n = size();
    k_quotient<<<div_up(n, 256), 256, 0, stream>>>(a, b, c, h, zinv_, (uint32_t)n);
    NKR_LAUNCH_CHECK();
}

}  // namespace nkr

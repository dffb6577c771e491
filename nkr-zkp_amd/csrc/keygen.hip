// Groth16 key generation (Backend::setup) on one GPU.
//
// The CPU backend (protocols/groth16/backends/cpu/src/backend.cpp:41-154):
//   L = Lagrange basis of the size-N domain evaluated at t
//   at = A^T L, bt = B^T L, ct = C^T L                  (column sums)
//   l[i]  = (beta at_i + alpha bt_i + ct_i) / delta     for private i, else 0
//   ic[i] = (beta at_i + alpha bt_i + ct_i) / gamma     for i <= public inputs
//   h[i]  = t^i Z(t) / delta  for i < N - 1, h[N-1] = 0, Z(t) = t^N - 1
//   queries = those scalars times G1 (and bt times G2)
//   alpha_g1_beta_g2 = e(alpha G1, beta G2)
// Here: L, the query scalars and the powers of t are elementwise kernels; the
// column sums run the SpMV kernel over the transposed matrices (built on the
// host by a counting sort); every query is a fixed-base comb multiplication
// (FixedBaseMul); the pairing runs on the host.
#include "keygen.h"

#include <cstring>
#include <vector>

#include "ntt.h"
#include "pairing.h"
#include "prover.h"
#include "runtime.h"
#include "setup.h"
#include "spmv.h"

namespace nkr {

namespace {

// L_i(t) = Z(t)/N * w^i / (t - w^i), w a primitive N-th root of unity.
__global__ void k_lagrange(fr_t *out, fr_t omega, fr_t t, fr_t zt_over_n, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fr_t w = fr_t::one(), sq = omega;
    for (uint32_t e = i; e; e >>= 1) {
        if (e & 1) w = fp_mul(w, sq);
        sq = fp_sqr(sq);
    }
    out[i] = fp_mul(fp_mul(zt_over_n, w), fp_inv(fp_sub(t, w)));
}

__global__ void k_query_scalars(const fr_t *at, const fr_t *bt, const fr_t *ct, fr_t *l, fr_t *ic, uint32_t cols,
                                uint32_t public_inputs, fr_t alpha, fr_t beta, fr_t delta_inv, fr_t gamma_inv)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cols) return;
    const fr_t v = fp_add(fp_add(fp_mul(beta, at[i]), fp_mul(alpha, bt[i])), ct[i]);
    l[i] = (i > public_inputs) ? fp_mul(v, delta_inv) : fr_t::zero();
    if (i <= public_inputs) ic[i] = fp_mul(v, gamma_inv);
}

// h[i] = t^i * scale for i < n - 1, h[n - 1] = 0.
__global__ void k_h_scalars(fr_t *h, fr_t t, fr_t scale, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i + 1 == n) {
        h[i] = fr_t::zero();
        return;
    }
    fr_t acc = scale, sq = t;
    for (uint32_t e = i; e; e >>= 1) {
        if (e & 1) acc = fp_mul(acc, sq);
        sq = fp_sqr(sq);
    }
    h[i] = acc;
}

// Transpose of a CSR matrix (rows x cols) into CSR of cols x rows.
struct HostCsr {
    std::vector<uint32_t> row_ptr, col_idx;
    std::vector<fr_t> values;
};

HostCsr transpose(uint64_t rows, uint64_t cols, const uint32_t *rp, const uint32_t *ci, const void *vals_v,
                  uint64_t nnz)
{
    const fr_t *vals = static_cast<const fr_t *>(vals_v);
    HostCsr t;
    t.row_ptr.assign(cols + 1, 0);
    t.col_idx.resize(nnz);
    t.values.resize(nnz);
    for (uint64_t j = 0; j < nnz; ++j) ++t.row_ptr[ci[j] + 1];
    for (uint64_t c = 0; c < cols; ++c) t.row_ptr[c + 1] += t.row_ptr[c];
    std::vector<uint32_t> cursor(t.row_ptr.begin(), t.row_ptr.end() - 1);
    for (uint64_t r = 0; r < rows; ++r)
        for (uint32_t j = rp[r]; j < rp[r + 1]; ++j) {
            const uint32_t dst = cursor[ci[j]]++;
            t.col_idx[dst] = (uint32_t)r;
            t.values[dst] = vals[j];
        }
    return t;
}

template <class T>
void upload(DeviceBuffer<T> &buf, const T *src, size_t count, hipStream_t stream)
{
    buf.allocate(std::max<size_t>(count, 1));
    if (count) NKR_HIP(hipMemcpyAsync(buf.get(), src, count * sizeof(T), hipMemcpyHostToDevice, stream));
}

fq_t fq_limbs(const uint32_t (&v)[8]) { return fp_to_mont(fq_t{{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]}}); }

struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard()
    {
        if (s) (void)hipStreamDestroy(s);
    }
};

}  // namespace

g1_jac g1_generator()
{
    g1_jac g;
    g.X = fp_to_mont(fq_t{{1, 0, 0, 0, 0, 0, 0, 0}});
    g.Y = fp_to_mont(fq_t{{2, 0, 0, 0, 0, 0, 0, 0}});
    g.Z = fq_t::one();
    return g;
}

g2_jac g2_generator()
{
    static const uint32_t x0[8] = {0xd992f6edu, 0x46debd5cu, 0xf75edaddu, 0x674322d4u,
                                   0x5e5c4479u, 0x426a0066u, 0x121f1e76u, 0x1800deefu};
    static const uint32_t x1[8] = {0xaef312c2u, 0x97e485b7u, 0x35a9e712u, 0xf1aa4933u,
                                   0x31fb5d25u, 0x7260bfb7u, 0x920d483au, 0x198e9393u};
    static const uint32_t y0[8] = {0x66fa7daau, 0x4ce6cc01u, 0x0c43d37bu, 0xe3d1e769u,
                                   0x8dcb408fu, 0x4aab7180u, 0xdb8c6debu, 0x12c85ea5u};
    static const uint32_t y1[8] = {0xd122975bu, 0x55acdadcu, 0x70b38ef3u, 0xbc4b3133u,
                                   0x690c3395u, 0xec9e99adu, 0x585ff075u, 0x090689d0u};
    g2_jac g;
    g.X = {fq_limbs(x0), fq_limbs(x1)};
    g.Y = {fq_limbs(y0), fq_limbs(y1)};
    g.Z = fq2_t::one();
    return g;
}

void groth16_keygen(const nkr_g16_circuit &c, const Trapdoors &td, int device, const nkr_g16_pk_out &pk,
                    const nkr_g16_vk_out &vk)
{
    check_circuit(c);
    NKR_REQUIRE(pk.a_query && pk.b_g1_query && pk.b_g2_query && pk.l_query && pk.h_query && pk.alpha_g1 &&
                    pk.beta_g1 && pk.beta_g2 && pk.delta_g1 && pk.delta_g2,
                "proving key output has null members");
    NKR_REQUIRE(vk.alpha_g1_beta_g2 && vk.gamma_g2 && vk.delta_g2 && vk.ic, "verification key output has null members");
    NKR_REQUIRE(!td.gamma.is_zero() && !td.delta.is_zero(), "gamma and delta must be invertible");
    const uint32_t n = (uint32_t)c.rows, cols = (uint32_t)c.cols, pub = (uint32_t)c.public_inputs;
    const unsigned log_n = ceil_log2(n);
    const fr_t zt = fp_sub(fr_pow_u64(td.t, n), fr_t::one());
    NKR_REQUIRE(!zt.is_zero(), "trapdoor t lies in the evaluation domain");

    NKR_HIP(hipSetDevice(device));
    StreamGuard sg;
    NKR_HIP(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
    hipStream_t s = sg.s;

    // Lagrange basis at t
    DeviceBuffer<fr_t> lag(n);
    const fr_t zt_over_n = fp_mul(zt, fp_inv(fr_from_u64(n)));
    k_lagrange<<<div_up(n, 256), 256, 0, s>>>(lag.get(), fr_root_of_unity(log_n), td.t, zt_over_n, n);
    NKR_LAUNCH_CHECK();

    // column evaluations at, bt, ct = M^T L
    const uint32_t *rps[3] = {c.a_row_ptr, c.b_row_ptr, c.c_row_ptr};
    const uint32_t *cis[3] = {c.a_col_idx, c.b_col_idx, c.c_col_idx};
    const void *vs[3] = {c.a_values, c.b_values, c.c_values};
    const uint64_t nnz[3] = {c.a_nnz, c.b_nnz, c.c_nnz};
    DeviceBuffer<uint32_t> trp[3], tci[3], tlong[3];
    DeviceBuffer<fr_t> tval[3];
    CsrView views[3];
    for (int m = 0; m < 3; ++m) {
        const HostCsr t = transpose(n, cols, rps[m], cis[m], vs[m], nnz[m]);
        upload(trp[m], t.row_ptr.data(), t.row_ptr.size(), s);
        upload(tci[m], t.col_idx.data(), t.col_idx.size(), s);
        upload(tval[m], t.values.data(), t.values.size(), s);
        // the host vectors die at the end of this iteration
        NKR_HIP(hipStreamSynchronize(s));
        views[m] = CsrView{trp[m].get(), tci[m].get(), tval[m].get(), cols, (uint32_t)nnz[m]};
        views[m].n_long = spmv_long_rows(t.row_ptr.data(), cols, tlong[m]);
        views[m].long_rows = tlong[m].get();
    }
    DeviceBuffer<fr_t> at(cols), bt(cols), ct(cols), l(cols), ic(pub + 1), h(n);
    spmv3(views[0], views[1], views[2], lag.get(), at.get(), bt.get(), ct.get(), s);

    const fr_t delta_inv = fp_inv(td.delta), gamma_inv = fp_inv(td.gamma);
    k_query_scalars<<<div_up(cols, 256), 256, 0, s>>>(at.get(), bt.get(), ct.get(), l.get(), ic.get(), cols, pub,
                                                       td.alpha, td.beta, delta_inv, gamma_inv);
    NKR_LAUNCH_CHECK();
    k_h_scalars<<<div_up(n, 256), 256, 0, s>>>(h.get(), td.t, fp_mul(zt, delta_inv), n);
    NKR_LAUNCH_CHECK();

    // fixed-base multiples of the generators
    const FixedBaseMul<fq_t> fb1(g1_generator(), s);
    const FixedBaseMul<fq2_t> fb2(g2_generator(), s);
    const uint64_t big = std::max<uint64_t>(cols, n);
    DeviceBuffer<g1_jac> out1(big);
    DeviceBuffer<g2_jac> out2(cols);
    struct G1Query {
        const fr_t *scalars;
        uint64_t count;
        void *dst;
    };
    const G1Query g1q[5] = {{at.get(), cols, pk.a_query},
                            {bt.get(), cols, pk.b_g1_query},
                            {l.get(), cols, pk.l_query},
                            {h.get(), n, pk.h_query},
                            {ic.get(), pub + 1ull, vk.ic}};
    for (const G1Query &q : g1q) {
        fb1.run(q.scalars, q.count, out1.get(), s);
        NKR_HIP(hipMemcpyAsync(q.dst, out1.get(), q.count * sizeof(g1_jac), hipMemcpyDeviceToHost, s));
        NKR_HIP(hipStreamSynchronize(s));  // out1 is reused by the next query
    }
    fb2.run(bt.get(), cols, out2.get(), s);
    NKR_HIP(hipMemcpyAsync(pk.b_g2_query, out2.get(), cols * sizeof(g2_jac), hipMemcpyDeviceToHost, s));

    // alpha, beta, delta in G1; beta, delta, gamma in G2
    const fr_t k1[3] = {td.alpha, td.beta, td.delta}, k2[3] = {td.beta, td.delta, td.gamma};
    DeviceBuffer<fr_t> kd(6);
    NKR_HIP(hipMemcpyAsync(kd.get(), k1, sizeof(k1), hipMemcpyHostToDevice, s));
    NKR_HIP(hipMemcpyAsync(kd.get() + 3, k2, sizeof(k2), hipMemcpyHostToDevice, s));
    DeviceBuffer<g1_jac> p1(3);
    DeviceBuffer<g2_jac> p2(3);
    fb1.run(kd.get(), 3, p1.get(), s);
    fb2.run(kd.get() + 3, 3, p2.get(), s);
    g1_jac h1[3];
    g2_jac h2[3];
    NKR_HIP(hipMemcpyAsync(h1, p1.get(), sizeof(h1), hipMemcpyDeviceToHost, s));
    NKR_HIP(hipMemcpyAsync(h2, p2.get(), sizeof(h2), hipMemcpyDeviceToHost, s));
    NKR_HIP(hipStreamSynchronize(s));
    std::memcpy(pk.alpha_g1, &h1[0], sizeof(g1_jac));
    std::memcpy(pk.beta_g1, &h1[1], sizeof(g1_jac));
    std::memcpy(pk.delta_g1, &h1[2], sizeof(g1_jac));
    std::memcpy(pk.beta_g2, &h2[0], sizeof(g2_jac));
    std::memcpy(pk.delta_g2, &h2[1], sizeof(g2_jac));
    std::memcpy(vk.delta_g2, &h2[1], sizeof(g2_jac));
    std::memcpy(vk.gamma_g2, &h2[2], sizeof(g2_jac));
    const fq12_t ab = bn254_pairing(h1[0], h2[0]);
    std::memcpy(vk.alpha_g1_beta_g2, &ab, sizeof(ab));
}

}  // namespace nkr

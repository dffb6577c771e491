// H query in the coset-evaluation basis (see hquery.h): a radix-2
// decimation-in-frequency transform over G1, one launch per stage, whose
// twiddle "products" are scalar multiplications of the group elements.
#include "hquery.h"

namespace nkr {

namespace {

__device__ __forceinline__ uint32_t rev_bits(uint32_t x, uint32_t bits)
{
    return bits == 0 ? 0u : (__brev(x) >> (32 - bits));
}

// P_k = (g^-k / n) * H_k; icoset_rev[j] holds g^-bitrev(j) / n.
__global__ __launch_bounds__(128) void k_hq_scale(const Jacobian<fq_t> *__restrict__ in,
                                                  const fr_t *__restrict__ icoset_rev, uint32_t log_n,
                                                  Xyzz<fq_t> *__restrict__ P)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (1u << log_n)) return;
    const fr_t c = fp_from_mont(icoset_rev[rev_bits(k, log_n)]);
    P[k] = xyzz_mul_u256(jacobian_to_xyzz(in[k]), c.l);
}

// w[e] = base^e for e < count (one exponentiation per thread).
__global__ void k_hq_roots(fr_t *__restrict__ w, fr_t base, uint32_t count)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    fr_t acc = fr_t::one(), sq = base;
    for (uint32_t x = e; x; x >>= 1) {
        if (x & 1u) acc = fp_mul(acc, sq);
        sq = fp_sqr(sq);
    }
    w[e] = acc;
}

// Stage s of the DIF transform (half-length h = n / 2^(s+1)): pair
// (i0, i0 + h) becomes (X + Y, (X - Y) * w^(j * 2^s)) for j = i0 mod h.
__global__ __launch_bounds__(128) void k_hq_stage(Xyzz<fq_t> *__restrict__ P, const fr_t *__restrict__ w,
                                                  uint32_t log_n, uint32_t s)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (1u << (log_n - 1))) return;
    const uint32_t h = 1u << (log_n - s - 1);
    const uint32_t j = t & (h - 1u);
    const uint32_t i0 = ((t >> (log_n - s - 1)) << (log_n - s)) | j;
    const Xyzz<fq_t> X = P[i0], Y = P[i0 + h];
    P[i0] = xyzz_add(X, Y);
    Xyzz<fq_t> D = xyzz_add(X, xyzz_neg(Y));
    const uint32_t e = j << s;
    if (e) {
        const fr_t c = fp_from_mont(w[e]);
        D = xyzz_mul_u256(D, c.l);
    }
    P[i0 + h] = D;
}

// XYZZ to Jacobian without an inversion: Z = ZZ*ZZZ, X = X*ZZ*ZZZ^2,
// Y = Y*ZZ^3*ZZZ^2 (x = X/ZZ = X'/Z'^2, y = Y/ZZZ = Y'/Z'^3; infinity keeps Z = 0).
__global__ __launch_bounds__(128) void k_hq_jacobian(const Xyzz<fq_t> *__restrict__ P, uint32_t n,
                                                     Jacobian<fq_t> *__restrict__ out)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Xyzz<fq_t> p = P[k];
    const fq_t zzz2 = fp_sqr(p.ZZZ);
    const fq_t zz3 = fp_mul(fp_sqr(p.ZZ), p.ZZ);
    Jacobian<fq_t> r;
    r.Z = fp_mul(p.ZZ, p.ZZZ);
    r.X = fp_mul(fp_mul(p.X, p.ZZ), zzz2);
    r.Y = fp_mul(fp_mul(p.Y, zz3), zzz2);
    out[k] = r;
}

}  // namespace

void h_query_coset_basis(const Jacobian<fq_t> *in, const NttDomain &domain, Jacobian<fq_t> *out,
                         hipStream_t stream)
{
    const uint32_t log_n = domain.log_n();
    NKR_REQUIRE(log_n >= 1, "the H query needs a domain of at least 2 points");
    const uint32_t n = 1u << log_n, half = n / 2;
    DeviceBuffer<Xyzz<fq_t>> P(n);
    DeviceBuffer<fr_t> w(half);
    k_hq_scale<<<div_up(n, 128), 128, 0, stream>>>(in, domain.icoset_rev_scaled(), log_n, P.get());
    NKR_LAUNCH_CHECK();
    k_hq_roots<<<div_up(half, 256), 256, 0, stream>>>(w.get(), fp_inv(domain.omega()), half);
    NKR_LAUNCH_CHECK();
    for (uint32_t s = 0; s < log_n; ++s) {
        k_hq_stage<<<div_up(half, 128), 128, 0, stream>>>(P.get(), w.get(), log_n, s);
        NKR_LAUNCH_CHECK();
    }
    k_hq_jacobian<<<div_up(n, 128), 128, 0, stream>>>(P.get(), n, out);
    NKR_LAUNCH_CHECK();
    // the temporaries die with this scope: finish before returning
    NKR_HIP(hipStreamSynchronize(stream));
}

}  // namespace nkr

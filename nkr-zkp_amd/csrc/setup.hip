// Fixed-base batch scalar multiplication on the GPU: out_i = k_i * P for one
// base P.  This is the building block of Groth16 key generation (the
// reference's FixedBaseMsm, operators/include/nkr/operators/fixed_base_msm.hpp:
// 14-39, and its HIP fixmsm_kernel, operators/hip/include/msm/fixmsm_kernel.hip:
// 6-20) and produces the synthetic proving-key vectors (k_i * G) the
// benchmarks and tests use.
//
// A table T[j][d] = d * 2^(8j) * P (32 windows of 8 bits, affine) is built
// once; each thread then needs 32 mixed additions and one inversion to emit a
// normalised point (Z = 1, or (0, 1, 0) for k = 0).
#include "setup.h"

namespace nkr {

namespace {

// T[j*256 + d] = d * 2^(8j) * P, affine (infinity flagged for d = 0).
template <class F>
__global__ void k_fixed_table(Jacobian<F> base, Affine<F> *table)
{
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 32 * 256) return;
    const uint32_t j = idx >> 8, d = idx & 255u;
    Xyzz<F> p = jacobian_to_xyzz(base);
    for (uint32_t t = 0; t < 8 * j; ++t) p = xyzz_dbl(p);
    Xyzz<F> acc = Xyzz<F>::infinity();
    for (int b = 7; b >= 0; --b) {
        acc = xyzz_dbl(acc);
        if ((d >> b) & 1u) acc = xyzz_add(acc, p);
    }
    table[idx] = xyzz_to_affine(acc);
}

template <class F>
__global__ __launch_bounds__(128) void k_fixed_mul(const Affine<F> *__restrict__ table,
                                                   const fr_t *__restrict__ scalars, uint64_t n,
                                                   Jacobian<F> *__restrict__ out)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const fr_t k = fp_from_mont(scalars[i]);
    Xyzz<F> acc = Xyzz<F>::infinity();
    for (uint32_t j = 0; j < 32; ++j) {
        const uint32_t d = (k.l[j >> 2] >> ((j & 3u) * 8)) & 255u;
        if (d == 0) continue;
        const Affine<F> t = table[j * 256 + d];
        if (affine_is_inf(t)) continue;
        acc = xyzz_add_affine(acc, t.x, t.y);
    }
    out[i] = xyzz_to_jacobian_affine(acc);
}

}  // namespace

template <class F>
FixedBaseMul<F>::FixedBaseMul(const Jacobian<F> &base, hipStream_t stream)
{
    table_.allocate(32 * 256);
    k_fixed_table<F><<<32, 256, 0, stream>>>(base, table_.get());
    NKR_LAUNCH_CHECK();
}

template <class F>
void FixedBaseMul<F>::run(const fr_t *scalars_dev, uint64_t n, Jacobian<F> *out_dev, hipStream_t stream) const
{
    if (n == 0) return;
    k_fixed_mul<F><<<(uint32_t)((n + 127) / 128), 128, 0, stream>>>(table_.get(), scalars_dev, n, out_dev);
    NKR_LAUNCH_CHECK();
}

template <class F>
std::vector<Affine<F>> FixedBaseMul<F>::table_host(hipStream_t stream) const
{
    std::vector<Affine<F>> t(32 * 256);
    NKR_HIP(hipMemcpyAsync(t.data(), table_.get(), t.size() * sizeof(Affine<F>), hipMemcpyDeviceToHost, stream));
    NKR_HIP(hipStreamSynchronize(stream));
    return t;
}

// Host: k * base from a comb table (32 mixed additions, no doublings).
template <class F>
Xyzz<F> fixed_base_mul_host(const std::vector<Affine<F>> &table, const fr_t &k_mont)
{
    const fr_t k = fp_from_mont(k_mont);
    Xyzz<F> acc = Xyzz<F>::infinity();
    for (uint32_t j = 0; j < 32; ++j) {
        const uint32_t d = (k.l[j >> 2] >> ((j & 3u) * 8)) & 255u;
        if (d == 0) continue;
        const Affine<F> &t = table[j * 256 + d];
        if (affine_is_inf(t)) continue;
        acc = xyzz_add_affine(acc, t.x, t.y);
    }
    return acc;
}

template Xyzz<fq_t> fixed_base_mul_host(const std::vector<Affine<fq_t>> &, const fr_t &);
template Xyzz<fq2_t> fixed_base_mul_host(const std::vector<Affine<fq2_t>> &, const fr_t &);

template class FixedBaseMul<fq_t>;
template class FixedBaseMul<fq2_t>;

}  // namespace nkr

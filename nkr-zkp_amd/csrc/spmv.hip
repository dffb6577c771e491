// R1CS sparse matrix-vector products A*z, B*z, C*z over Fr.
//
// Reference: protocols/groth16/backends/cpu/src/backend.cpp:25-35 (row sums in
// row order from zero) and the HIP twin backends/hip/include/spmvm/spmvm.hip:
// 9-21.  Field addition is exact and commutative, so any summation order gives
// the reference's element.  One launch covers all three matrices (grid.y).
#include "spmv.h"

#include <algorithm>
#include <vector>

namespace nkr {

namespace {

__device__ __forceinline__ fr_t shfl_xor_fr(const fr_t &x, int mask)
{
    fr_t r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.l[i] = (uint32_t)__shfl_xor((int)x.l[i], mask, 64);
    return r;
}

// Workgroups [0, short_blocks): one thread per row (rows on the long list
// are left to the others); the rest: one wave per listed long row, lanes
// striding over its non-zeros, then a shuffle reduction.
__device__ __forceinline__ void spmv_rows(const CsrView &m, const fr_t *__restrict__ z, fr_t *__restrict__ y,
                                          uint32_t short_blocks)
{
    if (blockIdx.x < short_blocks) {
        const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
        if (row >= m.rows) return;
        const uint32_t beg = m.row_ptr[row], end = m.row_ptr[row + 1];
        if (m.long_rows && end - beg > kSpmvShortRow) return;
        fr_t acc = fr_t::zero();
        for (uint32_t j = beg; j < end; ++j) acc = fp_add(acc, fp_mul(m.values[j], z[m.col_idx[j]]));
        y[row] = acc;
        return;
    }
    const uint32_t w = (blockIdx.x - short_blocks) * (blockDim.x / 64) + threadIdx.x / 64;
    if (w >= m.n_long) return;  // uniform across the wave
    const uint32_t row = m.long_rows[w], lane = threadIdx.x & 63u;
    const uint32_t beg = m.row_ptr[row], end = m.row_ptr[row + 1];
    fr_t acc = fr_t::zero();
    for (uint32_t j = beg + lane; j < end; j += 64) acc = fp_add(acc, fp_mul(m.values[j], z[m.col_idx[j]]));
    for (int off = 32; off >= 1; off >>= 1) acc = fp_add(acc, shfl_xor_fr(acc, off));
    if (lane == 0) y[row] = acc;
}

__global__ __launch_bounds__(256) void k_spmv3(CsrView m0, CsrView m1, CsrView m2, const fr_t *__restrict__ z,
                                               fr_t *__restrict__ y0, fr_t *__restrict__ y1, fr_t *__restrict__ y2,
                                               uint32_t short_blocks)
{
    const CsrView m = blockIdx.y == 0 ? m0 : (blockIdx.y == 1 ? m1 : m2);
    fr_t *y = blockIdx.y == 0 ? y0 : (blockIdx.y == 1 ? y1 : y2);
    if (m.values == nullptr) return;
    spmv_rows(m, z, y, short_blocks);
}

// Job blockIdx.z, matrix blockIdx.y.
__global__ __launch_bounds__(256) void k_spmv3_batch(const SpmvJob *__restrict__ jobs, uint32_t short_blocks)
{
    const SpmvJob &jb = jobs[blockIdx.z];
    spmv_rows(jb.m[blockIdx.y], jb.z, jb.y[blockIdx.y], short_blocks);
}

uint32_t long_blocks(uint32_t n_long) { return div_up(n_long, 256u / 64u); }

}  // namespace

uint32_t spmv_long_rows(const uint32_t *row_ptr_host, uint32_t rows, DeviceBuffer<uint32_t> &out)
{
    std::vector<uint32_t> list;
    for (uint32_t r = 0; r < rows; ++r)
        if (row_ptr_host[r + 1] - row_ptr_host[r] > kSpmvShortRow) list.push_back(r);
    out.allocate(std::max<size_t>(list.size(), 1));
    if (!list.empty())
        NKR_HIP(hipMemcpy(out.get(), list.data(), list.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    return (uint32_t)list.size();
}

void spmv3_batch(const SpmvJob *jobs_dev, uint32_t count, uint32_t max_rows, uint32_t max_long, hipStream_t stream)
{
    if (count == 0 || max_rows == 0) return;
    NKR_REQUIRE(count <= 65535, "SpMV batch must hold at most 65535 jobs");
    const uint32_t sb = div_up(max_rows, 256u);
    k_spmv3_batch<<<dim3(sb + long_blocks(max_long), 3, count), 256, 0, stream>>>(jobs_dev, sb);
    NKR_LAUNCH_CHECK();
}

void spmv3(const CsrView &a, const CsrView &b, const CsrView &c, const fr_t *z, fr_t *ya, fr_t *yb, fr_t *yc,
           hipStream_t stream)
{
    const uint32_t rows = std::max(a.rows, std::max(b.rows, c.rows));
    if (rows == 0) return;
    const uint32_t sb = div_up(rows, 256u);
    const uint32_t nl = std::max(a.n_long, std::max(b.n_long, c.n_long));
    k_spmv3<<<dim3(sb + long_blocks(nl), 3), 256, 0, stream>>>(a, b, c, z, ya, yb, yc, sb);
    NKR_LAUNCH_CHECK();
}

void spmv(const CsrView &a, const fr_t *z, fr_t *y, hipStream_t stream)
{
    if (a.rows == 0) return;
    CsrView none{};
    const uint32_t sb = div_up(a.rows, 256u);
    k_spmv3<<<dim3(sb + long_blocks(a.n_long), 1), 256, 0, stream>>>(a, none, none, z, y, nullptr, nullptr, sb);
    NKR_LAUNCH_CHECK();
}

}  // namespace nkr

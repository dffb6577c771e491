// R1CS sparse matrix-vector products A*z, B*z, C*z over Fr.
//
// Reference: protocols/groth16/backends/cpu/src/backend.cpp:25-35 (row sums in
// row order from zero) and the HIP twin backends/hip/include/spmvm/spmvm.hip:
// 9-21.  Field addition is exact and commutative, so any summation order gives
// the reference's element.  One launch covers all three matrices (grid.y).
#include "spmv.h"

namespace nkr {

namespace {

__global__ __launch_bounds__(256) void k_spmv3(CsrView m0, CsrView m1, CsrView m2, const fr_t *__restrict__ z,
                                               fr_t *__restrict__ y0, fr_t *__restrict__ y1, fr_t *__restrict__ y2)
{
    const CsrView m = blockIdx.y == 0 ? m0 : (blockIdx.y == 1 ? m1 : m2);
    fr_t *y = blockIdx.y == 0 ? y0 : (blockIdx.y == 1 ? y1 : y2);
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (m.values == nullptr || row >= m.rows) return;
    const uint32_t beg = m.row_ptr[row], end = m.row_ptr[row + 1];
    fr_t acc = fr_t::zero();
    for (uint32_t j = beg; j < end; ++j) acc = fp_add(acc, fp_mul(m.values[j], z[m.col_idx[j]]));
    y[row] = acc;
}

// Job blockIdx.z, matrix blockIdx.y.
__global__ __launch_bounds__(256) void k_spmv3_batch(const SpmvJob *__restrict__ jobs)
{
    const SpmvJob &jb = jobs[blockIdx.z];
    const CsrView &m = jb.m[blockIdx.y];
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= m.rows) return;
    const uint32_t beg = m.row_ptr[row], end = m.row_ptr[row + 1];
    const fr_t *__restrict__ z = jb.z;
    fr_t acc = fr_t::zero();
    for (uint32_t j = beg; j < end; ++j) acc = fp_add(acc, fp_mul(m.values[j], z[m.col_idx[j]]));
    jb.y[blockIdx.y][row] = acc;
}

}  // namespace

void spmv3_batch(const SpmvJob *jobs_dev, uint32_t count, uint32_t max_rows, hipStream_t stream)
{
    if (count == 0 || max_rows == 0) return;
    NKR_REQUIRE(count <= 65535, "SpMV batch must hold at most 65535 jobs");
    k_spmv3_batch<<<dim3(div_up(max_rows, 256), 3, count), 256, 0, stream>>>(jobs_dev);
    NKR_LAUNCH_CHECK();
}

void spmv3(const CsrView &a, const CsrView &b, const CsrView &c, const fr_t *z, fr_t *ya, fr_t *yb, fr_t *yc,
           hipStream_t stream)
{
    const uint32_t rows = std::max(a.rows, std::max(b.rows, c.rows));
    if (rows == 0) return;
    dim3 grid(div_up(rows, 256), 3);
    k_spmv3<<<grid, 256, 0, stream>>>(a, b, c, z, ya, yb, yc);
    NKR_LAUNCH_CHECK();
}

void spmv(const CsrView &a, const fr_t *z, fr_t *y, hipStream_t stream)
{
    if (a.rows == 0) return;
    CsrView none{};
    dim3 grid(div_up(a.rows, 256), 1);
    k_spmv3<<<grid, 256, 0, stream>>>(a, none, none, z, y, nullptr, nullptr);
    NKR_LAUNCH_CHECK();
}

}  // namespace nkr

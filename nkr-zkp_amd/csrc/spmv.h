#pragma once

#include "field.h"
#include "runtime.h"

namespace nkr {

// Device view of one CSR matrix (u32 row_ptr[rows + 1], u32 col_idx[nnz],
// Fr values[nnz]): the reference's CsrMatrix layout
// (protocols/groth16/common/include/r1cs/spmat.hpp:37-45).
struct CsrView {
    const uint32_t *row_ptr;
    const uint32_t *col_idx;
    const fr_t *values;
    uint32_t rows;
    uint32_t nnz;
    // Rows with more than kSpmvShortRow non-zeros, each summed by one wave
    // (64 lanes and a shuffle reduction) instead of one thread; without the
    // list every row is summed by one thread.
    const uint32_t *long_rows = nullptr;
    uint32_t n_long = 0;
};

constexpr uint32_t kSpmvShortRow = 16;

// Uploads the list of rows of a host CSR offset array with more than
// kSpmvShortRow non-zeros into `out` (synchronously) and returns its length.
uint32_t spmv_long_rows(const uint32_t *row_ptr_host, uint32_t rows, DeviceBuffer<uint32_t> &out);

void spmv3(const CsrView &a, const CsrView &b, const CsrView &c, const fr_t *z, fr_t *ya, fr_t *yb, fr_t *yc,
           hipStream_t stream);
void spmv(const CsrView &a, const fr_t *z, fr_t *y, hipStream_t stream);

// A*z, B*z, C*z for several circuits in one launch (jobs in device memory).
struct SpmvJob {
    CsrView m[3];
    const fr_t *z;
    fr_t *y[3];
};
void spmv3_batch(const SpmvJob *jobs_dev, uint32_t count, uint32_t max_rows, uint32_t max_long, hipStream_t stream);

}  // namespace nkr

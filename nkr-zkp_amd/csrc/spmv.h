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
};

void spmv3(const CsrView &a, const CsrView &b, const CsrView &c, const fr_t *z, fr_t *ya, fr_t *yb, fr_t *yc,
           hipStream_t stream);
void spmv(const CsrView &a, const fr_t *z, fr_t *y, hipStream_t stream);

// A*z, B*z, C*z for several circuits in one launch (jobs in device memory).
struct SpmvJob {
    CsrView m[3];
    const fr_t *z;
    fr_t *y[3];
};
void spmv3_batch(const SpmvJob *jobs_dev, uint32_t count, uint32_t max_rows, hipStream_t stream);

}  // namespace nkr

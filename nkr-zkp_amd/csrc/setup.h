#pragma once

#include <vector>

#include "curve.h"
#include "runtime.h"

namespace nkr {

template <class F>
class FixedBaseMul;

// Host: k * base from a table made by FixedBaseMul<F>::table_host.
template <class F>
Xyzz<F> fixed_base_mul_host(const std::vector<Affine<F>> &table, const fr_t &k_mont);

// out[i] = scalars[i] * base for a fixed base, normalised to the reference's
// Jacobian layout (Z = 1, or (0, 1, 0)).
template <class F>
class FixedBaseMul {
public:
    FixedBaseMul(const Jacobian<F> &base, hipStream_t stream);
    void run(const fr_t *scalars_dev, uint64_t n, Jacobian<F> *out_dev, hipStream_t stream) const;
    // The comb table T[j*256 + d] = d * 2^(8j) * base, affine, copied to the host.
    std::vector<Affine<F>> table_host(hipStream_t stream) const;

private:
    DeviceBuffer<Affine<F>> table_;
};

}  // namespace nkr

#pragma once

#include "curve.h"
#include "runtime.h"

namespace nkr {

// out[i] = scalars[i] * base for a fixed base, normalised to the reference's
// Jacobian layout (Z = 1, or (0, 1, 0)).
template <class F>
class FixedBaseMul {
public:
    FixedBaseMul(const Jacobian<F> &base, hipStream_t stream);
    void run(const fr_t *scalars_dev, uint64_t n, Jacobian<F> *out_dev, hipStream_t stream) const;

private:
    DeviceBuffer<Affine<F>> table_;
};

}  // namespace nkr

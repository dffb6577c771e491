// Pippenger multi-scalar multiplication over G1 and G2 for gfx950.
//
// Design (MI355X-first):
//  * Proving-key bases are fixed per circuit, so `prepare` stores, for every
//    signed-digit window w, the shifted copy 2^(c*w) * P_i (affine, 64 B for
//    G1 / 128 B for G2).  All windows then share ONE set of B = 2^(c-1)
//    buckets: no per-window bucket reduction and no doubling chain at the end.
//    The copies cost W*n bases of HBM (about 1 GiB for a 2^20-point G1 key),
//    which a 288 GB device affords.
//  * Scalars are recoded into signed c-bit digits; a counting sort groups the
//    (window, point, sign) entries by bucket using block-local LDS histograms,
//    one global atomic per (block, bucket) to reserve ranges, and an LDS
//    cursor scatter -- no inter-workgroup waiting anywhere.
//  * Buckets are summed by fixed-size pieces (at most T entries per thread,
//    mixed affine additions into XYZZ), and the partial sums of each bucket
//    are folded by further levels of the same segmented reduction, so skewed
//    digit distributions (witness scalars that are mostly 0 or 1) stay
//    balanced.
//  * sum_m m * S_m is formed as c bit-planes U_j = sum_{m has bit j} S_m on the
//    device; the host combines sum_j 2^j U_j (c doublings) and normalises.
// The reference computes the same sums with libff multi_exp<BDLO12>
// (third_party/.../libff/algebra/scalar_multiplication/multiexp.tcc:165-282);
// the affine result is unique, so it matches the reference bit for bit.
#pragma once

#include <vector>

#include "curve.h"
#include "runtime.h"

namespace nkr {

struct MsmPlan {
    uint32_t n = 0;      // number of bases / scalars
    uint32_t c = 0;      // window width in bits
    uint32_t W = 0;      // windows (signed digits), c*W >= 255
    uint32_t B = 0;      // buckets 1..B, B = 2^(c-1)
    static MsmPlan choose(uint32_t n, uint32_t c_override = 0);
};

// Window-shifted affine bases of one proving-key vector.
template <class F>
class MsmBases {
public:
    // `jacobian` is a device array of n reference-layout Jacobian points
    // (any Z, infinity as Z = 0); `perm` optionally reorders them
    // (base i = jacobian[perm[i]]).
    void prepare(const Jacobian<F> *jacobian, const uint32_t *perm, const MsmPlan &plan, hipStream_t stream);
    const Affine<F> *data() const { return pts_.get(); }
    const MsmPlan &plan() const { return plan_; }
    size_t bytes() const { return pts_.bytes(); }

private:
    MsmPlan plan_;
    DeviceBuffer<Affine<F>> pts_;
};

// Signed-digit recoding and bucket sort of one scalar vector; shared by every
// MSM that uses the same scalars (the witness drives four of the five).
class MsmSchedule {
public:
    void build(const fr_t *scalars_mont, const MsmPlan &plan, hipStream_t stream, StageTimer *timer = nullptr);
    const MsmPlan &plan() const { return plan_; }
    uint32_t entries() const { return host_total_; }

    MsmPlan plan_;
    uint32_t blocks_ = 0;
    DeviceBuffer<uint32_t> sorted_;        // entries: sign << 31 | (w * n + i)
    DeviceBuffer<uint32_t> count_;         // per bucket (B)
    DeviceBuffer<uint32_t> bucket_start_;  // B + 1
    DeviceBuffer<uint32_t> block_base_;    // blocks * B
    DeviceBuffer<uint32_t> max_count_;     // 1
    uint32_t host_max_count_ = 0;
    uint32_t host_total_ = 0;
};

// Stage indices used with StageTimer by the MSM code.
enum MsmStage { kStageSort = 3, kStageAccumulate = 4, kStageFold = 5 };

// Evaluates one MSM given prepared bases and a schedule of the same n and c.
// Writes the c bit-plane sums U_j (XYZZ) to `planes` (device, c points).
template <class F>
class MsmRunner {
public:
    void run(const MsmBases<F> &bases, const MsmSchedule &sched, Xyzz<F> *planes, hipStream_t stream,
             StageTimer *timer = nullptr);

private:
    DeviceBuffer<Xyzz<F>> part_[2];
    DeviceBuffer<uint32_t> start_[2];
    DeviceBuffer<Xyzz<F>> sums_;
    DeviceBuffer<uint32_t> plane_start_;
    std::vector<uint32_t> plane_levels_;
};

// Host: sum_j 2^j U_j, then the reference's affine form (Z = 1, or (0, 1, 0)).
template <class F>
Jacobian<F> msm_finish(const Xyzz<F> *planes_host, uint32_t c);

// Whole-vector convenience used by the standalone operator API.
template <class F>
class MsmContext {
public:
    MsmContext(const Jacobian<F> *bases_dev, uint32_t n, uint32_t c_override, hipStream_t stream);
    Jacobian<F> run(const fr_t *scalars_dev, hipStream_t stream);
    const MsmPlan &plan() const { return bases_.plan(); }

private:
    MsmBases<F> bases_;
    MsmSchedule sched_;
    MsmRunner<F> runner_;
    DeviceBuffer<Xyzz<F>> planes_;
};

}  // namespace nkr

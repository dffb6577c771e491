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
//    cursor scatter -- no inter-workgroup waiting and no host round trip.
//  * Several MSMs run as one batch ("jobs"): every stage below is a single
//    launch over all jobs, so the latency of each reduction level is paid
//    once per batch, not once per MSM.
//      1. accumulate: one thread sums a piece of 4..64 entries of one bucket
//         (mixed affine additions into XYZZ); the launch holds only the
//         resident waves, which take pieces by grid stride;
//      2. fold: buckets left with more than 256 partial sums (skewed digit
//         distributions, e.g. witnesses that are mostly 0 or 1) are folded
//         16:1 until they are not; other buckets are untouched;
//      3. complete: 16 lanes per bucket, strided partial sums then a 4-step
//         shuffle tree, give the bucket sums S_m;
//      4. bit planes: a log-depth tree merges bucket blocks into
//         U_j = sum over m with bit j of S_m (c points per job).
//    The host combines sum_j 2^j U_j (c doublings) and normalises.
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
    uint32_t piece = 0;  // entries summed serially by one accumulation thread
    static MsmPlan choose(uint32_t n, uint32_t c_override = 0);
    uint64_t max_entries() const { return (uint64_t)n * W; }
};

// Largest accumulation piece, fold fan-in and the completion limit.  Small
// MSMs use shorter pieces (down to kMsmPieceMin) so that the serial chain of
// one thread, not the amount of work, stops dominating their latency.
constexpr uint32_t kMsmPiece = 64;
constexpr uint32_t kMsmPieceMin = 4;
constexpr uint32_t kMsmFoldIn = 16;
constexpr uint32_t kMsmCompleteMax = 256;
constexpr uint32_t kMaxFoldLevels = 8;

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
// Entirely asynchronous on `stream`.
class MsmSchedule {
public:
    void build(const fr_t *scalars_mont, const MsmPlan &plan, hipStream_t stream, StageTimer *timer = nullptr);
    // Sizes the buffers for `plan`; returns the scalars per sort workgroup.
    uint32_t prepare(const MsmPlan &plan);
    const MsmPlan &plan() const { return plan_; }
    uint64_t max_pieces() const { return plan_.max_entries() / plan_.piece + plan_.B + 1; }
    // Device counters: [0] entries (non-zero digits), [1] level-0 pieces.
    const uint32_t *totals() const { return bucket_start_.get() + plan_.B; }

    MsmPlan plan_;
    uint32_t blocks_ = 0;
    DeviceBuffer<uint32_t> sorted_;        // entries: sign << 31 | (w * n + i)
    DeviceBuffer<uint32_t> count_;         // per bucket (B)
    DeviceBuffer<uint32_t> bucket_start_;  // B + 1, then the piece total at [B + 1]
    DeviceBuffer<uint32_t> piece_start_;   // B + 1
    DeviceBuffer<uint32_t> block_base_;    // blocks * B
};

// Many schedules with one launch per sort phase (zero, histogram, scan,
// scatter) instead of four launches per schedule.  Same result as
// MsmSchedule::build for each of them.
class MsmScheduleBatch {
public:
    struct Item {
        MsmSchedule *sched;
        const fr_t *scalars_mont;
        MsmPlan plan;
    };
    void build(const std::vector<Item> &items, hipStream_t stream);

private:
    PinnedStaging<uint8_t> stage_;
    DeviceBuffer<uint8_t> table_;
};

// Stage indices used with StageTimer by the MSM code.
enum MsmStage { kStageSort = 3, kStageAccumulate = 4, kStageFold = 5 };

template <class F>
struct MsmJob {
    const MsmBases<F> *bases;
    const MsmSchedule *sched;
    Xyzz<F> *planes;  // device output: c points U_0 .. U_{c-1}
};

// Runs a batch of MSMs of one group (all with the same window width).
template <class F>
class MsmBatch {
public:
    void run(const std::vector<MsmJob<F>> &jobs, hipStream_t stream, StageTimer *timer = nullptr);

private:
    void layout(const std::vector<MsmJob<F>> &jobs, hipStream_t stream);
    std::vector<const void *> key_;
    uint32_t fold_levels_ = 0;
    std::vector<uint64_t> fold_cap_;  // largest partial count of any job at each fold level
    uint64_t thread_total_ = 0;
    uint64_t complete_bound_ = 0;                      // completion threads at 16 lanes per bucket
    DeviceBuffer<uint32_t> complete_base_, complete_lanes_;
    std::vector<uint64_t> thread_base_;
    DeviceBuffer<uint8_t> jobs_dev_;
    DeviceBuffer<uint32_t> thread_base_dev_;
    PinnedStaging<uint8_t> jobs_stage_;
    DeviceBuffer<Xyzz<F>> pool_, sums_, tree_[2];
    DeviceBuffer<uint32_t> reg_start_, reg_count_, fold_start_;
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
    MsmBatch<F> batch_;
    DeviceBuffer<Xyzz<F>> planes_;
};

}  // namespace nkr

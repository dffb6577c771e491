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
//    (window, point, sign) entries by bucket: about one workgroup per CU
//    histograms a contiguous slice of scalars in LDS and writes its counts as
//    one row of a (workgroups x B) matrix, a column scan turns the matrix into
//    per-(workgroup, bucket) offsets, and an LDS-cursor scatter places the
//    entries -- no global atomics, no inter-workgroup waiting and no host
//    round trip.
//  * Several MSMs run as one batch ("jobs"): every stage below is a single
//    launch over all jobs, so the latency of each reduction level is paid
//    once per batch, not once per MSM.
//      1. accumulate: one thread sums a piece of 4..64 entries of one bucket
//         (mixed affine additions into XYZZ); the launch holds only the
//         resident waves, which take pieces by grid stride;
//      2. fold: buckets left with more than 64 partial sums (skewed digit
//         distributions, e.g. witnesses that are mostly 0 or 1) are folded
//         4:1 until they are not; other buckets are untouched;
//      3. complete: 1..16 lanes per bucket (about 8 partials per lane),
//         strided partial sums then a shuffle tree, give the bucket sums S_m;
//      4. bit planes: a log-depth tree merges bucket blocks into
//         U_j = sum over m with bit j of S_m (c points per job), several
//         levels per launch inside LDS.
//    The host combines sum_j 2^j U_j (c doublings) and normalises.
// The reference computes the same sums with libff multi_exp<BDLO12>
// (third_party/.../libff/algebra/scalar_multiplication/multiexp.tcc:165-282);
// the affine result is unique, so it matches the reference bit for bit.
#pragma once

#include <algorithm>
#include <vector>

#include "curve.h"
#include "curve29.h"
#include "msm_plan.h"
#include "runtime.h"

namespace nkr {

// Process-wide MSM tuning knobs: the target piece count (2^target_log, 10 ..
// 22) and the most lanes that complete one bucket (1 .. 16, a power of two).
// Set them before proving, not while proofs run.
void msm_set_shape(uint32_t target_pieces_log, uint32_t complete_lanes_cap);
uint32_t msm_target_pieces_log();
uint32_t msm_complete_lanes_cap();

// Device bytes of one schedule (MsmSchedule) and of one MSM job's partial
// sums and per-bucket arrays (MsmBatch) for a plan.
uint64_t msm_schedule_bytes(const MsmPlan &plan);
template <class F>
uint64_t msm_job_bytes(const MsmPlan &plan);

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
    uint64_t max_pieces() const { return plan_.max_pieces(target_log_); }
    // Device counters: [0] entries (non-zero digits), [1] level-0 pieces,
    // [2] largest piece count of a bucket, [3] piece length.
    const uint32_t *totals() const { return bucket_start_.get() + plan_.B; }
    // The same four counters on the host; blocks until the sort that writes
    // them has run.
    const uint32_t *totals_host() const
    {
        host_stage_->wait();
        return host_totals_;
    }
    // Sort workgroups for n scalars (each histograms a contiguous slice).
    static uint32_t sort_blocks(uint32_t n);

    // A second split of the same sorted entries into pieces of at least
    // `piece_lo` entries (chosen on the device as in build, within
    // [piece_lo, max(piece_lo, plan piece_hi)]), computed by the next builds:
    // one scalar vector feeds MSMs in different groups, and a G1 MSM
    // completes its buckets faster from fewer, longer pieces while the G2
    // MSM, at one wave per SIMD, needs the short ones.  0 turns it off.
    // Jobs select it with MsmJob::alt.
    void set_alt_pieces(uint32_t piece_lo) { alt_piece_lo_ = piece_lo; }
    // 2^target_log level-0 pieces per build instead of the process-wide
    // setting (0 restores it).
    void set_target_log(uint32_t target_log) { target_override_ = target_log; }
    uint64_t alt_max_pieces() const
    {
        NKR_REQUIRE(alt_piece_lo_ > 0, "an MSM job reads the second piece split of a schedule that has none");
        MsmPlan p = plan_;
        p.piece_lo = alt_piece_lo_;
        p.piece_hi = std::max(alt_piece_lo_, p.piece_hi);
        return p.max_pieces(target_log_);
    }
    const uint32_t *alt_totals_host() const
    {
        NKR_REQUIRE(alt_piece_lo_ > 0, "an MSM job reads the second piece split of a schedule that has none");
        return totals_host() + 4;
    }

    // Device bytes held, and their release (the next prepare reallocates).
    uint64_t device_bytes() const
    {
        return sorted_.bytes() + count_.bytes() + bucket_start_.bytes() + piece_start_.bytes() +
               block_count_.bytes() + alt_piece_start_.bytes();
    }
    void release_device()
    {
        sorted_.release();
        count_.release();
        bucket_start_.release();
        piece_start_.release();
        block_count_.release();
        alt_piece_start_.release();
    }

    MsmPlan plan_;
    uint32_t blocks_ = 0;
    uint32_t target_log_ = kMsmTargetPiecesLog;  // as of the last prepare
    DeviceBuffer<uint32_t> sorted_;        // entries: sign << 31 | (w * n + i)
    DeviceBuffer<uint32_t> count_;         // per bucket (B)
    DeviceBuffer<uint32_t> bucket_start_;  // B + 1, then the totals above, then the second split's
    DeviceBuffer<uint32_t> piece_start_;   // B + 1
    DeviceBuffer<uint32_t> block_count_;   // blocks x B: per-workgroup counts, then offsets
    PinnedStaging<uint32_t> own_totals_;   // host copy of the counters (build)
    uint32_t alt_piece_lo_ = 0;
    uint32_t target_override_ = 0;
    DeviceBuffer<uint32_t> alt_piece_start_;  // B + 1 (its counters follow the first split's)
    const uint32_t *host_totals_ = nullptr;
    const PinnedStaging<uint32_t> *host_stage_ = nullptr;
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
    // graph: the launches are being captured into a graph that is replayed
    // later, so nothing here may wait for or read device results on the host
    // (the schedules' counters are not copied back and no staging event is
    // recorded; the caller synchronises its streams before the staging
    // buffers are written again).
    void build(const std::vector<Item> &items, hipStream_t stream, bool graph = false);
    uint64_t device_bytes() const { return table_.bytes() + totals_dev_.bytes(); }
    void release_device()
    {
        table_.release();
        totals_dev_.release();
    }

private:
    PinnedStaging<uint8_t> stage_;
    DeviceBuffer<uint8_t> table_;
    DeviceBuffer<uint32_t> totals_dev_;
    PinnedStaging<uint32_t> totals_host_;
};

// The form of the partial sums between the accumulation and the bit planes:
// the reduced-radix form (curve29.h) through fold, completion and tree, so
// no addition converts.  The bit planes handed to the host are canonical.
template <class F>
struct MsmPartial;
template <>
struct MsmPartial<fq_t> {
    using type = Xyzz29<FqParams>;
};
template <>
struct MsmPartial<fq2_t> {
    using type = Xyzz2_29<FqParams>;
};
template <class F>
using partial_t = typename MsmPartial<F>::type;

// Stage indices used with StageTimer by the MSM code.
enum MsmStage { kStageSort = 3, kStageAccumulate = 4, kStageFold = 5, kStageAccumulateG2 = 7 };

template <class F>
struct MsmJob {
    const MsmBases<F> *bases;
    const MsmSchedule *sched;
    Xyzz<F> *planes;   // device output: c points U_0 .. U_{c-1}
    bool alt = false;  // use the schedule's second piece split (set_alt_pieces)
};

// Runs a batch of MSMs of one group (all with the same window width).
template <class F>
class MsmBatch {
public:
    // graph: as MsmScheduleBatch::build; the fold levels are then launched
    // for the largest bucket the partial-sum capacity allows (levels a batch
    // does not need find nothing to fold) instead of for the counters read
    // back from the schedules.
    void run(const std::vector<MsmJob<F>> &jobs, hipStream_t stream, StageTimer *timer = nullptr,
             bool graph = false);
    // Device bytes held by the batch's work buffers, and their release (the
    // next run reallocates what it needs).
    uint64_t device_bytes() const
    {
        uint64_t b = complete_base_.bytes() + class_count_.bytes() + class_list_.bytes() + jobs_dev_.bytes() +
                     thread_base_dev_.bytes() + pool_.bytes() + sums_.bytes() + reg_start_.bytes() +
                     reg_count_.bytes() + fold_start_.bytes();
        return b + tree_[0].bytes() + tree_[1].bytes();
    }
    void release_device()
    {
        complete_base_.release();
        class_count_.release();
        class_list_.release();
        jobs_dev_.release();
        thread_base_dev_.release();
        pool_.release();
        sums_.release();
        tree_[0].release();
        tree_[1].release();
        reg_start_.release();
        reg_count_.release();
        fold_start_.release();
        key_.clear();  // the layout cache refers to the released buffers
    }

    // The jobs' layout is the one the batch's buffers already hold.
    bool layout_matches(const std::vector<MsmJob<F>> &jobs) const;

private:
    void layout(const std::vector<MsmJob<F>> &jobs, hipStream_t stream);
    std::vector<const void *> key_;
    uint32_t fold_levels_ = 0;
    std::vector<uint64_t> fold_cap_;  // largest partial count of any job at each fold level
    uint64_t thread_total_ = 0;
    uint64_t complete_bound_ = 0;  // completion lanes at most
    DeviceBuffer<uint32_t> complete_base_, class_count_, class_list_;
    std::vector<uint64_t> thread_base_;
    DeviceBuffer<uint8_t> jobs_dev_;
    DeviceBuffer<uint32_t> thread_base_dev_;
    PinnedStaging<uint8_t> jobs_stage_;
    DeviceBuffer<partial_t<F>> pool_, sums_, tree_[2];
    DeviceBuffer<uint32_t> reg_start_, reg_count_, fold_start_;
};

// Host: sum_j 2^j U_j, then the reference's affine form (Z = 1, or (0, 1, 0)).
template <class F>
Jacobian<F> msm_finish(const Xyzz<F> *planes_host, uint32_t c);
// The same sum left in XYZZ form (not normalised).
template <class F>
Xyzz<F> msm_sum(const Xyzz<F> *planes_host, uint32_t c);

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

// Batched proving: many Groth16 proofs (different circuits and sizes, or
// the same circuit with different witnesses) as one launch sequence.
//
// The plain path pays every kernel's latency once per proof; small proofs
// (2^10 .. 2^18 constraints) are dominated by those latencies, not by work.
// A batch runs each stage once over a work list covering every proof:
//   SpMV        one launch over (job, matrix, row)
//   NTT chain   one launch per pass per size class (all A, B, C of the class)
//   quotient    one launch per size class
//   sort        one launch per sort phase for all witness schedules, and
//               one for all H schedules
//   MSM         one MsmBatch for all G1 MSMs (4 per proof) and one for the
//               G2 MSMs: each accumulation / fold / completion / tree level is
//               one launch for the whole batch
// Two streams, as in the plain prover: the witness schedules and the G2 MSMs
// need only the witnesses, so they run beside the SpMV, transforms, H
// schedules and G1 MSMs; each side's latency-bound reduction levels then
// overlap the other side's accumulation.  (Moving the G1 MSMs over the
// witnesses to the side stream as well, with the accumulations holding part
// of the resident slots, measured no faster on MI355X: the batch is bound by
// the accumulations' total work.)
//   witnesses   packed by the plain path's rule (values other than 0 and 1,
//               bit masks for the rest) on host threads in 1 MiB chunks, each
//               chunk's host-to-device copy issued as soon as it is packed,
//               and expanded into every slot by one kernel.
//   bit planes  three batch-wide regions (G1 over the witness, H, G2), one
//               device-to-host copy each.
//   graphs      a batch of the same contexts in the same order as the
//               previous call is captured into a HIP graph (one per prover)
//               and replayed while the composition repeats: the fold levels
//               are then launched for the capacity's worst case and nothing
//               is read back mid-sequence.  Any other batch runs eagerly and
//               drops the graph.
//   host tail   the r,s-only blinding terms while the GPU works, then the
//               final sums and blinding on a pool of host threads.
// Each proof's result is bit-identical to proving it alone.
#pragma once

#include <memory>
#include <vector>

#include "msm.h"
#include "prover.h"

namespace nkr {

// The MSM plan a batch uses for a context's plan: a fixed piece length of
// 32 entries.  The many jobs of a batch already fill the device, so long
// pieces (fewer partial sums to complete) pay; one proof alone lets the
// device choose shorter ones.  Results do not depend on the piece length.
MsmPlan batch_plan(const MsmPlan &plan);

struct BatchItem {
    Groth16Prover *ctx;
    const void *witness;  // cols Fr, host memory
    fr_t r, s;
    void *raw_out;    // 576 B
    void *proof_out;  // 384 B
};

class BatchProver {
public:
    explicit BatchProver(int device);
    ~BatchProver();
    BatchProver(const BatchProver &) = delete;
    BatchProver &operator=(const BatchProver &) = delete;

    void prove(const std::vector<BatchItem> &items);
    int device() const { return device_; }
    // Device bytes the per-proof slots and the batch's work buffers hold
    // (they grow to the largest batch run), and their release: the next
    // prove reallocates what it needs.
    uint64_t device_bytes() const;
    void trim();
    // Whether a batch repeating the previous call's contexts (same order) is
    // captured into a HIP graph and replayed (default on).
    void set_graphs(bool enabled);
    bool graph_active() const { return graph_exec_ != nullptr; }

private:
    void release() noexcept;  // streams and events (destructor, failed constructor)
    void drop_graph() noexcept;
    struct Slot {
        DeviceBuffer<fr_t> z, pa, pb, pc;
        MsmSchedule wsched, hsched;
    };
    int device_;
    hipStream_t stream_ = nullptr;   // SpMV, transforms, H schedules, G1 MSMs
    hipStream_t stream2_ = nullptr;  // witness schedules and G2 MSMs, beside them
    hipEvent_t ev_z_ = nullptr, ev_wsched_ = nullptr, ev_side_ = nullptr;
    std::vector<std::unique_ptr<Slot>> slots_;
    // the captured launch sequence, the contexts it was captured for, and
    // the previous call's contexts
    hipGraphExec_t graph_exec_ = nullptr;
    std::vector<const void *> graph_key_, last_key_;
    bool graphs_enabled_ = true;
    MsmScheduleBatch sort_w_, sort_h_;
    MsmBatch<fq_t> batch_g1_;
    MsmBatch<fq2_t> batch_g2_;
    WitnessBatchUpload wupload_;
    PinnedStaging<uint8_t> table_stage_;
    DeviceBuffer<uint8_t> tables_;
    DeviceBuffer<Xyzz<fq_t>> planes_w_, planes_h_;  // bit planes: ar, bs1, zk of every proof; H of every proof
    DeviceBuffer<Xyzz<fq2_t>> planes_2_;            // bs2 of every proof
    PinnedStaging<uint8_t> planes_host_;
};

}  // namespace nkr

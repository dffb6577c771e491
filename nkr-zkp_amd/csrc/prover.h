// Groth16 prove orchestration on one GPU: the MI355X counterpart of the
// reference's HipBackend (protocols/groth16/backends/hip/src/backend.hip:
// 270-335) and of CpuBackend::prove (backends/cpu/src/backend.cpp:156-201).
//
// prepare (constructor): CSR matrices and proving key to HBM, key points
// normalised and window-expanded once (MsmBases), NTT tables built, and the
// H query moved to the basis of coset evaluations (hquery.h).
// prove: witness H2D -> A*z, B*z, C*z -> iNTT x3 -> coset NTT x3 -> (ab - c)/Z
// on the coset, which the H MSM takes as its scalars (no seventh transform)
// -> one signed-digit sort of the witness shared by the ar / bs1 / zk / bs2
// MSMs -> one for H -> bit planes D2H -> host finish.
#pragma once

#include <memory>
#include <vector>

#include "../../include/nkr_g16.h"
#include "host_pool.h"
#include "msm.h"
#include "ntt.h"
#include "spmv.h"
#include "upload.h"

namespace nkr {

// validate_circuit's preconditions on a circuit view; throws Error.
void check_circuit(const nkr_g16_circuit &c);

class Groth16Prover {
public:
    // tuning: optional per-context choices (nkr_g16_tuning; null = defaults)
    Groth16Prover(const nkr_g16_circuit &circuit, const nkr_g16_pk &pk, int device,
                  const nkr_g16_tuning *tuning = nullptr);
    ~Groth16Prover();
    Groth16Prover(const Groth16Prover &) = delete;
    Groth16Prover &operator=(const Groth16Prover &) = delete;

    // raw_out: 576 B.  az/bz/cz/h: optional host outputs (rows Fr, natural order).
    void prove(const void *witness, void *raw_out, void *az, void *bz, void *cz, void *h);
    // witness already in device memory
    void prove_dev(const fr_t *witness_dev, void *raw_out);
    // prove followed by blind_proof with the given r, s; the parts of the
    // blinding that depend only on r and s run on the host while the GPU
    // works.  witness is host memory, or device memory when on_device.
    void prove_blind(const void *witness, bool on_device, const fr_t &r, const fr_t &s, void *raw_out,
                     void *proof_out);

    static constexpr int kStages = 8;
    void set_profiling(bool on);
    void profile_read(double *stage_ms, uint64_t *counters) const;

    uint64_t device_bytes() const;
    int device() const { return device_; }
    // prepared with alpha/beta/delta, which blinding needs
    bool can_blind() const { return have_blind_; }
    hipStream_t stream() const { return stream_; }
    uint32_t rows() const { return rows_; }
    uint32_t cols() const { return cols_; }
    const MsmPlan &witness_plan() const { return wplan_; }
    const MsmPlan &h_plan() const { return hplan_; }

private:
    void release() noexcept;  // streams, events and pinned memory (destructor, failed constructor)
    int device_;
    hipStream_t stream_ = nullptr;
    uint32_t rows_, cols_, log_n_;
    DeviceBuffer<uint32_t> row_ptr_[3], col_idx_[3], long_rows_[3];
    DeviceBuffer<fr_t> values_[3];
    CsrView views_[3];
    std::unique_ptr<NttDomain> domain_;
    MsmPlan wplan_, hplan_;
    MsmBases<fq_t> a_, b1_, l_, h_;
    MsmBases<fq2_t> b2_;
    MsmSchedule wsched_, hsched_;
    MsmBatch<fq_t> batch_g1_;            // ar, bs1, zk and H as one batch
    MsmBatch<fq2_t> batch_g2_;
    hipStream_t stream2_ = nullptr;      // witness schedule and the G2 MSM, beside the NTT chain
    hipEvent_t ev_start_ = nullptr, ev_wsched_ = nullptr;
    DeviceBuffer<fr_t> z_, pa_, pb_, pc_;
    DeviceBuffer<Xyzz<fq_t>> planes1_;
    DeviceBuffer<Xyzz<fq2_t>> planes2_;
    PinnedStaging<Xyzz<fq_t>> planes1_host_;
    PinnedStaging<Xyzz<fq2_t>> planes2_host_;
    std::unique_ptr<WitnessUpload> upload_;  // host witnesses: packed, chunked upload

    DeviceBuffer<fr_t *> poly_table_;  // {pa, pb, pc} for the batched transforms

    struct BlindJob {
        fr_t r, s;
        void *proof_out;
    };
    struct BlindPre {
        g1_xyzz A0, C0;
        g2_xyzz B0;
    };
    void run(const fr_t *z, void *raw_out, void *az, void *bz, void *cz, void *h, const BlindJob *blind = nullptr);
    BlindPre blind_pre(const fr_t &r, const fr_t &s) const;
    // p1w: the bit planes of ar, bs1 and zk (3 x wplan c), p1h: those of H
    // (hplan c), p2: those of bs2 (wplan c).
    void finish(const Xyzz<fq_t> *p1w, const Xyzz<fq_t> *p1h, const Xyzz<fq2_t> *p2, const BlindJob *blind,
                const BlindPre *pre, void *raw_out) const;
    // The two halves of finish: bs2 (and B) from the G2 planes, the rest from the G1 planes.
    void finish_g2(const Xyzz<fq2_t> *p2, const BlindJob *blind, const BlindPre *pre, void *raw_out) const;
    // helpers: two host threads to split the G1 tail over (null: all on
    // the calling thread, as the batch path's per-proof workers do).
    void finish_g1(const Xyzz<fq_t> *p1w, const Xyzz<fq_t> *p1h, const BlindJob *blind, const BlindPre *pre,
                   void *raw_out, HostHelpers *helpers) const;
    // finish on the calling thread with less arithmetic, for callers that
    // finish many proofs over their own threads: the same outputs (the same
    // points, whose canonical affine coordinates do not depend on how they
    // were reached), with every point kept in XYZZ form until one inversion
    // per group (ar, bs1, zk, qz, A, C; bs2, B) and s*ar + r*bs1 as one
    // multiplication sharing its doublings (g1_mul2_glv).
    void finish_serial(const Xyzz<fq_t> *p1w, const Xyzz<fq_t> *p1h, const Xyzz<fq2_t> *p2, const BlindJob *blind,
                       const BlindPre *pre, void *raw_out) const;
    friend class BatchProver;

    // Fixed-base comb tables of alpha_1, beta_1, delta_1 (G1) and delta_2 (G2)
    // and the key's group elements, for blinding.
    bool have_blind_ = false;
    std::vector<Affine<fq_t>> tab_alpha1_, tab_beta1_, tab_delta1_;
    std::vector<Affine<fq2_t>> tab_delta2_;
    g1_jac alpha1_, beta1_;
    g2_jac beta2_;
    std::unique_ptr<HostHelpers> helpers_;  // the plain path's host tail
    StageTimer timer_;
    double stage_ms_[kStages] = {};
    uint64_t counters_[4] = {};
};

// blind_proof (protocols/groth16/common/src/protocol.cpp:85-94) with given r, s.
void groth16_blind(const nkr_g16_pk &pk, const void *raw, const fr_t &r, const fr_t &s, void *proof_out);

}  // namespace nkr

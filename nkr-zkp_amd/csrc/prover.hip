#include "prover.h"

#include <chrono>
#include <cstring>
#include <exception>

#include "glv.h"
#include "hquery.h"
#include "setup.h"

namespace nkr {

namespace {

std::vector<uint32_t> bitrev_permutation(unsigned log_n)
{
    const size_t n = size_t(1) << log_n;
    std::vector<uint32_t> p(n);
    for (size_t i = 0; i < n; ++i) {
        uint32_t r = 0, x = (uint32_t)i;
        for (unsigned b = 0; b < log_n; ++b) {
            r = (r << 1) | (x & 1u);
            x >>= 1;
        }
        p[i] = r;
    }
    return p;
}

void check_csr(uint64_t rows, uint64_t cols, const uint32_t *rp, const uint32_t *ci, const void *vals, uint64_t nnz,
               const char *name)
{
    const std::string m(name);
    NKR_REQUIRE(rp && (nnz == 0 || (ci && vals)), m + ": null CSR array");
    NKR_REQUIRE(rp[0] == 0 && rp[rows] == nnz, m + " has invalid CSR offsets");
    for (uint64_t r = 0; r < rows; ++r) NKR_REQUIRE(rp[r + 1] >= rp[r], m + " has non-monotonic CSR offsets");
    for (uint64_t j = 0; j < nnz; ++j) NKR_REQUIRE(ci[j] < cols, m + " contains an out-of-range column");
}

template <class T>
void upload(DeviceBuffer<T> &buf, const void *src, size_t count, hipStream_t stream)
{
    buf.allocate(std::max<size_t>(count, 1));
    if (count) NKR_HIP(hipMemcpyAsync(buf.get(), src, count * sizeof(T), hipMemcpyHostToDevice, stream));
}

}  // namespace

void check_circuit(const nkr_g16_circuit &c)
{
    // Same preconditions as validate_circuit (protocols/groth16/common/src/protocol.cpp:9-52).
    NKR_REQUIRE(c.rows >= 2 && is_pow2(c.rows), "constraint count must be a power of two (>= 2)");
    NKR_REQUIRE(c.rows <= (1ull << 28), "constraint count exceeds the 2-adicity of Fr");
    NKR_REQUIRE(c.cols >= 1 && c.cols < (1ull << 31), "invalid variable count");
    NKR_REQUIRE(c.public_inputs + 1 <= c.cols, "public input count exceeds the variable count");
    check_csr(c.rows, c.cols, c.a_row_ptr, c.a_col_idx, c.a_values, c.a_nnz, "A");
    check_csr(c.rows, c.cols, c.b_row_ptr, c.b_col_idx, c.b_values, c.b_nnz, "B");
    check_csr(c.rows, c.cols, c.c_row_ptr, c.c_col_idx, c.c_values, c.c_nnz, "C");
    const uint64_t nnz[3] = {c.a_nnz, c.b_nnz, c.c_nnz};
    for (int m = 0; m < 3; ++m) NKR_REQUIRE(nnz[m] < (1ull << 32), "matrix has too many non-zeros");
}

Groth16Prover::Groth16Prover(const nkr_g16_circuit &c, const nkr_g16_pk &pk, int device,
                             const nkr_g16_tuning *tuning)
    : device_(device)
{
    nkr_g16_tuning tn{};
    if (tuning) tn = *tuning;
    NKR_REQUIRE(tn.witness_window_bits == 0 || (tn.witness_window_bits >= 2 && tn.witness_window_bits <= 15),
                "tuning: witness window must be 2..15 bits");
    NKR_REQUIRE(tn.h_window_bits == 0 || (tn.h_window_bits >= 2 && tn.h_window_bits <= 16),
                "tuning: H window must be 2..16 bits");
    NKR_REQUIRE(tn.ntt_split_policy <= 2, "tuning: NTT split policy must be 0 (default), 1 or 2");
    NKR_REQUIRE(tn.msm_target_pieces_log == 0 || (tn.msm_target_pieces_log >= 12 && tn.msm_target_pieces_log <= 24),
                "tuning: MSM target pieces must be 2^12 .. 2^24");
    NKR_REQUIRE(tn.witness_g1_piece_min <= kMsmPiece, "tuning: witness G1 piece minimum must be 0..64");
    // validate_circuit / validate_proving_key (protocols/groth16/common/src/protocol.cpp:9-67)
    check_circuit(c);
    NKR_REQUIRE(pk.a_query && pk.b_g1_query && pk.b_g2_query && pk.l_query && pk.h_query,
                "proving key query vectors missing");

    rows_ = (uint32_t)c.rows;
    cols_ = (uint32_t)c.cols;
    log_n_ = ceil_log2(rows_);
    NKR_HIP(hipSetDevice(device_));
    // A constructor that throws never runs the destructor: hand the streams,
    // events and pinned memory back if a later step (an upload, a prepare
    // that runs out of device memory) throws.
    struct ReleaseOnThrow {
        Groth16Prover *self;
        bool armed = true;
        ~ReleaseOnThrow()
        {
            if (armed) self->release();
        }
    } release_on_throw{this};
    NKR_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    NKR_HIP(hipStreamCreateWithFlags(&stream2_, hipStreamNonBlocking));
    NKR_HIP(hipEventCreateWithFlags(&ev_start_, hipEventDisableTiming));
    NKR_HIP(hipEventCreateWithFlags(&ev_wsched_, hipEventDisableTiming));

    const uint32_t *rps[3] = {c.a_row_ptr, c.b_row_ptr, c.c_row_ptr};
    const uint32_t *cis[3] = {c.a_col_idx, c.b_col_idx, c.c_col_idx};
    const void *vs[3] = {c.a_values, c.b_values, c.c_values};
    const uint64_t nnz[3] = {c.a_nnz, c.b_nnz, c.c_nnz};
    for (int m = 0; m < 3; ++m) {
        NKR_REQUIRE(nnz[m] < (1ull << 32), "matrix has too many non-zeros");
        upload(row_ptr_[m], rps[m], rows_ + 1, stream_);
        upload(col_idx_[m], cis[m], nnz[m], stream_);
        upload(values_[m], vs[m], nnz[m], stream_);
        views_[m] = CsrView{row_ptr_[m].get(), col_idx_[m].get(), values_[m].get(), rows_, (uint32_t)nnz[m]};
        views_[m].n_long = spmv_long_rows(rps[m], rows_, long_rows_[m]);
        views_[m].long_rows = long_rows_[m].get();
    }

    NttShape shape;
    shape.lds_log = tn.ntt_lds_log;
    shape.radix_log = tn.ntt_radix_log;
    shape.split_set = tn.ntt_split_policy != 0;
    shape.split_policy = tn.ntt_split_policy ? tn.ntt_split_policy - 1 : 0;
    domain_ = std::make_unique<NttDomain>(log_n_, stream_, &shape);
    // The witness schedule is sorted on the side stream beside the transforms:
    // at most 15-bit windows, whose 64 KiB histograms share a CU with them.
    wplan_ = MsmPlan::choose(cols_, tn.witness_window_bits);
    if (wplan_.c > 15) wplan_ = MsmPlan::choose(cols_, 15);
    hplan_ = MsmPlan::choose(rows_, tn.h_window_bits);
    wsched_.set_target_log(tn.msm_target_pieces_log);
    hsched_.set_target_log(tn.msm_target_pieces_log);

    {
        DeviceBuffer<g1_jac> tmp1(std::max(cols_, rows_));
        const void *g1_vecs[3] = {pk.a_query, pk.b_g1_query, pk.l_query};
        MsmBases<fq_t> *g1_sets[3] = {&a_, &b1_, &l_};
        for (int k = 0; k < 3; ++k) {
            NKR_HIP(hipMemcpyAsync(tmp1.get(), g1_vecs[k], (size_t)cols_ * sizeof(g1_jac), hipMemcpyHostToDevice,
                                   stream_));
            g1_sets[k]->prepare(tmp1.get(), nullptr, wplan_, stream_);
            NKR_HIP(hipStreamSynchronize(stream_));
        }
        // H query in the basis of the quotient's coset evaluations (the prove
        // chain then skips its last transform), which the group transform
        // leaves in bit-reversed order: MSM term i uses position bitrev(i).
        const std::vector<uint32_t> perm = bitrev_permutation(log_n_);
        DeviceBuffer<uint32_t> dperm;
        upload(dperm, perm.data(), perm.size(), stream_);
        NKR_HIP(hipMemcpyAsync(tmp1.get(), pk.h_query, (size_t)rows_ * sizeof(g1_jac), hipMemcpyHostToDevice,
                               stream_));
        DeviceBuffer<g1_jac> hq(rows_);
        h_query_coset_basis(tmp1.get(), *domain_, hq.get(), stream_);
        h_.prepare(hq.get(), dperm.get(), hplan_, stream_);
        NKR_HIP(hipStreamSynchronize(stream_));
    }
    {
        DeviceBuffer<g2_jac> tmp2(cols_);
        NKR_HIP(hipMemcpyAsync(tmp2.get(), pk.b_g2_query, (size_t)cols_ * sizeof(g2_jac), hipMemcpyHostToDevice,
                               stream_));
        b2_.prepare(tmp2.get(), nullptr, wplan_, stream_);
        NKR_HIP(hipStreamSynchronize(stream_));
    }

    if (pk.alpha_g1 && pk.beta_g1 && pk.beta_g2 && pk.delta_g1 && pk.delta_g2) {
        std::memcpy(&alpha1_, pk.alpha_g1, sizeof(alpha1_));
        std::memcpy(&beta1_, pk.beta_g1, sizeof(beta1_));
        std::memcpy(&beta2_, pk.beta_g2, sizeof(beta2_));
        g1_jac d1;
        g2_jac d2;
        std::memcpy(&d1, pk.delta_g1, sizeof(d1));
        std::memcpy(&d2, pk.delta_g2, sizeof(d2));
        tab_alpha1_ = FixedBaseMul<fq_t>(alpha1_, stream_).table_host(stream_);
        tab_beta1_ = FixedBaseMul<fq_t>(beta1_, stream_).table_host(stream_);
        tab_delta1_ = FixedBaseMul<fq_t>(d1, stream_).table_host(stream_);
        tab_delta2_ = FixedBaseMul<fq2_t>(d2, stream_).table_host(stream_);
        have_blind_ = true;
    }

    z_.allocate(cols_);
    pa_.allocate(rows_);
    pb_.allocate(rows_);
    pc_.allocate(rows_);
    fr_t *const table[3] = {pa_.get(), pb_.get(), pc_.get()};
    poly_table_.allocate(3);
    NKR_HIP(hipMemcpyAsync(poly_table_.get(), table, sizeof(table), hipMemcpyHostToDevice, stream_));
    planes1_.allocate(3 * wplan_.c + hplan_.c);
    planes2_.allocate(wplan_.c);
    upload_ = std::make_unique<WitnessUpload>(cols_);
    NKR_HIP(hipStreamSynchronize(stream_));
    helpers_ = std::make_unique<HostHelpers>(2);
    // The witness G1 MSMs read the schedule in pieces of at least 32 entries
    // (a few partial sums per bucket to complete); bs2 keeps the short
    // device-chosen pieces its one-wave-per-SIMD accumulation needs.
    // Measured on the 2^20 prove: 4.13 -> 3.98 ms.
    wsched_.set_alt_pieces(tn.witness_g1_piece_min ? tn.witness_g1_piece_min : 32);
    release_on_throw.armed = false;
}

Groth16Prover::~Groth16Prover()
{
    release();
}

void Groth16Prover::release() noexcept
{
    if (stream2_) {
        (void)hipStreamSynchronize(stream2_);
        (void)hipStreamDestroy(stream2_);
        stream2_ = nullptr;
    }
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
        stream_ = nullptr;
    }
    if (ev_start_) (void)hipEventDestroy(ev_start_);
    if (ev_wsched_) (void)hipEventDestroy(ev_wsched_);
    ev_start_ = ev_wsched_ = nullptr;
    upload_.reset();
}

uint64_t Groth16Prover::device_bytes() const
{
    uint64_t b = a_.bytes() + b1_.bytes() + l_.bytes() + h_.bytes() + b2_.bytes();
    for (int m = 0; m < 3; ++m) b += row_ptr_[m].bytes() + col_idx_[m].bytes() + values_[m].bytes();
    b += z_.bytes() + pa_.bytes() + pb_.bytes() + pc_.bytes();
    if (upload_) b += upload_->device_bytes();
    return b;
}

void Groth16Prover::prove(const void *witness, void *raw_out, void *az, void *bz, void *cz, void *h)
{
    NKR_REQUIRE(witness && raw_out, "null witness or output");
    NKR_HIP(hipSetDevice(device_));
    upload_->upload(static_cast<const fr_t *>(witness), z_.get(), stream_);
    run(z_.get(), raw_out, az, bz, cz, h);
}

void Groth16Prover::prove_dev(const fr_t *witness_dev, void *raw_out)
{
    NKR_REQUIRE(witness_dev && raw_out, "null witness or output");
    NKR_HIP(hipSetDevice(device_));
    run(witness_dev, raw_out, nullptr, nullptr, nullptr, nullptr);
}

void Groth16Prover::set_profiling(bool on)
{
    timer_.enable(on);
    for (double &v : stage_ms_) v = 0;
    for (uint64_t &v : counters_) v = 0;
}

void Groth16Prover::profile_read(double *stage_ms, uint64_t *counters) const
{
    for (int i = 0; i < kStages; ++i) stage_ms[i] = stage_ms_[i];
    for (int i = 0; i < 4; ++i) counters[i] = counters_[i];
}

void Groth16Prover::prove_blind(const void *witness, bool on_device, const fr_t &r, const fr_t &s, void *raw_out,
                                void *proof_out)
{
    NKR_REQUIRE(witness && raw_out && proof_out, "null witness or output");
    NKR_REQUIRE(have_blind_, "the proving key given to prepare lacks alpha/beta/delta");
    NKR_HIP(hipSetDevice(device_));
    const fr_t *z = static_cast<const fr_t *>(witness);
    if (!on_device) {
        upload_->upload(z, z_.get(), stream_);
        z = z_.get();
    }
    const BlindJob job{r, s, proof_out};
    run(z, raw_out, nullptr, nullptr, nullptr, nullptr, &job);
}

// A0 = alpha_1 + r delta_1, B0 = beta_2 + s delta_2 and
// C0 = s alpha_1 + r beta_1 + (r s) delta_1, using
// s*A + r*B_1 - (r s) delta_1 = s alpha_1 + s ar + r beta_1 + r bs1 + (r s) delta_1.
Groth16Prover::BlindPre Groth16Prover::blind_pre(const fr_t &r, const fr_t &s) const
{
    BlindPre p;
    p.A0 = xyzz_add(jacobian_to_xyzz(alpha1_), fixed_base_mul_host(tab_delta1_, r));
    p.B0 = xyzz_add(jacobian_to_xyzz(beta2_), fixed_base_mul_host(tab_delta2_, s));
    p.C0 = xyzz_add(fixed_base_mul_host(tab_alpha1_, s), fixed_base_mul_host(tab_beta1_, r));
    p.C0 = xyzz_add(p.C0, fixed_base_mul_host(tab_delta1_, fp_mul(r, s)));
    return p;
}

// Host tail: sum_j 2^j U_j for the five MSMs, the raw proof, then the
// blinded proof from the precomputed r,s-only terms.
void Groth16Prover::finish(const Xyzz<fq_t> *p1w, const Xyzz<fq_t> *p1h, const Xyzz<fq2_t> *p2,
                           const BlindJob *blind, const BlindPre *pre, void *raw_out) const
{
    finish_g2(p2, blind, pre, raw_out);
    finish_g1(p1w, p1h, blind, pre, raw_out, nullptr);
}

void Groth16Prover::finish_serial(const Xyzz<fq_t> *p1w, const Xyzz<fq_t> *p1h, const Xyzz<fq2_t> *p2,
                                  const BlindJob *blind, const BlindPre *pre, void *raw_out) const
{
    const uint32_t cw = wplan_.c;
    // raw proof: ar, bs1, zk, qz (G1) and bs2 (G2); blinded: A, C (G1), B (G2)
    g1_xyzz g1[6];
    g2_xyzz g2[2];
    g1[0] = msm_sum(p1w + 0 * cw, cw);
    g1[1] = msm_sum(p1w + 1 * cw, cw);
    g1[2] = msm_sum(p1w + 2 * cw, cw);
    g1[3] = msm_sum(p1h, hplan_.c);
    g2[0] = msm_sum(p2, cw);
    int n1 = 4, n2 = 1;
    if (blind) {
        g1[4] = xyzz_add(pre->A0, g1[0]);
        g2[1] = xyzz_add(pre->B0, g2[0]);
        g1_xyzz C = xyzz_add(g1[3], g1[2]);
        C = xyzz_add(C, pre->C0);
        C = xyzz_add(C, g1_mul2_glv(g1[0], blind->s, g1[1], blind->r));
        g1[5] = C;
        n1 = 6;
        n2 = 2;
    }
    g1_jac a1[6];
    g2_jac a2[2];
    xyzz_to_jacobian_affine_batch(g1, a1, n1);
    xyzz_to_jacobian_affine_batch(g2, a2, n2);
    uint8_t *out = static_cast<uint8_t *>(raw_out);
    for (int i = 0; i < 4; ++i) std::memcpy(out + 96 * i, &a1[i], 96);
    std::memcpy(out + 384, &a2[0], 192);
    if (!blind) return;
    uint8_t *proof = static_cast<uint8_t *>(blind->proof_out);
    std::memcpy(proof, &a1[4], 96);
    std::memcpy(proof + 96, &a2[1], 192);
    std::memcpy(proof + 288, &a1[5], 96);
}

void Groth16Prover::finish_g2(const Xyzz<fq2_t> *p2, const BlindJob *blind, const BlindPre *pre,
                              void *raw_out) const
{
    const g2_jac bs2 = msm_finish(p2, wplan_.c);
    std::memcpy(static_cast<uint8_t *>(raw_out) + 384, &bs2, 192);
    if (!blind) return;
    const g2_jac B = xyzz_to_jacobian_affine(xyzz_add(pre->B0, jacobian_to_xyzz(bs2)));
    std::memcpy(static_cast<uint8_t *>(blind->proof_out) + 96, &B, 192);
}

void Groth16Prover::finish_g1(const Xyzz<fq_t> *p1, const Xyzz<fq_t> *p1h, const BlindJob *blind,
                              const BlindPre *pre, void *raw_out, HostHelpers *helpers) const
{
    const uint32_t cw = wplan_.c;
    uint8_t *out = static_cast<uint8_t *>(raw_out);
    // Three independent parts: zk and H; bs1 and r*bs1; ar, A and s*ar.  The
    // points stay in XYZZ form until one inversion shared by all six (ar,
    // bs1, zk, qz, A, C).
    g1_xyzz pts[6];
    g1_xyzz s_ar = g1_xyzz::infinity(), r_bs1 = g1_xyzz::infinity();
    auto part_zq = [&] {
        pts[2] = msm_sum(p1 + 2 * cw, cw);
        pts[3] = msm_sum(p1h, hplan_.c);
    };
    auto part_b = [&] {
        pts[1] = msm_sum(p1 + 1 * cw, cw);
        if (blind) r_bs1 = g1_mul_glv(pts[1], blind->r);
    };
    // the helpers' tasks refer to this frame: if anything below throws, the
    // guard waits for them before the frame goes away
    struct Join {
        HostHelpers *h;
        ~Join()
        {
            if (!h) return;
            for (int i = 1; i >= 0; --i) try {
                    h->wait(i);
                } catch (...) {
                }
        }
    } join{nullptr};
    if (helpers) {
        helpers->run(0, part_zq);
        helpers->run(1, part_b);
        join.h = helpers;
    }
    pts[0] = msm_sum(p1 + 0 * cw, cw);
    if (blind) {
        pts[4] = xyzz_add(pre->A0, pts[0]);
        s_ar = g1_mul_glv(pts[0], blind->s);
    }
    if (helpers) {
        join.h = nullptr;
        std::exception_ptr err;
        for (int i = 1; i >= 0; --i) try {
                helpers->wait(i);
            } catch (...) {
                if (!err) err = std::current_exception();
            }
        if (err) std::rethrow_exception(err);
    } else {
        part_b();
        part_zq();
    }
    int n = 4;
    if (blind) {
        g1_xyzz C = xyzz_add(pts[3], pts[2]);
        C = xyzz_add(C, pre->C0);
        C = xyzz_add(C, xyzz_add(s_ar, r_bs1));
        pts[5] = C;
        n = 6;
    }
    g1_jac aff[6];
    xyzz_to_jacobian_affine_batch(pts, aff, n);
    for (int i = 0; i < 4; ++i) std::memcpy(out + 96 * i, &aff[i], 96);
    if (!blind) return;
    std::memcpy(static_cast<uint8_t *>(blind->proof_out), &aff[4], 96);
    std::memcpy(static_cast<uint8_t *>(blind->proof_out) + 288, &aff[5], 96);
}

void Groth16Prover::run(const fr_t *z, void *raw_out, void *az, void *bz, void *cz, void *h, const BlindJob *blind)
{
    hipStream_t s = stream_;
    StageTimer *t = timer_.enabled() ? &timer_ : nullptr;

    const uint32_t cw = wplan_.c;
    helpers_->arm();  // the host tail's helpers spin from here instead of sleeping
    NKR_HIP(hipEventRecord(ev_start_, s));

    // The quotient chain is the critical path: its launches are queued
    // first, the side stream's after them.
    if (t) t->begin(0, s);
    spmv3(views_[0], views_[1], views_[2], z, pa_.get(), pb_.get(), pc_.get(), s);
    if (t) t->end(0, s);
    const size_t poly_bytes = (size_t)rows_ * sizeof(fr_t);
    if (az) NKR_HIP(hipMemcpyAsync(az, pa_.get(), poly_bytes, hipMemcpyDeviceToHost, s));
    if (bz) NKR_HIP(hipMemcpyAsync(bz, pb_.get(), poly_bytes, hipMemcpyDeviceToHost, s));
    if (cz) NKR_HIP(hipMemcpyAsync(cz, pc_.get(), poly_bytes, hipMemcpyDeviceToHost, s));

    // quotient chain, no bit-reversal passes; one launch per pass covers A, B and C
    if (t) t->begin(1, s);
    domain_->dif_batch(poly_table_.get(), 3, true, nullptr, nullptr, nullptr, s);
    domain_->dit_batch(poly_table_.get(), 3, false, domain_->coset_rev_scaled(), nullptr, s);
    if (t) {
        t->end(1, s);
        t->begin(2, s);
    }
    domain_->quotient(pa_.get(), pb_.get(), pc_.get(), pa_.get(), s);
    if (t) t->end(2, s);
    // pa now holds the quotient's coset evaluations q, which the H MSM takes
    // directly (its bases are in the coset-evaluation basis).  The
    // coefficients h = icosetFFT(q) are formed only when asked for.
    if (h) {
        NKR_HIP(hipMemcpyAsync(pb_.get(), pa_.get(), poly_bytes, hipMemcpyDeviceToDevice, s));
        domain_->dif(pb_.get(), true, nullptr, domain_->icoset_rev_scaled(), nullptr, s);
        domain_->bitrev(pb_.get(), pc_.get(), s);
        NKR_HIP(hipMemcpyAsync(h, pc_.get(), poly_bytes, hipMemcpyDeviceToHost, s));
    }

    // The witness schedule and the G2 MSM (bs2) depend only on z: they run on
    // the side stream while the main stream computes H.
    NKR_HIP(hipStreamWaitEvent(stream2_, ev_start_, 0));
    wsched_.build(z, wplan_, stream2_, t);
    NKR_HIP(hipEventRecord(ev_wsched_, stream2_));
    batch_g2_.run({MsmJob<fq2_t>{&b2_, &wsched_, planes2_.get()}}, stream2_, t);
    Xyzz<fq2_t> *h2 = planes2_host_.reuse(cw);
    NKR_HIP(hipMemcpyAsync(h2, planes2_.get(), cw * sizeof(Xyzz<fq2_t>), hipMemcpyDeviceToHost, stream2_));
    planes2_host_.uploaded(stream2_);

    // The G1 MSMs ar, bs1, zk (witness) and H as one batch on the main
    // stream: the witness pieces fill the H accumulation launch, and the
    // latency-bound reduction levels are paid once for all four.
    hsched_.build(pa_.get(), hplan_, s, t);
    NKR_HIP(hipStreamWaitEvent(s, ev_wsched_, 0));
    batch_g1_.run({MsmJob<fq_t>{&a_, &wsched_, planes1_.get() + 0 * cw, true},
                   MsmJob<fq_t>{&b1_, &wsched_, planes1_.get() + 1 * cw, true},
                   MsmJob<fq_t>{&l_, &wsched_, planes1_.get() + 2 * cw, true},
                   MsmJob<fq_t>{&h_, &hsched_, planes1_.get() + 3 * cw}},
                  s, t);
    const size_t n1 = 3 * cw + hplan_.c;
    Xyzz<fq_t> *h1 = planes1_host_.reuse(n1);
    NKR_HIP(hipMemcpyAsync(h1, planes1_.get(), n1 * sizeof(Xyzz<fq_t>), hipMemcpyDeviceToHost, s));
    planes1_host_.uploaded(s);

    // blinding terms that need only r and s, then the G2 tail as soon as the
    // side stream is done, both while the GPU works on the G1 batch
    BlindPre pre;
    if (blind) pre = blind_pre(blind->r, blind->s);
    double host_ms = 0;
    planes2_host_.wait();
    auto host0 = std::chrono::steady_clock::now();
    finish_g2(h2, blind, blind ? &pre : nullptr, raw_out);
    host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - host0).count();
    planes1_host_.wait();
    host0 = std::chrono::steady_clock::now();
    finish_g1(h1, h1 + 3 * cw, blind, blind ? &pre : nullptr, raw_out, helpers_.get());
    host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - host0).count();

    if (t) {
        NKR_HIP(hipStreamSynchronize(s));
        NKR_HIP(hipStreamSynchronize(stream2_));
        t->collect(stage_ms_);
        stage_ms_[6] += host_ms;
        counters_[0] += 1;
        counters_[1] += 2;  // k_accumulate launches: G1 batch (witness and H), G2
        uint32_t tw = 0, th = 0;
        NKR_HIP(hipMemcpy(&tw, wsched_.totals(), sizeof(tw), hipMemcpyDeviceToHost));
        NKR_HIP(hipMemcpy(&th, hsched_.totals(), sizeof(th), hipMemcpyDeviceToHost));
        counters_[2] += tw;
        counters_[3] += th;
    }
}

namespace {

template <class F>
Xyzz<F> jmul(const Jacobian<F> &p, const fr_t &k_mont)
{
    const fr_t k = fp_from_mont(k_mont);
    return xyzz_mul_u256_w4(jacobian_to_xyzz(p), k.l);
}

template <class F>
Jacobian<F> load_point(const void *p)
{
    Jacobian<F> j;
    std::memcpy(&j, p, sizeof(j));
    return j;
}

}  // namespace

void groth16_blind(const nkr_g16_pk &pk, const void *raw, const fr_t &r, const fr_t &s, void *proof_out)
{
    NKR_REQUIRE(pk.alpha_g1 && pk.beta_g1 && pk.beta_g2 && pk.delta_g1 && pk.delta_g2,
                "proving key group elements missing");
    const uint8_t *rw = static_cast<const uint8_t *>(raw);
    const g1_xyzz ar = jacobian_to_xyzz(load_point<fq_t>(rw + 0));
    const g1_xyzz bs1 = jacobian_to_xyzz(load_point<fq_t>(rw + 96));
    const g1_xyzz zk = jacobian_to_xyzz(load_point<fq_t>(rw + 192));
    const g1_xyzz qz = jacobian_to_xyzz(load_point<fq_t>(rw + 288));
    const g2_xyzz bs2 = jacobian_to_xyzz(load_point<fq2_t>(rw + 384));
    const g1_jac alpha = load_point<fq_t>(pk.alpha_g1);
    const g1_jac beta1 = load_point<fq_t>(pk.beta_g1);
    const g1_jac delta1 = load_point<fq_t>(pk.delta_g1);
    const g2_jac beta2 = load_point<fq2_t>(pk.beta_g2);
    const g2_jac delta2 = load_point<fq2_t>(pk.delta_g2);

    // A = alpha + ar + r*delta1
    const g1_xyzz a = xyzz_add(xyzz_add(jacobian_to_xyzz(alpha), ar), jmul(delta1, r));
    // B1 = beta1 + bs1 + s*delta1 ; B = beta2 + bs2 + s*delta2
    const g1_xyzz b1 = xyzz_add(xyzz_add(jacobian_to_xyzz(beta1), bs1), jmul(delta1, s));
    const g2_xyzz b2 = xyzz_add(xyzz_add(jacobian_to_xyzz(beta2), bs2), jmul(delta2, s));
    // C = qz + zk + s*A + r*B1 - (r*s)*delta1
    const fr_t sk = fp_from_mont(s), rk = fp_from_mont(r);
    g1_xyzz cpt = xyzz_add(qz, zk);
    cpt = xyzz_add(cpt, xyzz_mul_u256_w4(a, sk.l));
    cpt = xyzz_add(cpt, xyzz_mul_u256_w4(b1, rk.l));
    cpt = xyzz_add(cpt, xyzz_neg(jmul(delta1, fp_mul(r, s))));

    const g1_jac A = xyzz_to_jacobian_affine(a);
    const g2_jac B = xyzz_to_jacobian_affine(b2);
    const g1_jac C = xyzz_to_jacobian_affine(cpt);
    uint8_t *o = static_cast<uint8_t *>(proof_out);
    std::memcpy(o, &A, 96);
    std::memcpy(o + 96, &B, 192);
    std::memcpy(o + 288, &C, 96);
}

}  // namespace nkr

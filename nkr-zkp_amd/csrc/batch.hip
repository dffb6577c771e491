// Batched Groth16 proving on one GPU (see batch.h for the stage plan).
#include "batch.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <thread>

namespace nkr {

namespace {

// fn(i) for i < n on up to 16 host threads.
template <class Fn>
void parallel_for(size_t n, Fn fn)
{
    const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
    const size_t workers = std::min<size_t>({n, hw, 16});
    if (workers <= 1) {
        for (size_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    std::exception_ptr err;
    std::atomic<bool> failed{false};
    for (size_t w = 0; w < workers; ++w)
        pool.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < n;) {
                try {
                    fn(i);
                } catch (...) {
                    if (!failed.exchange(true)) err = std::current_exception();
                }
            }
        });
    for (auto &t : pool) t.join();
    if (err) std::rethrow_exception(err);
}

size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

constexpr uint32_t kBatchPiece = 32;

}  // namespace

MsmPlan batch_plan(const MsmPlan &plan) { return plan.with_piece(kBatchPiece); }

BatchProver::BatchProver(int device) : device_(device)
{
    NKR_HIP(hipSetDevice(device_));
    struct Undo {
        BatchProver *p;
        bool armed = true;
        ~Undo()
        {
            if (armed) p->release();
        }
    } undo{this};
    NKR_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    NKR_HIP(hipStreamCreateWithFlags(&stream2_, hipStreamNonBlocking));
    NKR_HIP(hipEventCreateWithFlags(&ev_z_, hipEventDisableTiming));
    NKR_HIP(hipEventCreateWithFlags(&ev_wsched_, hipEventDisableTiming));
    NKR_HIP(hipEventCreateWithFlags(&ev_side_, hipEventDisableTiming));
    undo.armed = false;
}

BatchProver::~BatchProver()
{
    release();
}

void BatchProver::release() noexcept
{
    drop_graph();
    last_key_.clear();
    for (hipStream_t *s : {&stream2_, &stream_})
        if (*s) {
            (void)hipStreamSynchronize(*s);
            (void)hipStreamDestroy(*s);
            *s = nullptr;
        }
    for (hipEvent_t *e : {&ev_z_, &ev_wsched_, &ev_side_})
        if (*e) {
            (void)hipEventDestroy(*e);
            *e = nullptr;
        }
}

void BatchProver::drop_graph() noexcept
{
    if (graph_exec_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipGraphExecDestroy(graph_exec_);
        graph_exec_ = nullptr;
    }
    graph_key_.clear();
}

void BatchProver::set_graphs(bool enabled)
{
    graphs_enabled_ = enabled;
    if (!enabled) drop_graph();
}

uint64_t BatchProver::device_bytes() const
{
    uint64_t b = tables_.bytes() + sort_w_.device_bytes() + sort_h_.device_bytes() + batch_g1_.device_bytes() +
                 batch_g2_.device_bytes() + planes_w_.bytes() + planes_h_.bytes() +
                 planes_2_.bytes();
    for (const auto &sl : slots_)
        b += sl->z.bytes() + sl->pa.bytes() + sl->pb.bytes() + sl->pc.bytes() + sl->wsched.device_bytes() +
             sl->hsched.device_bytes();
    return b;
}

void BatchProver::trim()
{
    NKR_HIP(hipSetDevice(device_));
    NKR_HIP(hipStreamSynchronize(stream_));
    NKR_HIP(hipStreamSynchronize(stream2_));
    drop_graph();  // its launches refer to the buffers released here
    last_key_.clear();
    slots_.clear();
    tables_.release();
    planes_w_.release();
    planes_h_.release();
    planes_2_.release();
    sort_w_.release_device();
    sort_h_.release_device();
    batch_g1_.release_device();
    batch_g2_.release_device();
    wupload_.release_device();
}

void BatchProver::prove(const std::vector<BatchItem> &items)
{
    const size_t K = items.size();
    if (K == 0) return;
    NKR_REQUIRE(K <= 16384, "a batch holds at most 16384 proofs");
    for (const BatchItem &it : items) {
        NKR_REQUIRE(it.ctx && it.witness && it.raw_out && it.proof_out, "batch item with a null member");
        NKR_REQUIRE(it.ctx->device_ == device_, "batch item prepared on another device");
        NKR_REQUIRE(it.ctx->have_blind_, "the proving key given to prepare lacks alpha/beta/delta");
    }
    NKR_HIP(hipSetDevice(device_));
    while (slots_.size() < K) slots_.push_back(std::make_unique<Slot>());

    // Launch-sequence reuse: a batch of the same contexts in the same order as
    // the previous call (the same sizes, slots and tables) is captured into a
    // HIP graph on its second occurrence and replayed from then on; any other
    // batch runs eagerly and drops the graph, since it may move what the
    // graph's launches refer to.
    std::vector<const void *> key(K);
    for (size_t i = 0; i < K; ++i) key[i] = items[i].ctx;
    const bool replay = graph_exec_ && key == graph_key_;
    const bool capture = !replay && graphs_enabled_ && key == last_key_;
    if (!replay) drop_graph();
    last_key_ = key;
    const bool graph = replay || capture;

    // bit planes of every proof in three batch-wide regions (G1 over the
    // witness, H, G2) so that each reaches the host in one copy
    uint32_t max_rows = 0, max_long = 0;
    std::vector<size_t> pw_off(K + 1, 0), ph_off(K + 1, 0), p2_off(K + 1, 0);
    for (size_t i = 0; i < K; ++i) {
        const Groth16Prover &g = *items[i].ctx;
        for (int m = 0; m < 3; ++m) max_long = std::max(max_long, g.views_[m].n_long);
        max_rows = std::max(max_rows, g.rows_);
        pw_off[i + 1] = pw_off[i] + 3 * g.wplan_.c;
        ph_off[i + 1] = ph_off[i] + g.hplan_.c;
        p2_off[i + 1] = p2_off[i] + g.wplan_.c;
    }
    planes_w_.ensure(pw_off[K]);
    planes_h_.ensure(ph_off[K]);
    planes_2_.ensure(p2_off[K]);
    for (size_t i = 0; i < K; ++i) {
        const Groth16Prover &g = *items[i].ctx;
        Slot &sl = *slots_[i];
        sl.z.ensure(g.cols_);
        sl.pa.ensure(g.rows_);
        sl.pb.ensure(g.rows_);
        sl.pc.ensure(g.rows_);
    }
    // Witnesses: packed on host threads (values other than 0 and 1 travel,
    // the rest as bit masks), each 1 MiB chunk's copy issued as soon as it is
    // packed, then one expansion kernel into every slot.  The upload is
    // issued eagerly in every mode; a captured graph holds only the GPU work
    // that follows it (stream order puts that work after the expansion).
    {
        std::vector<WitnessBatchUpload::Item> wi(K);
        for (size_t i = 0; i < K; ++i)
            wi[i] = {static_cast<const fr_t *>(items[i].witness), slots_[i]->z.get(), items[i].ctx->cols_};
        wupload_.upload(wi, stream_);
    }

    if (replay) {
        NKR_HIP(hipGraphLaunch(graph_exec_, stream_));
    } else {
        // An exception between the capture's begin and end must end the
        // capture and drop what it recorded.
        struct CaptureGuard {
            hipStream_t s;
            bool open = false;
            ~CaptureGuard()
            {
                if (!open) return;
                hipGraph_t g = nullptr;
                if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
            }
        } guard{stream_};
        if (capture) {
            // The staging buffers the captured work uses are taken without
            // waiting on their events: the previous call ended with both
            // streams idle, and the witness copies finish here first.
            NKR_HIP(hipStreamSynchronize(stream_));
            NKR_HIP(hipStreamSynchronize(stream2_));
            NKR_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
            guard.open = true;
        }

        // Side stream: the witness schedules and the G2 MSMs need only z.
        NKR_HIP(hipEventRecord(ev_z_, stream_));
        NKR_HIP(hipStreamWaitEvent(stream2_, ev_z_, 0));
        {
            std::vector<MsmScheduleBatch::Item> wsorts;
            std::vector<MsmJob<fq2_t>> g2jobs;
            for (size_t i = 0; i < K; ++i) {
                Groth16Prover &g = *items[i].ctx;
                Slot &sl = *slots_[i];
                wsorts.push_back({&sl.wsched, sl.z.get(), batch_plan(g.wplan_)});
                g2jobs.push_back({&g.b2_, &sl.wsched, planes_2_.get() + p2_off[i]});
            }
            sort_w_.build(wsorts, stream2_, graph);
            NKR_HIP(hipEventRecord(ev_wsched_, stream2_));
            batch_g2_.run(g2jobs, stream2_, nullptr, graph);
        }

        // device tables: SpMV jobs, then per size class the transform and quotient pointer lists
        std::map<uint32_t, std::vector<size_t>> classes;
        for (size_t i = 0; i < K; ++i) classes[items[i].ctx->log_n_].push_back(i);
        size_t tbytes = align16(K * sizeof(SpmvJob));
        for (const auto &cl : classes) tbytes += align16(cl.second.size() * 7 * sizeof(fr_t *));
        uint8_t *tst = graph ? table_stage_.reuse_unsynced(tbytes) : table_stage_.reuse(tbytes);
        tables_.ensure(tbytes);
        SpmvJob *jobs = reinterpret_cast<SpmvJob *>(tst);
        for (size_t i = 0; i < K; ++i) {
            const Groth16Prover &g = *items[i].ctx;
            Slot &sl = *slots_[i];
            for (int m = 0; m < 3; ++m) jobs[i].m[m] = g.views_[m];
            jobs[i].z = sl.z.get();
            jobs[i].y[0] = sl.pa.get();
            jobs[i].y[1] = sl.pb.get();
            jobs[i].y[2] = sl.pc.get();
        }
        struct ClassTables {
            const NttDomain *domain;
            uint32_t count;
            size_t all3, quot;  // byte offsets into tables_
        };
        std::vector<ClassTables> ctabs;
        size_t off = align16(K * sizeof(SpmvJob));
        for (const auto &cl : classes) {
            const size_t k = cl.second.size();
            fr_t **p = reinterpret_cast<fr_t **>(tst + off);
            ClassTables ct{items[cl.second[0]].ctx->domain_.get(), (uint32_t)k, off, off + 3 * k * sizeof(fr_t *)};
            for (size_t q = 0; q < k; ++q) {
                Slot &sl = *slots_[cl.second[q]];
                p[3 * q + 0] = sl.pa.get();
                p[3 * q + 1] = sl.pb.get();
                p[3 * q + 2] = sl.pc.get();
                p[3 * k + 4 * q + 0] = sl.pa.get();
                p[3 * k + 4 * q + 1] = sl.pb.get();
                p[3 * k + 4 * q + 2] = sl.pc.get();
                p[3 * k + 4 * q + 3] = sl.pa.get();
            }
            ctabs.push_back(ct);
            off += align16(k * 7 * sizeof(fr_t *));
        }
        NKR_HIP(hipMemcpyAsync(tables_.get(), tst, tbytes, hipMemcpyHostToDevice, stream_));
        if (!graph) table_stage_.uploaded(stream_);

        // Main stream: A*z, B*z, C*z, then the quotient chain per size class
        spmv3_batch(reinterpret_cast<const SpmvJob *>(tables_.get()), (uint32_t)K, max_rows, max_long, stream_);
        for (const ClassTables &ct : ctabs) {
            fr_t *const *all3 = reinterpret_cast<fr_t *const *>(tables_.get() + ct.all3);
            fr_t *const *quot = reinterpret_cast<fr_t *const *>(tables_.get() + ct.quot);
            ct.domain->dif_batch(all3, 3 * ct.count, true, nullptr, nullptr, nullptr, stream_);
            ct.domain->dit_batch(all3, 3 * ct.count, false, ct.domain->coset_rev_scaled(), nullptr, stream_);
            ct.domain->quotient_batch(quot, ct.count, stream_);
        }
        // digit schedules of every H, then one G1 batch over all four G1 MSMs of
        // every proof (the witness schedules from the side stream)
        std::vector<MsmJob<fq_t>> g1jobs;
        std::vector<MsmScheduleBatch::Item> hsorts;
        for (size_t i = 0; i < K; ++i) {
            Groth16Prover &g = *items[i].ctx;
            Slot &sl = *slots_[i];
            hsorts.push_back({&sl.hsched, sl.pa.get(), batch_plan(g.hplan_)});
        }
        sort_h_.build(hsorts, stream_, graph);
        NKR_HIP(hipStreamWaitEvent(stream_, ev_wsched_, 0));
        for (size_t i = 0; i < K; ++i) {
            Groth16Prover &g = *items[i].ctx;
            Slot &sl = *slots_[i];
            const uint32_t cw = g.wplan_.c;
            Xyzz<fq_t> *pw = planes_w_.get() + pw_off[i];
            g1jobs.push_back({&g.a_, &sl.wsched, pw + 0 * cw});
            g1jobs.push_back({&g.b1_, &sl.wsched, pw + 1 * cw});
            g1jobs.push_back({&g.l_, &sl.wsched, pw + 2 * cw});
            g1jobs.push_back({&g.h_, &sl.hsched, planes_h_.get() + ph_off[i]});
        }
        batch_g1_.run(g1jobs, stream_, nullptr, graph);

        // bit planes of every proof to the host: three copies
        const size_t nw = pw_off[K] * sizeof(Xyzz<fq_t>), nh = ph_off[K] * sizeof(Xyzz<fq_t>);
        const size_t n2 = p2_off[K] * sizeof(Xyzz<fq2_t>);
        uint8_t *ph = graph ? planes_host_.reuse_unsynced(nw + nh + n2) : planes_host_.reuse(nw + nh + n2);
        NKR_HIP(hipMemcpyAsync(ph, planes_w_.get(), nw, hipMemcpyDeviceToHost, stream_));
        NKR_HIP(hipMemcpyAsync(ph + nw, planes_h_.get(), nh, hipMemcpyDeviceToHost, stream_));
        NKR_HIP(hipMemcpyAsync(ph + nw + nh, planes_2_.get(), n2, hipMemcpyDeviceToHost, stream2_));
        NKR_HIP(hipEventRecord(ev_side_, stream2_));
        NKR_HIP(hipStreamWaitEvent(stream_, ev_side_, 0));
        if (capture) {
            hipGraph_t g = nullptr;
            guard.open = false;
            NKR_HIP(hipStreamEndCapture(stream_, &g));
            const hipError_t ie = hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                graph_exec_ = nullptr;
                NKR_HIP(ie);
            }
            graph_key_ = key;
            NKR_HIP(hipGraphLaunch(graph_exec_, stream_));
        }
    }
    const size_t nw = pw_off[K] * sizeof(Xyzz<fq_t>), nh = ph_off[K] * sizeof(Xyzz<fq_t>);
    const uint8_t *ph = planes_host_.get();
    const Xyzz<fq_t> *hw = reinterpret_cast<const Xyzz<fq_t> *>(ph);
    const Xyzz<fq_t> *hh = reinterpret_cast<const Xyzz<fq_t> *>(ph + nw);
    const Xyzz<fq2_t> *h2 = reinterpret_cast<const Xyzz<fq2_t> *>(ph + nw + nh);

    // r,s-only blinding terms while the GPU works, then the host tails
    std::vector<Groth16Prover::BlindPre> pre(K);
    parallel_for(K, [&](size_t i) { pre[i] = items[i].ctx->blind_pre(items[i].r, items[i].s); });
    NKR_HIP(hipStreamSynchronize(stream_));
    parallel_for(K, [&](size_t i) {
        const BatchItem &it = items[i];
        const Groth16Prover &g = *it.ctx;
        const Groth16Prover::BlindJob job{it.r, it.s, it.proof_out};
        g.finish_serial(hw + pw_off[i], hh + ph_off[i], h2 + p2_off[i], &job, &pre[i], it.raw_out);
    });
}

}  // namespace nkr

// Continuous-batching request loop (see serve.h).
#include "serve.h"

#include <algorithm>
#include <chrono>

namespace nkr {

uint64_t Server::request_bytes(const Groth16Prover &g)
{
    const uint64_t polys = (uint64_t)g.cols() * sizeof(fr_t) + 3ull * g.rows() * sizeof(fr_t);
    // the packed witness on the device (every value, at worst) and its masks,
    // offsets and chunk entry
    const uint64_t packed = (uint64_t)g.cols() * sizeof(fr_t) + ((uint64_t)g.cols() / 64 + 1) * 20 +
                            ((uint64_t)g.cols() / WitnessUpload::kChunk + 1) * sizeof(WitnessBatchUpload::Chunk) + 16;
    const MsmPlan w = batch_plan(g.witness_plan()), h = batch_plan(g.h_plan());
    const uint64_t planes = (3ull * w.c + h.c) * sizeof(Xyzz<fq_t>) + (uint64_t)w.c * sizeof(Xyzz<fq2_t>);
    return polys + packed + planes + msm_schedule_bytes(w) + msm_schedule_bytes(h) + 3 * msm_job_bytes<fq_t>(w) +
           msm_job_bytes<fq_t>(h) + msm_job_bytes<fq2_t>(w);
}

Server::Server(int device, uint32_t workers, uint32_t max_batch, uint64_t work_budget_bytes)
    : device_(device), max_batch_(max_batch)
{
    NKR_REQUIRE(workers >= 1 && workers <= 8, "a server runs 1 .. 8 workers");
    NKR_REQUIRE(max_batch >= 1, "max_batch must be at least 1");
    NKR_REQUIRE(work_budget_bytes > 0, "the work budget must be positive");
    NKR_HIP(hipSetDevice(device_));
    size_t free_b = 0, total_b = 0;
    NKR_HIP(hipMemGetInfo(&free_b, &total_b));
    const uint64_t reserve = std::max<uint64_t>(1ull << 30, total_b / 32);
    const uint64_t avail = free_b > reserve ? free_b - reserve : 0;
    budget_ = std::min<uint64_t>(work_budget_bytes, avail);
    NKR_REQUIRE(budget_ > 0, "no device memory free for proving work (" + std::to_string(free_b) +
                                 " bytes free, " + std::to_string(reserve) + " kept in reserve)");
    share_ = budget_ / workers;
    held_.assign(workers, 0);
    for (uint32_t i = 0; i < workers; ++i) provers_.push_back(std::make_unique<BatchProver>(device));
    for (uint32_t i = 0; i < workers; ++i) threads_.emplace_back([this, i] { worker_loop(i); });
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (auto &t : threads_) t.join();
}

uint64_t Server::submit(Groth16Prover *ctx, const void *witness, const fr_t &r, const fr_t &s, void *raw_out,
                        void *proof_out)
{
    NKR_REQUIRE(ctx && witness && raw_out && proof_out, "null request member");
    NKR_REQUIRE(ctx->device() == device_, "context prepared on another device");
    const uint64_t bytes = request_bytes(*ctx);
    NKR_REQUIRE(bytes <= share_, "request needs " + std::to_string(bytes) +
                                     " bytes of work memory, more than a worker's share of " +
                                     std::to_string(share_) + " (a budget of " + std::to_string(budget_) +
                                     " over " + std::to_string(provers_.size()) + " workers)");
    std::lock_guard<std::mutex> lk(mu_);
    NKR_REQUIRE(!stop_, "server is shutting down");
    Request q;
    q.ticket = next_ticket_++;
    q.item = BatchItem{ctx, witness, r, s, raw_out, proof_out};
    q.bytes = bytes;
    q.t_submit = std::chrono::steady_clock::now();
    pending_.push_back(q);
    done_[q.ticket] = Done{};
    cv_work_.notify_one();
    return q.ticket;
}

double Server::wait(uint64_t ticket)
{
    std::unique_lock<std::mutex> lk(mu_);
    auto it = done_.find(ticket);
    NKR_REQUIRE(it != done_.end(), "unknown ticket " + std::to_string(ticket));
    cv_done_.wait(lk, [&] { return done_[ticket].finished; });
    const Done d = done_[ticket];
    done_.erase(ticket);
    if (!d.error.empty()) throw Error(d.error);
    return d.latency_ms;
}

void Server::warm(const std::vector<BatchItem> &items)
{
    NKR_REQUIRE(!items.empty(), "nothing to warm with");
    NKR_REQUIRE(items.size() <= max_batch_, "more warm-up items than max_batch");
    uint64_t bytes = 0;
    for (const BatchItem &it : items) {
        NKR_REQUIRE(it.ctx && it.witness && it.raw_out && it.proof_out, "null warm-up item member");
        NKR_REQUIRE(it.ctx->device() == device_, "context prepared on another device");
        NKR_REQUIRE(it.ctx->can_blind(), "the context was prepared without alpha/beta/delta, which blinding needs");
        bytes += request_bytes(*it.ctx);
    }
    NKR_REQUIRE(bytes <= share_, "warm-up items need " + std::to_string(bytes) +
                                     " bytes of work memory, more than a worker's share of " +
                                     std::to_string(share_));
    {
        std::unique_lock<std::mutex> lk(mu_);
        NKR_REQUIRE(!stop_, "server is shutting down");
        NKR_REQUIRE(!warming_, "a warm-up is already running");
        warming_ = true;
        cv_done_.wait(lk, [&] { return in_use_ == 0; });
    }
    struct Resume {
        Server *s;
        ~Resume()
        {
            {
                std::lock_guard<std::mutex> lk(s->mu_);
                s->warming_ = false;
            }
            s->cv_work_.notify_all();
        }
    } resume{this};
    // the workers are idle and admit nothing: their batch provers are free
    for (size_t i = 0; i < provers_.size(); ++i) {
        BatchProver &p = *provers_[i];
        if (p.device_bytes() + bytes > share_) p.trim();
        p.prove(items);
        uint64_t held = p.device_bytes();
        const uint64_t peak = held;
        if (held > share_) {
            p.trim();
            held = p.device_bytes();
        }
        std::lock_guard<std::mutex> lk(mu_);
        held_[i] = peak;
        uint64_t sum = 0;
        for (uint64_t h : held_) sum += h;
        peak_held_ = std::max(peak_held_, sum);
        held_[i] = held;
    }
}

Server::Stats Server::stats() const
{
    std::lock_guard<std::mutex> lk(mu_);
    return stats_;
}

Server::Memory Server::memory() const
{
    std::lock_guard<std::mutex> lk(mu_);
    Memory m;
    m.budget = budget_;
    m.share = share_;
    for (uint64_t h : held_) m.held += h;
    m.peak_held = peak_held_;
    return m;
}

std::vector<std::string> Server::prove_batch(BatchProver &prover, const std::vector<Request> &batch)
{
    std::vector<std::string> errors(batch.size());
    // requests that cannot be proved fail alone, before the batch runs
    std::vector<BatchItem> items;
    std::vector<size_t> at;
    for (size_t i = 0; i < batch.size(); ++i) {
        const BatchItem &it = batch[i].item;
        if (!it.ctx->can_blind())
            errors[i] = "the context was prepared without alpha/beta/delta, which blinding needs";
        else if (it.ctx->device() != device_)
            errors[i] = "context prepared on another device";
        else {
            items.push_back(it);
            at.push_back(i);
        }
    }
    if (items.empty()) return errors;
    try {
        prover.prove(items);
        return errors;
    } catch (const std::exception &e) {
        if (items.size() == 1) {
            errors[at[0]] = e.what();
            return errors;
        }
    }
    // the batch failed as a whole: one by one, so that only failing requests fail
    for (size_t k = 0; k < items.size(); ++k) {
        try {
            prover.prove({items[k]});
        } catch (const std::exception &e) {
            errors[at[k]] = e.what();
        }
    }
    return errors;
}

void Server::worker_loop(uint32_t index)
{
    BatchProver &prover = *provers_[index];
    for (;;) {
        std::vector<Request> batch;
        uint64_t bytes = 0;
        {
            std::unique_lock<std::mutex> lk(mu_);
            // admit pending requests in arrival order while their charges fit
            // this worker's share (every request fits a share: submit checks)
            cv_work_.wait(lk, [&] { return (!pending_.empty() && !warming_) || stop_; });
            if (stop_ && pending_.empty()) return;
            while (!pending_.empty() && batch.size() < max_batch_ && bytes + pending_.front().bytes <= share_) {
                bytes += pending_.front().bytes;
                batch.push_back(pending_.front());
                pending_.pop_front();
            }
            if (batch.empty()) continue;
            in_use_ += bytes;
            stats_.peak_work_bytes = std::max(stats_.peak_work_bytes, in_use_);
            stats_.batches += 1;
            stats_.proofs += batch.size();
            stats_.max_batch_seen = std::max<uint64_t>(stats_.max_batch_seen, batch.size());
        }
        // The prover's buffers from earlier batches stay unless this batch's
        // charges would take it over the share.
        std::vector<std::string> errors(batch.size());
        uint64_t held = 0;
        try {
            if (prover.device_bytes() + bytes > share_) prover.trim();
            errors = prove_batch(prover, batch);
            held = prover.device_bytes();
        } catch (const std::exception &e) {
            for (std::string &err : errors) err = e.what();
        }
        uint64_t kept = held;
        if (held > share_) {
            try {
                prover.trim();
                kept = prover.device_bytes();
            } catch (...) {
            }
        }
        const auto t_end = std::chrono::steady_clock::now();
        {
            std::lock_guard<std::mutex> lk(mu_);
            in_use_ -= bytes;
            held_[index] = held;
            uint64_t sum = 0;
            for (uint64_t h : held_) sum += h;
            peak_held_ = std::max(peak_held_, sum);
            held_[index] = kept;
            for (size_t i = 0; i < batch.size(); ++i) {
                Done &d = done_[batch[i].ticket];
                d.finished = true;
                d.error = errors[i];
                d.latency_ms = std::chrono::duration<double, std::milli>(t_end - batch[i].t_submit).count();
            }
        }
        cv_done_.notify_all();
        cv_work_.notify_all();  // memory freed: other workers may admit more
    }
}

}  // namespace nkr

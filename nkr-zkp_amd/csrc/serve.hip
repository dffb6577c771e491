// Continuous-batching request loop (see serve.h).
#include "serve.h"

#include <algorithm>
#include <chrono>

namespace nkr {

uint64_t Server::request_bytes(const Groth16Prover &g)
{
    const uint64_t polys = (uint64_t)g.cols() * sizeof(fr_t) + 3ull * g.rows() * sizeof(fr_t);
    const MsmPlan w = batch_plan(g.witness_plan()), h = batch_plan(g.h_plan());
    return polys + msm_schedule_bytes(w) + msm_schedule_bytes(h) + 3 * msm_job_bytes<fq_t>(w) +
           msm_job_bytes<fq_t>(h) + msm_job_bytes<fq2_t>(w);
}

Server::Server(int device, uint32_t workers, uint32_t max_batch, uint64_t work_budget_bytes)
    : device_(device), max_batch_(max_batch), budget_(work_budget_bytes)
{
    NKR_REQUIRE(workers >= 1 && workers <= 8, "a server runs 1 .. 8 workers");
    NKR_REQUIRE(max_batch >= 1, "max_batch must be at least 1");
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
    NKR_REQUIRE(bytes <= budget_, "request needs " + std::to_string(bytes) +
                                      " bytes of work memory, more than the server's budget of " +
                                      std::to_string(budget_));
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

Server::Stats Server::stats() const
{
    std::lock_guard<std::mutex> lk(mu_);
    return stats_;
}

void Server::worker_loop(uint32_t index)
{
    BatchProver &prover = *provers_[index];
    for (;;) {
        std::vector<Request> batch;
        uint64_t bytes = 0;
        {
            std::unique_lock<std::mutex> lk(mu_);
            // admit pending requests in arrival order while memory allows
            cv_work_.wait(lk, [&] {
                return pending_.empty() ? stop_ : in_use_ + pending_.front().bytes <= budget_;
            });
            if (stop_ && pending_.empty()) return;
            while (!pending_.empty() && batch.size() < max_batch_ && in_use_ + bytes + pending_.front().bytes <= budget_) {
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
        std::vector<BatchItem> items;
        for (const Request &q : batch) items.push_back(q.item);
        std::string error;
        try {
            prover.prove(items);
        } catch (const std::exception &e) {
            error = e.what();
        }
        const auto t_end = std::chrono::steady_clock::now();
        {
            std::lock_guard<std::mutex> lk(mu_);
            in_use_ -= bytes;
            for (const Request &q : batch) {
                Done &d = done_[q.ticket];
                d.finished = true;
                d.error = error;
                d.latency_ms = std::chrono::duration<double, std::milli>(t_end - q.t_submit).count();
            }
        }
        cv_done_.notify_all();
        cv_work_.notify_all();  // memory freed: other workers may admit more
    }
}

}  // namespace nkr

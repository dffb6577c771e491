// Serving: a request loop with continuous batching on one GPU.
//
// Proving keys stay resident in HBM as prepared contexts (Groth16Prover);
// requests name a context and carry a witness.  W worker threads each own a
// BatchProver.  A worker that is free takes every pending request that fits
// the work-memory budget (up to max_batch), proves them as one batch and
// completes them; meanwhile the other workers' batches are on the GPU on
// their own streams, so requests that arrive while the GPU is busy are
// admitted as soon as a worker and memory free up and their work overlaps
// the batches already running.
//
// Device memory: the work budget is clamped at creation to the device's
// free memory minus a reserve (the larger of 1 GiB and 1/32 of the device),
// and each worker gets an equal share of it.  A request is charged the device
// bytes its batch slot needs (witness, three domain polynomials, the two
// digit schedules, MSM partial sums and per-bucket arrays); a worker admits
// requests while their charges fit its share, and a request larger than a
// share is rejected at submission.  A worker's batch prover keeps its
// buffers between batches (they only grow), so before a batch whose charges
// would take it over its share, and after any batch that left it above its
// share, the worker releases them: the work memory all workers hold stays
// within the budget.
//
// Failures: requests that cannot be proved (a context prepared without the
// blinding elements, a context on another device) fail alone, at
// admission; if a whole batch fails, its requests are retried one by one so
// that only the ones that fail again report an error.
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <string>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.h"

namespace nkr {

class Server {
public:
    Server(int device, uint32_t workers, uint32_t max_batch, uint64_t work_budget_bytes);
    ~Server();
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    // Queues one proof; the caller keeps witness and outputs alive until
    // wait() returns for the ticket.
    uint64_t submit(Groth16Prover *ctx, const void *witness, const fr_t &r, const fr_t &s, void *raw_out,
                    void *proof_out);
    // Blocks until the request is done; returns its end-to-end latency in ms
    // (submission to completion) or throws the error its batch raised.
    double wait(uint64_t ticket);

    // Warm-up before traffic: every worker proves `items` once, as one batch
    // (the outputs are written to the items' buffers), so that each worker's
    // buffers already have the size those requests need; otherwise a worker
    // grows them on its first such request, which then pays the device
    // allocations.  Waits until no batch is in flight; requests submitted
    // meanwhile wait for the warm-up.  The items' charges must fit one
    // worker's share and max_batch; calls must not overlap each other or the
    // server's destruction.
    void warm(const std::vector<BatchItem> &items);

    // Device bytes a request on this context is charged at admission.
    static uint64_t request_bytes(const Groth16Prover &ctx);

    struct Stats {
        uint64_t batches = 0, proofs = 0, max_batch_seen = 0, peak_work_bytes = 0;
    };
    Stats stats() const;

    // Work memory: the effective budget, one worker's share, the bytes the
    // workers' batch provers hold now, and the most they held at once.
    struct Memory {
        uint64_t budget = 0, share = 0, held = 0, peak_held = 0;
    };
    Memory memory() const;

private:
    struct Request {
        uint64_t ticket;
        BatchItem item;
        uint64_t bytes;
        std::chrono::steady_clock::time_point t_submit;
    };
    struct Done {
        bool finished = false;
        double latency_ms = 0;
        std::string error;
    };
    void worker_loop(uint32_t index);
    // Proves a batch; on failure retries its requests one by one.  Returns
    // one error message per request (empty on success).
    std::vector<std::string> prove_batch(BatchProver &prover, const std::vector<Request> &batch);

    int device_;
    uint32_t max_batch_;
    uint64_t budget_ = 0, share_ = 0;
    std::vector<uint64_t> held_;  // per worker: bytes its batch prover holds
    uint64_t peak_held_ = 0;
    mutable std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<Request> pending_;
    std::map<uint64_t, Done> done_;
    uint64_t next_ticket_ = 1;
    uint64_t in_use_ = 0;
    bool stop_ = false;
    bool warming_ = false;  // workers admit nothing while warm() runs
    Stats stats_;
    std::vector<std::unique_ptr<BatchProver>> provers_;
    std::vector<std::thread> threads_;
};

}  // namespace nkr

// Serving: a request loop with continuous batching on one GPU.
//
// Proving keys stay resident in HBM as prepared contexts (Groth16Prover);
// requests name a context and carry a witness.  W worker threads each own a
// BatchProver.  A worker that is free takes every pending request that fits
// the work-memory budget (up to max_batch), proves them as one batch and
// completes them; meanwhile the other workers' batches are on the GPU on
// their own streams, so requests that arrive while the GPU is busy are
// admitted as soon as a worker and memory free up and their work overlaps
// the batches already running.  Admission charges each request the device
// bytes its batch slot needs (witness, three domain polynomials, the two
// digit schedules and MSM partial sums); a request that alone exceeds the
// budget is rejected at submission.
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

    // Device bytes a request on this context is charged at admission.
    static uint64_t request_bytes(const Groth16Prover &ctx);

    struct Stats {
        uint64_t batches = 0, proofs = 0, max_batch_seen = 0, peak_work_bytes = 0;
    };
    Stats stats() const;

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

    int device_;
    uint32_t max_batch_;
    uint64_t budget_;
    mutable std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<Request> pending_;
    std::map<uint64_t, Done> done_;
    uint64_t next_ticket_ = 1;
    uint64_t in_use_ = 0;
    bool stop_ = false;
    Stats stats_;
    std::vector<std::unique_ptr<BatchProver>> provers_;
    std::vector<std::thread> threads_;
};

}  // namespace nkr

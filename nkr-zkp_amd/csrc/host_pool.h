// Persistent host helper threads for the proof's host tail.
//
// A proof ends with a few hundred microseconds of host curve arithmetic that
// splits into independent parts.  Starting a thread per part costs more than
// some parts take, so the helpers live as long as their owner: arm() (at the
// start of a proof) lets them spin until a task comes or about 100 ms pass;
// idle helpers sleep on a condition variable.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "runtime.h"

namespace nkr {

class HostHelpers {
public:
    explicit HostHelpers(int count)
    {
        for (int i = 0; i < count; ++i) slots_.push_back(std::make_unique<Slot>());
        for (auto &s : slots_) {
            Slot *p = s.get();
            p->th = std::thread([p] { loop(*p); });
        }
    }
    ~HostHelpers()
    {
        for (auto &s : slots_) {
            {
                std::lock_guard<std::mutex> lk(s->m);
                s->state.store(kStop, std::memory_order_release);
            }
            s->cv.notify_one();
        }
        for (auto &s : slots_) s->th.join();
    }
    HostHelpers(const HostHelpers &) = delete;
    HostHelpers &operator=(const HostHelpers &) = delete;

    int size() const { return (int)slots_.size(); }

    // Helpers spin instead of sleeping until a task arrives (bounded).
    void arm()
    {
        const auto until = std::chrono::steady_clock::now() + std::chrono::milliseconds(100);
        for (auto &s : slots_) {
            {
                std::lock_guard<std::mutex> lk(s->m);
                s->spin_until = until;
                s->armed.store(true, std::memory_order_release);
            }
            s->cv.notify_one();
        }
    }

    // Hands fn to helper i (which must be free).
    void run(int i, std::function<void()> fn)
    {
        Slot &s = *slots_[i];
        s.fn = std::move(fn);
        {
            std::lock_guard<std::mutex> lk(s.m);
            s.state.store(kReady, std::memory_order_release);
        }
        s.cv.notify_one();
    }

    // Waits for helper i's task; rethrows what it threw.
    void wait(int i)
    {
        Slot &s = *slots_[i];
        while (s.state.load(std::memory_order_acquire) != kDone) std::this_thread::yield();
        s.state.store(kIdle, std::memory_order_release);
        if (s.err) {
            std::exception_ptr e = s.err;
            s.err = nullptr;
            std::rethrow_exception(e);
        }
    }

private:
    enum : int { kIdle = 0, kReady = 1, kDone = 2, kStop = 3 };
    struct Slot {
        std::thread th;
        std::atomic<int> state{kIdle};
        std::atomic<bool> armed{false};
        std::chrono::steady_clock::time_point spin_until{};
        std::function<void()> fn;
        std::exception_ptr err;
        std::mutex m;
        std::condition_variable cv;
    };

    static void loop(Slot &s)
    {
        for (;;) {
            int st = s.state.load(std::memory_order_acquire);
            if (st != kReady && st != kStop) {
                bool spin = s.armed.load(std::memory_order_acquire);
                if (spin) {
                    std::chrono::steady_clock::time_point until;
                    {
                        std::lock_guard<std::mutex> lk(s.m);
                        until = s.spin_until;
                    }
                    while ((st = s.state.load(std::memory_order_acquire)) != kReady && st != kStop &&
                           std::chrono::steady_clock::now() < until)
                        cpu_relax();
                    s.armed.store(false, std::memory_order_release);
                }
                if (st != kReady && st != kStop) {
                    std::unique_lock<std::mutex> lk(s.m);
                    s.cv.wait(lk, [&] {
                        const int x = s.state.load(std::memory_order_acquire);
                        return x == kReady || x == kStop || s.armed.load(std::memory_order_acquire);
                    });
                    continue;
                }
            }
            if (st == kStop) return;
            try {
                s.fn();
            } catch (...) {
                s.err = std::current_exception();
            }
            s.fn = nullptr;
            s.state.store(kDone, std::memory_order_release);
        }
    }

    std::vector<std::unique_ptr<Slot>> slots_;
};

}  // namespace nkr

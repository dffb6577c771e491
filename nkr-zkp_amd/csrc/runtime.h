// Host-side runtime helpers: HIP error checking that throws (never aborts),
// owned device buffers, and launch-shape checks.  Exceptions are converted to
// status codes at the C ABI boundary.
#pragma once

#include <hip/hip_runtime.h>

#include "error.h"

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace nkr {

#define NKR_HIP(call)                                                                         \
    do {                                                                                      \
        const hipError_t nkr_e_ = (call);                                                     \
        if (nkr_e_ != hipSuccess)                                                             \
            throw ::nkr::Error(std::string(#call) + " failed: " + hipGetErrorString(nkr_e_)); \
    } while (0)

#define NKR_LAUNCH_CHECK() NKR_HIP(hipGetLastError())

// Device allocation owned by one object; moves, never copies.
template <class T>
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t count) { allocate(count); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : ptr_(o.ptr_), count_(o.count_)
    {
        o.ptr_ = nullptr;
        o.count_ = 0;
    }
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept
    {
        if (this != &o) {
            release();
            ptr_ = o.ptr_;
            count_ = o.count_;
            o.ptr_ = nullptr;
            o.count_ = 0;
        }
        return *this;
    }

    void allocate(size_t count)
    {
        release();
        if (count == 0) return;
        void *p = nullptr;
        const hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess)
            throw Error("device allocation of " + std::to_string(count * sizeof(T)) +
                        " bytes failed: " + hipGetErrorString(e));
        ptr_ = static_cast<T *>(p);
        count_ = count;
    }
    // Grow-only reallocation; contents are not preserved.
    void ensure(size_t count)
    {
        if (count > count_) allocate(count);
    }
    void release()
    {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        count_ = 0;
    }

    T *get() const { return ptr_; }
    size_t size() const { return count_; }
    size_t bytes() const { return count_ * sizeof(T); }

private:
    T *ptr_ = nullptr;
    size_t count_ = 0;
};

// Pinned host staging buffer for asynchronous uploads.  `reuse` waits for the
// upload recorded by `uploaded` before the host writes the buffer again.
// A short pause for host spin loops.
inline void cpu_relax()
{
#if defined(__x86_64__) || defined(__i386__)
    for (int k = 0; k < 32; ++k) __builtin_ia32_pause();
#else
    std::this_thread::yield();
#endif
}

template <class T>
class PinnedStaging {
public:
    PinnedStaging() = default;
    ~PinnedStaging()
    {
        if (event_) (void)hipEventDestroy(event_);
        if (ptr_) (void)hipHostFree(ptr_);
    }
    PinnedStaging(const PinnedStaging &) = delete;
    PinnedStaging &operator=(const PinnedStaging &) = delete;

    // Room for `count` elements, writable now.
    T *reuse(size_t count)
    {
        if (event_) NKR_HIP(hipEventSynchronize(event_));
        return reuse_unsynced(count);
    }
    // reuse() without the wait, for a caller that knows the buffer's
    // transfers are done -- e.g. while capturing a graph after synchronising
    // its streams, where waiting on an event last recorded on a stream that
    // is now capturing is not permitted.
    T *reuse_unsynced(size_t count)
    {
        if (count > count_) {
            if (ptr_) NKR_HIP(hipHostFree(ptr_));
            ptr_ = nullptr;
            count_ = 0;
            void *p = nullptr;
            NKR_HIP(hipHostMalloc(&p, count * sizeof(T), hipHostMallocDefault));
            ptr_ = static_cast<T *>(p);
            count_ = count;
        }
        return ptr_;
    }
    // Marks the end of the transfers from or into the buffer on `stream`.
    void uploaded(hipStream_t stream)
    {
        if (!event_) NKR_HIP(hipEventCreateWithFlags(&event_, hipEventDisableTiming));
        NKR_HIP(hipEventRecord(event_, stream));
    }
    // Blocks until those transfers are done (a download is then readable).
    // Polls instead of sleeping in the runtime: the proof's host tail waits
    // here for a few kilobytes, and a sleeping thread's wake-up would add
    // tens of microseconds to every proof.
    void wait() const
    {
        if (!event_) return;
        for (;;) {
            const hipError_t e = hipEventQuery(event_);
            if (e == hipSuccess) return;
            if (e != hipErrorNotReady) NKR_HIP(e);
            cpu_relax();
        }
    }
    T *get() const { return ptr_; }

private:
    T *ptr_ = nullptr;
    size_t count_ = 0;
    hipEvent_t event_ = nullptr;
};

// Device-time accounting per stage with HIP events recorded on the stream
// that runs the work.  Disabled timers record nothing.
class StageTimer {
public:
    StageTimer() = default;
    StageTimer(const StageTimer &) = delete;
    StageTimer &operator=(const StageTimer &) = delete;
    ~StageTimer()
    {
        for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
    }
    void enable(bool on) { on_ = on; }
    bool enabled() const { return on_; }
    void begin(int stage, hipStream_t s)
    {
        if (!on_) return;
        recs_.push_back({stage, take(s), nullptr});
    }
    void end(int stage, hipStream_t s)
    {
        if (!on_) return;
        for (auto it = recs_.rbegin(); it != recs_.rend(); ++it)
            if (it->stage == stage && it->b == nullptr) {
                it->b = take(s);
                return;
            }
    }
    // Call after the stream has drained: adds elapsed ms per stage.
    void collect(double *acc_ms)
    {
        for (const Rec &r : recs_) {
            if (!r.b) continue;
            float ms = 0.f;
            NKR_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            acc_ms[r.stage] += ms;
        }
        recs_.clear();
        used_ = 0;
    }

private:
    struct Rec {
        int stage;
        hipEvent_t a, b;
    };
    hipEvent_t take(hipStream_t s)
    {
        if (used_ == pool_.size()) {
            hipEvent_t e;
            NKR_HIP(hipEventCreate(&e));
            pool_.push_back(e);
        }
        hipEvent_t e = pool_[used_++];
        NKR_HIP(hipEventRecord(e, s));
        return e;
    }
    bool on_ = false;
    std::vector<hipEvent_t> pool_;
    size_t used_ = 0;
    std::vector<Rec> recs_;
};

}  // namespace nkr

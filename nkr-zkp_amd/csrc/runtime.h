// Host-side runtime helpers: HIP error checking that throws (never aborts),
// owned device buffers, and launch-shape checks.  Exceptions are converted to
// status codes at the C ABI boundary.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

namespace nkr {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define NKR_HIP(call)                                                                         \
    do {                                                                                      \
        const hipError_t nkr_e_ = (call);                                                     \
        if (nkr_e_ != hipSuccess)                                                             \
            throw ::nkr::Error(std::string(#call) + " failed: " + hipGetErrorString(nkr_e_)); \
    } while (0)

#define NKR_REQUIRE(cond, msg)                                      \
    do {                                                            \
        if (!(cond)) throw ::nkr::Error(std::string(msg));          \
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

inline unsigned ceil_log2(uint64_t n)
{
    unsigned r = 0;
    while ((uint64_t(1) << r) < n) ++r;
    return r;
}

inline bool is_pow2(uint64_t n) { return n != 0 && (n & (n - 1)) == 0; }

inline uint32_t div_up(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }

}  // namespace nkr

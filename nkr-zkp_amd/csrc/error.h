// Errors and small integer helpers shared by device and host-only code.  No
// HIP dependency, so host tools and CPU unit tests can include it.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

namespace nkr {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define NKR_REQUIRE(cond, msg)                                      \
    do {                                                            \
        if (!(cond)) throw ::nkr::Error(std::string(msg));          \
    } while (0)

inline unsigned ceil_log2(uint64_t n)
{
    unsigned r = 0;
    while ((uint64_t(1) << r) < n) ++r;
    return r;
}

inline bool is_pow2(uint64_t n) { return n != 0 && (n & (n - 1)) == 0; }

// ceil(a / b); b must not be 0 (checked: a zero divisor is a sizing bug, and
// the integer division would otherwise end the process with SIGFPE).
inline uint32_t div_up(uint64_t a, uint64_t b)
{
    NKR_REQUIRE(b != 0, "internal sizing error: division by zero");
    return static_cast<uint32_t>((a + b - 1) / b);
}

}  // namespace nkr

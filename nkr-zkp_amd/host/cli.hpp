// Shared command-line driver of the nkr-g16-* programs: the option loop,
// usage errors (exit status 2) and the "[profile] <label> = <ms> ms" timing
// lines the reference's drivers print under --profile (the labels are the
// reference harness's contract, scripts/bench/protocols.py:87-105).
#pragma once

#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <filesystem>
#include <functional>
#include <iostream>
#include <string>
#include <vector>

namespace nkr::cli {

inline bool &profile_enabled()
{
    static bool on = false;
    return on;
}

// Prints the time spent in its lifetime when profiling is enabled.
class Scope {
public:
    explicit Scope(const char *label) : label_(label), t0_(std::chrono::steady_clock::now()) {}
    ~Scope()
    {
        if (!profile_enabled()) return;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
        std::cout << "[profile] " << label_ << " = " << ms << " ms\n";
    }
    Scope(const Scope &) = delete;
    Scope &operator=(const Scope &) = delete;

private:
    const char *label_;
    std::chrono::steady_clock::time_point t0_;
};

// Options every driver may accept.
enum Common : unsigned { kPath = 1, kBackend = 2, kSeed = 4, kProfile = 8 };

struct Args {
    std::filesystem::path directory = "app/data";
    std::string backend = "hip";
    bool seeded = false;
    uint64_t seed = 0;
    std::vector<std::string> positional;
};

// Parses argv: the common options named in `accept`, then the driver's own
// through `extra(flag, value)`, which returns how many arguments it took
// (1 for a switch, 2 for a flag with a value, 0 when it does not know the
// flag; value is null when the flag is last).  Any other argument starting
// with '-' prints `usage` and exits with status 2.
inline Args parse(int argc, char **argv, unsigned accept, const std::string &usage,
                  const std::function<int(const std::string &, const char *)> &extra = {})
{
    auto fail = [&] {
        std::cerr << "Usage: " << argv[0] << ' ' << usage << '\n';
        std::exit(2);
    };
    Args a;
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        const char *next = i + 1 < argc ? argv[i + 1] : nullptr;
        if ((accept & kPath) && arg == "--path" && next) {
            a.directory = argv[++i];
        } else if ((accept & kBackend) && arg == "--backend" && next) {
            a.backend = argv[++i];
        } else if ((accept & kSeed) && arg == "--seed" && next) {
            try {
                a.seed = std::stoull(argv[++i], nullptr, 0);
            } catch (const std::exception &) {
                fail();
            }
            a.seeded = true;
        } else if ((accept & kProfile) && arg == "--profile") {
            profile_enabled() = true;
        } else if (arg == "--help" || arg == "-h") {
            fail();
        } else if (int took = extra ? extra(arg, next) : 0) {
            i += took - 1;
        } else if (!arg.empty() && arg.front() == '-') {
            fail();
        } else {
            a.positional.push_back(arg);
        }
    }
    return a;
}

}  // namespace nkr::cli

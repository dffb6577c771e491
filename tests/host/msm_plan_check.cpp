// CPU check of the MSM sizing arithmetic (nkr-zkp_amd/csrc/msm_plan.h).
// stdin: one request per line,
//   plan n c_override target_log piece_lo piece_hi
//       -> "c W B max_pieces capacity levels" or "error <message>"
//   folds partials        -> fold levels for one bucket with that many partials
//   divup a b             -> ceil(a / b) or "error <message>"
// Errors are reported, never turned into a crash: a zero divisor must raise
// nkr::Error instead of ending the process with SIGFPE.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>

#include "msm_plan.h"

using namespace nkr;

int main()
{
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string op;
        in >> op;
        try {
            if (op == "plan") {
                uint64_t n, c, tl, lo, hi;
                in >> n >> c >> tl >> lo >> hi;
                MsmPlan p = MsmPlan::choose((uint32_t)n, (uint32_t)c);
                p.piece_lo = (uint32_t)lo;
                p.piece_hi = (uint32_t)hi;
                const uint64_t pieces = p.max_pieces((uint32_t)tl);
                uint32_t levels = 0;
                const uint64_t cap = msm_partial_capacity(p, pieces, &levels, nullptr);
                std::printf("%u %u %u %llu %llu %u\n", p.c, p.W, p.B, (unsigned long long)pieces,
                            (unsigned long long)cap, levels);
            } else if (op == "folds") {
                uint64_t k;
                in >> k;
                std::printf("%u\n", msm_fold_levels(k));
            } else if (op == "divup") {
                uint64_t a, b;
                in >> a >> b;
                std::printf("%u\n", div_up(a, b));
            } else {
                std::printf("error unknown request\n");
            }
        } catch (const Error &e) {
            std::printf("error %s\n", e.what());
        }
        std::fflush(stdout);
    }
    return 0;
}

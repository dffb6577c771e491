// The packing rule of the witness upload (upload.h), free of HIP so that the
// CPU tests compile it alone.  Elements are 32-byte Fr values (4 little-endian
// words, Montgomery form).
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>

namespace nkr {

constexpr uint32_t kWitnessGroup = 64;

// Packs the elements [begin, end) of host (begin a multiple of 64): for each
// group g of 64 they cover, masks[2g] marks the values other than 0 and
// `one`, masks[2g + 1] the values equal to `one`, and the marked values are
// copied to vals[offsets[g]], vals[offsets[g] + 1], ..., where offsets[g] =
// begin + the number of values the range packed before group g.  Returns the
// number of values packed.  host holds the elements from index host_first on
// (element e is host[e - host_first]), so several witnesses laid end to end
// share one index space.
inline uint32_t witness_pack(const void *host, uint32_t begin, uint32_t end, const uint64_t one[4], void *vals,
                             uint64_t *masks, uint32_t *offsets, uint32_t host_first = 0)
{
    const uint8_t *src = static_cast<const uint8_t *>(host);
    uint8_t *dst = static_cast<uint8_t *>(vals);
    uint32_t packed = 0;
    for (uint32_t g0 = begin; g0 < end; g0 += kWitnessGroup) {
        const uint32_t g = g0 / kWitnessGroup;
        const uint32_t g1 = std::min(end, g0 + kWitnessGroup);
        offsets[g] = begin + packed;
        uint64_t other = 0, ones = 0;
        for (uint32_t e = g0; e < g1; ++e) {
            uint64_t w[4];
            std::memcpy(w, src + (size_t)(e - host_first) * 32, sizeof(w));
            if ((w[0] | w[1] | w[2] | w[3]) == 0) continue;
            const uint64_t bit = 1ull << (e - g0);
            if (((w[0] ^ one[0]) | (w[1] ^ one[1]) | (w[2] ^ one[2]) | (w[3] ^ one[3])) == 0) {
                ones |= bit;
            } else {
                other |= bit;
                std::memcpy(dst + (size_t)(begin + packed) * 32, w, sizeof(w));
                ++packed;
            }
        }
        masks[2 * g] = other;
        masks[2 * g + 1] = ones;
    }
    return packed;
}

}  // namespace nkr

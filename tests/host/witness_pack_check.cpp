// CPU check of the witness packing rule (nkr-zkp_amd/csrc/witness_pack.h).
// argv: input file (n x 32 B elements), n, chunk (elements per packing
// range), output file.  The ranges [c * chunk, min(n, (c + 1) * chunk)) are
// packed one by one as the upload's workers pack them; the output holds the
// packed values (n x 32 B, unused slots zero), then masks (2 x u64 per group),
// offsets (u32 per group) and the packed count of each range (u32).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "witness_pack.h"

int main(int argc, char **argv)
{
    if (argc != 5) return 2;
    const uint32_t n = (uint32_t)std::strtoul(argv[2], nullptr, 10);
    const uint32_t chunk = (uint32_t)std::strtoul(argv[3], nullptr, 10);
    if (chunk == 0 || chunk % nkr::kWitnessGroup) return 2;
    std::vector<uint8_t> in((size_t)n * 32), vals((size_t)n * 32, 0);
    FILE *f = std::fopen(argv[1], "rb");
    if (!f || std::fread(in.data(), 1, in.size(), f) != in.size()) return 3;
    std::fclose(f);
    const uint32_t groups = (n + 63) / 64;
    std::vector<uint64_t> masks(2 * (size_t)groups, ~0ull);
    std::vector<uint32_t> offsets(groups, ~0u), counts;
    // Fr one in Montgomery form: 2^256 mod r
    const uint64_t one[4] = {0xac96341c4ffffffbull, 0x36fc76959f60cd29ull, 0x666ea36f7879462eull,
                             0x0e0a77c19a07df2full};
    for (uint32_t b = 0; b < n; b += chunk)
        counts.push_back(nkr::witness_pack(in.data(), b, b + chunk < n ? b + chunk : n, one, vals.data(),
                                           masks.data(), offsets.data()));
    FILE *o = std::fopen(argv[4], "wb");
    if (!o) return 3;
    std::fwrite(vals.data(), 1, vals.size(), o);
    std::fwrite(masks.data(), 8, masks.size(), o);
    std::fwrite(offsets.data(), 4, offsets.size(), o);
    std::fwrite(counts.data(), 4, counts.size(), o);
    std::fclose(o);
    return 0;
}

// MSM sizing: window choice, piece bounds and the partial-sum capacity of
// the fold levels.  Host arithmetic only (no HIP), so that a CPU unit test
// can check every bound the device buffers are sized from.
#pragma once

#include <algorithm>
#include <cstdint>

#include "error.h"

namespace nkr {

// Largest accumulation piece, fold fan-in and the completion limit.  Small
// MSMs use shorter pieces (down to kMsmPieceMin) so that the serial chain of
// one thread, not the amount of work, stops dominating their latency.
constexpr uint32_t kMsmPiece = 64;
constexpr uint32_t kMsmPieceMin = 4;
constexpr uint32_t kMsmFoldIn = 4;
constexpr uint32_t kMsmCompleteMax = 64;
constexpr uint32_t kMaxFoldLevels = 12;
// The piece length is chosen on the device once the number of non-zero
// digits E is known: the largest power of two <= E / 2^target_log,
// clamped to [piece_lo, piece_hi], so a schedule has about 2^target_log
// pieces (default 2^19: every resident accumulation thread busy).
constexpr uint32_t kMsmTargetPiecesLog = 19;
// Largest bucket count (c <= 16): the histograms and bucket scans hold B
// counts in one workgroup's LDS (128 KiB of gfx950's 160 KiB).
constexpr uint32_t kMsmMaxBuckets = 1u << 15;

// Fold levels (4:1 each) that bring `partials` partial sums of one bucket
// down to at most kMsmCompleteMax; 0 for an empty or short bucket.
inline uint32_t msm_fold_levels(uint64_t partials)
{
    uint32_t levels = 0;
    while (partials > kMsmCompleteMax) {
        partials = (partials + kMsmFoldIn - 1) / kMsmFoldIn;
        ++levels;
    }
    return levels;
}

struct MsmPlan {
    uint32_t n = 0;         // number of bases / scalars
    uint32_t c = 0;         // window width in bits
    uint32_t W = 0;         // windows (signed digits), c*W >= 255
    uint32_t B = 0;         // buckets 1..B, B = 2^(c-1)
    uint32_t piece_lo = 0;  // bounds of the accumulation piece length
    uint32_t piece_hi = 0;

    // ~3/4 log2 n bits (16 from 2^20 terms up), or c_override when non-zero.
    static MsmPlan choose(uint32_t n, uint32_t c_override = 0)
    {
        NKR_REQUIRE(n > 0, "MSM needs at least one term");
        MsmPlan p;
        p.n = n;
        uint32_t c = c_override;
        if (c == 0) {
            const unsigned lg = ceil_log2(n);
            // 16 bits from 2^20 terms up: 16 windows instead of 17, and its
            // 128 KiB bucket histogram still fits one workgroup's LDS.
            c = lg >= 20 ? 16u : std::max(4u, std::min(15u, (lg * 3 + 2) / 4));
            // Scalars have 254 bits; a top window with only a few of them sends
            // a whole window's entries to a handful of buckets (long serial
            // chains), so move to a neighbouring width whose top window keeps
            // at least 4 bits.
            if (top_bits(c) < 4) {
                if (c < 16 && top_bits(c + 1) >= 4)
                    c += 1;
                else if (c > 4 && top_bits(c - 1) >= 4)
                    c -= 1;
            }
        }
        NKR_REQUIRE(c >= 2 && c <= 16, "MSM window must be 2..16 bits");
        p.c = c;
        p.W = (255 + c - 1) / c;
        p.B = 1u << (c - 1);
        // The piece length follows the actual number of non-zero digits
        // (chosen on the device): sparse witnesses get short serial chains.
        p.piece_lo = kMsmPieceMin;
        p.piece_hi = kMsmPiece;
        NKR_REQUIRE((uint64_t)p.W * n < (1ull << 31), "MSM too large for 31-bit entry indices");
        return p;
    }
    // Bits of a 254-bit scalar left for the top window of width cc.
    static uint32_t top_bits(uint32_t cc) { return 254 - cc * ((255 + cc - 1) / cc - 1); }

    uint64_t max_entries() const { return (uint64_t)n * W; }
    // Upper bound of level-0 pieces for any scalars when the device targets
    // 2^target_log pieces: the piece p it picks is piece_hi, or
    // p > E / 2^(target_log + 1).  Every bucket may add one short piece.
    uint64_t max_pieces(uint32_t target_log) const
    {
        NKR_REQUIRE(piece_lo >= 1 && piece_lo <= piece_hi, "MSM piece bounds must satisfy 1 <= lo <= hi");
        NKR_REQUIRE(target_log < 40, "MSM target piece count out of range");
        const uint64_t by_lo = max_entries() / piece_lo;
        const uint64_t by_target = std::max<uint64_t>(2ull << target_log, max_entries() / piece_hi);
        return std::min(by_lo, by_target) + B + 1;
    }
    // The same plan with a fixed piece length.
    MsmPlan with_piece(uint32_t piece) const
    {
        NKR_REQUIRE(piece >= 1, "MSM piece length must be at least 1");
        MsmPlan p = *this;
        p.piece_lo = p.piece_hi = piece;
        return p;
    }
};

// Partial sums one job needs room for: the level-0 pieces, then each fold
// level's worst case (a level keeps at most a quarter of the previous one,
// plus one rounded-up group per bucket).  levels receives the fold levels
// the capacity covers; level_base[L] and level_cap[L] (if non-null,
// kMaxFoldLevels entries each) the offset and size of level L's output.
inline uint64_t msm_partial_capacity(const MsmPlan &p, uint64_t pieces, uint32_t *levels, uint64_t *level_base,
                                     uint64_t *level_cap = nullptr)
{
    uint64_t cap = pieces, cnt = pieces, base = pieces;
    uint32_t L = 0;
    while (cnt > kMsmCompleteMax) {
        NKR_REQUIRE(L < kMaxFoldLevels, "MSM needs too many fold levels");
        cnt = (cnt + kMsmFoldIn - 1) / kMsmFoldIn;
        cap = cap / kMsmFoldIn + p.B + 1;
        if (level_base) level_base[L] = base;
        if (level_cap) level_cap[L] = cap;
        base += cap;
        ++L;
    }
    for (uint32_t k = L; k < kMaxFoldLevels; ++k) {
        if (level_base) level_base[k] = base;
        if (level_cap) level_cap[k] = 0;
    }
    if (levels) *levels = L;
    return base;
}

}  // namespace nkr

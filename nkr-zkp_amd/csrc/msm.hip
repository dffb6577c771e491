// Pippenger MSM kernels for gfx950 (see msm.h for the design).
#include "msm.h"

#include "curve29.h"

#include <algorithm>
#include <array>
#include <cstring>
#include <mutex>
#include <type_traits>

namespace nkr {

namespace {

constexpr uint32_t kEntryMask = 0x7fffffffu;
// Sort workgroups: about one per CU, each histogramming a contiguous slice of
// at least kSortMinPerBlock scalars with up to kSortThreads threads.
constexpr uint32_t kSortBlocksTarget = 256;
constexpr uint32_t kSortMinPerBlock = 256;
constexpr uint32_t kSortThreads = 1024;
// Column scan tile: 64 buckets x kScanGroups row groups per workgroup.
constexpr uint32_t kScanGroups = 16;
// LDS per tree workgroup (two buffers of 2^h blocks): 64 KiB for G1 (two
// workgroups per CU); 32 KiB for G2, whose trees run beside the main
// stream's transforms and histograms and must leave them room on the CU.
template <class F>
constexpr size_t tree_lds_bytes()
{
    return std::is_same<F, fq2_t>::value ? 32 * 1024 : 64 * 1024;
}
// Completion lane classes (1, 2, 4, 8 lanes per bucket).
constexpr uint32_t kCompleteClasses = 4;

template <class F>
struct JobDev {
    const Affine<F> *bases;
    const uint32_t *sorted;
    const uint32_t *bucket_start;  // B + 1
    const uint32_t *piece_start;   // B + 1
    const uint32_t *piece_max;     // largest level-0 piece count of any bucket
    const uint32_t *piece_len;     // entries per level-0 piece (chosen on the device)
    partial_t<F> *pool;            // partial sums (level 0 region, then fold levels)
    uint32_t *reg_start;           // B: where bucket k's partials live in `pool`
    uint32_t *reg_count;           // B
    uint32_t *fold_start;          // B + 1
    partial_t<F> *sums;            // B: S_{k+1}
    partial_t<F> *tree[2];         // B each
    Xyzz<F> *planes;               // c (canonical)
    uint32_t B, c;
    uint32_t list_off;                    // this job's completion lists start at 4 * list_off
    uint32_t level_base[kMaxFoldLevels];  // where each fold level's partials start in `pool`
};

// Sum of two partial sums, and a partial sum as a canonical bit plane.
__device__ __forceinline__ Xyzz29<FqParams> partial_add(const Xyzz29<FqParams> &a, const Xyzz29<FqParams> &b)
{
    return xyzz29_add(a, b);
}
__device__ __forceinline__ Xyzz2_29<FqParams> partial_add(const Xyzz2_29<FqParams> &a, const Xyzz2_29<FqParams> &b)
{
    return xyzz2_29_add(a, b);
}
__device__ __forceinline__ Xyzz<fq_t> partial_plane(const Xyzz29<FqParams> &a) { return xyzz29_to(a); }
__device__ __forceinline__ Xyzz<fq2_t> partial_plane(const Xyzz2_29<FqParams> &a) { return xyzz2_29_to(a); }

// ------------------------------------------------------------ digit recoding
// Signed c-bit digits of an integer scalar (< 2^254): d_w in [-2^(c-1), 2^(c-1)],
// top window unsigned so that c*W >= 255 suffices.
template <class Fn>
__device__ __forceinline__ void for_each_digit(const fr_t &s, uint32_t c, uint32_t W, Fn fn)
{
    const uint32_t half = 1u << (c - 1);
    const uint32_t full = 1u << c;
    const uint32_t mask = full - 1u;
    uint32_t carry = 0;
    for (uint32_t w = 0; w < W; ++w) {
        const uint32_t off = w * c;
        const uint32_t word = off >> 5, sh = off & 31u;
        // selects instead of s.l[word]: a dynamically indexed register array
        // would live in scratch memory
        uint64_t lo = 0, hi = 0;
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) {
            lo = k == word ? s.l[k] : lo;
            hi = k == word + 1 ? s.l[k] : hi;
        }
        uint32_t v = (uint32_t)(((hi << 32) | lo) >> sh) & mask;
        v += carry;
        int32_t d;
        if (w + 1 < W && v >= half) {
            d = (int32_t)v - (int32_t)full;
            carry = 1;
        } else {
            d = (int32_t)v;
            carry = 0;
        }
        if (d != 0) {
            // Canonical scalars (< r) never exceed 2^(c-1); the clamp keeps a
            // malformed input inside the bucket arrays instead of faulting.
            const int32_t lim = (int32_t)half;
            fn(w, d > lim ? lim : (d < -lim ? -lim : d));
        }
    }
}

// Histogram of the bucket magnitudes of scalars [b * per_block, ...) in LDS,
// written as row b of the (blocks x B) count matrix.
__device__ __forceinline__ void hist_block(uint32_t *hist, const fr_t *__restrict__ scalars, uint32_t n, uint32_t c,
                                          uint32_t W, uint32_t B, uint32_t per_block, uint32_t b,
                                          uint32_t *__restrict__ block_count)
{
    for (uint32_t k = threadIdx.x; k < B; k += blockDim.x) hist[k] = 0;
    __syncthreads();
    const uint32_t beg = b * per_block;
    const uint32_t end = min(n, beg + per_block);
    for (uint32_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
        const fr_t s = fp_from_mont(scalars[i]);
        for_each_digit(s, c, W, [&](uint32_t, int32_t d) {
            const uint32_t m = (uint32_t)(d < 0 ? -d : d);
            atomicAdd(&hist[m - 1], 1u);
        });
    }
    __syncthreads();
    uint32_t *row = block_count + (size_t)b * B;
    for (uint32_t k = threadIdx.x; k < B; k += blockDim.x) row[k] = hist[k];
}

// Column scan of the count matrix for buckets [k0, k0 + 64): row b of
// block_count becomes the offset of workgroup b's entries inside each
// bucket (the sum of the rows above it) and count[k] the column total.
// 64 x kScanGroups threads: each row group sums its rows, the groups'
// sums are scanned in LDS, then each group rewrites its rows as offsets.
__device__ __forceinline__ void colscan_tile(uint32_t *__restrict__ block_count, uint32_t blocks, uint32_t B,
                                            uint32_t k0, uint32_t *__restrict__ count)
{
    __shared__ uint32_t part[kScanGroups][64];
    const uint32_t kk = threadIdx.x & 63u, g = threadIdx.x >> 6;
    const uint32_t k = k0 + kk;
    const uint32_t rows = (blocks + kScanGroups - 1) / kScanGroups;
    const uint32_t r0 = min(blocks, g * rows), r1 = min(blocks, r0 + rows);
    uint32_t sum = 0;
    if (k < B)
        for (uint32_t r = r0; r < r1; ++r) sum += block_count[(size_t)r * B + k];
    part[g][kk] = sum;
    __syncthreads();
    uint32_t run = 0;
    for (uint32_t h = 0; h < g; ++h) run += part[h][kk];
    if (k < B) {
        for (uint32_t r = r0; r < r1; ++r) {
            uint32_t *p = block_count + (size_t)r * B + k;
            const uint32_t v = *p;
            *p = run;
            run += v;
        }
        if (g == kScanGroups - 1) count[k] = run;
    }
}

__device__ __forceinline__ void scatter_block(uint32_t *cursor, const fr_t *__restrict__ scalars, uint32_t n,
                                             uint32_t c, uint32_t W, uint32_t B, uint32_t per_block, uint32_t b,
                                             const uint32_t *__restrict__ bucket_start,
                                             const uint32_t *__restrict__ block_off, uint32_t *__restrict__ sorted)
{
    const uint32_t *row = block_off + (size_t)b * B;
    for (uint32_t k = threadIdx.x; k < B; k += blockDim.x) cursor[k] = bucket_start[k] + row[k];
    __syncthreads();
    const uint32_t beg = b * per_block;
    const uint32_t end = min(n, beg + per_block);
    for (uint32_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
        const fr_t s = fp_from_mont(scalars[i]);
        for_each_digit(s, c, W, [&](uint32_t w, int32_t d) {
            const uint32_t m = (uint32_t)(d < 0 ? -d : d);
            const uint32_t pos = atomicAdd(&cursor[m - 1], 1u);
            sorted[pos] = (d < 0 ? 0x80000000u : 0u) | (w * n + i);
        });
    }
}

__global__ __launch_bounds__(kSortThreads) void k_hist(const fr_t *__restrict__ scalars, uint32_t n, uint32_t c,
                                                       uint32_t W, uint32_t B, uint32_t per_block,
                                                       uint32_t *__restrict__ block_count)
{
    extern __shared__ uint32_t lds[];
    hist_block(lds, scalars, n, c, W, B, per_block, blockIdx.x, block_count);
}

__global__ __launch_bounds__(64 * kScanGroups) void k_colscan(uint32_t *__restrict__ block_count, uint32_t blocks,
                                                              uint32_t B, uint32_t *__restrict__ count)
{
    colscan_tile(block_count, blocks, B, blockIdx.x * 64, count);
}

__global__ __launch_bounds__(kSortThreads) void k_scatter(const fr_t *__restrict__ scalars, uint32_t n, uint32_t c,
                                                          uint32_t W, uint32_t B, uint32_t per_block,
                                                          const uint32_t *__restrict__ bucket_start,
                                                          const uint32_t *__restrict__ block_off,
                                                          uint32_t *__restrict__ sorted)
{
    extern __shared__ uint32_t lds[];
    scatter_block(lds, scalars, n, c, W, B, per_block, blockIdx.x, bucket_start, block_off, sorted);
}


// Exclusive prefix of `len` (<= kMsmMaxBuckets) values into out[0..len]
// within one workgroup; returns the total.  The values are staged in LDS
// with coalesced loads, each thread scans a run of `per` of them, and the
// runs' sums are scanned across the workgroup.  Runs sit at an odd LDS
// stride, so the threads of a wave walking their runs hit distinct banks.
//   mode 0: v[k] = src[k]
//   mode 2: v[k] = ceil(src[k] / T)
//   mode 3: v[k] = src[k] > lim ? ceil(src[k] / T) : 0
__device__ uint32_t block_scan(const uint32_t *__restrict__ src, uint32_t len, uint32_t mode, uint32_t T,
                               uint32_t lim, uint32_t *__restrict__ out, uint32_t *__restrict__ max_out = nullptr)
{
    __shared__ uint32_t vals[kMsmMaxBuckets + 1024];
    __shared__ uint32_t part[1024];
    __shared__ uint32_t pmax;
    const uint32_t per = (len + blockDim.x - 1) / blockDim.x;
    const uint32_t stride = per | 1u;
    auto at = [&](uint32_t k) -> uint32_t { return (k / per) * stride + k % per; };
    if (threadIdx.x == 0) pmax = 0;
    for (uint32_t k = threadIdx.x; k < len; k += blockDim.x) {
        const uint32_t x = src[k];
        vals[at(k)] = mode == 0 ? x : (mode == 2 || x > lim) ? (x + T - 1) / T : 0u;
    }
    __syncthreads();
    const uint32_t beg = min(len, threadIdx.x * per);
    const uint32_t end = min(len, beg + per);
    uint32_t *run_vals = vals + threadIdx.x * stride;
    uint32_t sum = 0, mx = 0;
    for (uint32_t k = 0; k < end - beg; ++k) {
        const uint32_t v = run_vals[k];
        sum += v;
        mx = max(mx, v);
    }
    if (max_out && mx) atomicMax(&pmax, mx);
    // the runs' sums scanned across the workgroup: within each wave by
    // shuffles, then the waves' totals by the first wave (two barriers)
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    uint32_t incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, off);
        if (lane >= (uint32_t)off) incl += y;
    }
    if (lane == 63) part[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < waves ? part[lane] : 0u;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)w, off);
            if (lane >= (uint32_t)off) w += y;
        }
        if (lane < waves) part[lane] = w;
    }
    __syncthreads();
    uint32_t run = (wave ? part[wave - 1] : 0u) + incl - sum;
    for (uint32_t k = 0; k < end - beg; ++k) {
        const uint32_t v = run_vals[k];
        run_vals[k] = run;
        run += v;
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < len; k += blockDim.x) out[k] = vals[at(k)];
    const uint32_t total = part[waves - 1];
    if (threadIdx.x == 0) {
        out[len] = total;
        if (max_out) *max_out = pmax;
    }
    __syncthreads();
    return total;
}

// Bucket starts (prefix of counts), the piece length for the number of
// entries, and level-0 piece starts (prefix of ceil(count / piece)):
// bucket_start[B + 1] = pieces, [B + 2] = largest piece count of a bucket,
// [B + 3] = piece length.
__device__ void schedule_scan(const uint32_t *__restrict__ count, uint32_t B, uint32_t piece_lo, uint32_t piece_hi,
                              uint32_t target_log, uint32_t *__restrict__ bucket_start,
                              uint32_t *__restrict__ piece_start)
{
    const uint32_t entries = block_scan(count, B, 0, 1, 0, bucket_start);
    const uint32_t q = entries >> target_log;
    // Longer pieces also while an average bucket would be left with more
    // than kMsmCompleteMax partials: small MSMs (few entries against the
    // piece target) then complete without fold levels.  Longer pieces only
    // lower the piece count, so the host's max_pieces bound still holds.
    uint32_t piece = piece_lo;
    while (piece < piece_hi && (2 * piece <= q || (uint64_t)piece * B * kMsmCompleteMax < entries)) piece <<= 1;
    const uint32_t pieces = block_scan(count, B, 2, piece, 0, piece_start, bucket_start + B + 2);
    if (threadIdx.x == 0) {
        bucket_start[B + 1] = pieces;
        bucket_start[B + 3] = piece;
    }
}

// One schedule's bucket starts and piece split; with alt_lo > 0 also the
// second split (pieces of at least alt_lo entries) into alt_piece_start and
// the counters bucket_start[B + 4 .. B + 8).
__global__ __launch_bounds__(1024) void k_schedule_scan(const uint32_t *__restrict__ count, uint32_t B,
                                                        uint32_t piece_lo, uint32_t piece_hi, uint32_t target_log,
                                                        uint32_t *__restrict__ bucket_start,
                                                        uint32_t *__restrict__ piece_start, uint32_t alt_lo,
                                                        uint32_t *__restrict__ alt_piece_start)
{
    schedule_scan(count, B, piece_lo, piece_hi, target_log, bucket_start, piece_start);
    if (alt_lo == 0) return;
    const uint32_t entries = bucket_start[B];
    const uint32_t q = entries >> target_log;
    const uint32_t hi = max(alt_lo, piece_hi);
    uint32_t piece = alt_lo;
    while (piece < hi && 2 * piece <= q) piece <<= 1;
    uint32_t *ctr = bucket_start + B + 4;
    const uint32_t pieces = block_scan(count, B, 2, piece, 0, alt_piece_start, ctr + 2);
    if (threadIdx.x == 0) {
        ctr[0] = entries;
        ctr[1] = pieces;
        ctr[3] = piece;
    }
}

// Last index k in [0, len) with starts[k] <= p (starts[0] <= p < starts[len]).
__device__ __forceinline__ uint32_t find_segment(const uint32_t *starts, uint32_t len, uint32_t p)
{
    uint32_t lo = 0, hi = len;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (starts[mid] <= p)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// One schedule of a batched sort.
struct SortJob {
    const fr_t *scalars;
    uint32_t n, c, W, B, blocks, per_block, piece_lo, piece_hi, target_log;
    uint32_t *count, *block_count, *bucket_start, *piece_start, *sorted;
};

// Workgroup blockIdx.x is sort block (blockIdx.x - block_off[j]) of job j.
__global__ __launch_bounds__(kSortThreads) void k_hist_batch(const SortJob *__restrict__ jobs,
                                                             const uint32_t *__restrict__ block_off, uint32_t J)
{
    extern __shared__ uint32_t lds[];
    const uint32_t j = find_segment(block_off, J, blockIdx.x);
    const SortJob &jb = jobs[j];
    hist_block(lds, jb.scalars, jb.n, jb.c, jb.W, jb.B, jb.per_block, blockIdx.x - block_off[j], jb.block_count);
}

__global__ __launch_bounds__(64 * kScanGroups) void k_colscan_batch(const SortJob *__restrict__ jobs)
{
    const SortJob &jb = jobs[blockIdx.y];
    if (blockIdx.x * 64 >= jb.B) return;
    colscan_tile(jb.block_count, jb.blocks, jb.B, blockIdx.x * 64, jb.count);
}

__global__ __launch_bounds__(1024) void k_schedule_scan_batch(const SortJob *__restrict__ jobs)
{
    const SortJob &jb = jobs[blockIdx.x];
    schedule_scan(jb.count, jb.B, jb.piece_lo, jb.piece_hi, jb.target_log, jb.bucket_start, jb.piece_start);
}

// out[4j + i] = schedule j's counter i (entries, pieces, largest piece count, piece length).
__global__ __launch_bounds__(256) void k_gather_totals(const SortJob *__restrict__ jobs, uint32_t J,
                                                       uint32_t *__restrict__ out)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 4 * J) return;
    const SortJob &jb = jobs[g >> 2];
    out[g] = jb.bucket_start[jb.B + (g & 3u)];
}

__global__ __launch_bounds__(kSortThreads) void k_scatter_batch(const SortJob *__restrict__ jobs,
                                                                const uint32_t *__restrict__ block_off, uint32_t J)
{
    extern __shared__ uint32_t lds[];
    const uint32_t j = find_segment(block_off, J, blockIdx.x);
    const SortJob &jb = jobs[j];
    scatter_block(lds, jb.scalars, jb.n, jb.c, jb.W, jb.B, jb.per_block, blockIdx.x - block_off[j],
                  jb.bucket_start, jb.block_count, jb.sorted);
}

template <class F>
__global__ __launch_bounds__(256) void k_regions_init(const JobDev<F> *__restrict__ jobs,
                                                      uint32_t *__restrict__ class_count)
{
    const JobDev<F> &jb = jobs[blockIdx.y];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < kCompleteClasses) class_count[kCompleteClasses * blockIdx.y + k] = 0;
    if (k >= jb.B) return;
    jb.reg_start[k] = jb.piece_start[k];
    jb.reg_count[k] = jb.piece_start[k + 1] - jb.piece_start[k];
}

// piece_base[j] = level-0 pieces of jobs before j (as scheduled on the
// device), piece_base[J] = all of them.  One workgroup.
template <class F>
__global__ __launch_bounds__(1024) void k_piece_scan(const JobDev<F> *__restrict__ jobs, uint32_t J,
                                                     uint32_t *__restrict__ piece_base)
{
    __shared__ uint32_t part[1024];
    const uint32_t per = (J + blockDim.x - 1) / blockDim.x;
    const uint32_t beg = min(J, threadIdx.x * per), end = min(J, beg + per);
    uint32_t sum = 0;
    for (uint32_t j = beg; j < end; ++j) sum += jobs[j].piece_start[jobs[j].B];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
        const uint32_t t = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum;
    for (uint32_t j = beg; j < end; ++j) {
        piece_base[j] = run;
        run += jobs[j].piece_start[jobs[j].B];
    }
    if (threadIdx.x == blockDim.x - 1) piece_base[J] = part[threadIdx.x];
}

// The G2 accumulator of one lane with X and ZZ in registers and Y and ZZZ in
// LDS: each of Y and ZZZ is used twice per addition, far apart, and holding
// them (36 registers) across the products would push the addition's working
// set past the 256 registers a lane has at two waves per SIMD.  LDS holds
// limb i of a coordinate's 18 at [i * kAccThreads + lane], so a wave's access
// to one limb is 64 consecutive words (no bank conflicts).  The accesses are
// volatile so that the compiler neither merges the two reads of a coordinate
// nor keeps it in registers between them.
// The pointers carry the LDS address space so that the accesses compile to
// ds_read/ds_write rather than flat instructions.
constexpr uint32_t kAccThreads = 256;
typedef __attribute__((address_space(3))) volatile uint32_t lds_word_t;

template <class P>
struct Xyzz2_29Lds {
    Fq2_29<P> x, zz;
    lds_word_t *y, *zzz;  // this lane's column of each coordinate's 18 x kAccThreads words

    __device__ static Fq2_29<P> load(const lds_word_t *s)
    {
        Fq2_29<P> r;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            r.c0.l[i] = s[i * kAccThreads];
            r.c1.l[i] = s[(9 + i) * kAccThreads];
        }
        return r;
    }
    __device__ static void store(lds_word_t *s, const Fq2_29<P> &v)
    {
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            s[i * kAccThreads] = v.c0.l[i];
            s[(9 + i) * kAccThreads] = v.c1.l[i];
        }
    }
    __device__ bool is_inf() const { return zz.is_exact_zero(); }
    __device__ Fq2_29<P> X() const { return x; }
    __device__ Fq2_29<P> Y() const { return load(y); }
    __device__ Fq2_29<P> ZZ() const { return zz; }
    __device__ Fq2_29<P> ZZZ() const { return load(zzz); }
    __device__ void set(const Fq2_29<P> &X, const Fq2_29<P> &Y, const Fq2_29<P> &ZZ, const Fq2_29<P> &ZZZ)
    {
        x = X;
        zz = ZZ;
        store(y, Y);
        store(zzz, ZZZ);
    }
    __device__ void set(const Xyzz2_29<P> &q) { set(q.X, q.Y, q.ZZ, q.ZZZ); }
    __device__ Xyzz2_29<P> get() const { return {x, load(y), zz, load(zzz)}; }
};

// Level 0 over all jobs: a thread sums one piece of signed bases of one
// bucket into one partial (the loop strides by the grid, so any grid size
// covers every piece).
template <class F>
__device__ __forceinline__ void accumulate_body(const JobDev<F> *__restrict__ jobs,
                                                const uint32_t *__restrict__ piece_base, uint32_t J)
{
    const uint32_t total = piece_base[J];
#pragma unroll 1
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
        const uint32_t j = find_segment(piece_base, J, g);
        const JobDev<F> &jb = jobs[j];
        const uint32_t p = g - piece_base[j];
        const uint32_t k = find_segment(jb.piece_start, jb.B, p);
        const uint32_t q = p - jb.piece_start[k];
        const uint32_t piece = *jb.piece_len;
        const uint32_t beg = jb.bucket_start[k] + q * piece;
        const uint32_t end = min(beg + piece, jb.bucket_start[k + 1]);
        const Affine<F> *__restrict__ bases = jb.bases;
        const uint32_t *__restrict__ sorted = jb.sorted;
        partial_t<F> acc = partial_t<F>::infinity();
        if constexpr (std::is_same<F, fq2_t>::value) {
            // G2: the accumulator in the reduced-radix form as for G1
            // (curve29.h, Fq2 bounds), stored in it; Y and ZZZ in LDS
            // between their uses; no prefetch.
            __shared__ uint32_t side[36 * kAccThreads];
            Xyzz2_29Lds<FqParams> acc29{Fq2_29<FqParams>{}, Fq2_29<FqParams>{},
                                        (lds_word_t *)(side + threadIdx.x),
                                        (lds_word_t *)(side + 18 * kAccThreads + threadIdx.x)};
            acc29.set(Xyzz2_29<FqParams>::infinity());
#pragma unroll 1
            for (uint32_t e = beg; e < end; ++e) {
                const uint32_t v = sorted[e];
                const Affine<F> P = bases[v & kEntryMask];
                if (!affine_is_inf(P)) {
                    const F y = (v >> 31) ? FieldOps<F>::neg(P.y) : P.y;
                    xyzz2_29_madd<FqParams>(acc29, fq2_29_from<FqParams>(P.x), fq2_29_from<FqParams>(y));
                }
            }
            acc = acc29.get();
        } else if (beg < end) {
            // G1: the accumulator stays in the reduced-radix form
            // (curve29.h: one v_mad_u64_u32 per product term, no carries)
            // for the whole piece and is stored in it; bases are converted
            // as they arrive.  The next sorted index is
            // fetched one entry ahead; the base itself is not (its 16
            // registers cost more occupancy than its latency costs three
            // resident waves, measured on MI355X).
            Xyzz29<FqParams> acc29 = Xyzz29<FqParams>::infinity();
            uint32_t v = sorted[beg];
#pragma unroll 1
            for (uint32_t e = beg; e < end; ++e) {
                const Affine<F> P = bases[v & kEntryMask];
                const uint32_t vn = e + 1 < end ? sorted[e + 1] : 0u;
                if (!affine_is_inf(P)) {
                    const F y = (v >> 31) ? FieldOps<F>::neg(P.y) : P.y;
                    acc29 = xyzz29_add_affine(acc29, fp29_from_mont(P.x), fp29_from_mont(y));
                }
                v = vn;
            }
            acc = acc29;
        }
        jb.pool[p] = acc;
    }
}

template <class F>
__global__ __launch_bounds__(256) void k_accumulate(const JobDev<F> *__restrict__ jobs,
                                                    const uint32_t *__restrict__ piece_base, uint32_t J)
{
    accumulate_body<F>(jobs, piece_base, J);
}

// The G2 addition's working set exceeds 256 registers, which would leave one
// wave per SIMD; held to two waves, its spills cost less than the second wave
// gains (G2 accumulation of the dense 2^20 Poseidon-style proof 10.3 -> 8.5
// ms on MI355X), and with the accumulator's Y and ZZZ in LDS (36 KiB per
// workgroup) the spills fall from 280 to 80 bytes per lane.
template <>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_accumulate<fq2_t>(
    const JobDev<fq2_t> *__restrict__ jobs, const uint32_t *__restrict__ piece_base, uint32_t J)
{
    accumulate_body<fq2_t>(jobs, piece_base, J);
}

// Fold level L: buckets with more than kMsmCompleteMax partials get
// ceil(count / kMsmFoldIn) new partials written at pool[level_base[L] ...].
template <class F>
// 256 threads (one wave per SIMD): on the side stream this launch runs beside
// a resident-sized accumulation on the main stream, and a smaller workgroup
// needs less free room on a CU to start.
__global__ __launch_bounds__(256) void k_fold_scan(const JobDev<F> *__restrict__ jobs)
{
    const JobDev<F> &jb = jobs[blockIdx.x];
    if (*jb.piece_max <= kMsmCompleteMax) return;  // no skewed bucket: nothing to fold
    block_scan(jb.reg_count, jb.B, 3, kMsmFoldIn, kMsmCompleteMax, jb.fold_start);
}

template <class F>
__global__ __launch_bounds__(256) void k_fold(const JobDev<F> *__restrict__ jobs, uint32_t L)
{
    const JobDev<F> &jb = jobs[blockIdx.y];
    if (*jb.piece_max <= kMsmCompleteMax) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= jb.fold_start[jb.B]) return;
    const uint32_t k = find_segment(jb.fold_start, jb.B, p);
    const uint32_t q = p - jb.fold_start[k];
    const uint32_t beg = jb.reg_start[k] + q * kMsmFoldIn;
    const uint32_t end = min(beg + kMsmFoldIn, jb.reg_start[k] + jb.reg_count[k]);
    partial_t<F> acc = jb.pool[beg];
    for (uint32_t e = beg + 1; e < end; ++e) acc = partial_add(acc, jb.pool[e]);
    jb.pool[jb.level_base[L] + p] = acc;
}

template <class F>
__global__ __launch_bounds__(256) void k_fold_retarget(const JobDev<F> *__restrict__ jobs, uint32_t L)
{
    const JobDev<F> &jb = jobs[blockIdx.y];
    if (*jb.piece_max <= kMsmCompleteMax) return;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= jb.B) return;
    if (jb.reg_count[k] > kMsmCompleteMax) {
        jb.reg_start[k] = jb.level_base[L] + jb.fold_start[k];
        jb.reg_count[k] = jb.fold_start[k + 1] - jb.fold_start[k];
    }
}

template <class T>
__device__ __forceinline__ T shfl_xor_point(const T &p, int mask)
{
    T r;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(&p);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&r);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = (uint32_t)__shfl_xor((int)src[i], mask, 16);
    return r;
}

// Completion classes: a bucket with m partials (m <= kMsmCompleteMax = 64
// after folding) is summed by L = nextpow2(ceil(m / 8)) lanes (1, 2, 4 or 8:
// at most 8 serial additions, then log2 L shuffle levels), so light buckets
// do not drag idle lanes through a shuffle tree sized for the heaviest one.
// Class c = log2 L; segment 4j + c lists job j's buckets of class c.
__device__ __forceinline__ uint32_t complete_class(uint32_t m, uint32_t max_class)
{
    uint32_t c = 0;
    while (c < max_class && (8u << c) < m) ++c;
    return c;
}

// Buckets with no partial get S = infinity here; the others are appended to
// their class list (one atomic per wave and class: a ballot ranks the lanes).
template <class F>
__global__ __launch_bounds__(256) void k_classify(const JobDev<F> *__restrict__ jobs, uint32_t max_class,
                                                  uint32_t *__restrict__ class_count,
                                                  uint32_t *__restrict__ class_list)
{
    const uint32_t j = blockIdx.y;
    const JobDev<F> &jb = jobs[j];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = k < jb.B;
    const uint32_t m = in ? jb.reg_count[k] : 0u;
    if (in && m == 0) jb.sums[k] = partial_t<F>::infinity();
    const uint32_t c = complete_class(m, max_class);
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t *list = class_list + (size_t)jb.list_off * kCompleteClasses;
    for (uint32_t cc = 0; cc <= max_class; ++cc) {
        const bool mine = in && m > 0 && c == cc;
        const uint64_t mask = __ballot(mine);
        if (mask == 0) continue;
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
        uint32_t first = 0;
        if (lane == leader) first = atomicAdd(&class_count[kCompleteClasses * j + cc], (uint32_t)__popcll(mask));
        first = (uint32_t)__shfl((int)first, (int)leader);
        if (mine) list[(size_t)cc * jb.B + first + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = k;
    }
}

// Lane offsets of the completion segments: segment s = 4j + c takes
// count * 2^c lanes, padded to whole waves (every wave completes buckets of
// one segment, in aligned groups of 2^c lanes).  One workgroup.
__global__ __launch_bounds__(1024) void k_complete_scan(const uint32_t *__restrict__ class_count, uint32_t S,
                                                        uint32_t *__restrict__ base)
{
    __shared__ uint32_t part[1024];
    auto span = [&](uint32_t s) -> uint32_t {
        return ((class_count[s] << (s % kCompleteClasses)) + 63u) & ~63u;
    };
    const uint32_t per = (S + blockDim.x - 1) / blockDim.x;
    const uint32_t beg = min(S, threadIdx.x * per), end = min(S, beg + per);
    uint32_t sum = 0;
    for (uint32_t s = beg; s < end; ++s) sum += span(s);
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
        const uint32_t t = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum;
    for (uint32_t s = beg; s < end; ++s) {
        base[s] = run;
        run += span(s);
    }
    if (threadIdx.x == blockDim.x - 1) base[S] = part[threadIdx.x];
}

// Bucket completion: the 2^c lanes of a class-c bucket stride over its
// partials, then a shuffle tree of c levels; grid-stride over all segments.
template <class F>
__global__ __launch_bounds__(256) void k_complete(const JobDev<F> *__restrict__ jobs,
                                                  const uint32_t *__restrict__ class_count,
                                                  const uint32_t *__restrict__ class_list,
                                                  const uint32_t *__restrict__ base, uint32_t S)
{
    const uint32_t total = base[S];
#pragma unroll 1
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
        const uint32_t s = find_segment(base, S, g);
        const uint32_t j = s / kCompleteClasses, c = s % kCompleteClasses;
        const JobDev<F> &jb = jobs[j];
        const uint32_t local = g - base[s];
        const uint32_t idx = local >> c, lane = local & ((1u << c) - 1u);
        const bool live = idx < class_count[s];
        const uint32_t k =
            live ? class_list[(size_t)jb.list_off * kCompleteClasses + (size_t)c * jb.B + idx] : 0u;
        partial_t<F> acc = partial_t<F>::infinity();
        if (live) {
            const uint32_t start = jb.reg_start[k];
            const uint32_t cnt = jb.reg_count[k];
            for (uint32_t i = lane; i < cnt; i += (1u << c)) acc = partial_add(acc, jb.pool[start + i]);
        }
        for (uint32_t off = (1u << c) >> 1; off >= 1; off >>= 1)
            acc = partial_add(acc, shfl_xor_point(acc, (int)off));
        if (live && lane == 0) jb.sums[k] = acc;
    }
}

// Bit-plane tree.  A level-t block covers buckets [b*2^t, (b+1)*2^t) and
// holds P_0..P_{t-1} (the sums of its buckets with that bit set) and T (its
// total); level 0 is S_m itself (S_0 = infinity).  Merging blocks 2b and
// 2b+1 into level t+1: P_e = P_e(2b) + P_e(2b+1) for e < t, P_t = T(2b+1),
// T = T(2b) + T(2b+1).  At the last level (t+1 = c-1) the merged P_e are the
// bit planes U_e, and U_{c-1} = S_B (bucket B = 2^(c-1) alone has bit c-1).
//
// One launch runs up to h levels from level t0: a workgroup loads 2^h
// level-t0 blocks of one job into LDS and merges them level by level with a
// barrier in between (one addition of latency per level, no launch or
// global round trip), then stores the merged block to the other global
// tree buffer (`src` names the one it reads).  Jobs whose tree ends earlier
// run fewer levels; jobs already done return.
template <class F>
__global__ __launch_bounds__(256) void k_tree_lds(const JobDev<F> *__restrict__ jobs, uint32_t t0, uint32_t h,
                                                  uint32_t src)
{
    extern __shared__ uint8_t tree_lds[];
    const JobDev<F> &jb = jobs[blockIdx.y];
    if (t0 + 1 > jb.c - 1) return;
    const uint32_t lv = min(h, jb.c - 1 - t0);
    if (blockIdx.x >= (jb.B >> (t0 + lv))) return;
    const uint32_t cap = (1u << h) * (t0 + 1);
    partial_t<F> *buf[2] = {reinterpret_cast<partial_t<F> *>(tree_lds),
                            reinterpret_cast<partial_t<F> *>(tree_lds) + cap};

    const uint32_t nin = (1u << lv) * (t0 + 1);
    const uint32_t blk0 = blockIdx.x << lv;
    for (uint32_t idx = threadIdx.x; idx < nin; idx += blockDim.x) {
        const uint32_t blk = blk0 + idx / (t0 + 1), ent = idx % (t0 + 1);
        if (t0 == 0)
            buf[0][idx] = blk == 0 ? partial_t<F>::infinity() : jb.sums[blk - 1];
        else
            buf[0][idx] = jb.tree[src][(size_t)blk * (t0 + 1) + ent];
    }
    __syncthreads();
    for (uint32_t s = 0; s < lv; ++s) {
        const uint32_t t = t0 + s, E = t + 2;
        const bool last = (t + 1 == jb.c - 1);
        const partial_t<F> *in = buf[s & 1];
        partial_t<F> *out = buf[(s + 1) & 1];
        const uint32_t nout = (1u << (lv - s - 1)) * E;
        for (uint32_t idx = threadIdx.x; idx < nout; idx += blockDim.x) {
            const uint32_t b = idx / E, e = idx % E;
            const partial_t<F> *l = in + (size_t)(2 * b) * (t + 1);
            const partial_t<F> *r = l + (t + 1);
            // one addition site for every lane: the P_e and T merges do not
            // diverge into two serial additions
            const uint32_t q = e < t ? e : t;
            partial_t<F> o = partial_add(l[q], r[q]);
            if (e == t)
                o = r[t];
            else if (e > t && last)
                o = jb.sums[jb.B - 1];
            if (last)
                jb.planes[e] = partial_plane(o);
            else
                out[idx] = o;
        }
        __syncthreads();
    }
    const uint32_t t1 = t0 + lv;
    if (t1 < jb.c - 1)
        for (uint32_t e = threadIdx.x; e <= t1; e += blockDim.x)
            jb.tree[src ^ 1u][(size_t)blockIdx.x * (t1 + 1) + e] = buf[lv & 1][e];
}

// Base preparation: normalise the reference-layout Jacobian point and store
// 2^(c*w) * P for every window w (affine, infinity flagged).
template <class F>
__global__ __launch_bounds__(128) void k_prepare_bases(const Jacobian<F> *__restrict__ in,
                                                       const uint32_t *__restrict__ perm, uint32_t n, uint32_t c,
                                                       uint32_t W, Affine<F> *__restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Jacobian<F> J = in[perm ? perm[i] : i];
    Affine<F> a;
    if (!jacobian_to_affine(J, a)) {
        const Affine<F> inf = affine_infinity<F>();
        for (uint32_t w = 0; w < W; ++w) out[(size_t)w * n + i] = inf;
        return;
    }
    out[i] = a;
    Xyzz<F> P = Xyzz<F>::from_affine(a);
    for (uint32_t w = 1; w < W; ++w) {
        for (uint32_t t = 0; t < c; ++t) P = xyzz_dbl(P);
        out[(size_t)w * n + i] = xyzz_to_affine(P);
    }
}

// Workgroups of 256 threads of `kernel` the device keeps resident at once
// (all CUs at the kernel's occupancy), times two so that grid-stride loops
// end evenly.  Computed once per kernel.
template <auto Kernel>
uint64_t resident_blocks()
{
    static const uint64_t blocks = [] {
        int dev = 0, cus = 0, per_cu = 0;
        NKR_HIP(hipGetDevice(&dev));
        NKR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        NKR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, Kernel, 256, 0));
        return (uint64_t)std::max(1, cus) * std::max(1, per_cu) * 2;
    }();
    return blocks;
}

uint32_t g_target_log = kMsmTargetPiecesLog;
uint32_t g_lanes_cap = 16;

}  // namespace

void msm_set_shape(uint32_t target_pieces_log, uint32_t complete_lanes_cap)
{
    NKR_REQUIRE(target_pieces_log >= 10 && target_pieces_log <= 22, "MSM target pieces must be 2^10 .. 2^22");
    NKR_REQUIRE(complete_lanes_cap >= 1 && complete_lanes_cap <= 16 &&
                    (complete_lanes_cap & (complete_lanes_cap - 1)) == 0,
                "MSM completion lanes must be a power of two in 1 .. 16");
    g_target_log = target_pieces_log;
    g_lanes_cap = complete_lanes_cap;
}

uint32_t msm_target_pieces_log() { return g_target_log; }
uint32_t msm_complete_lanes_cap() { return g_lanes_cap; }

uint32_t MsmSchedule::sort_blocks(uint32_t n)
{
    return std::max<uint32_t>(1, std::min<uint32_t>(kSortBlocksTarget, div_up(n, kSortMinPerBlock)));
}

uint64_t msm_schedule_bytes(const MsmPlan &p)
{
    const uint64_t blocks = MsmSchedule::sort_blocks(p.n);
    return p.max_entries() * 4 + (uint64_t)p.B * 4 * 3 + blocks * p.B * 4 + 32;
}

template <class F>
uint64_t msm_job_bytes(const MsmPlan &p)
{
    const uint64_t total = msm_partial_capacity(p, p.max_pieces(msm_target_pieces_log()), nullptr, nullptr);
    // partial sums; per bucket: sums and two tree levels, region start,
    // count and fold start, the completion class lists; the bit planes; and
    // the job's fixed records (job table, class counters, piece offsets)
    return total * sizeof(partial_t<F>) + (uint64_t)p.B * (3 * sizeof(partial_t<F>) + 12 + 4 * kCompleteClasses) +
           p.c * sizeof(Xyzz<F>) + 1024;
}

template <class F>
void MsmBases<F>::prepare(const Jacobian<F> *jacobian, const uint32_t *perm, const MsmPlan &plan,
                          hipStream_t stream)
{
    plan_ = plan;
    pts_.allocate((size_t)plan.W * plan.n);
    k_prepare_bases<F><<<div_up(plan.n, 128), 128, 0, stream>>>(jacobian, perm, plan.n, plan.c, plan.W,
                                                                  pts_.get());
    NKR_LAUNCH_CHECK();
}

// The sort kernels hold B counts in dynamic LDS: up to 128 KiB for c = 16,
// above the 64 KiB a launch gets without asking.
static void allow_sort_lds()
{
    static std::once_flag once;
    std::call_once(once, [] {
        const int bytes = (int)(kMsmMaxBuckets * sizeof(uint32_t));
        NKR_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_hist),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        NKR_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_scatter),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        NKR_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_hist_batch),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        NKR_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_scatter_batch),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    });
}

uint32_t MsmSchedule::prepare(const MsmPlan &plan)
{
    allow_sort_lds();
    NKR_REQUIRE(plan.B <= kMsmMaxBuckets, "MSM window too wide for the bucket scans");
    NKR_REQUIRE(plan.piece_lo >= 1 && plan.piece_lo <= plan.piece_hi, "MSM piece bounds");
    plan_ = plan;
    target_log_ = target_override_ ? target_override_ : msm_target_pieces_log();
    const uint32_t n = plan.n, B = plan.B;
    blocks_ = sort_blocks(n);
    sorted_.ensure((size_t)n * plan.W);
    count_.ensure(B);
    bucket_start_.ensure(B + 8);  // + the four counters of each piece split
    piece_start_.ensure(B + 1);
    block_count_.ensure((size_t)blocks_ * B);
    return div_up(n, blocks_);
}

void MsmSchedule::build(const fr_t *scalars_mont, const MsmPlan &plan, hipStream_t stream, StageTimer *timer)
{
    const uint32_t per_block = prepare(plan);
    const uint32_t n = plan.n, B = plan.B;
    const size_t lds = (size_t)B * sizeof(uint32_t);
    const uint32_t threads = std::min<uint32_t>(kSortThreads, (per_block + 63) & ~63u);
    if (timer) timer->begin(kStageSort, stream);
    k_hist<<<blocks_, threads, lds, stream>>>(scalars_mont, n, plan.c, plan.W, B, per_block, block_count_.get());
    NKR_LAUNCH_CHECK();
    k_colscan<<<div_up(B, 64), 64 * kScanGroups, 0, stream>>>(block_count_.get(), blocks_, B, count_.get());
    NKR_LAUNCH_CHECK();
    if (alt_piece_lo_) alt_piece_start_.ensure(B + 1);
    k_schedule_scan<<<1, 1024, 0, stream>>>(count_.get(), B, plan.piece_lo, plan.piece_hi, target_log_,
                                            bucket_start_.get(), piece_start_.get(), alt_piece_lo_,
                                            alt_piece_lo_ ? alt_piece_start_.get() : nullptr);
    NKR_LAUNCH_CHECK();
    uint32_t *th = own_totals_.reuse(8);
    NKR_HIP(hipMemcpyAsync(th, bucket_start_.get() + B, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    own_totals_.uploaded(stream);
    host_totals_ = th;
    host_stage_ = &own_totals_;
    k_scatter<<<blocks_, threads, lds, stream>>>(scalars_mont, n, plan.c, plan.W, B, per_block,
                                                 bucket_start_.get(), block_count_.get(), sorted_.get());
    NKR_LAUNCH_CHECK();
    if (timer) timer->end(kStageSort, stream);
}

void MsmScheduleBatch::build(const std::vector<Item> &items, hipStream_t stream, bool graph)
{
    const uint32_t J = (uint32_t)items.size();
    if (J == 0) return;
    NKR_REQUIRE(J <= 65535, "a schedule batch holds at most 65535 schedules");
    const size_t jobs_bytes = (sizeof(SortJob) * J + 15) & ~size_t(15);
    const size_t total = jobs_bytes + sizeof(uint32_t) * (J + 1);
    uint8_t *st = graph ? stage_.reuse_unsynced(total) : stage_.reuse(total);
    table_.ensure(total);
    SortJob *jobs = reinterpret_cast<SortJob *>(st);
    uint32_t *off = reinterpret_cast<uint32_t *>(st + jobs_bytes);
    uint32_t Bmax = 0, threads = 64;
    uint64_t blocks = 0;
    for (uint32_t j = 0; j < J; ++j) {
        MsmSchedule &s = *items[j].sched;
        const MsmPlan &p = items[j].plan;
        NKR_REQUIRE(s.alt_piece_lo_ == 0, "batched schedules have one piece split");
        const uint32_t per_block = s.prepare(p);
        jobs[j] = SortJob{items[j].scalars_mont, p.n, p.c, p.W, p.B, s.blocks_, per_block, p.piece_lo, p.piece_hi,
                          s.target_log_, s.count_.get(), s.block_count_.get(), s.bucket_start_.get(), s.piece_start_.get(),
                          s.sorted_.get()};
        off[j] = (uint32_t)blocks;
        blocks += s.blocks_;
        Bmax = std::max(Bmax, p.B);
        threads = std::max(threads, std::min<uint32_t>(kSortThreads, (per_block + 63) & ~63u));
    }
    NKR_REQUIRE(blocks < (1ull << 31), "schedule batch too large");
    off[J] = (uint32_t)blocks;
    NKR_HIP(hipMemcpyAsync(table_.get(), st, total, hipMemcpyHostToDevice, stream));
    if (!graph) stage_.uploaded(stream);
    const SortJob *dj = reinterpret_cast<const SortJob *>(table_.get());
    const uint32_t *doff = reinterpret_cast<const uint32_t *>(table_.get() + jobs_bytes);
    const size_t lds = (size_t)Bmax * sizeof(uint32_t);
    k_hist_batch<<<(uint32_t)blocks, threads, lds, stream>>>(dj, doff, J);
    NKR_LAUNCH_CHECK();
    k_colscan_batch<<<dim3(div_up(Bmax, 64), J), 64 * kScanGroups, 0, stream>>>(dj);
    NKR_LAUNCH_CHECK();
    k_schedule_scan_batch<<<J, 1024, 0, stream>>>(dj);
    NKR_LAUNCH_CHECK();
    if (!graph) {
        // every schedule's four counters to the host in one copy
        totals_dev_.ensure(4 * J);
        k_gather_totals<<<div_up(4 * J, 256), 256, 0, stream>>>(dj, J, totals_dev_.get());
        NKR_LAUNCH_CHECK();
        uint32_t *th = totals_host_.reuse(4 * J);
        NKR_HIP(hipMemcpyAsync(th, totals_dev_.get(), 4 * J * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        totals_host_.uploaded(stream);
        for (uint32_t j = 0; j < J; ++j) {
            items[j].sched->host_totals_ = th + 4 * j;
            items[j].sched->host_stage_ = &totals_host_;
        }
    } else {
        for (uint32_t j = 0; j < J; ++j) {
            items[j].sched->host_totals_ = nullptr;
            items[j].sched->host_stage_ = nullptr;
        }
    }
    k_scatter_batch<<<(uint32_t)blocks, threads, lds, stream>>>(dj, doff, J);
    NKR_LAUNCH_CHECK();
}

// The piece split a job reads: its schedule's own, or the second one.
template <class F>
static const uint32_t *job_counters(const MsmJob<F> &jb)
{
    return jb.sched->bucket_start_.get() + jb.sched->plan().B + (jb.alt ? 4 : 0);
}
template <class F>
static const uint32_t *job_piece_start(const MsmJob<F> &jb)
{
    return jb.alt ? jb.sched->alt_piece_start_.get() : jb.sched->piece_start_.get();
}
template <class F>
static uint64_t job_max_pieces(const MsmJob<F> &jb)
{
    return jb.alt ? jb.sched->alt_max_pieces() : jb.sched->max_pieces();
}

template <class F>
static std::vector<const void *> layout_key(const std::vector<MsmJob<F>> &jobs)
{
    std::vector<const void *> key;
    for (const auto &jb : jobs) {
        key.push_back(jb.bases->data());
        key.push_back(jb.sched->sorted_.get());
        key.push_back(jb.sched->bucket_start_.get());
        key.push_back(job_piece_start(jb));
        key.push_back(jb.planes);
        key.push_back(reinterpret_cast<const void *>((uintptr_t)job_max_pieces(jb)));
    }
    return key;
}

template <class F>
bool MsmBatch<F>::layout_matches(const std::vector<MsmJob<F>> &jobs) const
{
    return layout_key(jobs) == key_;
}

template <class F>
void MsmBatch<F>::layout(const std::vector<MsmJob<F>> &jobs, hipStream_t stream)
{
    std::vector<const void *> key = layout_key(jobs);
    if (key == key_) return;
    const uint32_t J = (uint32_t)jobs.size();
    // Per job, sized by its own plan: the partial-sum pool (level-0 pieces,
    // then each fold level's worst case) and the per-bucket arrays.
    std::vector<uint64_t> pool_off(J + 1, 0), b_off(J + 1, 0), b1_off(J + 1, 0);
    std::vector<std::array<uint32_t, kMaxFoldLevels>> level_base(J);
    fold_levels_ = 0;
    fold_cap_.assign(kMaxFoldLevels, 0);
    thread_base_.assign(J + 1, 0);
    for (uint32_t j = 0; j < J; ++j) {
        const MsmPlan &p = jobs[j].sched->plan();
        NKR_REQUIRE(jobs[j].bases->plan().n == p.n && jobs[j].bases->plan().c == p.c,
                    "MSM bases and schedule disagree");
        const uint64_t pieces = job_max_pieces(jobs[j]);
        thread_base_[j + 1] = thread_base_[j] + pieces;
        uint64_t lb[kMaxFoldLevels], lc[kMaxFoldLevels];
        uint32_t levels = 0;
        const uint64_t base = msm_partial_capacity(p, pieces, &levels, lb, lc);
        NKR_REQUIRE(base < (1ull << 32), "MSM partial pool exceeds 32-bit indexing");
        for (uint32_t L = 0; L < kMaxFoldLevels; ++L) {
            level_base[j][L] = (uint32_t)lb[L];
            fold_cap_[L] = std::max<uint64_t>(fold_cap_[L], lc[L]);
        }
        fold_levels_ = std::max(fold_levels_, levels);
        pool_off[j + 1] = pool_off[j] + base;
        b_off[j + 1] = b_off[j] + p.B;
        b1_off[j + 1] = b1_off[j] + p.B + 1;
    }
    // grow-only: a batch whose composition changes from call to call reuses them
    pool_.ensure(pool_off[J]);
    sums_.ensure(b_off[J]);
    tree_[0].ensure(b_off[J]);
    tree_[1].ensure(b_off[J]);
    reg_start_.ensure(b_off[J]);
    reg_count_.ensure(b_off[J]);
    fold_start_.ensure(b1_off[J]);

    static_assert(std::is_trivially_copyable<JobDev<F>>::value, "job records are raw-copied");
    JobDev<F> *host = reinterpret_cast<JobDev<F> *>(jobs_stage_.reuse(sizeof(JobDev<F>) * J));
    for (uint32_t j = 0; j < J; ++j) {
        const MsmJob<F> &jb = jobs[j];
        JobDev<F> &d = host[j];
        d.bases = jb.bases->data();
        d.sorted = jb.sched->sorted_.get();
        d.bucket_start = jb.sched->bucket_start_.get();
        d.piece_start = job_piece_start(jb);
        d.piece_max = job_counters(jb) + 2;
        d.piece_len = job_counters(jb) + 3;
        d.pool = pool_.get() + pool_off[j];
        d.reg_start = reg_start_.get() + b_off[j];
        d.reg_count = reg_count_.get() + b_off[j];
        d.fold_start = fold_start_.get() + b1_off[j];
        d.sums = sums_.get() + b_off[j];
        d.tree[0] = tree_[0].get() + b_off[j];
        d.tree[1] = tree_[1].get() + b_off[j];
        d.planes = jb.planes;
        d.B = jb.sched->plan().B;
        d.c = jb.sched->plan().c;
        d.list_off = (uint32_t)b_off[j];
        for (uint32_t L = 0; L < kMaxFoldLevels; ++L) d.level_base[L] = level_base[j][L];
    }
    thread_total_ = thread_base_[J];
    // completion lanes: at most 8 per bucket, plus each segment's padding
    complete_bound_ = 0;
    for (uint32_t j = 0; j < J; ++j) complete_bound_ += (uint64_t)host[j].B * 8 + kCompleteClasses * 64;
    NKR_REQUIRE(complete_bound_ < (1ull << 32) && thread_total_ < (1ull << 32),
                "MSM batch too large for 32-bit work indices");
    complete_base_.ensure(kCompleteClasses * J + 1);
    class_count_.ensure(kCompleteClasses * J);
    class_list_.ensure(kCompleteClasses * b_off[J]);
    jobs_dev_.ensure(sizeof(JobDev<F>) * J);
    thread_base_dev_.ensure(J + 1);  // actual piece offsets, filled on the device per run
    NKR_HIP(hipMemcpyAsync(jobs_dev_.get(), host, sizeof(JobDev<F>) * J, hipMemcpyHostToDevice, stream));
    jobs_stage_.uploaded(stream);
    key_ = key;
}

template <class F>
void MsmBatch<F>::run(const std::vector<MsmJob<F>> &jobs, hipStream_t stream, StageTimer *timer, bool graph)
{
    if (jobs.empty()) return;
    NKR_REQUIRE(!graph || layout_matches(jobs), "a captured MSM batch must reuse the layout of its eager run");
    layout(jobs, stream);
    const uint32_t J = (uint32_t)jobs.size();
    uint32_t Bmax = 0, cmax = 0;
    for (const auto &jb : jobs) {
        Bmax = std::max(Bmax, jb.sched->plan().B);
        cmax = std::max(cmax, jb.sched->plan().c);
    }
    const JobDev<F> *jd = reinterpret_cast<const JobDev<F> *>(jobs_dev_.get());

    k_regions_init<F><<<dim3(div_up(Bmax, 256), J), 256, 0, stream>>>(jd, class_count_.get());
    NKR_LAUNCH_CHECK();
    k_piece_scan<F><<<1, 1024, 0, stream>>>(jd, J, thread_base_dev_.get());
    NKR_LAUNCH_CHECK();
    // the accumulation stage brackets exactly the k_accumulate launch
    const int acc_stage = std::is_same<F, fq2_t>::value ? kStageAccumulateG2 : kStageAccumulate;
    if (timer) timer->begin(acc_stage, stream);
    // One workgroup per kAccThreads pieces rather than a grid of resident
    // size: the accumulation's workgroups then retire all along the launch,
    // so the other stream's kernels (SpMV and transforms beside the G2
    // accumulation, the G2 bucket reduction beside the G1 one) get CU slots
    // as they free up instead of waiting for the whole launch (dense 2^20
    // Poseidon-style proof 12.0 -> 11.2 ms and 64-proof mix 33.5-34.0 ->
    // 31.2-32.4 ms with the witness in HBM, on MI355X).
    const uint64_t blocks = div_up(thread_total_, kAccThreads);
    k_accumulate<F><<<(uint32_t)std::max<uint64_t>(blocks, 1), kAccThreads, 0, stream>>>(jd, thread_base_dev_.get(), J);
    NKR_LAUNCH_CHECK();
    if (timer) {
        timer->end(acc_stage, stream);
        timer->begin(kStageFold, stream);
    }
    // Fold levels the largest bucket of any job needs: read from the host
    // copy of the schedules' counters.  The accumulation is already queued,
    // so this wait overlaps the GPU's work instead of stalling it.
    uint32_t folds = 0;
    if (graph) {
        folds = fold_levels_;
    } else {
        for (const auto &jb : jobs) {
            const uint64_t cnt = (jb.alt ? jb.sched->alt_totals_host() : jb.sched->totals_host())[2];
            folds = std::max(folds, msm_fold_levels(cnt));
        }
    }
    NKR_REQUIRE(folds <= fold_levels_, "MSM bucket exceeds the partial-sum capacity");
    for (uint32_t L = 0; L < folds; ++L) {
        k_fold_scan<F><<<J, 256, 0, stream>>>(jd);
        NKR_LAUNCH_CHECK();
        k_fold<F><<<dim3(div_up(fold_cap_[L], 256), J), 256, 0, stream>>>(jd, L);
        NKR_LAUNCH_CHECK();
        k_fold_retarget<F><<<dim3(div_up(Bmax, 256), J), 256, 0, stream>>>(jd, L);
        NKR_LAUNCH_CHECK();
    }
    uint32_t max_class = 0;
    while (max_class + 1 < kCompleteClasses && (2u << max_class) <= msm_complete_lanes_cap()) ++max_class;
    k_classify<F><<<dim3(div_up(Bmax, 256), J), 256, 0, stream>>>(jd, max_class, class_count_.get(),
                                                                   class_list_.get());
    NKR_LAUNCH_CHECK();
    const uint32_t S = kCompleteClasses * J;
    k_complete_scan<<<1, 1024, 0, stream>>>(class_count_.get(), S, complete_base_.get());
    NKR_LAUNCH_CHECK();
    {
        const uint64_t blocks = std::min<uint64_t>(div_up(complete_bound_, 256), resident_blocks<k_complete<F>>());
        k_complete<F><<<(uint32_t)std::max<uint64_t>(blocks, 1), 256, 0, stream>>>(
            jd, class_count_.get(), class_list_.get(), complete_base_.get(), S);
    }
    NKR_LAUNCH_CHECK();
    // bit-plane tree: as many levels per launch as two LDS buffers of 2^h
    // level-t blocks allow; a workgroup has as many lanes as its widest
    // level has merges (whole waves, at most 256), so no idle wave holds a
    // SIMD's registers
    for (uint32_t t = 0, src = 0; t + 1 <= cmax - 1; src ^= 1u) {
        uint32_t h = 1;
        while (t + h < cmax - 1 && 2 * (size_t(2) << h) * (t + 1) * sizeof(partial_t<F>) <= tree_lds_bytes<F>())
            ++h;
        const size_t lds = 2 * (size_t(1) << h) * (t + 1) * sizeof(partial_t<F>);
        const uint32_t widest = (1u << (h - 1)) * (t + 2);
        const uint32_t lanes = std::min<uint32_t>(256, (uint32_t)div_up(widest, 64) * 64);
        k_tree_lds<F><<<dim3(Bmax >> (t + h), J), lanes, lds, stream>>>(jd, t, h, src);
        NKR_LAUNCH_CHECK();
        t += h;
    }
    if (timer) timer->end(kStageFold, stream);
}

template <class F>
Xyzz<F> msm_sum(const Xyzz<F> *planes, uint32_t c)
{
    Xyzz<F> acc = planes[c - 1];
    for (int j = (int)c - 2; j >= 0; --j) {
        acc = xyzz_dbl(acc);
        acc = xyzz_add(acc, planes[j]);
    }
    return acc;
}

template <class F>
Jacobian<F> msm_finish(const Xyzz<F> *planes, uint32_t c)
{
    return xyzz_to_jacobian_affine(msm_sum(planes, c));
}

template <class F>
MsmContext<F>::MsmContext(const Jacobian<F> *bases_dev, uint32_t n, uint32_t c_override, hipStream_t stream)
{
    const MsmPlan plan = MsmPlan::choose(n, c_override);
    bases_.prepare(bases_dev, nullptr, plan, stream);
    planes_.allocate(plan.c);
}

template <class F>
Jacobian<F> MsmContext<F>::run(const fr_t *scalars_dev, hipStream_t stream)
{
    const MsmPlan &plan = bases_.plan();
    sched_.build(scalars_dev, plan, stream);
    batch_.run({MsmJob<F>{&bases_, &sched_, planes_.get()}}, stream);
    std::vector<Xyzz<F>> host(plan.c);
    NKR_HIP(hipMemcpyAsync(host.data(), planes_.get(), plan.c * sizeof(Xyzz<F>), hipMemcpyDeviceToHost, stream));
    NKR_HIP(hipStreamSynchronize(stream));
    return msm_finish(host.data(), plan.c);
}

template class MsmBases<fq_t>;
template class MsmBases<fq2_t>;
template class MsmBatch<fq_t>;
template class MsmBatch<fq2_t>;
template class MsmContext<fq_t>;
template class MsmContext<fq2_t>;
template Jacobian<fq_t> msm_finish(const Xyzz<fq_t> *, uint32_t);
template Jacobian<fq2_t> msm_finish(const Xyzz<fq2_t> *, uint32_t);
template Xyzz<fq_t> msm_sum(const Xyzz<fq_t> *, uint32_t);
template Xyzz<fq2_t> msm_sum(const Xyzz<fq2_t> *, uint32_t);
template uint64_t msm_job_bytes<fq_t>(const MsmPlan &);
template uint64_t msm_job_bytes<fq2_t>(const MsmPlan &);

}  // namespace nkr

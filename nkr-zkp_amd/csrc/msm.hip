// Pippenger MSM kernels for gfx950 (see msm.h for the design).
#include "msm.h"

#include <algorithm>
#include <array>
#include <cstring>
#include <type_traits>

namespace nkr {

namespace {

constexpr uint32_t kEntryMask = 0x7fffffffu;
constexpr uint32_t kSortBlocksMax = 1024;
constexpr uint32_t kTreeTailThreads = 256;

template <class F>
struct JobDev {
    const Affine<F> *bases;
    const uint32_t *sorted;
    const uint32_t *bucket_start;  // B + 1
    const uint32_t *piece_start;   // B + 1
    const uint32_t *piece_max;     // largest level-0 piece count of any bucket
    Xyzz<F> *pool;                 // partial sums (level 0 region, then fold levels)
    uint32_t *reg_start;           // B: where bucket k's partials live in `pool`
    uint32_t *reg_count;           // B
    uint32_t *fold_start;          // B + 1
    Xyzz<F> *sums;                 // B: S_{k+1}
    Xyzz<F> *tree[2];              // B each
    Xyzz<F> *planes;               // c
    uint32_t B, c, piece;
    uint32_t level_base[kMaxFoldLevels];  // where each fold level's partials start in `pool`
};

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
        const uint64_t lo = s.l[word];
        const uint64_t hi = (word + 1 < 8) ? s.l[word + 1] : 0u;
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

__global__ void k_zero_u32(uint32_t *p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Block-local histogram of bucket magnitudes, then one global atomic per
// (block, bucket) reserves the block's range inside each bucket.
__device__ __forceinline__ void hist_block(uint32_t *hist, const fr_t *__restrict__ scalars, uint32_t n, uint32_t c,
                                          uint32_t W, uint32_t B, uint32_t per_block, uint32_t b,
                                          uint32_t *__restrict__ count, uint32_t *__restrict__ block_base)
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
    uint32_t *mine = block_base + (size_t)b * B;
    for (uint32_t k = threadIdx.x; k < B; k += blockDim.x) {
        const uint32_t h = hist[k];
        mine[k] = h ? atomicAdd(&count[k], h) : 0u;
    }
}

__device__ __forceinline__ void scatter_block(uint32_t *cursor, const fr_t *__restrict__ scalars, uint32_t n,
                                             uint32_t c, uint32_t W, uint32_t B, uint32_t per_block, uint32_t b,
                                             const uint32_t *__restrict__ bucket_start,
                                             const uint32_t *__restrict__ block_base, uint32_t *__restrict__ sorted)
{
    const uint32_t *mine = block_base + (size_t)b * B;
    for (uint32_t k = threadIdx.x; k < B; k += blockDim.x) cursor[k] = bucket_start[k] + mine[k];
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

__global__ __launch_bounds__(256) void k_hist(const fr_t *__restrict__ scalars, uint32_t n, uint32_t c, uint32_t W,
                                              uint32_t B, uint32_t per_block, uint32_t *__restrict__ count,
                                              uint32_t *__restrict__ block_base)
{
    extern __shared__ uint32_t lds[];
    hist_block(lds, scalars, n, c, W, B, per_block, blockIdx.x, count, block_base);
}

__global__ __launch_bounds__(256) void k_scatter(const fr_t *__restrict__ scalars, uint32_t n, uint32_t c, uint32_t W,
                                                 uint32_t B, uint32_t per_block,
                                                 const uint32_t *__restrict__ bucket_start,
                                                 const uint32_t *__restrict__ block_base, uint32_t *__restrict__ sorted)
{
    extern __shared__ uint32_t lds[];
    scatter_block(lds, scalars, n, c, W, B, per_block, blockIdx.x, bucket_start, block_base, sorted);
}


// Exclusive prefix of `len` values into out[0..len] within one workgroup.
//   mode 0: v[k] = src[k]
//   mode 2: v[k] = ceil(src[k] / T)
//   mode 3: v[k] = src[k] > lim ? ceil(src[k] / T) : 0
__device__ void block_scan(const uint32_t *__restrict__ src, uint32_t len, uint32_t mode, uint32_t T, uint32_t lim,
                           uint32_t *__restrict__ out, uint32_t *__restrict__ max_out = nullptr)
{
    __shared__ uint32_t part[1024];
    __shared__ uint32_t pmax;
    if (threadIdx.x == 0) pmax = 0;
    __syncthreads();
    auto value = [&](uint32_t k) -> uint32_t {
        const uint32_t x = src[k];
        if (mode == 0) return x;
        if (mode == 2) return (x + T - 1) / T;
        return x > lim ? (x + T - 1) / T : 0u;
    };
    const uint32_t per = (len + blockDim.x - 1) / blockDim.x;
    const uint32_t beg = min(len, threadIdx.x * per);
    const uint32_t end = min(len, beg + per);
    uint32_t sum = 0, mx = 0;
    for (uint32_t k = beg; k < end; ++k) {
        const uint32_t v = value(k);
        sum += v;
        mx = max(mx, v);
    }
    part[threadIdx.x] = sum;
    if (max_out && mx) atomicMax(&pmax, mx);
    __syncthreads();
    for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
        const uint32_t t = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum;
    for (uint32_t k = beg; k < end; ++k) {
        out[k] = run;
        run += value(k);
    }
    if (threadIdx.x == blockDim.x - 1) out[len] = part[threadIdx.x];
    if (max_out && threadIdx.x == 0) *max_out = pmax;
    __syncthreads();
}

// Bucket starts (mode 0 over counts) and level-0 piece starts (ceil(count/piece)).
__global__ __launch_bounds__(1024) void k_schedule_scan(const uint32_t *__restrict__ count, uint32_t B,
                                                        uint32_t piece, uint32_t *__restrict__ bucket_start,
                                                        uint32_t *__restrict__ piece_start)
{
    block_scan(count, B, 0, 1, 0, bucket_start);
    block_scan(count, B, 2, piece, 0, piece_start, bucket_start + B + 2);
    if (threadIdx.x == 0) bucket_start[B + 1] = piece_start[B];
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
    uint32_t n, c, W, B, per_block, piece;
    uint32_t *count, *block_base, *bucket_start, *piece_start, *sorted;
};

__global__ __launch_bounds__(256) void k_zero_counts(const SortJob *__restrict__ jobs)
{
    const SortJob &jb = jobs[blockIdx.y];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < jb.B) jb.count[k] = 0;
}

// Workgroup blockIdx.x is sort block (blockIdx.x - block_off[j]) of job j.
__global__ __launch_bounds__(256) void k_hist_batch(const SortJob *__restrict__ jobs,
                                                    const uint32_t *__restrict__ block_off, uint32_t J)
{
    extern __shared__ uint32_t lds[];
    const uint32_t j = find_segment(block_off, J, blockIdx.x);
    const SortJob &jb = jobs[j];
    hist_block(lds, jb.scalars, jb.n, jb.c, jb.W, jb.B, jb.per_block, blockIdx.x - block_off[j], jb.count,
               jb.block_base);
}

__global__ __launch_bounds__(1024) void k_schedule_scan_batch(const SortJob *__restrict__ jobs)
{
    const SortJob &jb = jobs[blockIdx.x];
    block_scan(jb.count, jb.B, 0, 1, 0, jb.bucket_start);
    block_scan(jb.count, jb.B, 2, jb.piece, 0, jb.piece_start, jb.bucket_start + jb.B + 2);
    if (threadIdx.x == 0) jb.bucket_start[jb.B + 1] = jb.piece_start[jb.B];
}

__global__ __launch_bounds__(256) void k_scatter_batch(const SortJob *__restrict__ jobs,
                                                       const uint32_t *__restrict__ block_off, uint32_t J)
{
    extern __shared__ uint32_t lds[];
    const uint32_t j = find_segment(block_off, J, blockIdx.x);
    const SortJob &jb = jobs[j];
    scatter_block(lds, jb.scalars, jb.n, jb.c, jb.W, jb.B, jb.per_block, blockIdx.x - block_off[j],
                  jb.bucket_start, jb.block_base, jb.sorted);
}

template <class F>
__global__ __launch_bounds__(256) void k_regions_init(const JobDev<F> *__restrict__ jobs)
{
    const JobDev<F> &jb = jobs[blockIdx.y];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
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

// Level 0 over all jobs: a thread sums `piece` signed bases of one bucket
// into one partial, then moves on by the grid stride, so the launch can be
// sized to the resident waves and every wave gets the same share of pieces.
template <class F>
__global__ __launch_bounds__(256) void k_accumulate(const JobDev<F> *__restrict__ jobs,
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
        const uint32_t beg = jb.bucket_start[k] + q * jb.piece;
        const uint32_t end = min(beg + jb.piece, jb.bucket_start[k + 1]);
        const Affine<F> *__restrict__ bases = jb.bases;
        const uint32_t *__restrict__ sorted = jb.sorted;
        Xyzz<F> acc = Xyzz<F>::infinity();
        if (beg < end) {
            // Software pipeline: the gather of entry e + 1's base is issued
            // before entry e's addition, so its latency hides behind the
            // field arithmetic.
            uint32_t v = sorted[beg];
            Affine<F> P = bases[v & kEntryMask];
#pragma unroll 1
            for (uint32_t e = beg; e < end; ++e) {
                uint32_t vn = 0;
                Affine<F> Pn;
                if (e + 1 < end) {
                    vn = sorted[e + 1];
                    Pn = bases[vn & kEntryMask];
                }
                if (!affine_is_inf(P)) {
                    const F y = (v >> 31) ? FieldOps<F>::neg(P.y) : P.y;
                    acc = xyzz_add_affine(acc, P.x, y);
                }
                v = vn;
                P = Pn;
            }
        }
        jb.pool[p] = acc;
    }
}

// Fold level L: buckets with more than kMsmCompleteMax partials get
// ceil(count / kMsmFoldIn) new partials written at pool[level_base[L] ...].
template <class F>
__global__ __launch_bounds__(1024) void k_fold_scan(const JobDev<F> *__restrict__ jobs)
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
    Xyzz<F> acc = jb.pool[beg];
    for (uint32_t e = beg + 1; e < end; ++e) acc = xyzz_add(acc, jb.pool[e]);
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

template <class F>
__device__ __forceinline__ Xyzz<F> shfl_xor_point(const Xyzz<F> &p, int mask)
{
    Xyzz<F> r;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(&p);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&r);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(Xyzz<F>) / 4); ++i) dst[i] = (uint32_t)__shfl_xor((int)src[i], mask, 16);
    return r;
}

// Lanes per bucket for each job's completion and the job offsets of the
// completion index space.  A bucket with m partials summed by L lanes costs
// ceil(m / L) - 1 serial additions plus log2 L shuffle levels; L =
// nextpow2(m) / 2 (1 .. 16) minimises that for the job's largest bucket
// (m <= kMsmCompleteMax after folding).  Jobs are padded to whole waves so
// that every wave completes buckets of one job with one L.
template <class F>
__global__ __launch_bounds__(1024) void k_complete_scan(const JobDev<F> *__restrict__ jobs, uint32_t J,
                                                        uint32_t *__restrict__ base, uint32_t *__restrict__ lanes)
{
    __shared__ uint32_t part[1024];
    auto job_lanes = [&](uint32_t j) -> uint32_t {
        const uint32_t m = min(*jobs[j].piece_max, kMsmCompleteMax);
        uint32_t L = 1;
        while (2 * L < m && L < 16) L <<= 1;  // nextpow2(m) / 2
        return L;
    };
    auto job_span = [&](uint32_t j, uint32_t L) -> uint32_t { return (jobs[j].B * L + 63u) & ~63u; };
    const uint32_t per = (J + blockDim.x - 1) / blockDim.x;
    const uint32_t beg = min(J, threadIdx.x * per), end = min(J, beg + per);
    uint32_t sum = 0;
    for (uint32_t j = beg; j < end; ++j) sum += job_span(j, job_lanes(j));
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
        const uint32_t L = job_lanes(j);
        base[j] = run;
        lanes[j] = L;
        run += job_span(j, L);
    }
    if (threadIdx.x == blockDim.x - 1) base[J] = part[threadIdx.x];
}

// Bucket completion: L lanes per bucket stride over its partials, then a
// shuffle tree of log2 L levels; grid-stride over all jobs' buckets.
template <class F>
__global__ __launch_bounds__(256) void k_complete(const JobDev<F> *__restrict__ jobs,
                                                  const uint32_t *__restrict__ base,
                                                  const uint32_t *__restrict__ lanes, uint32_t J)
{
    const uint32_t total = base[J];
#pragma unroll 1
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
        const uint32_t j = find_segment(base, J, g);
        const JobDev<F> &jb = jobs[j];
        const uint32_t L = lanes[j];
        const uint32_t local = g - base[j];
        const uint32_t k = local / L, lane = local & (L - 1);
        Xyzz<F> acc = Xyzz<F>::infinity();
        if (k < jb.B) {
            const uint32_t start = jb.reg_start[k];
            const uint32_t cnt = jb.reg_count[k];
            for (uint32_t i = lane; i < cnt; i += L) acc = xyzz_add(acc, jb.pool[start + i]);
        }
        for (uint32_t off = L >> 1; off >= 1; off >>= 1) acc = xyzz_add(acc, shfl_xor_point(acc, (int)off));
        if (k < jb.B && lane == 0) jb.sums[k] = acc;
    }
}

// Bit-plane tree, level t -> t+1.  A level-t block covers buckets
// [b*2^t, (b+1)*2^t) and holds P_0..P_{t-1} (sums of its buckets with that
// bit set) and T (its total).  Level 0 reads S_m (S_0 = infinity) from sums.
// The last level writes U_j = P_j, and U_{c-1} = S_B, to the planes.
template <class F>
__global__ __launch_bounds__(256) void k_tree(const JobDev<F> *__restrict__ jobs, uint32_t t)
{
    const JobDev<F> &jb = jobs[blockIdx.y];
    if (t + 1 > jb.c - 1) return;
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (jb.B >> (t + 1)) * (t + 2)) return;
    tree_merge(jb, t, idx);
}

// One merge of the bit-plane tree (level t -> t+1) for output index idx.
template <class F>
__device__ __forceinline__ void tree_merge(const JobDev<F> &jb, uint32_t t, uint32_t idx)
{
    const uint32_t E = t + 2;
    const uint32_t b = idx / E, e = idx % E;
    const bool last = (t + 1 == jb.c - 1);
    auto in = [&](uint32_t blk, uint32_t ent) -> Xyzz<F> {
        if (t == 0) return blk == 0 ? Xyzz<F>::infinity() : jb.sums[blk - 1];
        return jb.tree[t & 1][(size_t)blk * (t + 1) + ent];
    };
    Xyzz<F> out;
    if (e < t) {
        out = xyzz_add(in(2 * b, e), in(2 * b + 1, e));
    } else if (e == t) {
        out = in(2 * b + 1, t);
    } else {
        if (last) {
            jb.planes[jb.c - 1] = jb.sums[jb.B - 1];
            return;
        }
        out = xyzz_add(in(2 * b, t), in(2 * b + 1, t));
    }
    if (last)
        jb.planes[e] = out;
    else
        jb.tree[(t + 1) & 1][(size_t)b * E + e] = out;
}

// The remaining tree levels t0 .. c-2 of one job in one workgroup (blockIdx.x
// = job), with a barrier between levels instead of a launch.  256 threads: a
// G2 addition needs more than the 128 registers a 1024-thread workgroup
// leaves per lane, and spilling it costs more than the launches it saves.
template <class F>
__global__ __launch_bounds__(kTreeTailThreads) void k_tree_tail(const JobDev<F> *__restrict__ jobs, uint32_t t0)
{
    const JobDev<F> &jb = jobs[blockIdx.x];
    for (uint32_t t = t0; t + 1 <= jb.c - 1; ++t) {
        const uint32_t work = (jb.B >> (t + 1)) * (t + 2);
        for (uint32_t idx = threadIdx.x; idx < work; idx += blockDim.x) tree_merge(jb, t, idx);
        __threadfence_block();
        __syncthreads();
    }
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

}  // namespace

MsmPlan MsmPlan::choose(uint32_t n, uint32_t c_override)
{
    NKR_REQUIRE(n > 0, "MSM needs at least one term");
    MsmPlan p;
    p.n = n;
    uint32_t c = c_override;
    if (c == 0) {
        const unsigned lg = ceil_log2(n);
        // ~3/4 log2 n, clamped so the bucket histogram fits 64 KiB of LDS.
        c = std::max(4u, std::min(15u, (lg * 3 + 2) / 4));
        // Scalars have 254 bits; a top window with only a few of them sends a
        // whole window's entries to a handful of buckets (long serial chains),
        // so move to a neighbouring width whose top window keeps >= 4 bits.
        auto top_bits = [](uint32_t cc) { return 254 - cc * ((255 + cc - 1) / cc - 1); };
        if (top_bits(c) < 4) {
            if (c < 15 && top_bits(c + 1) >= 4)
                c += 1;
            else if (c > 4 && top_bits(c - 1) >= 4)
                c -= 1;
        }
    }
    NKR_REQUIRE(c >= 2 && c <= 15, "MSM window must be 2..15 bits");
    p.c = c;
    p.W = (255 + c - 1) / c;
    p.B = 1u << (c - 1);
    // pieces of up to n*W / 2^19 entries: several per resident thread, so the
    // grid-stride accumulation stays balanced, and short chains for small MSMs
    const uint64_t want = p.max_entries() >> 19;
    p.piece = kMsmPieceMin;
    while (p.piece < kMsmPiece && 2 * p.piece <= want) p.piece <<= 1;
    NKR_REQUIRE((uint64_t)p.W * n < (1ull << 31), "MSM too large for 31-bit entry indices");
    return p;
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

uint32_t MsmSchedule::prepare(const MsmPlan &plan)
{
    plan_ = plan;
    const uint32_t n = plan.n, B = plan.B;
    blocks_ = std::min<uint32_t>(kSortBlocksMax, div_up(n, 256));
    sorted_.ensure((size_t)n * plan.W);
    count_.ensure(B);
    bucket_start_.ensure(B + 3);  // + total pieces, + largest piece count of a bucket
    piece_start_.ensure(B + 1);
    block_base_.ensure((size_t)blocks_ * B);
    return div_up(n, blocks_);
}

void MsmSchedule::build(const fr_t *scalars_mont, const MsmPlan &plan, hipStream_t stream, StageTimer *timer)
{
    const uint32_t per_block = prepare(plan);
    const uint32_t n = plan.n, B = plan.B;
    const size_t lds = (size_t)B * sizeof(uint32_t);
    if (timer) timer->begin(kStageSort, stream);
    k_zero_u32<<<div_up(B, 256), 256, 0, stream>>>(count_.get(), B);
    NKR_LAUNCH_CHECK();
    k_hist<<<blocks_, 256, lds, stream>>>(scalars_mont, n, plan.c, plan.W, B, per_block, count_.get(),
                                          block_base_.get());
    NKR_LAUNCH_CHECK();
    k_schedule_scan<<<1, 1024, 0, stream>>>(count_.get(), B, plan.piece, bucket_start_.get(), piece_start_.get());
    NKR_LAUNCH_CHECK();
    k_scatter<<<blocks_, 256, lds, stream>>>(scalars_mont, n, plan.c, plan.W, B, per_block, bucket_start_.get(),
                                             block_base_.get(), sorted_.get());
    NKR_LAUNCH_CHECK();
    if (timer) timer->end(kStageSort, stream);
}

void MsmScheduleBatch::build(const std::vector<Item> &items, hipStream_t stream)
{
    const uint32_t J = (uint32_t)items.size();
    if (J == 0) return;
    NKR_REQUIRE(J <= 65535, "a schedule batch holds at most 65535 schedules");
    const size_t jobs_bytes = (sizeof(SortJob) * J + 15) & ~size_t(15);
    const size_t total = jobs_bytes + sizeof(uint32_t) * (J + 1);
    uint8_t *st = stage_.reuse(total);
    table_.ensure(total);
    SortJob *jobs = reinterpret_cast<SortJob *>(st);
    uint32_t *off = reinterpret_cast<uint32_t *>(st + jobs_bytes);
    uint32_t Bmax = 0;
    uint64_t blocks = 0;
    for (uint32_t j = 0; j < J; ++j) {
        MsmSchedule &s = *items[j].sched;
        const MsmPlan &p = items[j].plan;
        const uint32_t per_block = s.prepare(p);
        jobs[j] = SortJob{items[j].scalars_mont, p.n, p.c, p.W, p.B, per_block, p.piece,
                          s.count_.get(), s.block_base_.get(), s.bucket_start_.get(), s.piece_start_.get(),
                          s.sorted_.get()};
        off[j] = (uint32_t)blocks;
        blocks += s.blocks_;
        Bmax = std::max(Bmax, p.B);
    }
    NKR_REQUIRE(blocks < (1ull << 31), "schedule batch too large");
    off[J] = (uint32_t)blocks;
    NKR_HIP(hipMemcpyAsync(table_.get(), st, total, hipMemcpyHostToDevice, stream));
    stage_.uploaded(stream);
    const SortJob *dj = reinterpret_cast<const SortJob *>(table_.get());
    const uint32_t *doff = reinterpret_cast<const uint32_t *>(table_.get() + jobs_bytes);
    const size_t lds = (size_t)Bmax * sizeof(uint32_t);
    k_zero_counts<<<dim3(div_up(Bmax, 256), J), 256, 0, stream>>>(dj);
    NKR_LAUNCH_CHECK();
    k_hist_batch<<<(uint32_t)blocks, 256, lds, stream>>>(dj, doff, J);
    NKR_LAUNCH_CHECK();
    k_schedule_scan_batch<<<J, 1024, 0, stream>>>(dj);
    NKR_LAUNCH_CHECK();
    k_scatter_batch<<<(uint32_t)blocks, 256, lds, stream>>>(dj, doff, J);
    NKR_LAUNCH_CHECK();
}

template <class F>
void MsmBatch<F>::layout(const std::vector<MsmJob<F>> &jobs, hipStream_t stream)
{
    std::vector<const void *> key;
    for (const auto &jb : jobs) {
        key.push_back(jb.bases->data());
        key.push_back(jb.sched->sorted_.get());
        key.push_back(jb.sched->bucket_start_.get());
        key.push_back(jb.sched->piece_start_.get());
        key.push_back(jb.planes);
    }
    if (key == key_) return;
    const uint32_t J = (uint32_t)jobs.size();
    // Per job, sized by its own plan: the partial-sum pool (level-0 pieces,
    // then each fold level's worst case) and the per-bucket arrays.
    std::vector<uint64_t> pool_off(J + 1, 0), b_off(J + 1, 0), b1_off(J + 1, 0);
    std::vector<std::array<uint32_t, kMaxFoldLevels>> level_base(J);
    fold_levels_ = 0;
    fold_cap_.assign(kMaxFoldLevels, 0);
    for (uint32_t j = 0; j < J; ++j) {
        const MsmPlan &p = jobs[j].sched->plan();
        NKR_REQUIRE(jobs[j].bases->plan().n == p.n && jobs[j].bases->plan().c == p.c,
                    "MSM bases and schedule disagree");
        uint64_t cap = jobs[j].sched->max_pieces(), cnt = cap, base = cap;
        uint32_t levels = 0;
        while (cnt > kMsmCompleteMax) {
            NKR_REQUIRE(levels < kMaxFoldLevels, "MSM needs too many fold levels");
            cnt = (cnt + kMsmFoldIn - 1) / kMsmFoldIn;
            cap = cap / kMsmFoldIn + p.B + 1;
            level_base[j][levels] = (uint32_t)base;
            fold_cap_[levels] = std::max<uint64_t>(fold_cap_[levels], cap);
            base += cap;
            ++levels;
        }
        for (uint32_t L = levels; L < kMaxFoldLevels; ++L) level_base[j][L] = (uint32_t)base;
        NKR_REQUIRE(base < (1ull << 32), "MSM partial pool exceeds 32-bit indexing");
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
    thread_base_.assign(J + 1, 0);
    for (uint32_t j = 0; j < J; ++j) {
        const MsmJob<F> &jb = jobs[j];
        JobDev<F> &d = host[j];
        d.bases = jb.bases->data();
        d.sorted = jb.sched->sorted_.get();
        d.bucket_start = jb.sched->bucket_start_.get();
        d.piece_start = jb.sched->piece_start_.get();
        d.piece_max = jb.sched->bucket_start_.get() + jb.sched->plan().B + 2;
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
        d.piece = jb.sched->plan().piece;
        for (uint32_t L = 0; L < kMaxFoldLevels; ++L) d.level_base[L] = level_base[j][L];
        thread_base_[j + 1] = thread_base_[j] + jb.sched->max_pieces();
    }
    thread_total_ = thread_base_[J];
    complete_bound_ = 0;
    for (uint32_t j = 0; j < J; ++j) complete_bound_ += ((uint64_t)host[j].B * 16 + 63) & ~uint64_t(63);
    NKR_REQUIRE(complete_bound_ < (1ull << 32) && thread_total_ < (1ull << 32),
                "MSM batch too large for 32-bit work indices");
    complete_base_.ensure(J + 1);
    complete_lanes_.ensure(J);
    jobs_dev_.ensure(sizeof(JobDev<F>) * J);
    thread_base_dev_.ensure(J + 1);  // actual piece offsets, filled on the device per run
    NKR_HIP(hipMemcpyAsync(jobs_dev_.get(), host, sizeof(JobDev<F>) * J, hipMemcpyHostToDevice, stream));
    jobs_stage_.uploaded(stream);
    key_ = key;
}

template <class F>
void MsmBatch<F>::run(const std::vector<MsmJob<F>> &jobs, hipStream_t stream, StageTimer *timer)
{
    if (jobs.empty()) return;
    layout(jobs, stream);
    const uint32_t J = (uint32_t)jobs.size();
    uint32_t Bmax = 0, cmax = 0;
    for (const auto &jb : jobs) {
        Bmax = std::max(Bmax, jb.sched->plan().B);
        cmax = std::max(cmax, jb.sched->plan().c);
    }
    const JobDev<F> *jd = reinterpret_cast<const JobDev<F> *>(jobs_dev_.get());

    k_regions_init<F><<<dim3(div_up(Bmax, 256), J), 256, 0, stream>>>(jd);
    NKR_LAUNCH_CHECK();
    if (timer) timer->begin(kStageAccumulate, stream);
    k_piece_scan<F><<<1, 1024, 0, stream>>>(jd, J, thread_base_dev_.get());
    NKR_LAUNCH_CHECK();
    const uint64_t blocks = std::min<uint64_t>(div_up(thread_total_, 256), resident_blocks<k_accumulate<F>>());
    k_accumulate<F><<<(uint32_t)std::max<uint64_t>(blocks, 1), 256, 0, stream>>>(jd, thread_base_dev_.get(), J);
    NKR_LAUNCH_CHECK();
    if (timer) {
        timer->end(kStageAccumulate, stream);
        timer->begin(kStageFold, stream);
    }
    for (uint32_t L = 0; L < fold_levels_; ++L) {
        k_fold_scan<F><<<J, 1024, 0, stream>>>(jd);
        NKR_LAUNCH_CHECK();
        k_fold<F><<<dim3(div_up(fold_cap_[L], 256), J), 256, 0, stream>>>(jd, L);
        NKR_LAUNCH_CHECK();
        k_fold_retarget<F><<<dim3(div_up(Bmax, 256), J), 256, 0, stream>>>(jd, L);
        NKR_LAUNCH_CHECK();
    }
    k_complete_scan<F><<<1, 1024, 0, stream>>>(jd, J, complete_base_.get(), complete_lanes_.get());
    NKR_LAUNCH_CHECK();
    {
        const uint64_t blocks = std::min<uint64_t>(div_up(complete_bound_, 256), resident_blocks<k_complete<F>>());
        k_complete<F><<<(uint32_t)std::max<uint64_t>(blocks, 1), 256, 0, stream>>>(
            jd, complete_base_.get(), complete_lanes_.get(), J);
    }
    NKR_LAUNCH_CHECK();
    uint32_t t = 0;
    for (; t + 1 <= cmax - 1; ++t) {
        const uint64_t work = (uint64_t)(Bmax >> (t + 1)) * (t + 2);
        if (work <= kTreeTailThreads) break;  // the rest fits one workgroup per job
        k_tree<F><<<dim3(div_up(work, 256), J), 256, 0, stream>>>(jd, t);
        NKR_LAUNCH_CHECK();
    }
    if (t + 1 <= cmax - 1) {
        k_tree_tail<F><<<J, kTreeTailThreads, 0, stream>>>(jd, t);
        NKR_LAUNCH_CHECK();
    }
    if (timer) timer->end(kStageFold, stream);
}

template <class F>
Jacobian<F> msm_finish(const Xyzz<F> *planes, uint32_t c)
{
    Xyzz<F> acc = planes[c - 1];
    for (int j = (int)c - 2; j >= 0; --j) {
        acc = xyzz_dbl(acc);
        acc = xyzz_add(acc, planes[j]);
    }
    return xyzz_to_jacobian_affine(acc);
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

}  // namespace nkr

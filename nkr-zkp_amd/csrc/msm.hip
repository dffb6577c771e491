// Pippenger MSM kernels for gfx950 (see msm.h for the design).
#include "msm.h"

#include <algorithm>

namespace nkr {

namespace {

constexpr uint32_t kEntryMask = 0x7fffffffu;
constexpr uint32_t kPieceT0 = 16;   // entries per thread when summing bases into buckets
constexpr uint32_t kPieceT1 = 8;    // partials per thread in the folding levels
constexpr uint32_t kPlaneT = 16;    // buckets per thread when forming bit planes
constexpr uint32_t kSortBlocksMax = 1024;

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
__global__ __launch_bounds__(256) void k_hist(const fr_t *__restrict__ scalars, uint32_t n, uint32_t c, uint32_t W,
                                              uint32_t B, uint32_t per_block, uint32_t *__restrict__ count,
                                              uint32_t *__restrict__ block_base)
{
    extern __shared__ uint32_t hist[];
    for (uint32_t k = threadIdx.x; k < B; k += blockDim.x) hist[k] = 0;
    __syncthreads();
    const uint32_t beg = blockIdx.x * per_block;
    const uint32_t end = min(n, beg + per_block);
    for (uint32_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
        const fr_t s = fp_from_mont(scalars[i]);
        for_each_digit(s, c, W, [&](uint32_t, int32_t d) {
            const uint32_t m = (uint32_t)(d < 0 ? -d : d);
            atomicAdd(&hist[m - 1], 1u);
        });
    }
    __syncthreads();
    uint32_t *mine = block_base + (size_t)blockIdx.x * B;
    for (uint32_t k = threadIdx.x; k < B; k += blockDim.x) {
        const uint32_t h = hist[k];
        mine[k] = h ? atomicAdd(&count[k], h) : 0u;
    }
}

__global__ __launch_bounds__(256) void k_scatter(const fr_t *__restrict__ scalars, uint32_t n, uint32_t c, uint32_t W,
                                                 uint32_t B, uint32_t per_block,
                                                 const uint32_t *__restrict__ bucket_start,
                                                 const uint32_t *__restrict__ block_base, uint32_t *__restrict__ sorted)
{
    extern __shared__ uint32_t cursor[];
    const uint32_t *mine = block_base + (size_t)blockIdx.x * B;
    for (uint32_t k = threadIdx.x; k < B; k += blockDim.x) cursor[k] = bucket_start[k] + mine[k];
    __syncthreads();
    const uint32_t beg = blockIdx.x * per_block;
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

// Exclusive prefix over `len` values into out[0..len]; optional maximum.
//   mode 0: v[k] = src[k]
//   mode 1: v[k] = ceil((src[k+1] - src[k]) / T)     (next folding level)
//   mode 2: v[k] = ceil(src[k] / T)                   (pieces of a bucket)
__global__ __launch_bounds__(1024) void k_scan(const uint32_t *__restrict__ src, uint32_t len, uint32_t mode,
                                               uint32_t T, uint32_t *__restrict__ out,
                                               uint32_t *__restrict__ max_out)
{
    __shared__ uint32_t part[1024];
    __shared__ uint32_t pmax[1024];
    auto value = [&](uint32_t k) -> uint32_t {
        if (mode == 0) return src[k];
        if (mode == 1) return (src[k + 1] - src[k] + T - 1) / T;
        return (src[k] + T - 1) / T;
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
    pmax[threadIdx.x] = mx;
    __syncthreads();
    // Hillis-Steele inclusive scan of the per-thread sums.
    for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
        const uint32_t t = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        const uint32_t m = threadIdx.x >= off ? pmax[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += t;
        pmax[threadIdx.x] = max(pmax[threadIdx.x], m);
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum;
    for (uint32_t k = beg; k < end; ++k) {
        out[k] = run;
        run += value(k);
    }
    if (threadIdx.x == blockDim.x - 1) {
        out[len] = part[threadIdx.x];
        if (max_out) *max_out = pmax[threadIdx.x];
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

// Level 0: one thread sums up to T signed bases of one bucket.
template <class F>
__global__ __launch_bounds__(256) void k_accumulate(const uint32_t *__restrict__ sorted,
                                                    const uint32_t *__restrict__ bucket_start,
                                                    const uint32_t *__restrict__ piece_start, uint32_t B, uint32_t T,
                                                    const Affine<F> *__restrict__ bases, Xyzz<F> *__restrict__ out)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= piece_start[B]) return;
    const uint32_t k = find_segment(piece_start, B, p);
    const uint32_t q = p - piece_start[k];
    const uint32_t beg = bucket_start[k] + q * T;
    const uint32_t end = min(beg + T, bucket_start[k + 1]);
    Xyzz<F> acc = Xyzz<F>::infinity();
    for (uint32_t e = beg; e < end; ++e) {
        const uint32_t v = sorted[e];
        const Affine<F> P = bases[v & kEntryMask];
        if (affine_is_inf(P)) continue;
        const F y = (v >> 31) ? FieldOps<F>::neg(P.y) : P.y;
        acc = xyzz_add_affine(acc, P.x, y);
    }
    out[p] = acc;
}

// Folding level: one thread sums up to T consecutive partials of one segment.
template <class F>
__global__ __launch_bounds__(256) void k_seg_reduce(const Xyzz<F> *__restrict__ in,
                                                    const uint32_t *__restrict__ in_start,
                                                    const uint32_t *__restrict__ out_start, uint32_t S, uint32_t T,
                                                    Xyzz<F> *__restrict__ out)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= out_start[S]) return;
    const uint32_t k = find_segment(out_start, S, p);
    const uint32_t q = p - out_start[k];
    const uint32_t beg = in_start[k] + q * T;
    const uint32_t end = min(beg + T, in_start[k + 1]);
    Xyzz<F> acc = in[beg];
    for (uint32_t e = beg + 1; e < end; ++e) acc = xyzz_add(acc, in[e]);
    out[p] = acc;
}

// Bit planes, first level: plane j gathers bucket sums S_m for m with bit j.
template <class F>
__global__ __launch_bounds__(256) void k_planes(const Xyzz<F> *__restrict__ part, const uint32_t *__restrict__ start,
                                                uint32_t c, uint32_t T, const uint32_t *__restrict__ pstart,
                                                Xyzz<F> *__restrict__ out)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pstart[c]) return;
    const uint32_t j = find_segment(pstart, c, p);
    const uint32_t r = p - pstart[j];
    const uint32_t cnt = (j == c - 1) ? 1u : (1u << (c - 2));
    Xyzz<F> acc = Xyzz<F>::infinity();
    for (uint32_t u = 0; u < T; ++u) {
        const uint32_t t = r * T + u;
        if (t >= cnt) break;
        const uint32_t m = (j == c - 1) ? (1u << (c - 1))
                                        : (((t >> j) << (j + 1)) | (1u << j) | (t & ((1u << j) - 1u)));
        const uint32_t k = m - 1;
        if (start[k + 1] > start[k]) acc = xyzz_add(acc, part[start[k]]);
    }
    out[p] = acc;
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
    }
    NKR_REQUIRE(c >= 2 && c <= 15, "MSM window must be 2..15 bits");
    p.c = c;
    p.W = (255 + c - 1) / c;
    p.B = 1u << (c - 1);
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

void MsmSchedule::build(const fr_t *scalars_mont, const MsmPlan &plan, hipStream_t stream, StageTimer *timer)
{
    if (timer) timer->begin(kStageSort, stream);
    plan_ = plan;
    const uint32_t n = plan.n, B = plan.B;
    blocks_ = std::min<uint32_t>(kSortBlocksMax, div_up(n, 256));
    const uint32_t per_block = div_up(n, blocks_);
    sorted_.ensure((size_t)n * plan.W);
    count_.ensure(B);
    bucket_start_.ensure(B + 1);
    block_base_.ensure((size_t)blocks_ * B);
    max_count_.ensure(1);
    const size_t lds = (size_t)B * sizeof(uint32_t);
    k_zero_u32<<<div_up(B, 256), 256, 0, stream>>>(count_.get(), B);
    NKR_LAUNCH_CHECK();
    k_hist<<<blocks_, 256, lds, stream>>>(scalars_mont, n, plan.c, plan.W, B, per_block, count_.get(),
                                          block_base_.get());
    NKR_LAUNCH_CHECK();
    k_scan<<<1, 1024, 0, stream>>>(count_.get(), B, 0, 1, bucket_start_.get(), max_count_.get());
    NKR_LAUNCH_CHECK();
    k_scatter<<<blocks_, 256, lds, stream>>>(scalars_mont, n, plan.c, plan.W, B, per_block, bucket_start_.get(),
                                             block_base_.get(), sorted_.get());
    NKR_LAUNCH_CHECK();
    if (timer) timer->end(kStageSort, stream);
    NKR_HIP(hipMemcpyAsync(&host_max_count_, max_count_.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    NKR_HIP(hipMemcpyAsync(&host_total_, bucket_start_.get() + B, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    NKR_HIP(hipStreamSynchronize(stream));
}

template <class F>
void MsmRunner<F>::run(const MsmBases<F> &bases, const MsmSchedule &sched, Xyzz<F> *planes, hipStream_t stream,
                       StageTimer *timer)
{
    const MsmPlan &p = sched.plan();
    NKR_REQUIRE(bases.plan().n == p.n && bases.plan().c == p.c, "MSM bases and schedule disagree");
    const uint32_t B = p.B;
    const uint64_t ub0 = (uint64_t)p.n * p.W / kPieceT0 + B + 1;
    const uint64_t ub1 = ub0 / kPieceT1 + B + 1;
    part_[0].ensure(ub0);
    part_[1].ensure(ub1);
    start_[0].ensure(B + 1);
    start_[1].ensure(B + 1);

    // level 0: bases into bucket pieces
    k_scan<<<1, 1024, 0, stream>>>(sched.count_.get(), B, 2, kPieceT0, start_[0].get(), nullptr);
    NKR_LAUNCH_CHECK();
    if (timer) timer->begin(kStageAccumulate, stream);
    k_accumulate<F><<<div_up(ub0, 256), 256, 0, stream>>>(sched.sorted_.get(), sched.bucket_start_.get(),
                                                          start_[0].get(), B, kPieceT0, bases.data(),
                                                          part_[0].get());
    NKR_LAUNCH_CHECK();
    if (timer) {
        timer->end(kStageAccumulate, stream);
        timer->begin(kStageFold, stream);
    }
    // folding levels until each bucket has one partial
    uint64_t maxp = (sched.host_max_count_ + kPieceT0 - 1) / kPieceT0;
    uint64_t ub = ub0;
    int cur = 0;
    while (maxp > 1) {
        const int nxt = cur ^ 1;
        const uint64_t ubn = ub / kPieceT1 + B + 1;
        NKR_REQUIRE(ubn <= part_[nxt].size(), "MSM folding buffer too small");
        k_scan<<<1, 1024, 0, stream>>>(start_[cur].get(), B, 1, kPieceT1, start_[nxt].get(), nullptr);
        NKR_LAUNCH_CHECK();
        k_seg_reduce<F><<<div_up(ubn, 256), 256, 0, stream>>>(part_[cur].get(), start_[cur].get(),
                                                              start_[nxt].get(), B, kPieceT1, part_[nxt].get());
        NKR_LAUNCH_CHECK();
        cur = nxt;
        ub = ubn;
        maxp = (maxp + kPieceT1 - 1) / kPieceT1;
    }

    // bit planes U_j, j < c: static segment layout, folded to one point each
    const uint32_t c = p.c;
    if (plane_levels_.empty() || plane_levels_[0] != c) {
        // layout: [c, levels, then (c+1) offsets per level]
        std::vector<std::vector<uint32_t>> lv;
        std::vector<uint32_t> cnt(c);
        for (uint32_t j = 0; j < c; ++j) cnt[j] = (j == c - 1) ? 1u : (1u << (c - 2));
        std::vector<uint32_t> pieces(c);
        for (uint32_t j = 0; j < c; ++j) pieces[j] = (cnt[j] + kPlaneT - 1) / kPlaneT;
        for (;;) {
            std::vector<uint32_t> st(c + 1, 0);
            for (uint32_t j = 0; j < c; ++j) st[j + 1] = st[j] + pieces[j];
            lv.push_back(st);
            if (*std::max_element(pieces.begin(), pieces.end()) <= 1) break;
            for (uint32_t j = 0; j < c; ++j) pieces[j] = (pieces[j] + kPieceT1 - 1) / kPieceT1;
        }
        plane_levels_.assign({c, (uint32_t)lv.size()});
        for (auto &st : lv) plane_levels_.insert(plane_levels_.end(), st.begin(), st.end());
        plane_start_.allocate(plane_levels_.size());
        NKR_HIP(hipMemcpyAsync(plane_start_.get(), plane_levels_.data(), plane_levels_.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice, stream));
        sums_.allocate(2 * (size_t)lv[0][c]);
    }
    const uint32_t levels = plane_levels_[1];
    const uint32_t *dev_levels = plane_start_.get() + 2;
    const uint32_t *host_levels = plane_levels_.data() + 2;
    Xyzz<F> *bufs[2] = {sums_.get(), sums_.get() + sums_.size() / 2};
    {
        Xyzz<F> *dst = (levels == 1) ? planes : bufs[0];
        const uint32_t total = host_levels[c];
        k_planes<F><<<div_up(total, 256), 256, 0, stream>>>(part_[cur].get(), start_[cur].get(), c, kPlaneT,
                                                            dev_levels, dst);
        NKR_LAUNCH_CHECK();
    }
    for (uint32_t L = 1; L < levels; ++L) {
        const uint32_t *in_st = dev_levels + (size_t)(L - 1) * (c + 1);
        const uint32_t *out_st = dev_levels + (size_t)L * (c + 1);
        const uint32_t total = host_levels[(size_t)L * (c + 1) + c];
        Xyzz<F> *src = bufs[(L - 1) & 1];
        Xyzz<F> *dst = (L == levels - 1) ? planes : bufs[L & 1];
        k_seg_reduce<F><<<div_up(total, 256), 256, 0, stream>>>(src, in_st, out_st, c, kPieceT1, dst);
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
    runner_.run(bases_, sched_, planes_.get(), stream);
    std::vector<Xyzz<F>> host(plan.c);
    NKR_HIP(hipMemcpyAsync(host.data(), planes_.get(), plan.c * sizeof(Xyzz<F>), hipMemcpyDeviceToHost, stream));
    NKR_HIP(hipStreamSynchronize(stream));
    return msm_finish(host.data(), plan.c);
}

template class MsmBases<fq_t>;
template class MsmBases<fq2_t>;
template class MsmRunner<fq_t>;
template class MsmRunner<fq2_t>;
template class MsmContext<fq_t>;
template class MsmContext<fq2_t>;
template Jacobian<fq_t> msm_finish(const Xyzz<fq_t> *, uint32_t);
template Jacobian<fq2_t> msm_finish(const Xyzz<fq2_t> *, uint32_t);

}  // namespace nkr

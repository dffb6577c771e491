// Packed witness upload: host packing on worker threads, chunked PCIe copies,
// device expansion.  The design notes are in upload.h.
#include "upload.h"

#include <sched.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>

#include "host_pool.h"

namespace nkr {

namespace {

// Element i of the packed index space: a lane finds its value by a popcount
// of its group's mask below it.
__device__ __forceinline__ fr_t expand_one(uint32_t i, const fr_t *__restrict__ pack,
                                           const uint64_t *__restrict__ masks, const uint32_t *__restrict__ offsets)
{
    const uint32_t g = i >> 6;
    const uint64_t bit = 1ull << (i & 63);
    const uint64_t other = masks[2 * g], one = masks[2 * g + 1];
    fr_t v = fr_t::zero();
    if (other & bit)
        v = pack[offsets[g] + __popcll(other & (bit - 1))];
    else if (one & bit)
        v = fr_t::one();
    return v;
}

// One wave expands one group of 64 elements.
__global__ void k_expand_witness(fr_t *__restrict__ z, const fr_t *__restrict__ pack,
                                 const uint64_t *__restrict__ masks, const uint32_t *__restrict__ offsets,
                                 uint32_t cols)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cols) return;
    z[i] = expand_one(i, pack, masks, offsets);
}

// Row y of the grid expands chunk y into its witness's device array.
__global__ void k_expand_witness_chunks(const WitnessBatchUpload::Chunk *__restrict__ chunks,
                                        const fr_t *__restrict__ pack, const uint64_t *__restrict__ masks,
                                        const uint32_t *__restrict__ offsets)
{
    const WitnessBatchUpload::Chunk ch = chunks[blockIdx.y];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ch.count) return;
    ch.z[i] = expand_one(ch.begin + i, pack, masks, offsets);
}

// Host threads that pack chunks, shared by every context of the process
// (uploads take turns).  At most 8 threads, fewer on a host that lets this
// process run on fewer CPUs.
struct PackPool {
    std::mutex m;
    std::unique_ptr<HostHelpers> helpers;
    PackPool()
    {
        int cpus = 1;
        cpu_set_t set;
        if (sched_getaffinity(0, sizeof(set), &set) == 0) cpus = CPU_COUNT(&set);
        helpers = std::make_unique<HostHelpers>(std::max(1, std::min(8, cpus - 1)));
    }
};

PackPool &pack_pool()
{
    // Never destroyed: its threads sleep until the process ends.
    static PackPool *pool = new PackPool();
    return *pool;
}

}  // namespace

WitnessUpload::WitnessUpload(uint32_t cols) : cols_(cols)
{
    NKR_REQUIRE(cols >= 1, "empty witness");
    groups_ = (cols + kGroup - 1) / kGroup;
    chunks_ = (cols + kChunk - 1) / kChunk;
    meta_bytes_ = (size_t)groups_ * (2 * sizeof(uint64_t) + sizeof(uint32_t));
    void *p = nullptr;
    NKR_HIP(hipHostMalloc(&p, (size_t)cols * sizeof(fr_t), hipHostMallocDefault));
    vals_host_ = static_cast<fr_t *>(p);
    p = nullptr;
    const hipError_t e = hipHostMalloc(&p, meta_bytes_, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipHostFree(vals_host_);
        NKR_HIP(e);
    }
    meta_host_ = static_cast<uint8_t *>(p);
    chunk_count_ = new std::atomic<uint32_t>[chunks_];
    pack_dev_.allocate(cols);
    meta_dev_.allocate(meta_bytes_);
    NKR_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
}

WitnessUpload::~WitnessUpload()
{
    if (done_) {
        (void)hipEventSynchronize(done_);
        (void)hipEventDestroy(done_);
    }
    if (vals_host_) (void)hipHostFree(vals_host_);
    if (meta_host_) (void)hipHostFree(meta_host_);
    delete[] chunk_count_;
}

void WitnessUpload::upload(const fr_t *host, fr_t *z_dev, hipStream_t stream)
{
    NKR_REQUIRE(host && z_dev, "null witness");
    // The pinned buffers are rewritten below: the previous upload's copies
    // must have read them.
    NKR_HIP(hipEventSynchronize(done_));
    for (uint32_t c = 0; c < chunks_; ++c) chunk_count_[c].store(0, std::memory_order_relaxed);
    uint64_t *masks = reinterpret_cast<uint64_t *>(meta_host_);
    uint32_t *offsets = reinterpret_cast<uint32_t *>(meta_host_ + (size_t)groups_ * 2 * sizeof(uint64_t));

    const fr_t one_fr = fr_t::one();
    uint64_t one[4];
    std::memcpy(one, &one_fr, sizeof(one));
    PackPool &pool = pack_pool();
    std::lock_guard<std::mutex> lock(pool.m);
    HostHelpers &h = *pool.helpers;
    std::atomic<uint32_t> next{0};
    auto work = [&] {
        for (;;) {
            const uint32_t c = next.fetch_add(1, std::memory_order_relaxed);
            if (c >= chunks_) return;
            const uint32_t b = c * kChunk, e = std::min(cols_, b + kChunk);
            const uint32_t n = witness_pack(host, b, e, one, vals_host_, masks, offsets);
            chunk_count_[c].store(n + 1, std::memory_order_release);
        }
    };
    const int workers = (int)std::min<uint32_t>((uint32_t)h.size(), chunks_);
    for (int i = 0; i < workers; ++i) h.run(i, work);

    // Each chunk's values go over PCIe as soon as it is packed (chunks are
    // taken in order, so they finish roughly in order).
    struct Join {
        HostHelpers &h;
        int n;
        ~Join()
        {
            for (int i = 0; i < n; ++i) {
                try {
                    h.wait(i);
                } catch (...) {
                }
            }
        }
    } join{h, workers};
    uint64_t bytes = meta_bytes_;
    for (uint32_t c = 0; c < chunks_; ++c) {
        uint32_t v;
        while ((v = chunk_count_[c].load(std::memory_order_acquire)) == 0) cpu_relax();
        const size_t n = v - 1, at = (size_t)c * kChunk;
        if (n) {
            NKR_HIP(hipMemcpyAsync(pack_dev_.get() + at, vals_host_ + at, n * sizeof(fr_t), hipMemcpyHostToDevice,
                                   stream));
            bytes += n * sizeof(fr_t);
        }
    }
    NKR_HIP(hipMemcpyAsync(meta_dev_.get(), meta_host_, meta_bytes_, hipMemcpyHostToDevice, stream));
    NKR_HIP(hipEventRecord(done_, stream));
    const uint64_t *masks_dev = reinterpret_cast<const uint64_t *>(meta_dev_.get());
    const uint32_t *offsets_dev =
        reinterpret_cast<const uint32_t *>(meta_dev_.get() + (size_t)groups_ * 2 * sizeof(uint64_t));
    k_expand_witness<<<div_up(cols_, 256), 256, 0, stream>>>(z_dev, pack_dev_.get(), masks_dev, offsets_dev, cols_);
    NKR_LAUNCH_CHECK();
    last_bytes_ = bytes;
}

void WitnessBatchUpload::release_device()
{
    pack_dev_.release();
    meta_dev_.release();
}

void WitnessBatchUpload::upload(const std::vector<Item> &items, hipStream_t stream)
{
    if (items.empty()) return;
    // Layout: each witness from a multiple of 64 in the shared index space,
    // cut into chunks of at most kChunk elements.
    constexpr uint32_t kChunk = WitnessUpload::kChunk;
    std::vector<Chunk> chunks;
    std::vector<std::pair<const fr_t *, uint32_t>> source;  // per chunk: the host witness and its first index
    uint64_t total = 0;
    for (const Item &it : items) {
        NKR_REQUIRE(it.host && it.z_dev && it.cols >= 1, "empty or null witness in a batch upload");
        const uint64_t base = total;
        for (uint32_t at = 0; at < it.cols; at += kChunk) {
            chunks.push_back({it.z_dev + at, (uint32_t)(base + at), std::min(kChunk, it.cols - at)});
            source.emplace_back(it.host, (uint32_t)base);
        }
        total = (base + it.cols + kWitnessGroup - 1) / kWitnessGroup * kWitnessGroup;
        NKR_REQUIRE(total < (1ull << 32), "the witnesses of a batch exceed 2^32 elements");
    }
    const uint32_t C = (uint32_t)chunks.size();
    NKR_REQUIRE(C <= 65535, "the witnesses of a batch exceed 65535 chunks");
    const size_t groups = total / kWitnessGroup;
    const size_t table_at = (groups * (2 * sizeof(uint64_t) + sizeof(uint32_t)) + 15) & ~size_t(15);
    const size_t meta_bytes = table_at + C * sizeof(Chunk);

    // The pinned buffers are rewritten below: reuse() waits for the previous
    // upload's copies to have read them.
    fr_t *vals = vals_.reuse(total);
    uint8_t *meta = meta_.reuse(meta_bytes);
    pack_dev_.ensure(total);
    meta_dev_.ensure(meta_bytes);
    if (chunk_cap_ < C) {
        chunk_count_.reset(new std::atomic<uint32_t>[C]);
        chunk_cap_ = C;
    }
    for (uint32_t c = 0; c < C; ++c) chunk_count_[c].store(0, std::memory_order_relaxed);
    uint64_t *masks = reinterpret_cast<uint64_t *>(meta);
    uint32_t *offsets = reinterpret_cast<uint32_t *>(meta + groups * 2 * sizeof(uint64_t));
    std::memcpy(meta + table_at, chunks.data(), C * sizeof(Chunk));

    const fr_t one_fr = fr_t::one();
    uint64_t one[4];
    std::memcpy(one, &one_fr, sizeof(one));
    PackPool &pool = pack_pool();
    std::lock_guard<std::mutex> lock(pool.m);
    HostHelpers &h = *pool.helpers;
    std::atomic<uint32_t> next{0};
    auto work = [&] {
        for (;;) {
            const uint32_t c = next.fetch_add(1, std::memory_order_relaxed);
            if (c >= C) return;
            const Chunk &ch = chunks[c];
            const uint32_t n = witness_pack(source[c].first, ch.begin, ch.begin + ch.count, one, vals, masks, offsets,
                                            source[c].second);
            chunk_count_[c].store(n + 1, std::memory_order_release);
        }
    };
    const int workers = (int)std::min<uint32_t>((uint32_t)h.size(), C);
    for (int i = 0; i < workers; ++i) h.run(i, work);
    struct Join {
        HostHelpers &h;
        int n;
        ~Join()
        {
            for (int i = 0; i < n; ++i) {
                try {
                    h.wait(i);
                } catch (...) {
                }
            }
        }
    } join{h, workers};
    // each chunk's values go over PCIe as soon as it is packed
    uint64_t bytes = meta_bytes;
    for (uint32_t c = 0; c < C; ++c) {
        uint32_t v;
        while ((v = chunk_count_[c].load(std::memory_order_acquire)) == 0) cpu_relax();
        const size_t n = v - 1, at = chunks[c].begin;
        if (n) {
            NKR_HIP(hipMemcpyAsync(pack_dev_.get() + at, vals + at, n * sizeof(fr_t), hipMemcpyHostToDevice, stream));
            bytes += n * sizeof(fr_t);
        }
    }
    NKR_HIP(hipMemcpyAsync(meta_dev_.get(), meta, meta_bytes, hipMemcpyHostToDevice, stream));
    vals_.uploaded(stream);
    meta_.uploaded(stream);
    const uint64_t *masks_dev = reinterpret_cast<const uint64_t *>(meta_dev_.get());
    const uint32_t *offsets_dev = reinterpret_cast<const uint32_t *>(meta_dev_.get() + groups * 2 * sizeof(uint64_t));
    const Chunk *table_dev = reinterpret_cast<const Chunk *>(meta_dev_.get() + table_at);
    k_expand_witness_chunks<<<dim3(kChunk / 256, C), 256, 0, stream>>>(table_dev, pack_dev_.get(), masks_dev,
                                                                      offsets_dev);
    NKR_LAUNCH_CHECK();
    last_bytes_ = bytes;
}

}  // namespace nkr

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

// One wave expands one group of 64 elements.
__global__ void k_expand_witness(fr_t *__restrict__ z, const fr_t *__restrict__ pack,
                                 const uint64_t *__restrict__ masks, const uint32_t *__restrict__ offsets,
                                 uint32_t cols)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cols) return;
    const uint32_t g = i >> 6;
    const uint64_t bit = 1ull << (i & 63);
    const uint64_t other = masks[2 * g], one = masks[2 * g + 1];
    fr_t v = fr_t::zero();
    if (other & bit)
        v = pack[offsets[g] + __popcll(other & (bit - 1))];
    else if (one & bit)
        v = fr_t::one();
    z[i] = v;
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

}  // namespace nkr

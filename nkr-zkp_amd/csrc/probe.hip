// Calibration kernels for the integer-VALU roof of the field arithmetic.
//
// There is no published integer multiply-add rate for MI355X, so the roof the
// benchmark grades its VALU-bound kernels against is measured here:
//   * k_mad_rate: eight independent v_mad_u64_u32 chains per lane, enough
//     waves for every SIMD -- the issue rate of the 32x32+64-bit multiply-add
//     that every Montgomery product is built from (136 per Fr/Fq product in
//     the textbook count: 2 x 8^2 + 8);
//   * k_mul_rate: two independent chains of this library's Fr product
//     (fp_mul, the product every kernel uses) per lane -- the rate the field
//     arithmetic actually sustains.
#include "probe.h"

#include <algorithm>

namespace nkr {
namespace {

constexpr uint32_t kProbeThreads = 256;

__global__ void __launch_bounds__(kProbeThreads) k_mad_rate(uint64_t *out, uint32_t iters)
{
    const uint32_t gid = blockIdx.x * kProbeThreads + threadIdx.x;
    const uint32_t x = gid * 2654435761u + 1u, y = (gid ^ 0x9E3779B9u) | 1u;
    uint64_t acc[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = gid + j;
    for (uint32_t k = 0; k < iters; ++k) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            uint64_t cc;
            asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(acc[j]), "=s"(cc) : "v"(x), "v"(y));
        }
    }
    uint64_t s = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) s ^= acc[j];
    out[gid] = s;
}

__global__ void __launch_bounds__(kProbeThreads) k_mul_rate(fr_t *out, uint32_t iters)
{
    const uint32_t gid = blockIdx.x * kProbeThreads + threadIdx.x;
    fr_t a = fr_t::one(), b = fr_t::one(), c = fr_t::one();
    a.l[0] ^= gid;
    b.l[1] ^= gid * 7u;
    c.l[2] ^= gid * 13u;
    for (uint32_t k = 0; k < iters; ++k) {
        a = fp_mul(a, c);
        b = fp_mul(b, c);
    }
    out[gid] = fp_add(a, b);
}

}  // namespace
}  // namespace nkr

namespace nkr {

void valu_probe(int device, double *mad_per_s, double *fr_mul_per_s)
{
    NKR_REQUIRE(mad_per_s && fr_mul_per_s, "null argument");
    int count = 0;
    NKR_HIP(hipGetDeviceCount(&count));
    NKR_REQUIRE(device >= 0 && device < count, "invalid device index");
    NKR_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    NKR_HIP(hipGetDeviceProperties(&prop, device));
    // 8 waves of 64 lanes per SIMD (4 SIMDs per CU): every SIMD has
    // independent work to issue every cycle.
    const uint32_t blocks = (uint32_t)prop.multiProcessorCount * 8;
    const size_t threads = (size_t)blocks * kProbeThreads;
    DeviceBuffer<fr_t> out(threads);
    hipEvent_t ev[2] = {};
    struct Events {
        hipEvent_t *e;
        ~Events()
        {
            for (int i = 0; i < 2; ++i)
                if (e[i]) (void)hipEventDestroy(e[i]);
        }
    } guard{ev};
    NKR_HIP(hipEventCreate(&ev[0]));
    NKR_HIP(hipEventCreate(&ev[1]));
    auto best_ms = [&](auto launch) {
        float best = 1e30f;
        for (int rep = 0; rep < 4; ++rep) {  // the first run warms up
            NKR_HIP(hipEventRecord(ev[0], nullptr));
            launch();
            NKR_LAUNCH_CHECK();
            NKR_HIP(hipEventRecord(ev[1], nullptr));
            NKR_HIP(hipEventSynchronize(ev[1]));
            float ms = 0;
            NKR_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            if (rep > 0) best = std::min(best, ms);
        }
        return (double)best;
    };
    const uint32_t mad_iters = 4096, mul_iters = 128;
    const double t_mad = best_ms(
        [&] { k_mad_rate<<<blocks, kProbeThreads>>>(reinterpret_cast<uint64_t *>(out.get()), mad_iters); });
    const double t_mul = best_ms([&] { k_mul_rate<<<blocks, kProbeThreads>>>(out.get(), mul_iters); });
    *mad_per_s = (double)threads * mad_iters * 8 / (t_mad * 1e-3);
    *fr_mul_per_s = (double)threads * mul_iters * 2 / (t_mul * 1e-3);
}

}  // namespace nkr

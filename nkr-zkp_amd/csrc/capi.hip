// extern "C" boundary (include/nkr_g16.h).  Every entry point converts
// exceptions into a non-zero status plus a thread-local message; nothing
// aborts across the ABI (the reference's HIP_CHECK aborts,
// operators/hip/include/utils.hip:22-31).
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nkr_g16.h"
#include "batch.h"
#include "keygen.h"
#include "msm.h"
#include "ntt.h"
#include "pairing.h"
#include "probe.h"
#include "prover.h"
#include "serve.h"
#include "setup.h"
#include "spmv.h"

namespace {

thread_local std::string g_last_error;

constexpr int kMaxDevices = 16;
std::mutex g_ntt_mutex;
nkr::NttDomain *g_ntt_cache[kMaxDevices][29] = {};
nkr::DeviceBuffer<nkr::fr_t> *g_ntt_scratch[kMaxDevices] = {};

// Device scratch of the allocation-free transform (nkr_fr_ntt_dev_noalloc):
// the table arena of a 2^20-point domain followed by a 2^20-element buffer
// for the final reordering -- 2^21 + 6144 elements of 32 B (64.2 MiB), one
// copy per device for the life of the process.  Host code never names
// `scratch` in a kernel argument or copy: in host code the name is a stand-in,
// not a device address.  Each device's address comes from
// hipGetSymbolAddress (noalloc_scratch) with that device current.
constexpr uint32_t kNoAllocMaxLog = 20;
constexpr size_t kNoAllocArenaElems = nkr::NttDomain::standalone_arena_elems(kNoAllocMaxLog);
constexpr size_t kNoAllocElems = kNoAllocArenaElems + (size_t(1) << kNoAllocMaxLog);
static __device__ __align__(256) uint4 scratch[kNoAllocElems * sizeof(nkr::fr_t) / sizeof(uint4)];
std::mutex g_noalloc_mutex;
nkr::fr_t *g_noalloc_ptr[kMaxDevices] = {};

template <class Fn>
int guarded(Fn &&fn)
{
    try {
        fn();
        g_last_error.clear();
        return 0;
    } catch (const std::exception &e) {
        g_last_error = e.what();
    } catch (...) {
        g_last_error = "unknown error";
    }
    return 1;
}

void set_device(int device)
{
    int count = 0;
    NKR_HIP(hipGetDeviceCount(&count));
    NKR_REQUIRE(device >= 0 && device < count, "invalid device index " + std::to_string(device));
    NKR_HIP(hipSetDevice(device));
}

using namespace nkr;

__global__ void k_field_op(int op, const uint8_t *a, const uint8_t *b, uint8_t *out, uint64_t n)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (op <= 2 || op == 8) {
        const fr_t x = reinterpret_cast<const fr_t *>(a)[i];
        const fr_t y = (op == 8) ? x : reinterpret_cast<const fr_t *>(b)[i];
        fr_t r;
        if (op == 0) r = fp_mul(x, y);
        else if (op == 1) r = fp_add(x, y);
        else if (op == 2) r = fp_sub(x, y);
        else r = fp_inv(x);
        reinterpret_cast<fr_t *>(out)[i] = r;
    } else if (op <= 5) {
        const fq_t x = reinterpret_cast<const fq_t *>(a)[i];
        const fq_t y = reinterpret_cast<const fq_t *>(b)[i];
        fq_t r;
        if (op == 3) r = fp_mul(x, y);
        else if (op == 4) r = fp_add(x, y);
        else r = fp_sub(x, y);
        reinterpret_cast<fq_t *>(out)[i] = r;
    } else {
        const fq2_t x = reinterpret_cast<const fq2_t *>(a)[i];
        if (op == 6) {
            const fq2_t y = reinterpret_cast<const fq2_t *>(b)[i];
            reinterpret_cast<fq2_t *>(out)[i] = fq2_mul(x, y);
        } else {
            reinterpret_cast<fq2_t *>(out)[i] = fq2_sqr(x);
        }
    }
}

__global__ void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

struct nkr_g16_ctx {
    std::unique_ptr<nkr::Groth16Prover> prover;
};

struct nkr_g16_batch {
    std::unique_ptr<nkr::BatchProver> prover;
};

struct nkr_g16_server {
    std::unique_ptr<nkr::Server> server;
};

struct nkr_msm {
    int device = 0;
    bool g2 = false;
    uint64_t n = 0;
    hipStream_t stream = nullptr;
    std::unique_ptr<nkr::MsmContext<nkr::fq_t>> g1;
    std::unique_ptr<nkr::MsmContext<nkr::fq2_t>> g2ctx;
    nkr::DeviceBuffer<nkr::fr_t> scalars;
    ~nkr_msm()
    {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

const char *nkr_last_error(void) { return g_last_error.c_str(); }

const char *nkr_version(void) { return "nkr-g16-mi355x 0.1 (gfx950)"; }

int nkr_device_count(int *count)
{
    return guarded([&] {
        NKR_REQUIRE(count, "null output");
        NKR_HIP(hipGetDeviceCount(count));
    });
}

int nkr_g16_ctx_create(const nkr_g16_circuit *circuit, const nkr_g16_pk *pk, int device, nkr_g16_ctx **out)
{
    return guarded([&] {
        NKR_REQUIRE(circuit && pk && out, "null argument");
        *out = nullptr;
        set_device(device);
        auto ctx = std::make_unique<nkr_g16_ctx>();
        ctx->prover = std::make_unique<Groth16Prover>(*circuit, *pk, device);
        *out = ctx.release();
    });
}

int nkr_g16_ctx_create_tuned(const nkr_g16_circuit *circuit, const nkr_g16_pk *pk, int device,
                             const nkr_g16_tuning *tuning, nkr_g16_ctx **out)
{
    return guarded([&] {
        NKR_REQUIRE(circuit && pk && out && tuning, "null argument");
        *out = nullptr;
        set_device(device);
        auto ctx = std::make_unique<nkr_g16_ctx>();
        ctx->prover = std::make_unique<Groth16Prover>(*circuit, *pk, device, tuning);
        *out = ctx.release();
    });
}

int nkr_g16_ctx_destroy(nkr_g16_ctx *ctx)
{
    return guarded([&] { delete ctx; });
}

int nkr_g16_ctx_device_bytes(const nkr_g16_ctx *ctx, uint64_t *bytes)
{
    return guarded([&] {
        NKR_REQUIRE(ctx && bytes, "null argument");
        *bytes = ctx->prover->device_bytes();
    });
}

int nkr_g16_prove(nkr_g16_ctx *ctx, const void *witness, void *raw_out)
{
    return guarded([&] {
        NKR_REQUIRE(ctx, "null context");
        ctx->prover->prove(witness, raw_out, nullptr, nullptr, nullptr, nullptr);
    });
}

int nkr_g16_prove_ex(nkr_g16_ctx *ctx, const void *witness, void *raw_out, void *az, void *bz, void *cz, void *h)
{
    return guarded([&] {
        NKR_REQUIRE(ctx, "null context");
        ctx->prover->prove(witness, raw_out, az, bz, cz, h);
    });
}

int nkr_g16_check_witness(const nkr_g16_circuit *c, const void *witness, int *ok)
{
    return guarded([&] {
        NKR_REQUIRE(c && witness && ok, "null argument");
        const fr_t *z = static_cast<const fr_t *>(witness);
        const uint32_t *rps[3] = {c->a_row_ptr, c->b_row_ptr, c->c_row_ptr};
        const uint32_t *cis[3] = {c->a_col_idx, c->b_col_idx, c->c_col_idx};
        const fr_t *vs[3] = {static_cast<const fr_t *>(c->a_values), static_cast<const fr_t *>(c->b_values),
                             static_cast<const fr_t *>(c->c_values)};
        const uint64_t nnz[3] = {c->a_nnz, c->b_nnz, c->c_nnz};
        for (int m = 0; m < 3; ++m) {
            NKR_REQUIRE(rps[m] && rps[m][0] == 0 && rps[m][c->rows] == nnz[m], "invalid CSR offsets");
            for (uint64_t j = 0; j < nnz[m]; ++j) NKR_REQUIRE(cis[m][j] < c->cols, "out-of-range column");
        }
        *ok = 1;
        for (uint64_t row = 0; row < c->rows && *ok; ++row) {
            fr_t v[3];
            for (int m = 0; m < 3; ++m) {
                v[m] = fr_t::zero();
                for (uint32_t j = rps[m][row]; j < rps[m][row + 1]; ++j)
                    v[m] = fp_add(v[m], fp_mul(vs[m][j], z[cis[m][j]]));
            }
            if (fp_mul(v[0], v[1]) != v[2]) *ok = 0;
        }
    });
}

int nkr_g16_prove_dev(nkr_g16_ctx *ctx, const void *witness_dev, void *raw_out)
{
    return guarded([&] {
        NKR_REQUIRE(ctx, "null context");
        ctx->prover->prove_dev(static_cast<const fr_t *>(witness_dev), raw_out);
    });
}

static int prove_blind(nkr_g16_ctx *ctx, const void *witness, bool dev, const void *r, const void *s,
                       void *raw_out, void *proof_out)
{
    return guarded([&] {
        NKR_REQUIRE(ctx && r && s, "null argument");
        fr_t rr, ss;
        std::memcpy(&rr, r, sizeof(rr));
        std::memcpy(&ss, s, sizeof(ss));
        ctx->prover->prove_blind(witness, dev, rr, ss, raw_out, proof_out);
    });
}

int nkr_g16_prove_blind(nkr_g16_ctx *ctx, const void *witness, const void *r, const void *s, void *raw_out,
                        void *proof_out)
{
    return prove_blind(ctx, witness, false, r, s, raw_out, proof_out);
}

int nkr_g16_prove_blind_dev(nkr_g16_ctx *ctx, const void *witness_dev, const void *r, const void *s,
                            void *raw_out, void *proof_out)
{
    return prove_blind(ctx, witness_dev, true, r, s, raw_out, proof_out);
}

int nkr_g16_batch_create(int device, nkr_g16_batch **out)
{
    return guarded([&] {
        NKR_REQUIRE(out, "null output");
        *out = nullptr;
        set_device(device);
        auto b = std::make_unique<nkr_g16_batch>();
        b->prover = std::make_unique<BatchProver>(device);
        *out = b.release();
    });
}

int nkr_g16_batch_prove(nkr_g16_batch *batch, const nkr_g16_batch_item *items, uint64_t count)
{
    return guarded([&] {
        NKR_REQUIRE(batch && (items || count == 0), "null argument");
        std::vector<BatchItem> v(count);
        for (uint64_t i = 0; i < count; ++i) {
            const nkr_g16_batch_item &it = items[i];
            NKR_REQUIRE(it.ctx && it.r && it.s, "batch item with a null member");
            v[i].ctx = it.ctx->prover.get();
            v[i].witness = it.witness;
            std::memcpy(&v[i].r, it.r, sizeof(fr_t));
            std::memcpy(&v[i].s, it.s, sizeof(fr_t));
            v[i].raw_out = it.raw_out;
            v[i].proof_out = it.proof_out;
        }
        batch->prover->prove(v);
    });
}

int nkr_g16_batch_destroy(nkr_g16_batch *batch)
{
    return guarded([&] { delete batch; });
}

int nkr_g16_batch_set_graphs(nkr_g16_batch *batch, int enabled)
{
    return guarded([&] {
        NKR_REQUIRE(batch, "null argument");
        batch->prover->set_graphs(enabled != 0);
    });
}

int nkr_g16_batch_graph_active(const nkr_g16_batch *batch, int *active)
{
    return guarded([&] {
        NKR_REQUIRE(batch && active, "null argument");
        *active = batch->prover->graph_active() ? 1 : 0;
    });
}

int nkr_g16_server_create(int device, uint32_t workers, uint32_t max_batch, uint64_t work_budget_bytes,
                          nkr_g16_server **out)
{
    return guarded([&] {
        NKR_REQUIRE(out, "null output");
        *out = nullptr;
        set_device(device);
        auto s = std::make_unique<nkr_g16_server>();
        s->server = std::make_unique<Server>(device, workers, max_batch, work_budget_bytes);
        *out = s.release();
    });
}

int nkr_g16_server_submit(nkr_g16_server *server, nkr_g16_ctx *ctx, const void *witness, const void *r,
                          const void *s, void *raw_out, void *proof_out, uint64_t *ticket)
{
    return guarded([&] {
        NKR_REQUIRE(server && ctx && r && s && ticket, "null argument");
        fr_t rr, ss;
        std::memcpy(&rr, r, sizeof(rr));
        std::memcpy(&ss, s, sizeof(ss));
        *ticket = server->server->submit(ctx->prover.get(), witness, rr, ss, raw_out, proof_out);
    });
}

int nkr_g16_server_wait(nkr_g16_server *server, uint64_t ticket, double *latency_ms)
{
    return guarded([&] {
        NKR_REQUIRE(server, "null server");
        const double ms = server->server->wait(ticket);
        if (latency_ms) *latency_ms = ms;
    });
}

int nkr_g16_server_warm(nkr_g16_server *server, const nkr_g16_batch_item *items, uint64_t count)
{
    return guarded([&] {
        NKR_REQUIRE(server && items, "null argument");
        NKR_REQUIRE(count >= 1 && count <= 16384, "a warm-up holds 1 .. 16384 items");
        std::vector<BatchItem> v(count);
        for (uint64_t i = 0; i < count; ++i) {
            const nkr_g16_batch_item &it = items[i];
            NKR_REQUIRE(it.ctx && it.r && it.s, "null warm-up item member");
            fr_t r, s;
            std::memcpy(&r, it.r, sizeof(r));
            std::memcpy(&s, it.s, sizeof(s));
            v[i] = BatchItem{it.ctx->prover.get(), it.witness, r, s, it.raw_out, it.proof_out};
        }
        server->server->warm(v);
    });
}

int nkr_g16_server_request_bytes(const nkr_g16_ctx *ctx, uint64_t *bytes)
{
    return guarded([&] {
        NKR_REQUIRE(ctx && bytes, "null argument");
        *bytes = Server::request_bytes(*ctx->prover);
    });
}

int nkr_g16_server_stats(const nkr_g16_server *server, uint64_t *stats)
{
    return guarded([&] {
        NKR_REQUIRE(server && stats, "null argument");
        const Server::Stats s = server->server->stats();
        stats[0] = s.batches;
        stats[1] = s.proofs;
        stats[2] = s.max_batch_seen;
        stats[3] = s.peak_work_bytes;
    });
}

int nkr_g16_server_memory(const nkr_g16_server *server, uint64_t *memory)
{
    return guarded([&] {
        NKR_REQUIRE(server && memory, "null argument");
        const Server::Memory m = server->server->memory();
        memory[0] = m.budget;
        memory[1] = m.share;
        memory[2] = m.held;
        memory[3] = m.peak_held;
    });
}

int nkr_g16_server_destroy(nkr_g16_server *server)
{
    return guarded([&] { delete server; });
}

int nkr_g16_profile(nkr_g16_ctx *ctx, int enable)
{
    return guarded([&] {
        NKR_REQUIRE(ctx, "null context");
        ctx->prover->set_profiling(enable != 0);
    });
}

int nkr_g16_profile_read(const nkr_g16_ctx *ctx, double *stage_ms, uint64_t *counters)
{
    return guarded([&] {
        NKR_REQUIRE(ctx && stage_ms && counters, "null argument");
        ctx->prover->profile_read(stage_ms, counters);
    });
}

int nkr_g16_blind(const nkr_g16_pk *pk, const void *raw, const void *r, const void *s, void *proof_out)
{
    return guarded([&] {
        NKR_REQUIRE(pk && raw && r && s && proof_out, "null argument");
        fr_t rr, ss;
        std::memcpy(&rr, r, sizeof(rr));
        std::memcpy(&ss, s, sizeof(ss));
        groth16_blind(*pk, raw, rr, ss, proof_out);
    });
}

int nkr_g16_setup(const nkr_g16_circuit *circuit, const void *trapdoors, int device, const nkr_g16_pk_out *pk,
                  const nkr_g16_vk_out *vk)
{
    return guarded([&] {
        NKR_REQUIRE(circuit && trapdoors && pk && vk, "null argument");
        set_device(device);
        Trapdoors td;
        std::memcpy(&td, trapdoors, sizeof(td));
        groth16_keygen(*circuit, td, device, *pk, *vk);
    });
}

int nkr_g16_verify(const nkr_g16_vk *vk, const void *public_input, uint64_t n_public, const void *proof,
                   int *accepted)
{
    return guarded([&] {
        NKR_REQUIRE(vk && proof && accepted, "null argument");
        NKR_REQUIRE(vk->alpha_g1_beta_g2 && vk->gamma_g2 && vk->delta_g2, "verification key has null members");
        NKR_REQUIRE(vk->ic || vk->ic_len == 0, "verification key has a null ic array");
        NKR_REQUIRE(public_input || n_public == 0, "null public input");
        *accepted = 0;
        // copies, so the host arithmetic never reads through unaligned pointers
        fq12_t ab;
        g2_jac gamma, delta;
        std::memcpy(&ab, vk->alpha_g1_beta_g2, sizeof(ab));
        std::memcpy(&gamma, vk->gamma_g2, sizeof(gamma));
        std::memcpy(&delta, vk->delta_g2, sizeof(delta));
        std::vector<g1_jac> ic(vk->ic_len);
        if (vk->ic_len) std::memcpy(ic.data(), vk->ic, vk->ic_len * sizeof(g1_jac));
        std::vector<fr_t> x(n_public);
        if (n_public) std::memcpy(x.data(), public_input, n_public * sizeof(fr_t));
        struct {
            g1_jac a;
            g2_jac b;
            g1_jac c;
        } pr;
        static_assert(sizeof(pr) == 384, "Proof layout");
        std::memcpy(&pr, proof, sizeof(pr));
        const VerifyingKeyView view{&ab, &gamma, &delta, ic.data(), ic.size()};
        *accepted = groth16_verify(view, x.data(), n_public, pr.a, pr.b, pr.c) ? 1 : 0;
    });
}

int nkr_bn254_pairing(const void *g1, const void *g2, void *gt_out)
{
    return nkr_bn254_pairing_product(g1, g2, 1, gt_out);
}

int nkr_bn254_pairing_product(const void *g1s, const void *g2s, uint64_t k, void *gt_out)
{
    return guarded([&] {
        NKR_REQUIRE(gt_out && ((g1s && g2s) || k == 0), "null argument");
        std::vector<g1_jac> p(k);
        std::vector<g2_jac> q(k);
        if (k) {
            std::memcpy(p.data(), g1s, k * sizeof(g1_jac));
            std::memcpy(q.data(), g2s, k * sizeof(g2_jac));
        }
        const fq12_t r = bn254_pairing_product(p.data(), q.data(), k);
        std::memcpy(gt_out, &r, sizeof(r));
    });
}

int nkr_bn254_miller_loop(const void *g1, const void *g2, void *f_out)
{
    return guarded([&] {
        NKR_REQUIRE(g1 && g2 && f_out, "null argument");
        g1_jac p;
        g2_jac q;
        std::memcpy(&p, g1, sizeof(p));
        std::memcpy(&q, g2, sizeof(q));
        const fq12_t f = bn254_miller_loop(p, q);
        std::memcpy(f_out, &f, sizeof(f));
    });
}

int nkr_bn254_final_exponentiation(const void *f, void *gt_out)
{
    return guarded([&] {
        NKR_REQUIRE(f && gt_out, "null argument");
        fq12_t x;
        std::memcpy(&x, f, sizeof(x));
        const fq12_t r = bn254_final_exponentiation(x);
        std::memcpy(gt_out, &r, sizeof(r));
    });
}

int nkr_g16_quotient(const void *az, const void *bz, const void *cz, uint64_t rows, void *h, int device)
{
    return guarded([&] {
        NKR_REQUIRE(az && bz && cz && h, "null argument");
        NKR_REQUIRE(rows >= 2 && is_pow2(rows) && rows <= (1ull << 28), "rows must be a power of two in 2..2^28");
        set_device(device);
        NttDomain dom(ceil_log2(rows), nullptr);
        const size_t bytes = rows * sizeof(fr_t);
        DeviceBuffer<fr_t> a(rows), b(rows), c(rows);
        NKR_HIP(hipMemcpy(a.get(), az, bytes, hipMemcpyHostToDevice));
        NKR_HIP(hipMemcpy(b.get(), bz, bytes, hipMemcpyHostToDevice));
        NKR_HIP(hipMemcpy(c.get(), cz, bytes, hipMemcpyHostToDevice));
        fr_t *polys[3] = {a.get(), b.get(), c.get()};
        for (fr_t *p : polys) dom.dif(p, true, nullptr, nullptr, nullptr, nullptr);
        for (fr_t *p : polys) dom.dit(p, false, dom.coset_rev_scaled(), nullptr, nullptr);
        dom.quotient(a.get(), b.get(), c.get(), a.get(), nullptr);
        dom.dif(a.get(), true, nullptr, dom.icoset_rev_scaled(), nullptr, nullptr);
        dom.bitrev(a.get(), b.get(), nullptr);
        NKR_HIP(hipMemcpy(h, b.get(), bytes, hipMemcpyDeviceToHost));
    });
}

int nkr_fr_ntt(void *data, uint32_t log_n, int kind, int device)
{
    return guarded([&] {
        NKR_REQUIRE(data, "null data");
        NKR_REQUIRE(log_n >= 1 && log_n <= 28, "log_n must be 1..28");
        set_device(device);
        NttDomain dom(log_n, nullptr);
        const size_t n = size_t(1) << log_n;
        DeviceBuffer<fr_t> d(n), scratch(n);
        NKR_HIP(hipMemcpy(d.get(), data, n * sizeof(fr_t), hipMemcpyHostToDevice));
        dom.standalone(d.get(), scratch.get(), kind, nullptr);
        NKR_HIP(hipMemcpy(data, d.get(), n * sizeof(fr_t), hipMemcpyDeviceToHost));
    });
}

int nkr_fr_ntt_dev(void *data_dev, uint32_t log_n, int kind, void *stream)
{
    return guarded([&] {
        NKR_REQUIRE(data_dev, "null data");
        NKR_REQUIRE(log_n >= 1 && log_n <= 28, "log_n must be 1..28");
        hipStream_t s = static_cast<hipStream_t>(stream);
        // Domains are cached per (device, size) for repeated device-side calls.
        // They live until the process ends: freeing device memory from static
        // destructors would race the HIP runtime's own teardown.
        int dev = 0;
        NKR_HIP(hipGetDevice(&dev));
        NKR_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
        std::lock_guard<std::mutex> lock(g_ntt_mutex);
        NttDomain *&dom = g_ntt_cache[dev][log_n];
        if (!dom) dom = new NttDomain(log_n, s);
        DeviceBuffer<fr_t> *&scratch = g_ntt_scratch[dev];
        if (!scratch) scratch = new DeviceBuffer<fr_t>();
        scratch->ensure(size_t(1) << log_n);
        dom->standalone(static_cast<fr_t *>(data_dev), scratch->get(), kind, s);
    });
}

// The current device's copy of `scratch`, looked up once per device.
static fr_t *noalloc_scratch(int dev)
{
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        NKR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        NKR_REQUIRE(p, "no device address for the transform scratch");
        g_noalloc_ptr[dev] = static_cast<fr_t *>(p);
    }
    return g_noalloc_ptr[dev];
}

int nkr_fr_ntt_dev_noalloc(void *data_dev, uint32_t log_n, int kind, void *stream)
{
    return guarded([&] {
        NKR_REQUIRE(data_dev, "null data");
        NKR_REQUIRE(log_n >= 1 && log_n <= kNoAllocMaxLog,
                    "nkr_fr_ntt_dev_noalloc transforms 2^1 .. 2^20 elements (the size of its fixed device "
                    "scratch); use nkr_fr_ntt_dev for larger transforms");
        NKR_REQUIRE(kind >= 0 && kind <= 3, "unknown transform kind (0 FFT, 1 iFFT, 2 cosetFFT, 3 icosetFFT)");
        hipStream_t s = static_cast<hipStream_t>(stream);
        int dev = 0;
        NKR_HIP(hipGetDevice(&dev));
        NKR_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
        fr_t *arena = noalloc_scratch(dev);
        static_assert(NttDomain::standalone_arena_elems(kNoAllocMaxLog) == kNoAllocArenaElems, "arena size");
        NKR_REQUIRE(NttDomain::standalone_arena_elems(log_n) <= kNoAllocArenaElems, "transform scratch too small");
        // Tables are rebuilt on the stream by every call (they are exact, so
        // the results equal nkr_fr_ntt_dev's); nothing needs clearing.
        const NttDomain dom(log_n, kind, arena, s);
        dom.standalone(static_cast<fr_t *>(data_dev), arena + kNoAllocArenaElems, kind, s);
    });
}

int nkr_msm_set_shape(uint32_t target_pieces_log, uint32_t complete_lanes_cap)
{
    return guarded([&] { msm_set_shape(target_pieces_log, complete_lanes_cap); });
}

int nkr_ntt_set_shape(uint32_t lds_log, uint32_t radix_log, uint32_t split_policy)
{
    return guarded([&] {
        ntt_set_shape(lds_log, radix_log, split_policy);
        // Cached standalone domains were built with the old shape.
        std::lock_guard<std::mutex> lock(g_ntt_mutex);
        bool any = false;
        for (auto &row : g_ntt_cache)
            for (NttDomain *d : row) any = any || d;
        if (any) NKR_HIP(hipDeviceSynchronize());
        for (auto &row : g_ntt_cache)
            for (NttDomain *&d : row) {
                delete d;
                d = nullptr;
            }
    });
}

// The CSR checks every SpMV entry point makes before touching the device.
static void check_csr(uint64_t rows, uint64_t cols, const uint32_t *row_ptr, const uint32_t *col_idx,
                      const void *values, uint64_t nnz)
{
    NKR_REQUIRE(row_ptr && (nnz == 0 || (col_idx && values)), "null argument");
    NKR_REQUIRE(rows < (1ull << 32) && nnz < (1ull << 32) && cols < (1ull << 32), "matrix too large");
    NKR_REQUIRE(row_ptr[0] == 0 && row_ptr[rows] == nnz, "invalid CSR offsets");
    for (uint64_t r = 0; r < rows; ++r) NKR_REQUIRE(row_ptr[r + 1] >= row_ptr[r], "non-monotonic CSR offsets");
    for (uint64_t j = 0; j < nnz; ++j) NKR_REQUIRE(col_idx[j] < cols, "out-of-range column");
}

int nkr_spmv(uint64_t rows, uint64_t cols, const uint32_t *row_ptr, const uint32_t *col_idx, const void *values,
             uint64_t nnz, const void *z, void *y, int device)
{
    return guarded([&] {
        NKR_REQUIRE(z && y, "null argument");
        check_csr(rows, cols, row_ptr, col_idx, values, nnz);
        set_device(device);
        DeviceBuffer<uint32_t> rp(rows + 1), ci(std::max<uint64_t>(nnz, 1));
        DeviceBuffer<fr_t> v(std::max<uint64_t>(nnz, 1)), dz(std::max<uint64_t>(cols, 1)), dy(std::max<uint64_t>(rows, 1));
        NKR_HIP(hipMemcpy(rp.get(), row_ptr, (rows + 1) * 4, hipMemcpyHostToDevice));
        if (nnz) {
            NKR_HIP(hipMemcpy(ci.get(), col_idx, nnz * 4, hipMemcpyHostToDevice));
            NKR_HIP(hipMemcpy(v.get(), values, nnz * sizeof(fr_t), hipMemcpyHostToDevice));
        }
        if (cols) NKR_HIP(hipMemcpy(dz.get(), z, cols * sizeof(fr_t), hipMemcpyHostToDevice));
        CsrView m{rp.get(), ci.get(), v.get(), (uint32_t)rows, (uint32_t)nnz};
        DeviceBuffer<uint32_t> lr;
        m.n_long = spmv_long_rows(row_ptr, (uint32_t)rows, lr);
        m.long_rows = lr.get();
        spmv(m, dz.get(), dy.get(), nullptr);
        if (rows) NKR_HIP(hipMemcpy(y, dy.get(), rows * sizeof(fr_t), hipMemcpyDeviceToHost));
    });
}

struct nkr_spmv_ctx {
    int device = 0;
    nkr::DeviceBuffer<uint32_t> row_ptr, col_idx, long_rows;
    nkr::DeviceBuffer<nkr::fr_t> values;
    nkr::CsrView view{};
    uint64_t cols = 0;
};

int nkr_spmv_create(uint64_t rows, uint64_t cols, const uint32_t *row_ptr, const uint32_t *col_idx,
                    const void *values, uint64_t nnz, int device, nkr_spmv_ctx **out)
{
    return guarded([&] {
        NKR_REQUIRE(out, "null argument");
        *out = nullptr;
        check_csr(rows, cols, row_ptr, col_idx, values, nnz);
        set_device(device);
        auto m = std::make_unique<nkr_spmv_ctx>();
        m->device = device;
        m->cols = cols;
        m->row_ptr.allocate(rows + 1);
        m->col_idx.allocate(std::max<uint64_t>(nnz, 1));
        m->values.allocate(std::max<uint64_t>(nnz, 1));
        NKR_HIP(hipMemcpy(m->row_ptr.get(), row_ptr, (rows + 1) * 4, hipMemcpyHostToDevice));
        if (nnz) {
            NKR_HIP(hipMemcpy(m->col_idx.get(), col_idx, nnz * 4, hipMemcpyHostToDevice));
            NKR_HIP(hipMemcpy(m->values.get(), values, nnz * sizeof(fr_t), hipMemcpyHostToDevice));
        }
        m->view = CsrView{m->row_ptr.get(), m->col_idx.get(), m->values.get(), (uint32_t)rows, (uint32_t)nnz};
        m->view.n_long = spmv_long_rows(row_ptr, (uint32_t)rows, m->long_rows);
        m->view.long_rows = m->long_rows.get();
        *out = m.release();
    });
}

int nkr_spmv_run_dev(nkr_spmv_ctx *m, const void *z_dev, void *y_dev, void *stream)
{
    return guarded([&] {
        NKR_REQUIRE(m && z_dev && y_dev, "null argument");
        NKR_HIP(hipSetDevice(m->device));
        spmv(m->view, static_cast<const fr_t *>(z_dev), static_cast<fr_t *>(y_dev),
             static_cast<hipStream_t>(stream));
    });
}

int nkr_spmv_destroy(nkr_spmv_ctx *m)
{
    return guarded([&] {
        if (m) NKR_HIP(hipSetDevice(m->device));
        delete m;
    });
}

static int msm_create(const void *bases, uint64_t n, uint32_t c, int device, bool g2, nkr_msm **out)
{
    return guarded([&] {
        NKR_REQUIRE(bases && out, "null argument");
        NKR_REQUIRE(n >= 1 && n < (1ull << 31), "MSM size must be 1 .. 2^31-1");
        *out = nullptr;
        set_device(device);
        auto m = std::make_unique<nkr_msm>();
        m->device = device;
        m->g2 = g2;
        m->n = n;
        NKR_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
        if (!g2) {
            DeviceBuffer<g1_jac> tmp(n);
            NKR_HIP(hipMemcpyAsync(tmp.get(), bases, n * sizeof(g1_jac), hipMemcpyHostToDevice, m->stream));
            m->g1 = std::make_unique<MsmContext<fq_t>>(tmp.get(), (uint32_t)n, c, m->stream);
            NKR_HIP(hipStreamSynchronize(m->stream));
        } else {
            DeviceBuffer<g2_jac> tmp(n);
            NKR_HIP(hipMemcpyAsync(tmp.get(), bases, n * sizeof(g2_jac), hipMemcpyHostToDevice, m->stream));
            m->g2ctx = std::make_unique<MsmContext<fq2_t>>(tmp.get(), (uint32_t)n, c, m->stream);
            NKR_HIP(hipStreamSynchronize(m->stream));
        }
        m->scalars.allocate(n);
        *out = m.release();
    });
}

int nkr_g1_msm_create(const void *bases, uint64_t n, uint32_t c, int device, nkr_msm **out)
{
    return msm_create(bases, n, c, device, false, out);
}

int nkr_g2_msm_create(const void *bases, uint64_t n, uint32_t c, int device, nkr_msm **out)
{
    return msm_create(bases, n, c, device, true, out);
}

static void msm_run_dev(nkr_msm *m, const fr_t *scalars_dev, void *out, hipStream_t s)
{
    if (!m->g2) {
        const g1_jac r = m->g1->run(scalars_dev, s);
        std::memcpy(out, &r, sizeof(r));
    } else {
        const g2_jac r = m->g2ctx->run(scalars_dev, s);
        std::memcpy(out, &r, sizeof(r));
    }
}

int nkr_msm_run(nkr_msm *m, const void *scalars, void *out)
{
    return guarded([&] {
        NKR_REQUIRE(m && scalars && out, "null argument");
        NKR_HIP(hipSetDevice(m->device));
        NKR_HIP(hipMemcpyAsync(m->scalars.get(), scalars, m->n * sizeof(fr_t), hipMemcpyHostToDevice, m->stream));
        msm_run_dev(m, m->scalars.get(), out, m->stream);
    });
}

int nkr_msm_run_dev(nkr_msm *m, const void *scalars_dev, void *out, void *stream)
{
    return guarded([&] {
        NKR_REQUIRE(m && scalars_dev && out, "null argument");
        NKR_HIP(hipSetDevice(m->device));
        msm_run_dev(m, static_cast<const fr_t *>(scalars_dev), out,
                    stream ? static_cast<hipStream_t>(stream) : m->stream);
    });
}

int nkr_msm_window_bits(const nkr_msm *m, uint32_t *c)
{
    return guarded([&] {
        NKR_REQUIRE(m && c, "null argument");
        *c = m->g2 ? m->g2ctx->plan().c : m->g1->plan().c;
    });
}

int nkr_msm_destroy(nkr_msm *m)
{
    return guarded([&] { delete m; });
}

int nkr_g1_msm(const void *bases, const void *scalars, uint64_t n, void *out, int device)
{
    nkr_msm *m = nullptr;
    int rc = nkr_g1_msm_create(bases, n, 0, device, &m);
    if (rc == 0) rc = nkr_msm_run(m, scalars, out);
    if (m) {
        const std::string keep = g_last_error;
        nkr_msm_destroy(m);
        g_last_error = keep;
    }
    return rc;
}

int nkr_g2_msm(const void *bases, const void *scalars, uint64_t n, void *out, int device)
{
    nkr_msm *m = nullptr;
    int rc = nkr_g2_msm_create(bases, n, 0, device, &m);
    if (rc == 0) rc = nkr_msm_run(m, scalars, out);
    if (m) {
        const std::string keep = g_last_error;
        nkr_msm_destroy(m);
        g_last_error = keep;
    }
    return rc;
}

}  // extern "C"

template <class F>
static int fixed_base_mul(const void *base, const void *scalars, uint64_t n, void *out, int device)
{
    return guarded([&] {
        NKR_REQUIRE(base && scalars && out, "null argument");
        if (n == 0) return;
        set_device(device);
        Jacobian<F> b;
        std::memcpy(&b, base, sizeof(b));
        FixedBaseMul<F> fbm(b, nullptr);
        DeviceBuffer<fr_t> ds(n);
        DeviceBuffer<Jacobian<F>> dout(n);
        NKR_HIP(hipMemcpy(ds.get(), scalars, n * sizeof(fr_t), hipMemcpyHostToDevice));
        fbm.run(ds.get(), n, dout.get(), nullptr);
        NKR_HIP(hipMemcpy(out, dout.get(), n * sizeof(Jacobian<F>), hipMemcpyDeviceToHost));
    });
}

template <class F>
static int fixed_base_mul_dev(const void *base, const void *scalars_dev, uint64_t n, void *out_dev, void *stream)
{
    return guarded([&] {
        NKR_REQUIRE(base && scalars_dev && out_dev, "null argument");
        if (n == 0) return;
        hipStream_t s = static_cast<hipStream_t>(stream);
        Jacobian<F> b;
        std::memcpy(&b, base, sizeof(b));
        FixedBaseMul<F> fbm(b, s);
        fbm.run(static_cast<const fr_t *>(scalars_dev), n, static_cast<Jacobian<F> *>(out_dev), s);
        NKR_HIP(hipStreamSynchronize(s));
    });
}

extern "C" {

int nkr_g1_fixed_base_mul(const void *base, const void *scalars, uint64_t n, void *out, int device)
{
    return fixed_base_mul<fq_t>(base, scalars, n, out, device);
}
int nkr_g2_fixed_base_mul(const void *base, const void *scalars, uint64_t n, void *out, int device)
{
    return fixed_base_mul<fq2_t>(base, scalars, n, out, device);
}
int nkr_g1_fixed_base_mul_dev(const void *base, const void *scalars_dev, uint64_t n, void *out_dev, void *stream)
{
    return fixed_base_mul_dev<fq_t>(base, scalars_dev, n, out_dev, stream);
}
int nkr_g2_fixed_base_mul_dev(const void *base, const void *scalars_dev, uint64_t n, void *out_dev, void *stream)
{
    return fixed_base_mul_dev<fq2_t>(base, scalars_dev, n, out_dev, stream);
}

int nkr_field_op(int op, const void *a, const void *b, void *out, uint64_t n, int device)
{
    return guarded([&] {
        NKR_REQUIRE(op >= 0 && op <= 8, "unknown field op");
        NKR_REQUIRE(a && out && (b || op == 7 || op == 8), "null argument");
        if (n == 0) return;
        set_device(device);
        const size_t esz = (op == 6 || op == 7) ? 64 : 32;
        DeviceBuffer<uint8_t> da(n * esz), db(n * esz), dout(n * esz);
        NKR_HIP(hipMemcpy(da.get(), a, n * esz, hipMemcpyHostToDevice));
        if (b && op != 7 && op != 8) NKR_HIP(hipMemcpy(db.get(), b, n * esz, hipMemcpyHostToDevice));
        k_field_op<<<div_up(n, 256), 256>>>(op, da.get(), db.get(), dout.get(), n);
        NKR_LAUNCH_CHECK();
        NKR_HIP(hipMemcpy(out, dout.get(), n * esz, hipMemcpyDeviceToHost));
    });
}

int nkr_dev_alloc(uint64_t bytes, void **ptr)
{
    return guarded([&] {
        NKR_REQUIRE(ptr, "null argument");
        NKR_HIP(hipMalloc(ptr, bytes));
    });
}

int nkr_dev_free(void *ptr)
{
    return guarded([&] { NKR_HIP(hipFree(ptr)); });
}

int nkr_memcpy_htod(void *dst, const void *src, uint64_t bytes)
{
    return guarded([&] { NKR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}

int nkr_memcpy_dtoh(void *dst, const void *src, uint64_t bytes)
{
    return guarded([&] { NKR_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}

int nkr_device_synchronize(void)
{
    return guarded([&] { NKR_HIP(hipDeviceSynchronize()); });
}

int nkr_device_pci_bus_id(int device, char *buf, int len)
{
    return guarded([&] {
        NKR_REQUIRE(buf && len >= 13, "buffer too small for a PCI address");
        NKR_HIP(hipDeviceGetPCIBusId(buf, len, device));
    });
}

int nkr_dev_copy(void *dst, const void *src, uint64_t bytes, void *stream)
{
    return guarded([&] {
        NKR_REQUIRE(dst && src, "null argument");
        NKR_REQUIRE(bytes % 16 == 0, "copy size must be a multiple of 16 bytes");
        const uint64_t n = bytes / 16;
        if (n == 0) return;
        const uint32_t blocks = (uint32_t)std::min<uint64_t>(div_up(n, 256), 8192);
        k_copy16<<<blocks, 256, 0, static_cast<hipStream_t>(stream)>>>(static_cast<const uint4 *>(src),
                                                                        static_cast<uint4 *>(dst), n);
        NKR_LAUNCH_CHECK();
    });
}

int nkr_valu_probe(int device, double *mad_per_s, double *fr_mul_per_s)
{
    return guarded([&] { valu_probe(device, mad_per_s, fr_mul_per_s); });
}

}  // extern "C"

// Radix-2 number-theoretic transforms over Fr for one power-of-two domain.
//
// A transform of N = 2^n points runs as 1-4 passes.  Each pass performs
// 2^np-point sub-transforms entirely in LDS (one workgroup holds up to 2048
// elements = 64 KiB) and multiplies by the inter-pass twiddles omega^(lo*k*2^S)
// (the four-step factorisation).  A domain that owns its tables holds them
// whole, one element per position of each twiddled pass (read like the data,
// one product per element); an arena domain, whose memory is bounded, makes
// each from two small tables (omega^e = hi[e >> m] * lo[e & (2^m - 1)]: one
// more product).
//
// dif(): natural order in, bit-reversed order out (decimation in frequency).
// dit(): bit-reversed order in, natural order out; it is the transpose of
//        dif() and uses the same twiddles.
// The Groth16 quotient chain composes the two so that no bit-reversal pass is
// ever needed; the standalone transforms add one gather pass to return
// natural order like libfqfft's basic_radix2_domain.
#pragma once

#include <vector>

#include "field.h"
#include "runtime.h"

namespace nkr {

// Workgroup shape of the transform passes: 2^lds_log elements per LDS tile,
// 2^radix_log per register round, pass split policy (0 balanced, 1 a full
// unit-stride last pass).  Zero fields take the process-wide setting.
struct NttShape {
    unsigned lds_log = 0, radix_log = 0, split_policy = 0;
    bool split_set = false;
};

class NttDomain {
public:
    NttDomain(unsigned log_n, hipStream_t stream, const NttShape *shape = nullptr);
    // A domain for standalone transforms of one `kind` (0 FFT, 1 iFFT,
    // 2 cosetFFT, 3 icosetFFT) whose tables are built in `arena`, caller-owned
    // device memory of standalone_arena_elems(log_n) elements, instead of
    // allocated: the construction and its transforms allocate no device
    // memory.  Only that kind's direction and coset table exist.
    NttDomain(unsigned log_n, int kind, fr_t *arena, hipStream_t stream);
    // Elements an arena needs for any kind at 2^log_n points: twiddles and
    // their two-level power tables (2 x (lo + hi)), sub-transform roots
    // (<= 2^11) and one coset table (2^log_n).
    static constexpr size_t standalone_arena_elems(unsigned log_n)
    {
        return 2 * ((size_t(1) << ((log_n + 1) / 2)) + (size_t(1) << (log_n / 2))) + 2048 + (size_t(1) << log_n);
    }

    unsigned log_n() const { return log_n_; }
    size_t size() const { return size_t(1) << log_n_; }

    // In-place transforms on device data.  `pre` / `post` are optional
    // per-position multipliers (indexed by storage position), `post_const` an
    // optional constant factor applied last.
    void dif(fr_t *data, bool inverse, const fr_t *pre, const fr_t *post, const fr_t *post_const,
             hipStream_t stream) const;
    void dit(fr_t *data, bool inverse, const fr_t *pre, const fr_t *post, hipStream_t stream) const;

    // The same transforms on `count` arrays of this size at once (one launch
    // per pass); batch_dev is a device array of the data pointers, and the
    // pre / post tables are shared by all of them.
    void dif_batch(fr_t *const *batch_dev, uint32_t count, bool inverse, const fr_t *pre, const fr_t *post,
                   const fr_t *post_const, hipStream_t stream) const;
    void dit_batch(fr_t *const *batch_dev, uint32_t count, bool inverse, const fr_t *pre, const fr_t *post,
                   hipStream_t stream) const;
    // quotient() for `count` jobs; abch_dev holds 4 device pointers per job: a, b, c, h.
    void quotient_batch(fr_t *const *abch_dev, uint32_t count, hipStream_t stream) const;

    // out[i] = in[bitrev(i)] (out of place).
    void bitrev(const fr_t *in, fr_t *out, hipStream_t stream) const;

    // libfqfft semantics, natural order in and out (kind: 0 FFT, 1 iFFT,
    // 2 cosetFFT, 3 icosetFFT), using `scratch` (n elements) for the final
    // reordering.
    void standalone(fr_t *data, fr_t *scratch, int kind, hipStream_t stream) const;

    // h[i] = (a[i]*b[i] - c[i]) / (g^N - 1) on the coset (divide_by_Z_on_coset).
    void quotient(const fr_t *a, const fr_t *b, const fr_t *c, fr_t *h, hipStream_t stream) const;

    // Tables the prove chain fuses into its passes.
    const fr_t *coset_rev_scaled() const { return coset_rev_p_; }      // g^rev(j) / N
    const fr_t *icoset_rev_scaled() const { return icoset_rev_p_; }    // g^-rev(j) / N

    fr_t omega() const { return omega_[0]; }

private:
    struct Pass {
        unsigned S, np, log_L, log_cols;
    };
    void run(fr_t *data, fr_t *const *batch, uint32_t count, bool dit, bool inverse, const fr_t *pre,
             const fr_t *post, const fr_t *post_const, hipStream_t stream) const;
    // Pass split, constants and table sizes (both constructors).
    void plan(unsigned log_n, const NttShape *shape);

    unsigned log_n_;
    unsigned lds_log_ = 11, radix_log_ = 3;
    unsigned m_lo_;        // bits of the low twiddle table index
    unsigned log_roots_;   // sub-transform roots are of order 2^log_roots_
    std::vector<Pass> passes_;
    fr_t omega_[2];        // omega, omega^-1
    fr_t n_inv_;
    fr_t zinv_;            // (g^N - 1)^-1
    uint32_t lo_count_ = 0, hi_count_ = 0, root_count_ = 0;
    int arena_kind_ = -1;  // the kind an arena domain serves; -1 when the tables are owned
    DeviceBuffer<fr_t> tw_lo_[2], tw_hi_[2], roots_[2];
    // whole inter-pass twiddle tables per direction and pass (none for the
    // last pass, which has no twiddle): element (j << log_L) + lo holds
    // omega^((lo * bitrev(j)) << S)
    std::vector<DeviceBuffer<fr_t>> tw_full_[2];
    DeviceBuffer<fr_t> coset_nat_, coset_rev_, icoset_rev_;
    // The tables in use: the buffers above, or places in an arena (null when absent).
    const fr_t *tw_lo_p_[2] = {}, *tw_hi_p_[2] = {}, *roots_p_[2] = {};
    const fr_t *coset_nat_p_ = nullptr, *coset_rev_p_ = nullptr, *icoset_rev_p_ = nullptr;
};

// Workgroup shape for domains built afterwards: 2^lds_log LDS elements per
// workgroup (10 or 11), 2^radix_log elements per thread per round (2 or 3)
// and the pass split policy (0 balanced, 1 unit-stride last pass full).
void ntt_set_shape(unsigned lds_log, unsigned radix_log, unsigned split_policy);

// Fr constants shared by host code.
fr_t fr_from_u64(uint64_t x);
fr_t fr_root_of_unity(unsigned log_n);   // primitive 2^log_n-th root (libff field_utils rule)
fr_t fr_pow_u64(const fr_t &x, uint64_t e);

}  // namespace nkr

/*
 * CPU oracle for the Groth16 prove path on BN254 -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference CPU backend's algorithm
 * (protocols/groth16/backends/cpu/src/backend.cpp:16-35,156-201) and of the
 * vendored libff / libfqfft routines it calls.  It is the checker for the HIP
 * backend and the CPU baseline timed beside it; it is never linked into, or
 * called by, the product library.  Only tests, the smoke check and the CPU
 * baseline leg of the benchmark load it.
 *
 * All buffers use the reference's in-memory layout: a field element is 4
 * little-endian u64 words of the Montgomery residue x*2^256 mod p (libff
 * Fp_model mont_repr); G1 is Jacobian X,Y,Z over Fq (96 B); G2 is Jacobian
 * X,Y,Z over Fq2 (192 B, c0 then c1).
 */
#ifndef NKR_BN254_ORACLE_H
#define NKR_BN254_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    size_t rows, cols, nnz;
    const uint32_t *row_ptr; /* rows + 1 */
    const uint32_t *col_idx; /* nnz */
    const void *values;      /* nnz Fr */
} or_csr;

/* --- pseudo-random source ------------------------------------------------ */
void or_mt19937_64(uint64_t seed, uint64_t *out, size_t n);
/* libff Fp_model::random_element rule on an mt19937_64(seed) stream. */
void or_fr_random(uint64_t seed, void *out, size_t n);

/* --- field element batches (n elements each) ---------------------------- */
void or_fr_mul(const void *a, const void *b, void *out, size_t n);
void or_fr_add(const void *a, const void *b, void *out, size_t n);
void or_fr_sub(const void *a, const void *b, void *out, size_t n);
void or_fr_inv(const void *a, void *out, size_t n);
void or_fr_to_mont(const void *in, void *out, size_t n);
void or_fr_from_mont(const void *in, void *out, size_t n);
void or_fq_mul(const void *a, const void *b, void *out, size_t n);
void or_fq_add(const void *a, const void *b, void *out, size_t n);
void or_fq_sub(const void *a, const void *b, void *out, size_t n);
void or_fq_to_mont(const void *in, void *out, size_t n);
void or_fq_from_mont(const void *in, void *out, size_t n);
void or_fq2_mul(const void *a, const void *b, void *out, size_t n);
void or_fq2_sqr(const void *a, void *out, size_t n);

/* --- groups --------------------------------------------------------------- */
void or_g1_generator(void *out96);
void or_g2_generator(void *out192);
/* to_affine_coordinates: Z = 1, or (0, 1, 0) for the point at infinity. */
void or_g1_to_affine(const void *in96, void *out96, size_t n);
void or_g2_to_affine(const void *in192, void *out192, size_t n);
void or_g1_add(const void *a96, const void *b96, void *out96);
void or_g2_add(const void *a192, const void *b192, void *out192);
void or_g1_dbl(const void *a96, void *out96);
void or_g1_neg(const void *a96, void *out96);
/* scalar (Montgomery Fr) times base, libff scalar_mul double-and-add. */
void or_g1_mul(const void *base96, const void *scalar, void *out96, size_t n);
void or_g2_mul(const void *base192, const void *scalar, void *out192, size_t n);
int or_g1_on_curve(const void *p96);
int or_g2_on_curve(const void *p192);

/* --- prove-path operators --------------------------------------------------- */
void or_spmv(const or_csr *m, const void *z, void *out);
/* kind: 0 FFT, 1 iFFT, 2 cosetFFT(g=5), 3 icosetFFT(g=5).  n a power of two >= 2. */
int or_fft(void *a, size_t n, int kind);
/* The quotient chain of CpuBackend::prove: returns H (n coefficients). */
int or_quotient(const void *az, const void *bz, const void *cz, size_t n, void *h);
/* libff multi_exp<BDLO12> with chunks = 1. */
int or_msm_g1(const void *bases96, const void *scalars, size_t n, void *out96);
int or_msm_g2(const void *bases192, const void *scalars, size_t n, void *out192);

/* --- Groth16 -------------------------------------------------------------- */
typedef struct {
    size_t public_inputs;
    or_csr a, b, c;
} or_circuit;

typedef struct {
    const void *a_query;    /* cols G1 */
    const void *b_g1_query; /* cols G1 */
    const void *b_g2_query; /* cols G2 */
    const void *l_query;    /* cols G1 */
    const void *h_query;    /* rows G1 */
    const void *alpha_g1, *beta_g1, *beta_g2, *delta_g1, *delta_g2;
} or_pk;

/* CpuBackend::prove.  raw_out: ar, bs1, zk, qz (G1) then bs2 (G2) = 576 B.
 * Optional az/bz/cz/h (n Fr each) receive the intermediates when non-NULL. */
int or_prove(const or_circuit *c, const or_pk *pk, const void *witness, void *raw_out,
             void *az, void *bz, void *cz, void *h);
/* The same on `threads` host threads (POSIX threads; 0 or 1 = one): the
 * MSMs split into min(n, threads) contiguous chunks, each summed by
 * BDLO12 and the partial sums added, as the reference's parallel
 * variable_base_msm (operators/include/nkr/operators/variable_base_msm.hpp:
 * 17-49).  Results equal the single-threaded ones (affine). */
int or_msm_g1_mt(const void *bases96, const void *scalars, size_t n, void *out96, size_t threads);
int or_msm_g2_mt(const void *bases192, const void *scalars, size_t n, void *out192, size_t threads);
int or_prove_mt(const or_circuit *c, const or_pk *pk, const void *witness, void *raw_out,
                void *az, void *bz, void *cz, void *h, size_t threads);
/* blind_proof with caller-given r, s.  proof_out: A (G1), B (G2), C (G1) = 384 B. */
void or_blind(const or_pk *pk, const void *raw, const void *r, const void *s, void *proof_out);

#ifdef __cplusplus
}
#endif

#endif

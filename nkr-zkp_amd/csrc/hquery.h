// The proving key's H query in the basis of coset evaluations.
//
// The quotient chain ends with the coset inverse transform
//   h_k = (g^-k / n) * sum_i omega^(-ik) * q_i
// of the quotient's coset evaluations q, and the H MSM is sum_k h_k H_k.
// Exchanging the sums, sum_k h_k H_k = sum_i q_i H'_i with
//   H'_i = sum_k omega^(-ik) * (g^-k / n) * H_k,
// the same inverse transform applied to the group elements.  Transforming the
// key once, when a circuit is prepared, lets every proof skip its seventh
// transform and feed q to the MSM directly; the MSM's sum, and so the proof,
// is unchanged.  (The reference computes h with libfqfft's icosetFFT,
// backends/cpu/src/backend.cpp:156-201, then multi_exp over h_query.)
#pragma once

#include "curve.h"
#include "ntt.h"

namespace nkr {

// in: the n = domain.size() reference-layout Jacobian points H_k (device).
// out: n Jacobian points (device) with H'_i at position bitrev(i) -- the
// order a decimation-in-frequency transform leaves them in.
void h_query_coset_basis(const Jacobian<fq_t> *in, const NttDomain &domain, Jacobian<fq_t> *out,
                         hipStream_t stream);

}  // namespace nkr

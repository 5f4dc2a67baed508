// Batched negacyclic NTT kernels (two passes over HBM, see mi_ntt_core.hpp).
//
// Forward = reference ntt_forward_dit2 (ntt.cpp:235-297): column pass runs
// global stages 0..n1-1, row pass runs n1..n-1 and reduces to [0, q).
// Inverse = reference ntt_inverse_dif2 (ntt.cpp:331-386): row pass runs the
// distance 1..R2/2 stages, column pass the rest and multiplies by N^-1.
//
// The kernels are templated on a Job that maps the batch's limb k to a view
// with basis(), load(i) and store(i, v); plain in-place NTTs and the fused
// rescale / key-switch steps (prologue computes the input, epilogue combines
// the output) share the same butterfly code.
#include "mi_ntt_kernels.hpp"

namespace mi {

void ntt_forward(const Context& c, const LimbBatch& b, hipStream_t s) {
    if (b.count <= 0) return;
    const PlainJob j = plain_job(b, c.N);
    run_ntt_fwd(c, j, j, b.count, s);
}

void ntt_inverse(const Context& c, const LimbBatch& b, hipStream_t s) {
    if (b.count <= 0) return;
    const PlainJob j = plain_job(b, c.N);
    run_ntt_inv(c, j, j, b.count, s);
}

}  // namespace mi

// Device view of the branch scorer's weights, as the CKF kernels and the
// scorer kernels take it by value.
#pragma once

#include <cstdint>

struct ckf_scorer;

namespace tamd {

struct mlp_dev {
    const short* w1t;      // [64 out][16 in] bf16 (inputs 12..15 zero)
    const short* w2;       // [64 in][64 out] bf16 (B of dH1 = dZ2 W2^T)
    const short* w2t;      // [64 out][64 in] bf16 (B of Z2 = H1 W2)
    const float* b1;       // [64]
    const float* b2;       // [64]
    const float* w3;       // [64] BF16-rounded, kept in FP32
    const float* b3;       // [1]
    const float* mean;     // [12] feature normalisation
    const float* inv_std;  // [12]
};

// The device view of a scorer created through the C ABI (scorer/mlp.hip).
mlp_dev scorer_view(const ckf_scorer* s);

}  // namespace tamd

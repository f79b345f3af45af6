// Branch-scorer MLP (12 -> 64 -> 64 -> 1, GELU) on CDNA matrix cores.
//
// One wave evaluates a tile of 64 rows (samples).  Layers 1 and 2 run on
// v_mfma_f32_16x16x16_bf16 (BF16 operands, FP32 accumulation) over 4 x 4
// output tiles of 16 x 16; layer 3 (64 -> 1) is a dot product reduced across
// the 16 lanes that share a row.  Fragment layout of the 16x16x16 MFMA, lane
// l of 64 (checked on the device by ckf_debug_mfma_layout):
//   A (M x K):  lane l holds A[l % 16][4 (l / 16) + e],  e = 0..3
//   B (K x N):  lane l holds B[4 (l / 16) + e][l % 16]
//   D (M x N):  lane l holds D[4 (l / 16) + e][l % 16]
// so a D tile read as B (same rows = K, columns = N) needs no data movement;
// D -> A needs a transpose, which goes through the wave's LDS area.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp_view.hpp"

namespace tamd {

constexpr int kMlpIn = 12;
constexpr int kMlpInPad = 16;
constexpr int kMlpHidden = 64;
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + kMlpHidden * kMlpHidden +
                           kMlpHidden + kMlpHidden + 1;  // 5057
// parameter offsets in the flat FP32 vector W1[12][64] b1 W2[64][64] b2 W3 b3
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpIn * kMlpHidden;
constexpr int kOffW2 = kOffB1 + kMlpHidden;
constexpr int kOffB2 = kOffW2 + kMlpHidden * kMlpHidden;
constexpr int kOffW3 = kOffB2 + kMlpHidden;
constexpr int kOffB3 = kOffW3 + kMlpHidden;
// LDS per wave: X tile [64][16] bf16 + H tile [64][kHStride] bf16
constexpr int kHStride = 72;  // 144 B rows spread the 16 rows of a fragment over the banks
constexpr int kMlpLdsPerWave = 64 * kMlpInPad + 64 * kHStride;  // shorts

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline short f2bf(float f) {
    // round to nearest even: gfx950's v_cvt_pk_bf16_f32
    return __builtin_bit_cast(short, static_cast<__bf16>(f));
}
__device__ inline float bf2f(short h) {
    return __uint_as_float(static_cast<uint32_t>(static_cast<uint16_t>(h)) << 16);
}

__device__ inline f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// erf by Abramowitz & Stegun, Handbook of Mathematical Functions, 7.1.26:
// erf(x) = 1 - (a1 t + a2 t^2 + a3 t^3 + a4 t^4 + a5 t^5) exp(-x^2),
// t = 1 / (1 + p |x|), odd in x; absolute error <= 1.5e-7 in exact
// arithmetic, 5.4e-7 evaluated in FP32 here -- far below the BF16 rounding of
// the activations.  Branch-free (one reciprocal, one
// exponential): the scorer's 128 GELUs per sample are most of its VALU work,
// and a library erf's range branches diverge across the wave.
__device__ inline float erf_as(float x) {
    const float ax = fabsf(x);
    const float t = __builtin_amdgcn_rcpf(1.f + 0.3275911f * ax);
    const float poly =
        ((((1.061405429f * t - 1.453152027f) * t + 1.421413741f) * t - 0.284496736f) * t +
         0.254829592f) * t;
    return copysignf(1.f - poly * __expf(-ax * ax), x);
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.f + erf_as(z * 0.70710678118654752f)); }

// GELU and its derivative 0.5 (1 + erf(z / sqrt 2)) + z phi(z) from one
// exponential: erf_as's exp(-x^2) at x = z / sqrt 2 is the exp(-z^2 / 2) of
// the normal density phi.
__device__ inline void gelu_and_grad(float z, float& g, float& gg) {
    const float x = z * 0.70710678118654752f;
    const float ax = fabsf(x);
    const float t = __builtin_amdgcn_rcpf(1.f + 0.3275911f * ax);
    const float poly =
        ((((1.061405429f * t - 1.453152027f) * t + 1.421413741f) * t - 0.284496736f) * t +
         0.254829592f) * t;
    const float ex = __expf(-ax * ax);
    const float one_erf = 1.f + copysignf(1.f - poly * ex, x);
    g = 0.5f * z * one_erf;
    gg = 0.5f * one_erf + z * 0.39894228040143268f * ex;
}
__device__ inline float gelu_grad(float z) {
    float g, gg;
    gelu_and_grad(z, g, gg);
    return gg;
}

// Wave-internal LDS hand-over: every lane's earlier LDS writes are visible to
// every lane's later LDS reads.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline bf16x4 load4(const short* p) { return *reinterpret_cast<const bf16x4*>(p); }

// Layer-1 and layer-2 pre-activations of a 64-row tile.  On entry each lane
// holds its row's 12 raw features (lane = row; valid = false gives a zero
// row).  z1 / z2 are in D layout: z[mt][nt][e] is row 16 mt + 4 (l / 16) + e,
// column 16 nt + l % 16.  Leaves the normalised X tile in lds_x and H1 =
// GELU(z1) in lds_h; with lds_xt / lds_ht also X^T [16][kHStride] and H1^T
// [64][kHStride] (the A operands of the backward pass's weight gradients).
__device__ inline void mlp_tile_layers(const mlp_dev& W, const float* feat, bool valid,
                                       short* lds_x, short* lds_h, float (&z1)[4][4][4],
                                       float (&z2)[4][4][4], short* lds_xt = nullptr,
                                       short* lds_ht = nullptr) {
    const uint32_t l = __lane_id();
    const uint32_t r16 = l % 16u;
    const uint32_t g4 = 4u * (l / 16u);
#pragma unroll
    for (int f = 0; f < kMlpInPad; ++f) {
        float v = 0.f;
        if (f < kMlpIn && valid) {
            v = (feat[f] - W.mean[f]) * W.inv_std[f];
        }
        lds_x[l * kMlpInPad + f] = f2bf(v);
        if (lds_xt != nullptr) lds_xt[f * kHStride + l] = f2bf(v);
    }
    wave_lds_sync();
    // layer 1: Z1 = X W1 + b1 (K = 16, one MFMA per tile)
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const bf16x4 a = load4(lds_x + (16 * mt + r16) * kMlpInPad + g4);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const bf16x4 b = load4(W.w1t + (16 * nt + r16) * kMlpInPad + g4);
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma16(a, b, acc);
            const float bias = W.b1[16 * nt + r16];
            bf16x4 hb;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                z1[mt][nt][e] = acc[e] + bias;
                hb[e] = f2bf(gelu(z1[mt][nt][e]));
                lds_h[(16 * mt + g4 + e) * kHStride + 16 * nt + r16] = hb[e];
            }
            if (lds_ht != nullptr) {
                *reinterpret_cast<bf16x4*>(lds_ht + (16 * nt + r16) * kHStride + 16 * mt + g4) = hb;
            }
        }
    }
    wave_lds_sync();
    // layer 2: Z2 = H1 W2 + b2 (K = 64, four MFMAs per tile)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        bf16x4 b[4];
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            b[kt] = load4(W.w2t + (16 * nt + r16) * kMlpHidden + 16 * kt + g4);
        }
        const float bias = W.b2[16 * nt + r16];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                const bf16x4 a = load4(lds_h + (16 * mt + r16) * kHStride + 16 * kt + g4);
                acc = mfma16(a, b[kt], acc);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) z2[mt][nt][e] = acc[e] + bias;
        }
    }
}

// Layer 3 from z2: logit of row 16 mt + 4 (l / 16) + e in s[mt][e] (every
// lane of the 16-lane group holds its group's rows).
__device__ inline void mlp_tile_output(const mlp_dev& W, const float (&z2)[4][4][4],
                                       float (&s)[4][4]) {
    const uint32_t r16 = __lane_id() % 16u;
    float w3[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) w3[nt] = W.w3[16 * nt + r16];
    const float b3 = W.b3[0];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            float acc = 0.f;
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc += gelu(z2[mt][nt][e]) * w3[nt];
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) acc += __shfl_xor(acc, d, 64);
            s[mt][e] = acc + b3;
        }
    }
}

// Inference without LDS transposes.  The layers run transposed,
// Z^T = W^T X^T, with the weights as the A operand: the D fragment of one
// layer (rows = hidden units = the next K, columns = samples) is then the
// B operand of the next as it stands.  A wave scores 16 samples per call;
// lane l holds features 4 (l / 16) .. 4 (l / 16) + 3 of sample l % 16, and the
// A fragments of W1^T / W2^T are the same registers as the B fragments of
// W1 / W2 in mlp_tile_layers, kept for the wave's lifetime.
struct mlp_frags {
    bf16x4 w1[4];     // W1^T rows 16 nt .., the 16 (padded) inputs
    bf16x4 w2[4][4];  // W2^T block (out tile nt, in tile kt)
};

// Per-block copy of the vectors a lane reads four at a time (hidden units
// 16 nt + 4 (l / 16) .. + 3): biases, layer-3 weights, normalisation.
struct alignas(16) mlp_vecs {
    float b1[kMlpHidden];
    float b2[kMlpHidden];
    float w3[kMlpHidden];
    float mean[kMlpInPad];
    float inv_std[kMlpInPad];  // 0 for the padded inputs
    float b3;
};

__device__ inline void mlp_load_frags(const mlp_dev& W, mlp_frags& F) {
    const uint32_t l = __lane_id();
    const uint32_t r16 = l % 16u;
    const uint32_t g4 = 4u * (l / 16u);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        F.w1[nt] = load4(W.w1t + (16 * nt + r16) * kMlpInPad + g4);
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            F.w2[nt][kt] = load4(W.w2t + (16 * nt + r16) * kMlpHidden + 16 * kt + g4);
        }
    }
}

// The same fragments kept once per block in LDS (20 per lane position),
// read at each use so they hold no registers between uses.
struct alignas(16) mlp_frags_lds {
    bf16x4 f[20][64];
};

__device__ inline void mlp_fill_frags(const mlp_dev& W, mlp_frags_lds& L, uint32_t tid,
                                      uint32_t n) {
    for (uint32_t i = tid; i < 20u * 64u; i += n) {
        const uint32_t k = i / 64u, l = i % 64u;
        const uint32_t r16 = l % 16u, g4 = 4u * (l / 16u);
        if (k < 4u) {
            L.f[k][l] = load4(W.w1t + (16u * k + r16) * kMlpInPad + g4);
        } else {
            const uint32_t nt = (k - 4u) / 4u, kt = (k - 4u) % 4u;
            L.f[k][l] = load4(W.w2t + (16u * nt + r16) * kMlpHidden + 16u * kt + g4);
        }
    }
}

// Fill V from the block's threads (tid of n); a block barrier must follow.
__device__ inline void mlp_fill_vecs(const mlp_dev& W, mlp_vecs& V, uint32_t tid, uint32_t n) {
    for (uint32_t k = tid; k < kMlpHidden; k += n) {
        V.b1[k] = W.b1[k];
        V.b2[k] = W.b2[k];
        V.w3[k] = W.w3[k];
    }
    for (uint32_t k = tid; k < kMlpInPad; k += n) {
        V.mean[k] = k < kMlpIn ? W.mean[k] : 0.f;
        V.inv_std[k] = k < kMlpIn ? W.inv_std[k] : 0.f;
    }
    if (tid == 0u) V.b3 = W.b3[0];
}

// Logit of sample l % 16 (returned in all four lanes of that column) from
// its normalised BF16 features: lane l holds features 4 (l / 16) .. + 3 of
// sample l % 16 (the B fragment of layer 1).
__device__ inline float mlp_logit16_bf(const mlp_frags& F, const mlp_vecs& V, bf16x4 xb) {
    const uint32_t g4 = 4u * (__lane_id() / 16u);
    // read the vectors from LDS at each use instead of holding 56 registers
    // across the caller's loop
    __asm__ volatile("" ::: "memory");
    // layer 1: H1^T tile nt = GELU(W1^T X^T + b1), straight into B fragments
    bf16x4 h[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma16(F.w1[nt], xb, acc);
        const f32x4 b = *reinterpret_cast<const f32x4*>(&V.b1[16 * nt + g4]);
#pragma unroll
        for (int e = 0; e < 4; ++e) h[nt][e] = f2bf(gelu(acc[e] + b[e]));
    }
    // layer 2 and the layer-3 dot product over the lane's hidden units
    float part = 0.f;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) acc = mfma16(F.w2[nt][kt], h[kt], acc);
        const f32x4 b = *reinterpret_cast<const f32x4*>(&V.b2[16 * nt + g4]);
        const f32x4 w = *reinterpret_cast<const f32x4*>(&V.w3[16 * nt + g4]);
#pragma unroll
        for (int e = 0; e < 4; ++e) part += gelu(acc[e] + b[e]) * w[e];
    }
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    return part + V.b3;
}

// mlp_logit16_bf with the weight fragments read from the block's LDS copy.
__device__ inline float mlp_logit16_lds(const mlp_frags_lds& L, const mlp_vecs& V, bf16x4 xb) {
    const uint32_t l = __lane_id();
    const uint32_t g4 = 4u * (l / 16u);
    bf16x4 h[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma16(L.f[nt][l], xb, acc);
        const f32x4 b = *reinterpret_cast<const f32x4*>(&V.b1[16 * nt + g4]);
#pragma unroll
        for (int e = 0; e < 4; ++e) h[nt][e] = f2bf(gelu(acc[e] + b[e]));
    }
    float part = 0.f;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) acc = mfma16(L.f[4 + 4 * nt + kt][l], h[kt], acc);
        const f32x4 b = *reinterpret_cast<const f32x4*>(&V.b2[16 * nt + g4]);
        const f32x4 w = *reinterpret_cast<const f32x4*>(&V.w3[16 * nt + g4]);
#pragma unroll
        for (int e = 0; e < 4; ++e) part += gelu(acc[e] + b[e]) * w[e];
    }
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    return part + V.b3;
}

// Normalised BF16 value of raw feature f (index k); padded inputs are zero.
__device__ inline short mlp_feature_bf(const mlp_vecs& V, float f, uint32_t k) {
    return f2bf(k < static_cast<uint32_t>(kMlpIn) ? (f - V.mean[k]) * V.inv_std[k] : 0.f);
}

// Logit of sample l % 16 from raw features.  x: the lane's four raw
// features (4 (l / 16) .. + 3); valid = false gives a zero row.
__device__ inline float mlp_logit16(const mlp_frags& F, const mlp_vecs& V, const float (&x)[4],
                                    bool valid) {
    const uint32_t g4 = 4u * (__lane_id() / 16u);
    __asm__ volatile("" ::: "memory");
    const f32x4 mu = *reinterpret_cast<const f32x4*>(&V.mean[g4]);
    const f32x4 is = *reinterpret_cast<const f32x4*>(&V.inv_std[g4]);
    bf16x4 xb;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const bool real = valid && g4 + static_cast<uint32_t>(e) < static_cast<uint32_t>(kMlpIn);
        xb[e] = f2bf(real ? (x[e] - mu[e]) * is[e] : 0.f);
    }
    return mlp_logit16_bf(F, V, xb);
}

// Scores of a 64-row tile: the logit of lane l's row returned in lane l.
__device__ inline float mlp_tile_logit(const mlp_dev& W, const float* feat, bool valid,
                                       short* lds_x, short* lds_h) {
    float z1[4][4][4], z2[4][4][4], s[4][4];
    mlp_tile_layers(W, feat, valid, lds_x, lds_h, z1, z2);
    mlp_tile_output(W, z2, s);
    // hand row r's logit (held by lanes of group (r % 16) / 4, tile r / 16)
    // to lane r through the LDS area
    const uint32_t l = __lane_id();
    float* lds_s = reinterpret_cast<float*>(lds_x);  // 64 floats fit in the X tile
    wave_lds_sync();
    if ((l % 16u) == 0u) {
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
            for (int e = 0; e < 4; ++e) lds_s[16 * mt + 4 * (l / 16u) + e] = s[mt][e];
        }
    }
    wave_lds_sync();
    return lds_s[l];
}

}  // namespace tamd

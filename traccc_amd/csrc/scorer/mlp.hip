// Branch-scorer kernels: inference, gradient (forward + backward), Adam, and
// the C ABI of include/traccc_amd/scorer.h.
//
// Training runs on BF16 matrix cores with FP32 accumulation: the forward
// pass as in mlp_device.hpp; the backward pass forms
//   dW3 = H2^T dS,  dZ2 = dS W3^T * GELU'(Z2),  dW2 = H1^T dZ2,
//   dZ1 = (dZ2 W2^T) * GELU'(Z1),  dW1 = X^T dZ1
// with dZ2 / dZ1 rounded to BF16 as MFMA operands.  Every wave writes its
// tile's full gradient to its own slot; the step kernel sums the slots in a
// fixed order (16 waves per 64 parameters, each over a fixed stride of
// tiles, combined in wave order), so training is bitwise reproducible, and
// applies AdamW and writes the BF16 views of its parameters in the same
// pass.  The batch index, the Adam step count and the epoch's permutation
// live on the device, so a run of steps is one captured graph replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../common/errors.hpp"
#include "mlp_device.hpp"
#include "traccc_amd/scorer.h"

namespace tamd {

namespace {

constexpr int kWavesPerBlock = 4;
// LDS per wave of the gradient kernel: the forward pass's X and H1 tiles,
// then H1^T [64][kHStride], X^T [16][kHStride] and the tile's 64 labels (FP32)
constexpr int kGradLdsPerWave = kMlpLdsPerWave + 64 * kHStride + 16 * kHStride + 128;
constexpr int kBlock = 64 * kWavesPerBlock;

// Pseudo-random bijection of [0, n): i -> (a i + b) mod n, gcd(a, n) = 1.
struct perm_map {
    uint64_t a, b, n;
    __device__ uint32_t operator()(uint64_t i) const {
        return static_cast<uint32_t>((a * i + b) % n);
    }
};

// Training state on the device: the epoch's sample order, the batch the next
// step trains on, Adam's step count and the blocks of the step kernel that
// have finished.
struct train_state {
    perm_map perm;
    uint32_t batch_idx;
    uint32_t adam_t;
    uint32_t done_blocks;
    uint32_t pad;
};

__device__ inline uint32_t batch_rows(const train_state& ts, uint32_t n, uint32_t batch) {
    const uint64_t row0 = static_cast<uint64_t>(ts.batch_idx) * batch;
    return row0 >= n ? 0u : static_cast<uint32_t>(n - row0 < batch ? n - row0 : batch);
}

__global__ __launch_bounds__(kBlock) void k_mlp_infer(mlp_dev W, const float* feats, uint32_t n,
                                                       float* out) {
    __shared__ mlp_vecs V;
    mlp_fill_vecs(W, V, threadIdx.x, blockDim.x);
    __syncthreads();
    mlp_frags F;
    mlp_load_frags(W, F);
    const uint32_t wave = threadIdx.x / 64u;
    const uint32_t l = __lane_id();
    const uint32_t r16 = l % 16u;
    const uint32_t g4 = 4u * (l / 16u);
    const uint32_t n_tiles = (n + 63u) / 64u;
    for (uint32_t t = blockIdx.x * kWavesPerBlock + wave; t < n_tiles;
         t += gridDim.x * kWavesPerBlock) {
        // sub-tile mt: lane l holds features g4 .. g4 + 3 of row 64 t + 16 mt + l % 16
        float x[4][4];
        bool valid[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const uint32_t r = t * 64u + 16u * mt + r16;
            valid[mt] = r < n;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint32_t k = g4 + static_cast<uint32_t>(e);
                x[mt][e] = (valid[mt] && k < static_cast<uint32_t>(kMlpIn))
                               ? feats[static_cast<size_t>(r) * kMlpIn + k] : 0.f;
            }
        }
        // one sub-tile at a time (the next ones shift down) keeps a single
        // MLP evaluation live, and the wave at 3 waves / SIMD
        float mine = 0.f;
#pragma unroll 1
        for (int mt = 0; mt < 4; ++mt) {
            const float s = mlp_logit16(F, V, x[0], valid[0]);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                valid[q] = valid[q + 1];
#pragma unroll
                for (int e = 0; e < 4; ++e) x[q][e] = x[q + 1][e];
            }
            if (l / 16u == static_cast<uint32_t>(mt)) mine = s;  // row 64 t + l
        }
        const uint32_t i = t * 64u + l;
        if (i < n) out[i] = mine;
    }
}

// Gradient of one 64-sample tile per wave into grad_part[tile][kMlpParams],
// for the batch the training state names.
__global__ __launch_bounds__(kBlock) void k_mlp_grad(mlp_dev W, const float* feats,
                                                      const uint8_t* labels,
                                                      const train_state* state, uint32_t n,
                                                      uint32_t batch, float pos_weight,
                                                      float* grad_part, float* loss_part) {
    __shared__ short lds[kWavesPerBlock * kGradLdsPerWave];
    const train_state ts = *state;
    const perm_map perm = ts.perm;
    const uint32_t row0 = ts.batch_idx * batch;
    const uint32_t rows = batch_rows(ts, n, batch);
    const float inv_norm = rows > 0u ? 1.f / static_cast<float>(rows) : 0.f;
    const uint32_t wave = threadIdx.x / 64u;
    short* lds_x = lds + wave * kGradLdsPerWave;
    short* lds_h = lds_x + 64 * kMlpInPad;
    short* lds_ht = lds_h + 64 * kHStride;
    short* lds_xt = lds_ht + 64 * kHStride;
    float* lds_y = reinterpret_cast<float*>(lds_xt + 16 * kHStride);
    const uint32_t tile = blockIdx.x * kWavesPerBlock + wave;
    const uint32_t l = __lane_id();
    const uint32_t r16 = l % 16u;
    const uint32_t g4 = 4u * (l / 16u);
    if (tile * 64u >= rows) {
        return;  // whole waves only: no wave-level operation is left half-done
    }
    float* gp = grad_part + static_cast<size_t>(tile) * kMlpParams;

    const uint32_t i = tile * 64u + l;
    const bool valid = i < rows;
    // the lane's row: its features, and its label into the wave's LDS for
    // the lanes that hold the row's logit below (the permutation is
    // evaluated once per row)
    float f[kMlpIn];
    {
        const uint32_t row = valid ? perm(row0 + i) : 0u;
#pragma unroll
        for (int k = 0; k < kMlpIn; ++k) f[k] = valid ? feats[static_cast<size_t>(row) * kMlpIn + k] : 0.f;
        lds_y[l] = (valid && labels[row]) ? 1.f : 0.f;
    }
    float z1[4][4][4], z2[4][4][4], s[4][4];
    mlp_tile_layers(W, f, valid, lds_x, lds_h, z1, z2, lds_xt, lds_ht);
    mlp_tile_output(W, z2, s);

    // loss and dL/ds for the rows this lane holds
    float ds[4][4];
    float loss = 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t pos = tile * 64u + 16u * mt + g4 + e;
            const float y = lds_y[16u * mt + g4 + e];
            ds[mt][e] = 0.f;
            if (pos < rows) {
                const float w = (y > 0.5f ? pos_weight : 1.f) * inv_norm;
                const float x = s[mt][e];
                const float sig = 1.f / (1.f + expf(-x));
                ds[mt][e] = w * (sig - y);
                if (r16 == 0u) {
                    loss += w * (fmaxf(x, 0.f) - x * y + log1pf(expf(-fabsf(x))));
                }
            }
        }
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) loss += __shfl_xor(loss, d, 64);
    if (l == 0u) loss_part[tile] = loss;

    float w3[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) w3[nt] = W.w3[16 * nt + r16];
    // db3 = sum_r dS[r]
    {
        float db3 = 0.f;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
            for (int e = 0; e < 4; ++e) db3 += ds[mt][e];
        }
        db3 = (r16 == 0u) ? db3 : 0.f;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) db3 += __shfl_xor(db3, d, 64);
        if (l == 0u) gp[kOffB3] = db3;
    }
    // dW3[c] = sum_r H2[r][c] dS[r]; dZ2 = dS W3^T * GELU'(Z2) (in place of
    // z2) and db2, from one GELU evaluation per entry
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        float accw = 0.f, accb = 0.f;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float h, gg;
                gelu_and_grad(z2[mt][nt][e], h, gg);
                accw += h * ds[mt][e];
                z2[mt][nt][e] = ds[mt][e] * w3[nt] * gg;
                accb += z2[mt][nt][e];
            }
        }
        accw += __shfl_xor(accw, 16, 64);
        accw += __shfl_xor(accw, 32, 64);
        accb += __shfl_xor(accb, 16, 64);
        accb += __shfl_xor(accb, 32, 64);
        if (l < 16u) {
            gp[kOffW3 + 16 * nt + r16] = accw;
            gp[kOffB2 + 16 * nt + r16] = accb;
        }
    }
    // dW2 = H1^T dZ2: A = H1^T (the transposed tile in LDS), B = dZ2 in D layout
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        bf16x4 b[4];
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
#pragma unroll
            for (int e = 0; e < 4; ++e) b[kt][e] = f2bf(z2[kt][nt][e]);
        }
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                const bf16x4 a = load4(lds_ht + (16 * mi + r16) * kHStride + 16 * kt + g4);
                acc = mfma16(a, b[kt], acc);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) gp[kOffW2 + (16 * mi + g4 + e) * kMlpHidden + 16 * nt + r16] = acc[e];
        }
    }
    // dZ2 -> LDS (A layout for dH1), replacing H1
    wave_lds_sync();
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                lds_h[(16 * mt + g4 + e) * kHStride + 16 * nt + r16] = f2bf(z2[mt][nt][e]);
            }
        }
    }
    wave_lds_sync();
    // dH1 = dZ2 W2^T, dZ1 = dH1 * GELU'(Z1) (in place of z1), db1
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        bf16x4 b[4];
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) b[kt] = load4(W.w2 + (16 * nt + r16) * kMlpHidden + 16 * kt + g4);
        float accb = 0.f;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                const bf16x4 a = load4(lds_h + (16 * mt + r16) * kHStride + 16 * kt + g4);
                acc = mfma16(a, b[kt], acc);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                z1[mt][nt][e] = acc[e] * gelu_grad(z1[mt][nt][e]);
                accb += z1[mt][nt][e];
            }
        }
        accb += __shfl_xor(accb, 16, 64);
        accb += __shfl_xor(accb, 32, 64);
        if (l < 16u) gp[kOffB1 + 16 * nt + r16] = accb;
    }
    // dW1 = X^T dZ1: A = X^T (the transposed tile in LDS), B = dZ1 in D layout
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            const bf16x4 a = load4(lds_xt + r16 * kHStride + 16 * kt + g4);
            bf16x4 b;
#pragma unroll
            for (int e = 0; e < 4; ++e) b[e] = f2bf(z1[kt][nt][e]);
            acc = mfma16(a, b, acc);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t fi = g4 + e;
            if (fi < static_cast<uint32_t>(kMlpIn)) gp[kOffW1 + fi * kMlpHidden + 16 * nt + r16] = acc[e];
        }
    }
}

struct adam_hyper {
    float lr, beta1, beta2, eps, wd;
};

// One optimiser step.  Block b owns parameters 64 b .. 64 b + 63 (one per
// lane): its 16 waves each sum a fixed stride of the batch's tile gradients
// (coalesced 256 B rows), the block combines them in wave order, then --
// with apply != 0 -- AdamW updates the parameter and writes its BF16 / FP32
// views (the k_mlp_pack layouts).  Block 0 also sums the tile losses into
// losses[batch_idx].  The last block to finish advances the training state.
constexpr int kStepWaves = 16;
constexpr int kStepLoads = 32;

__device__ inline void pack_param(uint32_t p, float w, short* w1t, short* w2, short* w2t,
                                  float* b1, float* b2, float* w3, float* b3) {
    if (p < static_cast<uint32_t>(kOffB1)) {
        const uint32_t in = p / kMlpHidden, o = p % kMlpHidden;
        w1t[o * kMlpInPad + in] = f2bf(w);
    } else if (p < static_cast<uint32_t>(kOffW2)) {
        b1[p - kOffB1] = w;
    } else if (p < static_cast<uint32_t>(kOffB2)) {
        const uint32_t r = (p - kOffW2) / kMlpHidden, c = (p - kOffW2) % kMlpHidden;
        w2[r * kMlpHidden + c] = f2bf(w);   // [in r][out c]
        w2t[c * kMlpHidden + r] = f2bf(w);  // [out c][in r]
    } else if (p < static_cast<uint32_t>(kOffW3)) {
        b2[p - kOffB2] = w;
    } else if (p < static_cast<uint32_t>(kOffB3)) {
        w3[p - kOffW3] = bf2f(f2bf(w));
    } else {
        b3[0] = w;
    }
}

__global__ __launch_bounds__(64 * kStepWaves) void k_mlp_step(
    const float* grad_part, const float* loss_part, uint32_t n, uint32_t batch,
    train_state* state, int apply, adam_hyper h, float* params, float* m, float* v,
    float* grad_out, float* losses, short* w1t, short* w2, short* w2t, float* b1, float* b2,
    float* w3, float* b3) {
    __shared__ float s_part[kStepWaves][64];
    const train_state ts = *state;
    const uint32_t rows = batch_rows(ts, n, batch);
    const uint32_t tiles = (rows + 63u) / 64u;
    const uint32_t wave = threadIdx.x / 64u;
    const uint32_t l = __lane_id();
    const uint32_t p = blockIdx.x * 64u + l;
    const bool real = p < static_cast<uint32_t>(kMlpParams);
    // the wave's tiles in batches of kStepLoads loads in flight, added in
    // tile order (the same sum as one load at a time)
    float acc = 0.f;
    for (uint32_t t0 = wave; t0 < tiles; t0 += kStepWaves * kStepLoads) {
        float g[kStepLoads];
#pragma unroll
        for (int k = 0; k < kStepLoads; ++k) {
            const uint32_t t = t0 + static_cast<uint32_t>(k) * kStepWaves;
            g[k] = (real && t < tiles) ? grad_part[static_cast<size_t>(t) * kMlpParams + p] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < kStepLoads; ++k) {
            if (t0 + static_cast<uint32_t>(k) * kStepWaves < tiles) acc += g[k];
        }
    }
    s_part[wave][l] = acc;
    if (blockIdx.x == 0u && wave == 1u) {
        float lo = 0.f;
        for (uint32_t t0 = l; t0 < tiles; t0 += 64u * kStepLoads) {
            float x[kStepLoads];
#pragma unroll
            for (int k = 0; k < kStepLoads; ++k) {
                const uint32_t t = t0 + 64u * static_cast<uint32_t>(k);
                x[k] = t < tiles ? loss_part[t] : 0.f;
            }
#pragma unroll
            for (int k = 0; k < kStepLoads; ++k) {
                if (t0 + 64u * static_cast<uint32_t>(k) < tiles) lo += x[k];
            }
        }
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) lo += __shfl_xor(lo, d, 64);
        if (l == 0u) losses[ts.batch_idx] = lo;
    }
    __syncthreads();
    if (wave == 0u && real) {
        float g = 0.f;
#pragma unroll
        for (int w = 0; w < kStepWaves; ++w) g += s_part[w][l];
        if (grad_out != nullptr) grad_out[p] = g;
        if (apply) {
            const float t = static_cast<float>(ts.adam_t + 1u);
            const float bc1 = 1.f - powf(h.beta1, t);
            const float bc2 = 1.f - powf(h.beta2, t);
            const float mm = h.beta1 * m[p] + (1.f - h.beta1) * g;
            const float vv = h.beta2 * v[p] + (1.f - h.beta2) * g * g;
            m[p] = mm;
            v[p] = vv;
            const float w = params[p] - h.lr * ((mm / bc1) / (sqrtf(vv / bc2) + h.eps) +
                                                h.wd * params[p]);
            params[p] = w;
            pack_param(p, w, w1t, w2, w2t, b1, b2, w3, b3);
        }
    }
    // the last block to finish moves the state to the next batch (every
    // block read the state at its start, before counting itself)
    __syncthreads();
    if (threadIdx.x == 0u) {
        __threadfence();
        const uint32_t before = atomicAdd(&state->done_blocks, 1u);
        if (before + 1u == gridDim.x) {
            state->done_blocks = 0u;
            state->batch_idx = ts.batch_idx + 1u;
            state->adam_t = ts.adam_t + (apply ? 1u : 0u);
        }
    }
}

// BF16 / FP32 views of the master parameters in the fragment layouts.
__global__ void k_mlp_pack(const float* params, short* w1t, short* w2, short* w2t, float* b1,
                           float* b2, float* w3, float* b3) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 64u * 16u) {
        const uint32_t o = t / 16u, in = t % 16u;
        w1t[t] = in < static_cast<uint32_t>(kMlpIn) ? f2bf(params[kOffW1 + in * kMlpHidden + o]) : short(0);
    }
    if (t < 64u * 64u) {
        const uint32_t r = t / 64u, c = t % 64u;
        w2[t] = f2bf(params[kOffW2 + r * kMlpHidden + c]);   // [in r][out c]
        w2t[t] = f2bf(params[kOffW2 + c * kMlpHidden + r]);  // [out r][in c]
    }
    if (t < 64u) {
        b1[t] = params[kOffB1 + t];
        b2[t] = params[kOffB2 + t];
        w3[t] = bf2f(f2bf(params[kOffW3 + t]));
    }
    if (t == 0u) b3[0] = params[kOffB3];
}

// Per-feature count / mean / M2 for the normalisation in FP32 with Welford's
// update per thread and Chan, Golub and LeVeque's pairwise combination
// across the block (fixed order); one (count, mean[12], M2[12]) record per
// block, combined the same way on the host.
struct moments {
    float n;
    float mean[kMlpIn];
    float m2[kMlpIn];
};

__device__ __host__ inline void moments_merge(moments& a, const moments& b) {
    const float n = a.n + b.n;
    if (b.n == 0.f) return;
    if (a.n == 0.f) {
        a = b;
        return;
    }
    const float wb = b.n / n;
    const float cross = a.n * b.n / n;
    for (int k = 0; k < kMlpIn; ++k) {
        const float d = b.mean[k] - a.mean[k];
        a.mean[k] += d * wb;
        a.m2[k] += b.m2[k] + d * d * cross;
    }
    a.n = n;
}

constexpr int kMomentThreads = 256;

__global__ __launch_bounds__(kMomentThreads) void k_mlp_moments(const float* feats, uint32_t n,
                                                                moments* out) {
    __shared__ moments s[kMomentThreads];
    moments me;
    me.n = 0.f;
    for (int k = 0; k < kMlpIn; ++k) me.mean[k] = me.m2[k] = 0.f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        me.n += 1.f;
        const float inv = 1.f / me.n;
        for (int k = 0; k < kMlpIn; ++k) {
            const float x = feats[static_cast<size_t>(i) * kMlpIn + k];
            const float d = x - me.mean[k];
            me.mean[k] += d * inv;
            me.m2[k] += d * (x - me.mean[k]);
        }
    }
    s[threadIdx.x] = me;
    __syncthreads();
    for (uint32_t w = kMomentThreads / 2; w >= 1u; w >>= 1) {
        if (threadIdx.x < w) {
            moments a = s[threadIdx.x];
            moments_merge(a, s[threadIdx.x + w]);
            s[threadIdx.x] = a;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0u) out[blockIdx.x] = s[0];
}

// Training state set on the stream (no host copy), for callers that record
// the call into a graph.
__global__ void k_set_state(train_state* s, perm_map perm, uint32_t batch_idx, uint32_t adam_t) {
    if (blockIdx.x == 0u && threadIdx.x == 0u) {
        s->perm = perm;
        s->batch_idx = batch_idx;
        s->adam_t = adam_t;
        s->done_blocks = 0u;
        s->pad = 0u;
    }
}

// One 16x16x16 product through the documented fragment layout (diagnostics).
__global__ void k_mfma_layout(const float* A, const float* B, float* D) {
    const uint32_t l = threadIdx.x;
    const uint32_t r16 = l % 16u, g4 = 4u * (l / 16u);
    bf16x4 a, b;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        a[e] = f2bf(A[r16 * 16u + g4 + e]);
        b[e] = f2bf(B[(g4 + e) * 16u + r16]);
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma16(a, b, acc);
#pragma unroll
    for (int e = 0; e < 4; ++e) D[(g4 + e) * 16u + r16] = acc[e];
}

}  // namespace

// Scratch of the allocation-free gradient (ckf_scorer_gradient_noalloc): the
// per-tile partial gradients and losses of up to kNoAllocMaxSamples samples
// (1,024 tiles x 5,057 parameters + 1,024 losses, 20.7 MB).  Every GPU the
// library runs on holds its own copy until the process ends.
constexpr uint32_t kNoAllocMaxSamples = CKF_SCORER_NOALLOC_MAX_SAMPLES;
constexpr uint32_t kNoAllocTiles = kNoAllocMaxSamples / 64u;
static __device__ __align__(256) float scratch[size_t(kNoAllocTiles) * kMlpParams + kNoAllocTiles];

}  // namespace tamd

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct ckf_scorer {
    float* d = nullptr;       // one allocation: params, m, v, grad, packed views
    float* params = nullptr;  // FP32 master
    float* m = nullptr;
    float* v = nullptr;
    float* grad = nullptr;
    float* loss = nullptr;
    short* w1t = nullptr;
    short* w2 = nullptr;
    short* w2t = nullptr;
    float* b1 = nullptr;
    float* b2 = nullptr;
    float* w3 = nullptr;
    float* b3 = nullptr;
    float* mean = nullptr;
    float* inv_std = nullptr;
    tamd::train_state* state = nullptr;  // device training state
    float h_mean[tamd::kMlpIn];
    float h_std[tamd::kMlpIn];
    uint64_t adam_t = 0;
    ~ckf_scorer() {
        if (d) (void)hipFree(d);
    }
    tamd::mlp_dev view() const {
        return tamd::mlp_dev{w1t, w2, w2t, b1, b2, w3, b3, mean, inv_std};
    }
};

namespace {

using tamd::guarded;
using tamd::hip_check;
using tamd::ckf_error;

void repack(ckf_scorer& S, hipStream_t st) {
    hipLaunchKernelGGL(tamd::k_mlp_pack, dim3(16), dim3(256), 0, st, S.params, S.w1t, S.w2, S.w2t,
                       S.b1, S.b2, S.w3, S.b3);
    hip_check(hipGetLastError(), "scorer pack");
}

void upload_norm(ckf_scorer& S, hipStream_t st) {
    float buf[2 * tamd::kMlpIn];
    for (int k = 0; k < tamd::kMlpIn; ++k) {
        buf[k] = S.h_mean[k];
        buf[tamd::kMlpIn + k] = 1.f / (S.h_std[k] > 1e-12f ? S.h_std[k] : 1.f);
    }
    hip_check(hipMemcpyAsync(S.mean, buf, sizeof(buf), hipMemcpyHostToDevice, st), "scorer norm");
    hip_check(hipStreamSynchronize(st), "scorer norm sync");
}

uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

tamd::perm_map make_perm(uint64_t n, uint64_t& rng, bool identity) {
    if (identity || n <= 1) return tamd::perm_map{1, 0, n ? n : 1};
    uint64_t a = 0;
    do {
        a = splitmix(rng) % n;
    } while (a == 0 || gcd64(a, n) != 1);
    return tamd::perm_map{a, splitmix(rng) % n, n};
}

// Device address of `scratch` on the current device, looked up (with that
// device current) at the first call on each device and kept: the array's name
// in host code is a host-side stand-in, not a device address.
float* noalloc_scratch() {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static float* ptrs[kMaxDevices] = {};
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw ckf_error(CKF_ERR_UNSUPPORTED, "device index");
    std::lock_guard<std::mutex> lk(mu);
    if (ptrs[dev] == nullptr) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(tamd::scratch)), "scratch address");
        ptrs[dev] = static_cast<float*>(p);
    }
    return ptrs[dev];
}

// Set the device training state (sample order, batch index, Adam steps).
void set_state(ckf_scorer& S, const tamd::perm_map& perm, uint32_t batch_idx, hipStream_t st) {
    tamd::train_state ts{perm, batch_idx, static_cast<uint32_t>(S.adam_t), 0u, 0u};
    hip_check(hipMemcpyAsync(S.state, &ts, sizeof(ts), hipMemcpyHostToDevice, st), "scorer state");
    hip_check(hipStreamSynchronize(st), "scorer state");
}

// One optimiser step on the batch the device state names: the tile
// gradients, then their fixed-order sum with (apply) the AdamW update and the
// BF16 views; the state moves to the next batch.
void launch_step(ckf_scorer& S, const float* feats, const uint8_t* labels, uint32_t n,
                 uint32_t batch, float pos_weight, int apply, const tamd::adam_hyper& h,
                 float* part, float* loss_part, float* losses, float* grad_out, hipStream_t st) {
    const uint32_t tiles = (batch + 63u) / 64u;
    const uint32_t blocks = (tiles + tamd::kWavesPerBlock - 1) / tamd::kWavesPerBlock;
    hipLaunchKernelGGL(tamd::k_mlp_grad, dim3(blocks), dim3(tamd::kBlock), 0, st, S.view(), feats,
                       labels, S.state, n, batch, pos_weight, part, loss_part);
    hip_check(hipGetLastError(), "scorer grad");
    hipLaunchKernelGGL(tamd::k_mlp_step, dim3((tamd::kMlpParams + 63) / 64),
                       dim3(64 * tamd::kStepWaves), 0, st, part, loss_part, n, batch, S.state,
                       apply, h, S.params, S.m, S.v, grad_out, losses, S.w1t, S.w2, S.w2t, S.b1,
                       S.b2, S.w3, S.b3);
    hip_check(hipGetLastError(), "scorer step");
}

}  // namespace

extern "C" {

uint32_t ckf_scorer_n_params(void) { return static_cast<uint32_t>(tamd::kMlpParams); }

void ckf_scorer_train_config_default(ckf_scorer_train_config* c) {
    if (!c) return;
    c->epochs = 3;
    c->batch = 16384;
    c->learning_rate = 2e-3f;
    c->beta1 = 0.9f;
    c->beta2 = 0.999f;
    c->eps = 1e-8f;
    c->weight_decay = 1e-4f;
    c->pos_weight = 1.f;
    c->shuffle_seed = 1;
}

ckf_status ckf_scorer_create(uint64_t seed, ckf_scorer** out) {
    return guarded([&] {
        if (!out) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        auto* S = new ckf_scorer();
        const size_t P = tamd::kMlpParams;
        // floats: params, m, v, grad, loss(4), b1, b2, w3, b3(4), mean, inv_std, the
        // training state (12 words, 8-byte aligned); shorts after
        const size_t nf = 4 * P + 4 + 64 * 3 + 4 + 2 * tamd::kMlpIn + 12;
        const size_t ns = 64 * 16 + 64 * 64 * 2;
        if (hipMalloc(&S->d, nf * 4 + ns * 2 + 256) != hipSuccess) {
            delete S;
            throw ckf_error(CKF_ERR_OUT_OF_MEMORY, "scorer allocation failed");
        }
        float* f = S->d;
        S->params = f; f += P;
        S->m = f; f += P;
        S->v = f; f += P;
        S->grad = f; f += P;
        S->loss = f; f += 4;
        S->b1 = f; f += 64;
        S->b2 = f; f += 64;
        S->w3 = f; f += 64;
        S->b3 = f; f += 4;
        S->mean = f; f += tamd::kMlpIn;
        S->inv_std = f; f += tamd::kMlpIn;
        S->state = reinterpret_cast<tamd::train_state*>(f); f += 12;
        static_assert(sizeof(tamd::train_state) <= 12 * sizeof(float), "state words");
        if (reinterpret_cast<uintptr_t>(S->state) % alignof(tamd::train_state) != 0u) {
            delete S;
            throw ckf_error(CKF_ERR_INTERNAL, "training state misaligned");
        }
        short* sh = reinterpret_cast<short*>(f);
        S->w1t = sh; sh += 64 * 16;
        S->w2 = sh; sh += 64 * 64;
        S->w2t = sh;
        // scaled uniform initialisation (fan-in), zero biases
        std::vector<float> h(P, 0.f);
        uint64_t rng = seed;
        auto uni = [&]() { return static_cast<float>((splitmix(rng) >> 11) * (1.0 / 9007199254740992.0)) * 2.f - 1.f; };
        for (int i = 0; i < tamd::kMlpIn * 64; ++i) h[tamd::kOffW1 + i] = uni() * std::sqrt(6.f / tamd::kMlpIn);
        for (int i = 0; i < 64 * 64; ++i) h[tamd::kOffW2 + i] = uni() * std::sqrt(6.f / 64.f);
        for (int i = 0; i < 64; ++i) h[tamd::kOffW3 + i] = uni() * std::sqrt(3.f / 64.f);
        hip_check(hipMemcpy(S->params, h.data(), P * 4, hipMemcpyHostToDevice), "scorer init");
        hip_check(hipMemset(S->m, 0, 2 * P * 4), "scorer init");
        for (int k = 0; k < tamd::kMlpIn; ++k) {
            S->h_mean[k] = 0.f;
            S->h_std[k] = 1.f;
        }
        upload_norm(*S, nullptr);
        repack(*S, nullptr);
        hip_check(hipDeviceSynchronize(), "scorer init");
        *out = S;
    });
}

ckf_status ckf_scorer_destroy(ckf_scorer* s) {
    return guarded([&] { delete s; });
}

ckf_status ckf_scorer_set_params(ckf_scorer* s, const float* params, const float* mean,
                                 const float* std_dev) {
    return guarded([&] {
        if (!s) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null scorer");
        if (params) {
            hip_check(hipMemcpy(s->params, params, tamd::kMlpParams * 4, hipMemcpyHostToDevice), "scorer set");
            hip_check(hipMemset(s->m, 0, 2 * tamd::kMlpParams * 4), "scorer set");
            s->adam_t = 0;
        }
        if (mean) std::memcpy(s->h_mean, mean, sizeof(s->h_mean));
        if (std_dev) std::memcpy(s->h_std, std_dev, sizeof(s->h_std));
        upload_norm(*s, nullptr);
        repack(*s, nullptr);
        hip_check(hipDeviceSynchronize(), "scorer set");
    });
}

ckf_status ckf_scorer_get_params(const ckf_scorer* s, float* params, float* mean, float* std_dev) {
    return guarded([&] {
        if (!s) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null scorer");
        if (params) hip_check(hipMemcpy(params, s->params, tamd::kMlpParams * 4, hipMemcpyDeviceToHost), "scorer get");
        if (mean) std::memcpy(mean, s->h_mean, sizeof(s->h_mean));
        if (std_dev) std::memcpy(std_dev, s->h_std, sizeof(s->h_std));
    });
}

ckf_status ckf_scorer_fit_normalisation(ckf_scorer* s, const float* d_features, uint32_t n,
                                        void* stream) {
    return guarded([&] {
        if (!s || (!d_features && n)) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        if (n == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        constexpr uint32_t kBlocks = 512;
        tamd::moments* part = nullptr;
        hip_check(hipMalloc(&part, kBlocks * sizeof(tamd::moments)), "scorer moments");
        hipLaunchKernelGGL(tamd::k_mlp_moments, dim3(kBlocks), dim3(tamd::kMomentThreads), 0, st,
                           d_features, n, part);
        std::vector<tamd::moments> h(kBlocks);
        const hipError_t e1 = hipGetLastError();
        const hipError_t e2 = hipMemcpyAsync(h.data(), part, kBlocks * sizeof(tamd::moments),
                                             hipMemcpyDeviceToHost, st);
        const hipError_t e3 = hipStreamSynchronize(st);
        (void)hipFree(part);
        hip_check(e1, "scorer moments");
        hip_check(e2, "scorer moments");
        hip_check(e3, "scorer moments");
        // pairwise (tree) combination of the block records, as on the device
        for (uint32_t w = kBlocks / 2; w >= 1u; w >>= 1) {
            for (uint32_t b = 0; b < w; ++b) tamd::moments_merge(h[b], h[b + w]);
        }
        for (int k = 0; k < tamd::kMlpIn; ++k) {
            const float var = h[0].n > 0.f ? std::max(h[0].m2[k] / h[0].n, 0.f) : 0.f;
            const float sd = std::sqrt(var);
            s->h_mean[k] = h[0].mean[k];
            s->h_std[k] = sd > 1e-6f ? sd : 1.f;
        }
        upload_norm(*s, st);
    });
}

ckf_status ckf_scorer_forward(const ckf_scorer* s, const float* d_features, uint32_t n,
                              float* d_scores, void* stream) {
    return guarded([&] {
        if (!s || ((!d_features || !d_scores) && n)) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        if (n == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const uint32_t tiles = (n + 63u) / 64u;
        const uint32_t blocks = std::min<uint32_t>((tiles + 3u) / 4u, 4096u);
        hipLaunchKernelGGL(tamd::k_mlp_infer, dim3(blocks), dim3(tamd::kBlock), 0, st, s->view(),
                           d_features, n, d_scores);
        hip_check(hipGetLastError(), "scorer forward");
        hip_check(hipStreamSynchronize(st), "scorer forward sync");
    });
}

ckf_status ckf_scorer_gradient(const ckf_scorer* s, const float* d_features,
                               const uint8_t* d_labels, uint32_t n, float pos_weight,
                               float* grad, float* loss, void* stream) {
    return guarded([&] {
        if (!s || !d_features || !d_labels || n == 0) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "invalid argument");
        hipStream_t st = static_cast<hipStream_t>(stream);
        const uint32_t tiles = (n + 63u) / 64u;
        float* part = nullptr;
        hip_check(hipMalloc(&part, (size_t(tiles) * tamd::kMlpParams + tiles) * 4), "scorer grad");
        ckf_scorer& S = const_cast<ckf_scorer&>(*s);
        try {
            uint64_t rng = 0;
            set_state(S, make_perm(n, rng, true), 0u, st);
            launch_step(S, d_features, d_labels, n, n, pos_weight, 0, tamd::adam_hyper{},
                        part, part + size_t(tiles) * tamd::kMlpParams, S.loss, S.grad, st);
            if (grad) hip_check(hipMemcpyAsync(grad, s->grad, tamd::kMlpParams * 4, hipMemcpyDeviceToHost, st), "grad");
            if (loss) hip_check(hipMemcpyAsync(loss, s->loss, 4, hipMemcpyDeviceToHost, st), "loss");
            hip_check(hipStreamSynchronize(st), "scorer grad sync");
        } catch (...) {
            (void)hipFree(part);
            throw;
        }
        (void)hipFree(part);
    });
}

ckf_status ckf_scorer_gradient_noalloc(const ckf_scorer* s, const float* d_features,
                                       const uint8_t* d_labels, uint32_t n, float pos_weight,
                                       float* d_grad, float* d_loss, void* stream) {
    return guarded([&] {
        if (!s || !d_features || !d_labels || !d_grad || !d_loss || n == 0) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "invalid argument");
        }
        if (n > tamd::kNoAllocMaxSamples) {
            throw ckf_error(CKF_ERR_INVALID_ARGUMENT,
                            "n exceeds CKF_SCORER_NOALLOC_MAX_SAMPLES (65536) samples");
        }
        hipStream_t st = static_cast<hipStream_t>(stream);
        float* part = noalloc_scratch();
        const uint32_t tiles = (n + 63u) / 64u;
        ckf_scorer& S = const_cast<ckf_scorer&>(*s);
        uint64_t rng = 0;
        hipLaunchKernelGGL(tamd::k_set_state, dim3(1), dim3(64), 0, st, S.state,
                           make_perm(n, rng, true), 0u, static_cast<uint32_t>(S.adam_t));
        hip_check(hipGetLastError(), "scorer state");
        launch_step(S, d_features, d_labels, n, n, pos_weight, 0, tamd::adam_hyper{}, part,
                    part + size_t(tiles) * tamd::kMlpParams, d_loss, d_grad, st);
    });
}

ckf_status ckf_scorer_train(ckf_scorer* s, const float* d_features, const uint8_t* d_labels,
                            uint32_t n, const ckf_scorer_train_config* cfg,
                            ckf_scorer_train_stats* stats, void* stream) {
    return guarded([&] {
        if (!s || !d_features || !d_labels || !cfg || n == 0) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "invalid argument");
        if (cfg->batch == 0 || cfg->batch % 64u != 0) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "batch must be a positive multiple of 64");
        // the steps are captured into a graph, which the legacy default
        // stream cannot be: without a caller stream the training runs on a
        // stream of its own, after everything queued on the default stream
        hipStream_t st = static_cast<hipStream_t>(stream);
        hipStream_t own = nullptr;
        if (st == nullptr) {
            hip_check(hipDeviceSynchronize(), "scorer train");
            hip_check(hipStreamCreateWithFlags(&own, hipStreamNonBlocking), "scorer stream");
            st = own;
        }
        struct stream_guard {
            hipStream_t s;
            ~stream_guard() {
                if (s) (void)hipStreamDestroy(s);
            }
        } own_guard{own};
        const uint32_t batch = std::min(cfg->batch, ((n + 63u) / 64u) * 64u);
        const uint32_t tiles = (batch + 63u) / 64u;
        const uint32_t n_batches = (n + batch - 1) / batch;
        // steps per captured graph: the step kernels read their batch from
        // the device state, so one capture serves every replay
        constexpr uint32_t kGraphSteps = 32;
        float* part = nullptr;
        hip_check(hipMalloc(&part, (size_t(tiles) * tamd::kMlpParams + tiles + n_batches) * 4),
                  "scorer train");
        float* loss_part = part + size_t(tiles) * tamd::kMlpParams;
        float* losses = loss_part + tiles;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipGraphExec_t graph = nullptr;
        try {
            hip_check(hipEventCreate(&e0), "event");
            hip_check(hipEventCreate(&e1), "event");
            const tamd::adam_hyper h{cfg->learning_rate, cfg->beta1, cfg->beta2, cfg->eps,
                                     cfg->weight_decay};
            if (n_batches >= kGraphSteps) {
                hipGraph_t g = nullptr;
                hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "capture");
                for (uint32_t k = 0; k < kGraphSteps; ++k) {
                    launch_step(*s, d_features, d_labels, n, batch, cfg->pos_weight, 1, h, part,
                                loss_part, losses, nullptr, st);
                }
                hip_check(hipStreamEndCapture(st, &g), "capture");
                const hipError_t e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                hip_check(e, "graph instantiate");
            }
            uint64_t rng = cfg->shuffle_seed;
            uint32_t steps = 0;
            float first = 0.f, last = 0.f;
            std::vector<float> h_losses(n_batches, 0.f);
            hip_check(hipEventRecord(e0, st), "event");
            for (uint32_t ep = 0; ep < cfg->epochs; ++ep) {
                set_state(*s, make_perm(n, rng, false), 0u, st);
                uint32_t done = 0;
                for (; graph != nullptr && done + kGraphSteps <= n_batches; done += kGraphSteps) {
                    hip_check(hipGraphLaunch(graph, st), "graph launch");
                }
                for (; done < n_batches; ++done) {
                    launch_step(*s, d_features, d_labels, n, batch, cfg->pos_weight, 1, h, part,
                                loss_part, losses, nullptr, st);
                }
                s->adam_t += n_batches;
                steps += n_batches;
                hip_check(hipMemcpyAsync(h_losses.data(), losses, 4u * n_batches,
                                         hipMemcpyDeviceToHost, st), "losses");
                hip_check(hipStreamSynchronize(st), "scorer epoch");
                double sum = 0.0, wsum = 0.0;
                for (uint32_t bi = 0; bi < n_batches; ++bi) {
                    const uint32_t rows = std::min(batch, n - bi * batch);
                    sum += double(h_losses[bi]) * rows;
                    wsum += rows;
                }
                const float mean_loss = static_cast<float>(sum / wsum);
                if (ep == 0) first = mean_loss;
                last = mean_loss;
            }
            hip_check(hipEventRecord(e1, st), "event");
            hip_check(hipEventSynchronize(e1), "event");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
            if (stats) {
                stats->steps = steps;
                stats->first_loss = first;
                stats->last_loss = last;
                stats->seconds = ms * 1e-3;
                // forward 2 * 4,928 MACs, backward twice that, per sample
                const double flop = 3.0 * 2.0 * (12.0 * 64 + 64.0 * 64 + 64.0) * double(n) * cfg->epochs;
                stats->tflops = ms > 0.f ? flop / (ms * 1e-3) / 1e12 : 0.0;
            }
        } catch (...) {
            if (graph) (void)hipGraphExecDestroy(graph);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            (void)hipFree(part);
            throw;
        }
        if (graph) (void)hipGraphExecDestroy(graph);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(part);
    });
}

ckf_status ckf_debug_mfma_layout(const float* A, const float* B, float* D) {
    return guarded([&] {
        if (!A || !B || !D) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "null argument");
        float* d = nullptr;
        hip_check(hipMalloc(&d, 3 * 256 * 4), "layout");
        hipError_t e = hipMemcpy(d, A, 256 * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d + 256, B, 256 * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(tamd::k_mfma_layout, dim3(1), dim3(64), 0, nullptr, d, d + 256, d + 512);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpy(D, d + 512, 256 * 4, hipMemcpyDeviceToHost);
        (void)hipFree(d);
        hip_check(e, "layout");
    });
}

}  // extern "C"

tamd::mlp_dev tamd::scorer_view(const ckf_scorer* s) { return s->view(); }

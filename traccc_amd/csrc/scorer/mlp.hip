// Branch-scorer kernels: inference, gradient (forward + backward), Adam, and
// the C ABI of include/traccc_amd/scorer.h.
//
// Training runs on BF16 matrix cores with FP32 accumulation: the forward
// pass as in mlp_device.hpp; the backward pass forms
//   dW3 = H2^T dS,  dZ2 = dS W3^T * GELU'(Z2),  dW2 = H1^T dZ2,
//   dZ1 = (dZ2 W2^T) * GELU'(Z1),  dW1 = X^T dZ1
// with dZ2 / dZ1 rounded to BF16 as MFMA operands.  Every wave writes its
// tile's full gradient to its own slot and a second kernel sums the slots in
// a fixed order, so training is bitwise reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../common/errors.hpp"
#include "mlp_device.hpp"
#include "traccc_amd/scorer.h"

namespace tamd {

namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kBlock = 64 * kWavesPerBlock;

// Pseudo-random bijection of [0, n): i -> (a i + b) mod n, gcd(a, n) = 1.
struct perm_map {
    uint64_t a, b, n;
    __device__ uint32_t operator()(uint64_t i) const {
        return static_cast<uint32_t>((a * i + b) % n);
    }
};

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

// Gradient of one 64-sample tile per wave into grad_part[tile][kMlpParams].
__global__ __launch_bounds__(kBlock) void k_mlp_grad(mlp_dev W, const float* feats,
                                                      const uint8_t* labels, perm_map perm,
                                                      uint32_t row0, uint32_t rows,
                                                      float pos_weight, float inv_norm,
                                                      float* grad_part, float* loss_part) {
    __shared__ short lds[kWavesPerBlock * kMlpLdsPerWave];
    const uint32_t wave = threadIdx.x / 64u;
    short* lds_x = lds + wave * kMlpLdsPerWave;
    short* lds_h = lds_x + 64 * kMlpInPad;
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
    float f[kMlpIn];
    {
        const uint32_t row = valid ? perm(row0 + i) : 0u;
#pragma unroll
        for (int k = 0; k < kMlpIn; ++k) f[k] = valid ? feats[static_cast<size_t>(row) * kMlpIn + k] : 0.f;
    }
    float z1[4][4][4], z2[4][4][4], s[4][4];
    mlp_tile_layers(W, f, valid, lds_x, lds_h, z1, z2);
    mlp_tile_output(W, z2, s);

    // loss and dL/ds for the rows this lane holds
    float ds[4][4];
    float loss = 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t pos = tile * 64u + 16u * mt + g4 + e;
            ds[mt][e] = 0.f;
            if (pos < rows) {
                const float y = labels[perm(row0 + pos)] ? 1.f : 0.f;
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
    // dW3[c] = sum_r H2[r][c] dS[r]; db3 = sum_r dS[r]
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
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            float acc = 0.f;
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
                for (int e = 0; e < 4; ++e) acc += gelu(z2[mt][nt][e]) * ds[mt][e];
            }
            acc += __shfl_xor(acc, 16, 64);
            acc += __shfl_xor(acc, 32, 64);
            if (l < 16u) gp[kOffW3 + 16 * nt + r16] = acc;
        }
    }
    // dZ2 (in place of z2) and db2
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        float acc = 0.f;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                z2[mt][nt][e] = ds[mt][e] * w3[nt] * gelu_grad(z2[mt][nt][e]);
                acc += z2[mt][nt][e];
            }
        }
        acc += __shfl_xor(acc, 16, 64);
        acc += __shfl_xor(acc, 32, 64);
        if (l < 16u) gp[kOffB2 + 16 * nt + r16] = acc;
    }
    // dW2 = H1^T dZ2: A = H1^T (from the H1 tile in LDS), B = dZ2 in D layout
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
                bf16x4 a;
#pragma unroll
                for (int e = 0; e < 4; ++e) a[e] = lds_h[(16 * kt + g4 + e) * kHStride + 16 * mi + r16];
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
    // dW1 = X^T dZ1: A = X^T (from the X tile in LDS), B = dZ1 in D layout
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            bf16x4 a, b;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                a[e] = lds_x[(16 * kt + g4 + e) * kMlpInPad + r16];
                b[e] = f2bf(z1[kt][nt][e]);
            }
            acc = mfma16(a, b, acc);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t fi = g4 + e;
            if (fi < static_cast<uint32_t>(kMlpIn)) gp[kOffW1 + fi * kMlpHidden + 16 * nt + r16] = acc[e];
        }
    }
}

// grad[p] = sum over tiles in order; loss = sum of the tile losses.
__global__ void k_mlp_reduce(const float* grad_part, const float* loss_part, uint32_t n_tiles,
                             float* grad, float* loss) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < static_cast<uint32_t>(kMlpParams)) {
        float acc = 0.f;
        for (uint32_t t = 0; t < n_tiles; ++t) acc += grad_part[static_cast<size_t>(t) * kMlpParams + p];
        grad[p] = acc;
    }
    if (p == 0u) {
        float acc = 0.f;
        for (uint32_t t = 0; t < n_tiles; ++t) acc += loss_part[t];
        *loss = acc;
    }
}

struct adam_hyper {
    float lr, beta1, beta2, eps, wd, bc1, bc2;
};

__global__ void k_mlp_adam(float* params, const float* grad, float* m, float* v, adam_hyper h) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= static_cast<uint32_t>(kMlpParams)) return;
    const float g = grad[p];
    const float mm = h.beta1 * m[p] + (1.f - h.beta1) * g;
    const float vv = h.beta2 * v[p] + (1.f - h.beta2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mhat = mm / h.bc1;
    const float vhat = vv / h.bc2;
    params[p] -= h.lr * (mhat / (sqrtf(vhat) + h.eps) + h.wd * params[p]);
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

// Per-feature sums of x and x^2 (double) for the normalisation.
__global__ void k_mlp_moments(const float* feats, uint32_t n, double* acc) {
    __shared__ double s[2 * kMlpIn];
    for (int k = threadIdx.x; k < 2 * kMlpIn; k += blockDim.x) s[k] = 0.0;
    __syncthreads();
    double a[kMlpIn], b[kMlpIn];
    for (int k = 0; k < kMlpIn; ++k) a[k] = b[k] = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int k = 0; k < kMlpIn; ++k) {
            const double x = feats[static_cast<size_t>(i) * kMlpIn + k];
            a[k] += x;
            b[k] += x * x;
        }
    }
    for (int k = 0; k < kMlpIn; ++k) {
        atomicAdd(&s[k], a[k]);
        atomicAdd(&s[kMlpIn + k], b[k]);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 2 * kMlpIn; k += blockDim.x) atomicAdd(&acc[k], s[k]);
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

// gradient of `rows` samples starting at position row0 of the permutation;
// result in S.grad / S.loss (device)
void grad_batch(ckf_scorer& S, const float* feats, const uint8_t* labels, tamd::perm_map perm,
                uint32_t row0, uint32_t rows, float pos_weight, float* part, float* loss_part,
                hipStream_t st) {
    const uint32_t tiles = (rows + 63u) / 64u;
    const uint32_t blocks = (tiles + tamd::kWavesPerBlock - 1) / tamd::kWavesPerBlock;
    hipLaunchKernelGGL(tamd::k_mlp_grad, dim3(blocks), dim3(tamd::kBlock), 0, st, S.view(), feats,
                       labels, perm, row0, rows, pos_weight, 1.f / static_cast<float>(rows), part,
                       loss_part);
    hip_check(hipGetLastError(), "scorer grad");
    hipLaunchKernelGGL(tamd::k_mlp_reduce, dim3((tamd::kMlpParams + 255) / 256), dim3(256), 0, st,
                       part, loss_part, tiles, S.grad, S.loss);
    hip_check(hipGetLastError(), "scorer reduce");
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
        // floats: params, m, v, grad, loss(4), b1, b2, w3, b3(4), mean, inv_std; shorts after
        const size_t nf = 4 * P + 4 + 64 * 3 + 4 + 2 * tamd::kMlpIn;
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
        double* acc = nullptr;
        hip_check(hipMalloc(&acc, 2 * tamd::kMlpIn * sizeof(double)), "scorer moments");
        hip_check(hipMemsetAsync(acc, 0, 2 * tamd::kMlpIn * sizeof(double), st), "scorer moments");
        hipLaunchKernelGGL(tamd::k_mlp_moments, dim3(512), dim3(256), 0, st, d_features, n, acc);
        double h[2 * tamd::kMlpIn];
        const hipError_t e1 = hipGetLastError();
        const hipError_t e2 = hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, st);
        const hipError_t e3 = hipStreamSynchronize(st);
        (void)hipFree(acc);
        hip_check(e1, "scorer moments");
        hip_check(e2, "scorer moments");
        hip_check(e3, "scorer moments");
        for (int k = 0; k < tamd::kMlpIn; ++k) {
            const double mu = h[k] / n;
            const double var = std::max(h[tamd::kMlpIn + k] / n - mu * mu, 0.0);
            s->h_mean[k] = static_cast<float>(mu);
            s->h_std[k] = static_cast<float>(std::sqrt(var) > 1e-6 ? std::sqrt(var) : 1.0);
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
        uint64_t rng = 0;
        ckf_scorer& S = const_cast<ckf_scorer&>(*s);
        try {
            grad_batch(S, d_features, d_labels, make_perm(n, rng, true), 0u, n, pos_weight, part,
                       part + size_t(tiles) * tamd::kMlpParams, st);
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

ckf_status ckf_scorer_train(ckf_scorer* s, const float* d_features, const uint8_t* d_labels,
                            uint32_t n, const ckf_scorer_train_config* cfg,
                            ckf_scorer_train_stats* stats, void* stream) {
    return guarded([&] {
        if (!s || !d_features || !d_labels || !cfg || n == 0) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "invalid argument");
        if (cfg->batch == 0 || cfg->batch % 64u != 0) throw ckf_error(CKF_ERR_INVALID_ARGUMENT, "batch must be a positive multiple of 64");
        hipStream_t st = static_cast<hipStream_t>(stream);
        const uint32_t batch = std::min(cfg->batch, ((n + 63u) / 64u) * 64u);
        const uint32_t tiles = (batch + 63u) / 64u;
        float* part = nullptr;
        hip_check(hipMalloc(&part, (size_t(tiles) * tamd::kMlpParams + tiles) * 4), "scorer train");
        hipEvent_t e0 = nullptr, e1 = nullptr;
        try {
            hip_check(hipEventCreate(&e0), "event");
            hip_check(hipEventCreate(&e1), "event");
            uint64_t rng = cfg->shuffle_seed;
            uint32_t steps = 0;
            float first = 0.f, last = 0.f;
            std::vector<float> losses;
            hip_check(hipEventRecord(e0, st), "event");
            for (uint32_t ep = 0; ep < cfg->epochs; ++ep) {
                const tamd::perm_map perm = make_perm(n, rng, false);
                const uint32_t n_batches = (n + batch - 1) / batch;
                losses.assign(n_batches, 0.f);
                for (uint32_t bi = 0; bi < n_batches; ++bi) {
                    const uint32_t row0 = bi * batch;
                    const uint32_t rows = std::min(batch, n - row0);
                    grad_batch(*s, d_features, d_labels, perm, row0, rows, cfg->pos_weight, part,
                               part + size_t(tiles) * tamd::kMlpParams, st);
                    s->adam_t += 1;
                    tamd::adam_hyper h{cfg->learning_rate, cfg->beta1, cfg->beta2, cfg->eps,
                                       cfg->weight_decay,
                                       1.f - std::pow(cfg->beta1, static_cast<float>(s->adam_t)),
                                       1.f - std::pow(cfg->beta2, static_cast<float>(s->adam_t))};
                    hipLaunchKernelGGL(tamd::k_mlp_adam, dim3((tamd::kMlpParams + 255) / 256), dim3(256), 0, st,
                                       s->params, s->grad, s->m, s->v, h);
                    hip_check(hipGetLastError(), "scorer adam");
                    repack(*s, st);
                    hip_check(hipMemcpyAsync(&losses[bi], s->loss, 4, hipMemcpyDeviceToHost, st), "loss");
                    ++steps;
                }
                hip_check(hipStreamSynchronize(st), "scorer epoch");
                double sum = 0.0, wsum = 0.0;
                for (uint32_t bi = 0; bi < n_batches; ++bi) {
                    const uint32_t rows = std::min(batch, n - bi * batch);
                    sum += double(losses[bi]) * rows;
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
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            (void)hipFree(part);
            throw;
        }
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

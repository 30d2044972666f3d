// gw_mlp.hip -- the learned block scheduler's network: an MLP in -> h1 -> h2 -> 1 with
// ReLU hidden layers, BF16 weights and FP32 accumulation, trained on the GPU.
//
// Forward:  z1 = W1 x + b1, a1 = relu(z1); z2 = W2 a1 + b2, a2 = relu(z2); y = w3 . a2 + b3
// Loss:     mean squared error 1/n sum (y - t)^2
// Backward: dy = 2 (y - t) / n; dz2 = w3 * dy * [z2 > 0]; dz1 = W2^T dz2 * [z1 > 0];
//           dW3 = sum dy a2, dW2 = sum dz2 a1^T, dW1 = sum dz1 x^T (biases: sums of deltas)
// Update:   Adam on FP32 master weights, re-rounded to BF16 (round to nearest even) for use.
// Parameter order in every flat array: W1 (h1 x in, row major), b1, W2 (h2 x h1), b2,
// w3 (h2), b3.
//
// Sizes are tiny (12-64-64-1: 5,057 parameters); each kernel keeps the BF16 weights in LDS
// and gives one thread to one sample (forward / deltas) or to one parameter (gradients).
#include "gw_device.h"

namespace gw {

struct MlpShape {
    uint32_t in, h1, h2;
    uint32_t o_w1, o_b1, o_w2, o_b2, o_w3, o_b3, nparams;
};

__host__ __device__ inline MlpShape mlp_shape(uint32_t in, uint32_t h1, uint32_t h2)
{
    MlpShape s{in, h1, h2, 0, 0, 0, 0, 0, 0, 0};
    s.o_w1 = 0;
    s.o_b1 = s.o_w1 + h1 * in;
    s.o_w2 = s.o_b1 + h1;
    s.o_b2 = s.o_w2 + h2 * h1;
    s.o_w3 = s.o_b2 + h2;
    s.o_b3 = s.o_w3 + h2;
    s.nparams = s.o_b3 + 1;
    return s;
}

constexpr uint32_t kMaxHidden = 128, kMaxIn = 32, kMaxParams = 24576;

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__host__ __device__ inline uint16_t f32_to_bf16(float f)
{
    uint32_t u;
#ifdef __HIP_DEVICE_COMPILE__
    u = __float_as_uint(f);
#else
    __builtin_memcpy(&u, &f, 4);
#endif
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) return (uint16_t)((u >> 16) | 0x40);   // quiet NaN
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}

// forward for n samples; stores activations (a1, a2 pre-ReLU kept as z for the masks)
__global__ void __launch_bounds__(256) mlp_forward_kernel(MlpShape s, const uint16_t* __restrict__ wbf,
                                                           const float* __restrict__ x, uint32_t n, float* __restrict__ y,
                                                           float* __restrict__ z1, float* __restrict__ z2)
{
    extern __shared__ float w[];   // fp32 copy of the bf16 weights
    for (uint32_t i = threadIdx.x; i < s.nparams; i += blockDim.x) w[i] = bf16_to_f32(wbf[i]);
    __syncthreads();
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    float xi[kMaxIn];
    for (uint32_t k = 0; k < s.in; ++k) xi[k] = x[(uint64_t)r * s.in + k];
    float a1[kMaxHidden];
    for (uint32_t j = 0; j < s.h1; ++j) {
        float acc = w[s.o_b1 + j];
        for (uint32_t k = 0; k < s.in; ++k) acc = fmaf(w[s.o_w1 + j * s.in + k], xi[k], acc);
        if (z1) z1[(uint64_t)r * s.h1 + j] = acc;
        a1[j] = acc > 0.f ? acc : 0.f;
    }
    float out = w[s.o_b3];
    for (uint32_t j = 0; j < s.h2; ++j) {
        float acc = w[s.o_b2 + j];
        for (uint32_t k = 0; k < s.h1; ++k) acc = fmaf(w[s.o_w2 + j * s.h1 + k], a1[k], acc);
        if (z2) z2[(uint64_t)r * s.h2 + j] = acc;
        out = fmaf(w[s.o_w3 + j], acc > 0.f ? acc : 0.f, out);
    }
    y[r] = out;
}

// per-sample deltas: dz2 (n x h2) and dz1 (n x h1), from dy = 2 (y - t) / n
__global__ void __launch_bounds__(256) mlp_delta_kernel(MlpShape s, const uint16_t* __restrict__ wbf,
                                                         const float* __restrict__ y, const float* __restrict__ t,
                                                         uint32_t n, const float* __restrict__ z1,
                                                         const float* __restrict__ z2, float* __restrict__ dy,
                                                         float* __restrict__ dz1, float* __restrict__ dz2)
{
    extern __shared__ float w[];
    for (uint32_t i = threadIdx.x; i < s.nparams; i += blockDim.x) w[i] = bf16_to_f32(wbf[i]);
    __syncthreads();
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float d = 2.f * (y[r] - t[r]) / (float)n;
    dy[r] = d;
    float g2[kMaxHidden];
    for (uint32_t j = 0; j < s.h2; ++j) {
        g2[j] = z2[(uint64_t)r * s.h2 + j] > 0.f ? w[s.o_w3 + j] * d : 0.f;
        dz2[(uint64_t)r * s.h2 + j] = g2[j];
    }
    for (uint32_t k = 0; k < s.h1; ++k) {
        float acc = 0.f;
        for (uint32_t j = 0; j < s.h2; ++j) acc = fmaf(w[s.o_w2 + j * s.h1 + k], g2[j], acc);
        dz1[(uint64_t)r * s.h1 + k] = z1[(uint64_t)r * s.h1 + k] > 0.f ? acc : 0.f;
    }
}

// one thread per parameter: sum over samples
__global__ void __launch_bounds__(256) mlp_grad_kernel(MlpShape s, const float* __restrict__ x, uint32_t n,
                                                        const float* __restrict__ z1, const float* __restrict__ z2,
                                                        const float* __restrict__ dy, const float* __restrict__ dz1,
                                                        const float* __restrict__ dz2, float* __restrict__ g)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= s.nparams) return;
    float acc = 0.f;
    if (p < s.o_b1) {                       // W1[j][k]
        const uint32_t j = p / s.in, k = p % s.in;
        for (uint32_t r = 0; r < n; ++r) acc = fmaf(dz1[(uint64_t)r * s.h1 + j], x[(uint64_t)r * s.in + k], acc);
    } else if (p < s.o_w2) {                // b1[j]
        const uint32_t j = p - s.o_b1;
        for (uint32_t r = 0; r < n; ++r) acc += dz1[(uint64_t)r * s.h1 + j];
    } else if (p < s.o_b2) {                // W2[j][k]
        const uint32_t q = p - s.o_w2, j = q / s.h1, k = q % s.h1;
        for (uint32_t r = 0; r < n; ++r) {
            const float a = z1[(uint64_t)r * s.h1 + k];
            acc = fmaf(dz2[(uint64_t)r * s.h2 + j], a > 0.f ? a : 0.f, acc);
        }
    } else if (p < s.o_w3) {                // b2[j]
        const uint32_t j = p - s.o_b2;
        for (uint32_t r = 0; r < n; ++r) acc += dz2[(uint64_t)r * s.h2 + j];
    } else if (p < s.o_b3) {                // w3[j]
        const uint32_t j = p - s.o_w3;
        for (uint32_t r = 0; r < n; ++r) {
            const float a = z2[(uint64_t)r * s.h2 + j];
            acc = fmaf(dy[r], a > 0.f ? a : 0.f, acc);
        }
    } else {                                // b3
        for (uint32_t r = 0; r < n; ++r) acc += dy[r];
    }
    g[p] = acc;
}

// Adam on the fp32 master weights; the bf16 working copy is re-rounded from them
__global__ void __launch_bounds__(256) mlp_adam_kernel(uint32_t np, float* __restrict__ master, uint16_t* __restrict__ wbf,
                                                        const float* __restrict__ g, float* __restrict__ m1,
                                                        float* __restrict__ m2, float lr, float b1, float b2, float eps,
                                                        float c1, float c2)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const float gi = g[p];
    const float a = b1 * m1[p] + (1.f - b1) * gi;
    const float b = b2 * m2[p] + (1.f - b2) * gi * gi;
    m1[p] = a;
    m2[p] = b;
    const float upd = lr * (a / c1) / (sqrtf(b / c2) + eps);
    const float v = master[p] - upd;
    master[p] = v;
    wbf[p] = f32_to_bf16(v);
}

// ------------------------------------------------------------------------------------
// Fused kernels for MLPs whose weights and a chunk of activations fit in 64 KB of LDS
// (the 12-64-64-1 block scorer uses 54 KB).  Workgroup b takes the chunks b, b + G, ...
// of R samples: forward, per-sample deltas and the per-parameter gradient sums of its
// chunks in LDS, each thread owning up to kFusedOwn parameters whose partial gradients
// stay in registers and are written to row b of a partial-gradient table.  A second,
// single-workgroup kernel sums the G rows in order and applies Adam.  Every sum runs in
// a fixed order, so results are deterministic.
// MODE 0: forward (y); 1: partial gradients and squared errors.
// ------------------------------------------------------------------------------------
constexpr uint32_t kFusedThreads = 1024, kFusedOwn = 16, kFusedLdsFloats = 16384, kFusedMaxGroups = 32;

struct AdamArgs {
    float* master;
    float* m1;
    float* m2;
    uint16_t* wbf_out;
    float lr, b1, b2, eps, c1, c2;
};

__device__ __forceinline__ float relu(float v) { return v > 0.f ? v : 0.f; }

__device__ __forceinline__ float wave_sum_f(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int MODE>
__global__ void __launch_bounds__(1024) mlp_fused_kernel(MlpShape s, uint32_t R, const uint16_t* __restrict__ wbf,
                                                          const float* __restrict__ x, const float* __restrict__ t,
                                                          uint32_t n, float* __restrict__ y, float* __restrict__ gpart,
                                                          float* __restrict__ lpart)
{
    // LDS: weights, then the chunk's activations with odd row strides (in+1, h+1) so that
    // lanes walking consecutive rows hit distinct banks; weight reads are broadcasts
    extern __shared__ float lds[];
    __shared__ float red[kFusedThreads / 64];
    const uint32_t xs = s.in + 1, z1s = s.h1 + 1, z2s = s.h2 + 1;
    float* sW = lds;
    float* sX = sW + s.nparams;
    float* sZ1 = sX + R * xs;
    float* sZ2 = sZ1 + R * z1s;
    float* sD1 = sZ2 + R * z2s;
    float* sD2 = sD1 + R * z1s;
    float* sDY = sD2 + R * z2s;
    const uint32_t tid = threadIdx.x, nt = blockDim.x, lane = tid % 64, wave = tid / 64;
    for (uint32_t i = tid; i < s.nparams; i += nt) sW[i] = bf16_to_f32(wbf[i]);
    float acc[kFusedOwn];
#pragma unroll
    for (uint32_t i = 0; i < kFusedOwn; ++i) acc[i] = 0.f;
    float lsum = 0.f;
    __syncthreads();
    for (uint32_t c0 = blockIdx.x * R; c0 < n; c0 += gridDim.x * R) {
        const uint32_t rows = n - c0 < R ? n - c0 : R;
        for (uint32_t i = tid; i < rows * s.in; i += nt) sX[(i / s.in) * xs + i % s.in] = x[(uint64_t)c0 * s.in + i];
        __syncthreads();
        // outputs (r, j) with r fastest across lanes: weight row j is a broadcast
        for (uint32_t o = tid; o < R * s.h1; o += nt) {
            const uint32_t r = o % R, j = o / R;
            if (r >= rows) continue;
            float a = sW[s.o_b1 + j];
            for (uint32_t k = 0; k < s.in; ++k) a = fmaf(sW[s.o_w1 + j * s.in + k], sX[r * xs + k], a);
            sZ1[r * z1s + j] = a;
        }
        __syncthreads();
        for (uint32_t o = tid; o < R * s.h2; o += nt) {
            const uint32_t r = o % R, j = o / R;
            if (r >= rows) continue;
            float a = sW[s.o_b2 + j];
            for (uint32_t k = 0; k < s.h1; ++k) a = fmaf(sW[s.o_w2 + j * s.h1 + k], relu(sZ1[r * z1s + k]), a);
            sZ2[r * z2s + j] = a;
        }
        __syncthreads();
        for (uint32_t r = wave; r < rows; r += nt / 64) {
            float v = 0.f;
            for (uint32_t j = lane; j < s.h2; j += 64) v = fmaf(sW[s.o_w3 + j], relu(sZ2[r * z2s + j]), v);
            v = wave_sum_f(v) + sW[s.o_b3];
            if (lane == 0) {
                if (MODE == 0) {
                    y[c0 + r] = v;
                } else {
                    const float d = v - t[c0 + r];
                    lsum += d * d;
                    sDY[r] = 2.f * d / (float)n;
                }
            }
        }
        __syncthreads();
        if (MODE == 0) continue;
        for (uint32_t o = tid; o < R * s.h2; o += nt) {
            const uint32_t r = o % R, j = o / R;
            if (r >= rows) continue;
            sD2[r * z2s + j] = sZ2[r * z2s + j] > 0.f ? sW[s.o_w3 + j] * sDY[r] : 0.f;
        }
        __syncthreads();
        for (uint32_t o = tid; o < R * s.h1; o += nt) {
            const uint32_t r = o % R, k = o / R;
            if (r >= rows) continue;
            float a = 0.f;
            for (uint32_t j = 0; j < s.h2; ++j) a = fmaf(sW[s.o_w2 + j * s.h1 + k], sD2[r * z2s + j], a);
            sD1[r * z1s + k] = sZ1[r * z1s + k] > 0.f ? a : 0.f;
        }
        __syncthreads();
        // gradient sums over the chunk's rows for the parameters this thread owns
#pragma unroll
        for (uint32_t i = 0; i < kFusedOwn; ++i) {
            const uint32_t p = tid + i * nt;
            if (p >= s.nparams) break;
            float a = 0.f;
            if (p < s.o_b1) {
                const uint32_t j = p / s.in, k = p % s.in;
                for (uint32_t r = 0; r < rows; ++r) a = fmaf(sD1[r * z1s + j], sX[r * xs + k], a);
            } else if (p < s.o_w2) {
                const uint32_t j = p - s.o_b1;
                for (uint32_t r = 0; r < rows; ++r) a += sD1[r * z1s + j];
            } else if (p < s.o_b2) {
                const uint32_t q = p - s.o_w2, j = q / s.h1, k = q % s.h1;
                for (uint32_t r = 0; r < rows; ++r) a = fmaf(sD2[r * z2s + j], relu(sZ1[r * z1s + k]), a);
            } else if (p < s.o_w3) {
                const uint32_t j = p - s.o_b2;
                for (uint32_t r = 0; r < rows; ++r) a += sD2[r * z2s + j];
            } else if (p < s.o_b3) {
                const uint32_t j = p - s.o_w3;
                for (uint32_t r = 0; r < rows; ++r) a = fmaf(sDY[r], relu(sZ2[r * z2s + j]), a);
            } else {
                for (uint32_t r = 0; r < rows; ++r) a += sDY[r];
            }
            acc[i] += a;
        }
        __syncthreads();
    }
    if (MODE == 0) return;
    // squared errors: lane 0 of every wavefront holds its rows'
    if (lane == 0) red[wave] = lsum;
    __syncthreads();
    if (tid == 0) {
        float tot = 0.f;
        for (uint32_t w = 0; w < nt / 64; ++w) tot += red[w];
        lpart[blockIdx.x] = tot;
    }
#pragma unroll
    for (uint32_t i = 0; i < kFusedOwn; ++i) {
        const uint32_t p = tid + i * nt;
        if (p >= s.nparams) break;
        gpart[(uint64_t)blockIdx.x * s.nparams + p] = acc[i];
    }
}

// Sums the G partial gradients in order into g and the loss; with Adam, updates the fp32
// master weights and re-rounds the bf16 working copy.
template <bool ADAM>
__global__ void __launch_bounds__(1024) mlp_finish_kernel(uint32_t np, uint32_t G, const float* __restrict__ gpart,
                                                           const float* __restrict__ lpart, uint32_t n,
                                                           float* __restrict__ g, float* __restrict__ loss, AdamArgs adam)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float tot = 0.f;
        for (uint32_t b = 0; b < G; ++b) tot += lpart[b];
        loss[0] = tot / (float)n;
    }
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        float gi = 0.f;
        for (uint32_t b = 0; b < G; ++b) gi += gpart[(uint64_t)b * np + p];
        g[p] = gi;
        if (ADAM) {
            const float a = adam.b1 * adam.m1[p] + (1.f - adam.b1) * gi;
            const float b = adam.b2 * adam.m2[p] + (1.f - adam.b2) * gi * gi;
            adam.m1[p] = a;
            adam.m2[p] = b;
            const float v = adam.master[p] - adam.lr * (a / adam.c1) / (sqrtf(b / adam.c2) + adam.eps);
            adam.master[p] = v;
            adam.wbf_out[p] = f32_to_bf16(v);
        }
    }
}

// chunk of samples that fits next to the weights in LDS (0: the shape needs the
// multi-kernel path)
__host__ inline uint32_t fused_rows(const MlpShape& s)
{
    if (s.nparams > kFusedThreads * kFusedOwn || s.nparams >= kFusedLdsFloats) return 0;
    const uint32_t per = (s.in + 1) + 2 * (s.h1 + 1) + 2 * (s.h2 + 1) + 1;
    uint32_t r = (kFusedLdsFloats - s.nparams) / per;
    if (r > 64) r = 64;
    return r >= 8 ? r : 0;
}

__global__ void mlp_sqerr_kernel(const float* y, const float* t, uint32_t n, float* out)
{
    __shared__ float red[256];
    float acc = 0.f;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const float d = y[i] - t[i];
        acc += d * d;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] / (float)n;
}

}  // namespace gw

// ------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "gw.h"

namespace gw {
gw_status gen_fail(gw_ctx* ctx, gw_status st, const char* msg);
hipStream_t ctx_stream(gw_ctx* ctx);
}  // namespace gw

using namespace gw;

struct gw_mlp {
    MlpShape s;
    uint32_t cap = 0;             // sample capacity of the scratch buffers
    uint16_t* wbf = nullptr;      // bf16 weights
    float* master = nullptr;      // fp32 master weights
    float *m1 = nullptr, *m2 = nullptr, *g = nullptr;
    float *x = nullptr, *y = nullptr, *t = nullptr, *z1 = nullptr, *z2 = nullptr;
    float *dy = nullptr, *dz1 = nullptr, *dz2 = nullptr, *loss = nullptr;
    float* gpart = nullptr;       // kFusedMaxGroups x nparams partial gradients (+ losses)
    uint64_t steps = 0;
};

static void mlp_free_scratch(gw_mlp* m)
{
    float** bufs[] = {&m->x, &m->y, &m->t, &m->z1, &m->z2, &m->dy, &m->dz1, &m->dz2};
    for (float** b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    m->cap = 0;
}

static gw_status mlp_reserve(gw_ctx* ctx, gw_mlp* m, uint32_t n)
{
    if (n <= m->cap) return GW_OK;
    mlp_free_scratch(m);
    const MlpShape& s = m->s;
    struct B { float** p; uint64_t n; } bs[] = {
        {&m->x, (uint64_t)n * s.in}, {&m->y, n}, {&m->t, n}, {&m->z1, (uint64_t)n * s.h1},
        {&m->z2, (uint64_t)n * s.h2}, {&m->dy, n}, {&m->dz1, (uint64_t)n * s.h1}, {&m->dz2, (uint64_t)n * s.h2}};
    for (auto& b : bs)
        if (hipMalloc(b.p, b.n * 4) != hipSuccess) {
            mlp_free_scratch(m);
            return gen_fail(ctx, GW_ERR_NOMEM, "MLP scratch allocation failed");
        }
    m->cap = n;
    return GW_OK;
}

extern "C" gw_status gw_mlp_create(gw_ctx* ctx, uint32_t in, uint32_t h1, uint32_t h2, uint64_t seed, gw_mlp** out)
{
    if (!ctx || !out) return gen_fail(ctx, GW_ERR_ARG, "null argument");
    if (in == 0 || in > kMaxIn || h1 == 0 || h1 > kMaxHidden || h2 == 0 || h2 > kMaxHidden)
        return gen_fail(ctx, GW_ERR_ARG, "MLP sizes must be in [1, 32] x [1, 128] x [1, 128]");
    (void)ctx_stream(ctx);
    gw_mlp* m = new gw_mlp();
    m->s = mlp_shape(in, h1, h2);
    const uint32_t np = m->s.nparams;
    if (np > kMaxParams) { delete m; return gen_fail(ctx, GW_ERR_ARG, "MLP too large"); }
    // He-uniform initialisation from a splitmix64 stream of the seed
    std::vector<float> w(np, 0.f);
    uint64_t z = seed;
    auto uni = [&]() {
        z += 0x9E3779B97F4A7C15ULL;
        uint64_t x = z;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
        x ^= x >> 31;
        return (double)(x >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    const MlpShape& s = m->s;
    for (uint32_t i = 0; i < h1 * in; ++i) w[s.o_w1 + i] = (float)(uni() * std::sqrt(6.0 / in));
    for (uint32_t i = 0; i < h2 * h1; ++i) w[s.o_w2 + i] = (float)(uni() * std::sqrt(6.0 / h1));
    for (uint32_t i = 0; i < h2; ++i) w[s.o_w3 + i] = (float)(uni() * std::sqrt(6.0 / h2));
    std::vector<uint16_t> wb(np);
    for (uint32_t i = 0; i < np; ++i) {
        wb[i] = f32_to_bf16(w[i]);
        uint32_t u = (uint32_t)wb[i] << 16;
        std::memcpy(&w[i], &u, 4);   // master starts at the bf16 value
    }
    bool ok = hipMalloc(&m->wbf, np * 2) == hipSuccess && hipMalloc(&m->master, np * 4) == hipSuccess &&
              hipMalloc(&m->m1, np * 4) == hipSuccess && hipMalloc(&m->m2, np * 4) == hipSuccess &&
              hipMalloc(&m->g, np * 4) == hipSuccess && hipMalloc(&m->loss, 16) == hipSuccess &&
              hipMalloc(&m->gpart, ((size_t)kFusedMaxGroups * np + kFusedMaxGroups) * 4) == hipSuccess;
    ok = ok && hipMemcpy(m->wbf, wb.data(), np * 2, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(m->master, w.data(), np * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemset(m->m1, 0, np * 4) == hipSuccess && hipMemset(m->m2, 0, np * 4) == hipSuccess;
    if (!ok) {
        extern gw_status gw_mlp_destroy(gw_ctx*, gw_mlp*);
        gw_mlp_destroy(ctx, m);
        return gen_fail(ctx, GW_ERR_NOMEM, "MLP allocation failed");
    }
    *out = m;
    return GW_OK;
}

extern "C" gw_status gw_mlp_destroy(gw_ctx* ctx, gw_mlp* m)
{
    (void)ctx;
    if (!m) return GW_OK;
    mlp_free_scratch(m);
    void* ps[] = {m->wbf, m->master, m->m1, m->m2, m->g, m->loss, m->gpart};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    delete m;
    return GW_OK;
}

extern "C" gw_status gw_mlp_num_params(const gw_mlp* m, uint32_t* n)
{
    if (!m || !n) return gen_fail(nullptr, GW_ERR_ARG, "null argument");
    *n = m->s.nparams;
    return GW_OK;
}

// params: fp32 host array in the flat order; the device keeps bf16 copies (rounded)
extern "C" gw_status gw_mlp_set_params(gw_ctx* ctx, gw_mlp* m, const float* params)
{
    if (!ctx || !m || !params) return gen_fail(ctx, GW_ERR_ARG, "null argument");
    const uint32_t np = m->s.nparams;
    std::vector<uint16_t> wb(np);
    std::vector<float> master(np);
    for (uint32_t i = 0; i < np; ++i) {
        wb[i] = f32_to_bf16(params[i]);
        uint32_t u = (uint32_t)wb[i] << 16;
        std::memcpy(&master[i], &u, 4);
    }
    hipStream_t st = ctx_stream(ctx);
    if (hipMemcpyAsync(m->wbf, wb.data(), np * 2, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(m->master, master.data(), np * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemsetAsync(m->m1, 0, np * 4, st) != hipSuccess || hipMemsetAsync(m->m2, 0, np * 4, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return gen_fail(ctx, GW_ERR_HIP, "MLP parameter upload failed");
    m->steps = 0;
    return GW_OK;
}

// the bf16 weights in use, widened to fp32
extern "C" gw_status gw_mlp_get_params(gw_ctx* ctx, const gw_mlp* m, float* params)
{
    if (!ctx || !m || !params) return gen_fail(ctx, GW_ERR_ARG, "null argument");
    const uint32_t np = m->s.nparams;
    std::vector<uint16_t> wb(np);
    hipStream_t st = ctx_stream(ctx);
    if (hipMemcpyAsync(wb.data(), m->wbf, np * 2, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return gen_fail(ctx, GW_ERR_HIP, "MLP parameter read failed");
    for (uint32_t i = 0; i < np; ++i) {
        uint32_t u = (uint32_t)wb[i] << 16;
        std::memcpy(&params[i], &u, 4);
    }
    return GW_OK;
}

// x == nullptr: the inputs are already in m->x
static gw_status mlp_fwd(gw_ctx* ctx, gw_mlp* m, const float* x, uint32_t n, bool keep)
{
    gw_status st;
    if ((st = mlp_reserve(ctx, m, n)) != GW_OK) return st;
    hipStream_t s = ctx_stream(ctx);
    if (x && gw::h2d_small(ctx, m->x, x, (uint64_t)n * m->s.in * 4) != hipSuccess)
        return gen_fail(ctx, GW_ERR_HIP, "MLP input upload failed");
    const uint32_t lds = m->s.nparams * 4;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((n + 255) / 256), dim3(256), lds, s, m->s, m->wbf, m->x, n, m->y,
                       keep ? m->z1 : nullptr, keep ? m->z2 : nullptr);
    if (hipGetLastError() != hipSuccess) return gen_fail(ctx, GW_ERR_HIP, "MLP forward launch failed");
    return GW_OK;
}

// Fused path: inputs (and targets) through the transfer ring (x == nullptr: the inputs are
// already in m->x), one single-workgroup launch.
static gw_status mlp_fused(gw_ctx* ctx, gw_mlp* m, uint32_t R, int mode, const float* x, const float* t, uint32_t n,
                           float lr)
{
    gw_status st;
    if ((st = mlp_reserve(ctx, m, n)) != GW_OK) return st;
    hipStream_t s = ctx_stream(ctx);
    if ((x && gw::h2d_small(ctx, m->x, x, (uint64_t)n * m->s.in * 4) != hipSuccess) ||
        (t && gw::h2d_small(ctx, m->t, t, n * 4ull) != hipSuccess))
        return gen_fail(ctx, GW_ERR_HIP, "MLP input upload failed");
    const MlpShape& sh = m->s;
    const size_t lds = ((size_t)sh.nparams + (size_t)R * ((sh.in + 1) + 2 * (sh.h1 + 1) + 2 * (sh.h2 + 1) + 1)) * 4;
    AdamArgs a{m->master, m->m1, m->m2, m->wbf, lr, 0.9f, 0.999f, 1e-8f, 1.f, 1.f};
    if (mode == 2) {
        m->steps++;
        a.c1 = 1.f - std::pow(a.b1, (float)m->steps);
        a.c2 = 1.f - std::pow(a.b2, (float)m->steps);
    }
    uint32_t G = (n + R - 1) / R;
    if (G > kFusedMaxGroups) G = kFusedMaxGroups;
    float* lpart = m->gpart + (size_t)kFusedMaxGroups * sh.nparams;
    if (mode == 0) {
        hipLaunchKernelGGL(mlp_fused_kernel<0>, dim3(G), dim3(kFusedThreads), lds, s, sh, R, m->wbf, m->x, m->t, n,
                           m->y, m->gpart, lpart);
    } else {
        hipLaunchKernelGGL(mlp_fused_kernel<1>, dim3(G), dim3(kFusedThreads), lds, s, sh, R, m->wbf, m->x, m->t, n,
                           m->y, m->gpart, lpart);
        const uint32_t fg = (sh.nparams + kFusedThreads - 1) / kFusedThreads;
        if (mode == 1)
            hipLaunchKernelGGL(mlp_finish_kernel<false>, dim3(fg), dim3(kFusedThreads), 0, s, sh.nparams, G, m->gpart,
                               lpart, n, m->g, m->loss, a);
        else
            hipLaunchKernelGGL(mlp_finish_kernel<true>, dim3(fg), dim3(kFusedThreads), 0, s, sh.nparams, G, m->gpart,
                               lpart, n, m->g, m->loss, a);
    }
    if (hipGetLastError() != hipSuccess) return gen_fail(ctx, GW_ERR_HIP, "MLP kernel launch failed");
    return GW_OK;
}

// x[i] = (f[i] - mu[j]) / sigma[j], j = i mod nin; musig holds nin means then nin deviations
__global__ void standardize_kernel(const float* __restrict__ f, const float* __restrict__ musig, uint32_t nin,
                                   uint64_t total, float* __restrict__ x)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint32_t j = (uint32_t)(i % nin);
    x[i] = (f[i] - musig[j]) / musig[nin + j];
}

namespace gw {
uint32_t mlp_num_inputs(const gw_mlp* m) { return m->s.in; }

gw_status mlp_forward_device(gw_ctx* ctx, gw_mlp* m, const float* f, const float* musig, uint32_t n, const float** y)
{
    gw_status st;
    if ((st = mlp_reserve(ctx, m, n)) != GW_OK) return st;
    const uint64_t total = (uint64_t)n * m->s.in;
    hipLaunchKernelGGL(standardize_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, ctx_stream(ctx), f, musig,
                       m->s.in, total, m->x);
    if (hipGetLastError() != hipSuccess) return gen_fail(ctx, GW_ERR_HIP, "MLP input standardization failed");
    const uint32_t R = fused_rows(m->s);
    if ((st = R ? mlp_fused(ctx, m, R, 0, nullptr, nullptr, n, 0.f) : mlp_fwd(ctx, m, nullptr, n, false)) != GW_OK)
        return st;
    *y = m->y;
    return GW_OK;
}
}  // namespace gw

extern "C" gw_status gw_mlp_forward(gw_ctx* ctx, gw_mlp* m, const float* x, uint32_t n, float* y)
{
    if (!ctx || !m || !x || !y || n == 0) return gen_fail(ctx, GW_ERR_ARG, "bad MLP forward arguments");
    gw_status st;
    if (const uint32_t R = fused_rows(m->s)) {
        if ((st = mlp_fused(ctx, m, R, 0, x, nullptr, n, 0.f)) != GW_OK) return st;
        if (gw::d2h_small(ctx, y, m->y, n * 4ull) != hipSuccess) return gen_fail(ctx, GW_ERR_HIP, "MLP forward failed");
        return GW_OK;
    }
    if ((st = mlp_fwd(ctx, m, x, n, false)) != GW_OK) return st;
    hipStream_t s = ctx_stream(ctx);
    (void)s;
    if (gw::d2h_small(ctx, y, m->y, n * 4ull) != hipSuccess) return gen_fail(ctx, GW_ERR_HIP, "MLP forward failed");
    return GW_OK;
}

static gw_status mlp_grad(gw_ctx* ctx, gw_mlp* m, const float* x, const float* t, uint32_t n)
{
    gw_status st;
    if ((st = mlp_fwd(ctx, m, x, n, true)) != GW_OK) return st;
    hipStream_t s = ctx_stream(ctx);
    if (gw::h2d_small(ctx, m->t, t, n * 4ull) != hipSuccess) return gen_fail(ctx, GW_ERR_HIP, "MLP target upload failed");
    const uint32_t lds = m->s.nparams * 4;
    hipLaunchKernelGGL(mlp_delta_kernel, dim3((n + 255) / 256), dim3(256), lds, s, m->s, m->wbf, m->y, m->t, n, m->z1,
                       m->z2, m->dy, m->dz1, m->dz2);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3((m->s.nparams + 255) / 256), dim3(256), 0, s, m->s, m->x, n, m->z1, m->z2,
                       m->dy, m->dz1, m->dz2, m->g);
    hipLaunchKernelGGL(mlp_sqerr_kernel, dim3(1), dim3(256), 0, s, m->y, m->t, n, m->loss);
    if (hipGetLastError() != hipSuccess) return gen_fail(ctx, GW_ERR_HIP, "MLP backward launch failed");
    return GW_OK;
}

// gradient of the mean squared error with respect to every parameter (fp32, flat order)
extern "C" gw_status gw_mlp_backward(gw_ctx* ctx, gw_mlp* m, const float* x, const float* t, uint32_t n,
                                     float* grad, float* loss)
{
    if (!ctx || !m || !x || !t || n == 0) return gen_fail(ctx, GW_ERR_ARG, "bad MLP backward arguments");
    gw_status st;
    const uint32_t R = fused_rows(m->s);
    if ((st = R ? mlp_fused(ctx, m, R, 1, x, t, n, 0.f) : mlp_grad(ctx, m, x, t, n)) != GW_OK) return st;
    if ((grad && gw::d2h_small(ctx, grad, m->g, m->s.nparams * 4ull) != hipSuccess) ||
        (loss && gw::d2h_small(ctx, loss, m->loss, 4) != hipSuccess) ||
        hipStreamSynchronize(ctx_stream(ctx)) != hipSuccess)
        return gen_fail(ctx, GW_ERR_HIP, "MLP gradient readback failed");
    return GW_OK;
}

// one Adam step on a batch; returns the batch loss before the step
extern "C" gw_status gw_mlp_step(gw_ctx* ctx, gw_mlp* m, const float* x, const float* t, uint32_t n, float lr,
                                 float* loss)
{
    if (!ctx || !m || !x || !t || n == 0 || !(lr > 0.f)) return gen_fail(ctx, GW_ERR_ARG, "bad MLP step arguments");
    gw_status st;
    hipStream_t s = ctx_stream(ctx);
    if (const uint32_t R = fused_rows(m->s)) {
        if ((st = mlp_fused(ctx, m, R, 2, x, t, n, lr)) != GW_OK) return st;
        if (loss && gw::d2h_small(ctx, loss, m->loss, 4) != hipSuccess) return gen_fail(ctx, GW_ERR_HIP, "MLP step failed");
        return GW_OK;
    }
    if ((st = mlp_grad(ctx, m, x, t, n)) != GW_OK) return st;
    m->steps++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)m->steps), c2 = 1.f - std::pow(b2, (float)m->steps);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((m->s.nparams + 255) / 256), dim3(256), 0, s, m->s.nparams, m->master,
                       m->wbf, m->g, m->m1, m->m2, lr, b1, b2, 1e-8f, c1, c2);
    if (hipGetLastError() != hipSuccess) return gen_fail(ctx, GW_ERR_HIP, "MLP update launch failed");
    if (loss && gw::d2h_small(ctx, loss, m->loss, 4) != hipSuccess) return gen_fail(ctx, GW_ERR_HIP, "MLP step failed");
    return GW_OK;
}

// The learned pruner's network: an MLP 64 -> 128 -> 128 -> 1 (ReLU, ReLU,
// logit), bf16 weights with fp32 accumulation, and its training (weighted
// binary cross-entropy, Adam on fp32 master weights that are rounded to bf16
// after every step, so the forward pass always sees the bf16 weights).
//
// Parameters, packed (kMlpParams values; the bf16 copy has the same layout):
//   W1 [128][64] | b1 [128] | W2 [128][128] | b2 [128] | w3 [128] | b3 [1]
//
// The work is small next to the clustering (24.6k multiply-adds per sample),
// so the kernels use VALU fmaf over 32-sample tiles held in LDS.  Every sum
// runs in a fixed order: the forward pass per sample over its inputs in
// ascending order, the gradient per block over its tiles and then over the
// blocks (k_mlp_reduce), so results do not depend on timing.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "inscy_internal.h"

#pragma clang fp contract(off)

namespace inscy {

namespace {

constexpr int kIn = kMlpIn, kH = kMlpHidden;
constexpr int kThreads = 256;
constexpr int kTile = 32;             // samples per tile
constexpr int kHalf = kTile / 2;      // samples per thread in the layer passes
constexpr int kW1 = 0, kB1 = kW1 + kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH, kB3 = kW3 + kH;
static_assert(kB3 + 1 == kMlpParams, "parameter layout");
static_assert(kThreads == 2 * kH, "two threads per hidden unit");

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t) h << 16); }

// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    const uint32_t r = 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t) ((u + r) >> 16);
}

struct Tile {
    float x[kTile][kIn];
    float h1[kTile][kH];
    float h2[kTile][kH];  // after the error at layer 2 is formed: the error at layer 1
    float d2[kTile][kH];
    float z[kTile];
    float dz[kTile];
};

__device__ void load_tile(const float *__restrict__ X, long long N, long long t0, Tile &T, int &ns) {
    ns = (int) min((long long) kTile, N - t0);
    for (int e = threadIdx.x; e < kTile * kIn; e += blockDim.x) {
        const int s = e / kIn, i = e % kIn;
        T.x[s][i] = s < ns ? X[(t0 + s) * kIn + i] : 0.0f;
    }
    __syncthreads();
}

// out[s][j] = relu(b[j] + sum_i in[s][i] W[j][i]) for the thread's unit j and
// its half of the tile's samples
template <int K>
__device__ __forceinline__ void layer(const uint16_t *__restrict__ W, const uint16_t *__restrict__ b,
                                      const float (*in)[K], float (*out)[kH], int ns) {
    const int j = threadIdx.x & (kH - 1), s0 = (threadIdx.x >> 7) * kHalf;
    float acc[kHalf];
    const float bj = bf16_to_f32(b[j]);
#pragma unroll
    for (int s = 0; s < kHalf; ++s) acc[s] = bj;
    for (int i0 = 0; i0 < K; i0 += 8) {
        float w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) w[u] = bf16_to_f32(W[j * K + i0 + u]);
#pragma unroll
        for (int s = 0; s < kHalf; ++s)
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[s] = fmaf(in[s0 + s][i0 + u], w[u], acc[s]);
    }
#pragma unroll
    for (int s = 0; s < kHalf; ++s) out[s0 + s][j] = s0 + s < ns ? fmaxf(acc[s], 0.0f) : 0.0f;
}

// The tile through the network: h1, h2 (after ReLU) and the logits z.
__device__ void tile_forward(const uint16_t *__restrict__ P, Tile &T, int ns) {
    layer<kIn>(P + kW1, P + kB1, T.x, T.h1, ns);
    __syncthreads();
    layer<kH>(P + kW2, P + kB2, T.h1, T.h2, ns);
    __syncthreads();
    // logits: 8 consecutive lanes per sample, 16 terms each, then a fixed
    // butterfly over the 8 lanes
    {
        const int s = threadIdx.x >> 3, q = threadIdx.x & 7;
        float acc = 0.0f;
        for (int i = q * 16; i < q * 16 + 16; ++i) acc = fmaf(T.h2[s][i], bf16_to_f32(P[kW3 + i]), acc);
        acc += __shfl_xor(acc, 1);
        acc += __shfl_xor(acc, 2);
        acc += __shfl_xor(acc, 4);
        if (q == 0) T.z[s] = acc + bf16_to_f32(P[kB3]);
    }
    __syncthreads();
}

__global__ __launch_bounds__(kThreads) void k_mlp_predict(const float *__restrict__ X, long long N,
                                                          const uint16_t *__restrict__ P, float *__restrict__ prob) {
    __shared__ Tile T;
    for (long long t0 = (long long) blockIdx.x * kTile; t0 < N; t0 += (long long) gridDim.x * kTile) {
        int ns;
        load_tile(X, N, t0, T, ns);
        tile_forward(P, T, ns);
        if ((int) threadIdx.x < ns) prob[t0 + threadIdx.x] = 1.0f / (1.0f + expf(-T.z[threadIdx.x]));
        __syncthreads();
    }
}

// Gradient of  inv_wsum * sum_s wt_s * (softplus(z_s) - y_s z_s)  over the
// samples of this block's tiles, written to this block's row of `part`
// ([gridDim.x][kMlpParams + 1], the last column the loss).
// Thread t owns half of row t/2 of dW2 (64 entries) and of dW1 (32 entries)
// in registers.
__global__ __launch_bounds__(kThreads) void k_mlp_grad(const float *__restrict__ X, const float *__restrict__ Y,
                                                       const float *__restrict__ WT, long long N, float inv_wsum,
                                                       const uint16_t *__restrict__ P, float *__restrict__ part) {
    __shared__ Tile T;
    __shared__ float red[3][kThreads];
    const int t = threadIdx.x;
    const int j = t & (kH - 1), s0 = (t >> 7) * kHalf;  // layer passes: unit j, samples s0..s0+15
    const int r = t >> 1, c2 = (t & 1) * (kH / 2), c1 = (t & 1) * (kIn / 2);  // weight rows: row r, column halves
    float g2[kH / 2], g1[kIn / 2];
#pragma unroll
    for (int i = 0; i < kH / 2; ++i) g2[i] = 0.0f;
#pragma unroll
    for (int c = 0; c < kIn / 2; ++c) g1[c] = 0.0f;
    float gb1 = 0.0f, gb2 = 0.0f, gw3 = 0.0f, gb3 = 0.0f, lsum = 0.0f;
    const float w3 = bf16_to_f32(P[kW3 + j]);
    for (long long t0 = (long long) blockIdx.x * kTile; t0 < N; t0 += (long long) gridDim.x * kTile) {
        int ns;
        load_tile(X, N, t0, T, ns);
        tile_forward(P, T, ns);
        if (t < kTile) {
            float g = 0.0f;
            if (t < ns) {
                const float zz = T.z[t], y = Y[t0 + t], w = WT[t0 + t] * inv_wsum;
                const float p = 1.0f / (1.0f + expf(-zz));
                g = w * (p - y);
                // softplus(z) - y z, in a form that does not overflow
                lsum += w * (fmaxf(zz, 0.0f) + log1pf(expf(-fabsf(zz))) - y * zz);
            }
            T.dz[t] = g;
            gb3 += g;
        }
        __syncthreads();
        // layer 3 and the error at layer 2 (ReLU mask)
#pragma unroll
        for (int s = 0; s < kHalf; ++s) {
            const float h = T.h2[s0 + s][j], dz = T.dz[s0 + s];
            gw3 = fmaf(dz, h, gw3);
            const float v = h > 0.0f ? dz * w3 : 0.0f;
            T.d2[s0 + s][j] = v;
            gb2 += v;
        }
        __syncthreads();
        // error at layer 1: sum_jj d2[s][jj] W2[jj][j], masked by h1 > 0,
        // into the h2 slots (h2 is no longer needed)
        {
            float d1[kHalf];
#pragma unroll
            for (int s = 0; s < kHalf; ++s) d1[s] = 0.0f;
            for (int jj = 0; jj < kH; ++jj) {
                const float w = bf16_to_f32(P[kW2 + jj * kH + j]);
#pragma unroll
                for (int s = 0; s < kHalf; ++s) d1[s] = fmaf(T.d2[s0 + s][jj], w, d1[s]);
            }
#pragma unroll
            for (int s = 0; s < kHalf; ++s) {
                const float v = T.h1[s0 + s][j] > 0.0f ? d1[s] : 0.0f;
                T.h2[s0 + s][j] = v;
                gb1 += v;
            }
        }
        // dW2[r][c2 + i] += sum_s d2[s][r] h1[s][c2 + i]
#pragma unroll 2
        for (int s = 0; s < kTile; ++s) {
            const float e = T.d2[s][r];
#pragma unroll
            for (int i = 0; i < kH / 2; ++i) g2[i] = fmaf(e, T.h1[s][c2 + i], g2[i]);
        }
        __syncthreads();  // the layer-1 errors are complete
        // dW1[r][c1 + c] += sum_s d1[s][r] x[s][c1 + c]
#pragma unroll 2
        for (int s = 0; s < kTile; ++s) {
            const float e = T.h2[s][r];
#pragma unroll
            for (int c = 0; c < kIn / 2; ++c) g1[c] = fmaf(e, T.x[s][c1 + c], g1[c]);
        }
        __syncthreads();  // the next tile overwrites the tile buffers
    }
    float *row = part + (size_t) blockIdx.x * (kMlpParams + 1);
#pragma unroll
    for (int c = 0; c < kIn / 2; ++c) row[kW1 + r * kIn + c1 + c] = g1[c];
#pragma unroll
    for (int i = 0; i < kH / 2; ++i) row[kW2 + r * kH + c2 + i] = g2[i];
    // per-unit sums: the two sample halves of unit j, lower half first
    red[0][t] = gb1;
    red[1][t] = gb2;
    red[2][t] = gw3;
    __syncthreads();
    if (t < kH) {
        row[kB1 + t] = red[0][t] + red[0][t + kH];
        row[kB2 + t] = red[1][t] + red[1][t + kH];
        row[kW3 + t] = red[2][t] + red[2][t + kH];
    }
    // b3 and the loss: lanes 0..31 hold one sample slot each
    float a = gb3, l = lsum;
    for (int o = 16; o > 0; o >>= 1) {
        a += __shfl_down(a, o);
        l += __shfl_down(l, o);
    }
    if (t == 0) {
        row[kB3] = a;
        row[kMlpParams] = l;
    }
}

// grad[i] = sum over blocks b (in order) of part[b][i], i <= kMlpParams
__global__ void k_mlp_reduce(const float *__restrict__ part, int nb, float *__restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > kMlpParams) return;
    float s = 0.0f;
    for (int b = 0; b < nb; ++b) s += part[(size_t) b * (kMlpParams + 1) + i];
    grad[i] = s;
}

__global__ void k_mlp_adam(float *__restrict__ master, const float *__restrict__ grad, float *__restrict__ m,
                           float *__restrict__ v, uint16_t *__restrict__ P, int n, float lr, float b1, float b2,
                           float c1, float c2, float eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * (g * g);
    m[i] = mi;
    v[i] = vi;
    const float w = master[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    master[i] = w;
    P[i] = f32_to_bf16(w);
}

__global__ void k_mlp_round(const float *__restrict__ master, uint16_t *__restrict__ P, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) P[i] = f32_to_bf16(master[i]);
}

unsigned param_blocks() { return (unsigned) ((kMlpParams + 1 + 255) / 256); }

}  // namespace

hipError_t launch_mlp_round(const float *master, uint16_t *P, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_round, dim3(param_blocks()), dim3(256), 0, st, master, P, kMlpParams);
    return hipGetLastError();
}

hipError_t launch_mlp_predict(const float *X, long long N, const uint16_t *P, float *prob, hipStream_t st) {
    if (N <= 0) return hipSuccess;
    const long long tiles = (N + kTile - 1) / kTile;
    const unsigned grid = (unsigned) std::min<long long>(tiles, 8192);
    hipLaunchKernelGGL(k_mlp_predict, dim3(grid), dim3(kThreads), 0, st, X, N, P, prob);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const float *X, const float *Y, const float *WT, long long N, float inv_wsum,
                           const uint16_t *P, float *part, int nblocks, float *grad, hipStream_t st) {
    if (N <= 0 || nblocks <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblocks), dim3(kThreads), 0, st, X, Y, WT, N, inv_wsum, P, part);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(param_blocks()), dim3(256), 0, st, part, nblocks, grad);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float *master, const float *grad, float *m, float *v, uint16_t *P, int step, float lr,
                           hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - std::pow(b1, (float) step), c2 = 1.0f - std::pow(b2, (float) step);
    hipLaunchKernelGGL(k_mlp_adam, dim3(param_blocks()), dim3(256), 0, st, master, grad, m, v, P, kMlpParams, lr,
                       b1, b2, c1, c2, eps);
    return hipGetLastError();
}

}  // namespace inscy

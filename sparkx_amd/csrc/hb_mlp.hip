// hb_mlp.hip -- the learned surrogate: an MLP 6 -> 64 -> 64 -> 1 with GELU that maps
// a node's degree, log-degree and log HyperBall estimates N_1..N_4 to its log
// harmonic centrality.  Matmuls run on BF16 MFMA with f32 accumulation: products
// over the 64 hidden units on gfx950's v_mfma_f32_16x16x32_bf16 (K = 32 per
// instruction), products over 16 samples or the 6 (padded) inputs on
// v_mfma_f32_16x16x16_bf16; GELU, biases, loss and the Adam update run in f32, and
// every rounding to bf16 is the hardware's round-to-nearest-even convert.  One
// wavefront handles a tile of 16 samples; activations move between MFMA output and
// operand layouts through LDS.
//
// MFMA lane layout (lane l, g = l / 16): D (M x N, both shapes): D[4 g + j][l % 16],
// j = 0..3.  An A (M x K) register j of lane l multiplies the B (K x N) register j
// of the lane with the same g, so any map (g, j) -> k used for both operands is a
// valid K order; the gathers here use k = 4 g + j (16x16x16, j < 4) and k = 8 g + j
// (16x16x32, j < 8), with A[l % 16][k] and B[k][l % 16].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/sparkx_hb.h"
#include "hb_kernels.h"

namespace hb {
namespace mlp {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kDIn = 6, kDH = 64;
constexpr int oW1 = 0, oB1 = 384, oW2 = 448, oB2 = 4544, oW3 = 4608, oB3 = 4672, kNParams = 4673;
constexpr int kBlock = 128;  // two wavefronts, one 16-sample tile each per round
constexpr int kWaves = kBlock / 64;

// round to nearest even (v_cvt_pk_bf16_f32); inputs are finite
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bfr(float f) { return (float)(__bf16)f; }

// erf by Abramowitz & Stegun, Handbook of Mathematical Functions, 7.1.26 (absolute error
// at most 1.5e-7): erf(z) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-z^2), with
// t = 1 / (1 + p z), for z >= 0, odd in z.  About half the instructions of erff; the
// activations it feeds are rounded to bf16 (8 bits) right after.  Training and inference
// both use it, so the network is evaluated with the function it was trained with.
__device__ __forceinline__ float erf_as(float x)
{
    const float z = fabsf(x);
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, z, 1.0f));
    float q = fmaf(t, 1.061405429f, -1.453152027f);
    q = fmaf(t, q, 1.421413741f);
    q = fmaf(t, q, -0.284496736f);
    q = fmaf(t, q, 0.254829592f);
    const float e = __builtin_amdgcn_exp2f(-(z * z) * 1.44269504088896341f);  // exp(-z^2)
    return copysignf(fmaf(-(t * q), e, 1.0f), x);
}
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erf_as(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x)
{
    const float cdf = 0.5f * (1.0f + erf_as(x * 0.70710678118654752f));
    return cdf + x * (__builtin_amdgcn_exp2f(-0.5f * x * x * 1.44269504088896341f) * 0.39894228040143268f);
}

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
__device__ __forceinline__ v4f mfma32(v8bf a, v8bf b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

struct Shared {
    uint16_t w1[16][kDH];  // bf16, input rows 6..15 zero
    uint16_t w2[kDH][kDH];
    float w3[kDH];         // bf16-rounded
    float b1[kDH], b2[kDH], b3;
    float mean[kDIn], sd[kDIn];
};

struct Wave {
    float xs[16][16];      // bf16(normalised input), padded
    float h1[16][kDH];     // bf16(gelu(z1))
    float g1[16][kDH];     // gelu'(z1)
    float h2[16][kDH];     // bf16(gelu(z2))
    float g2[16][kDH];     // gelu'(z2), then dz1
    float d[16][kDH];      // dz2
    float y[16], t[16], dy[16];
};

// operand gathers (see the layout above)
__device__ __forceinline__ v4s opA(const float *a, int ld, int kc, int l)
{
    const float *p = a + (l & 15) * ld + kc + 4 * (l >> 4);
    v4s r;
    r[0] = (short)f2bf(p[0]);
    r[1] = (short)f2bf(p[1]);
    r[2] = (short)f2bf(p[2]);
    r[3] = (short)f2bf(p[3]);
    return r;
}
__device__ __forceinline__ v4s opAT(const float *src, int ld, int mc, int l)  // A[m][k] = src[k][m]
{
    v4s r;
    for (int j = 0; j < 4; ++j) r[j] = (short)f2bf(src[(4 * (l >> 4) + j) * ld + mc + (l & 15)]);
    return r;
}
__device__ __forceinline__ v4s opBw(const uint16_t *w, int ld, int kc, int nc, int l)  // B = W
{
    v4s r;
    for (int j = 0; j < 4; ++j) r[j] = (short)w[(kc + 4 * (l >> 4) + j) * ld + nc + (l & 15)];
    return r;
}
__device__ __forceinline__ v4s opBf(const float *src, int ld, int nc, int l)  // B[k][n] = src[k][n]
{
    v4s r;
    for (int j = 0; j < 4; ++j) r[j] = (short)f2bf(src[(4 * (l >> 4) + j) * ld + nc + (l & 15)]);
    return r;
}

// K = 32 gathers (k = kc + 8 g + j)
__device__ __forceinline__ v8bf opA8(const float *a, int ld, int kc, int l)
{
    const float *p = a + (l & 15) * ld + kc + 8 * (l >> 4);
    v8bf r;
    for (int j = 0; j < 8; ++j) r[j] = (__bf16)p[j];
    return r;
}
__device__ __forceinline__ v8bf opBw8(const uint16_t *w, int ld, int kc, int nc, int l)  // B = W
{
    v8bf r;
    for (int j = 0; j < 8; ++j) r[j] = __builtin_bit_cast(__bf16, w[(kc + 8 * (l >> 4) + j) * ld + nc + (l & 15)]);
    return r;
}
__device__ __forceinline__ v8bf opBwT8(const uint16_t *w, int ld, int kc, int nc, int l)  // B = W^T
{
    const uint16_t *p = w + (nc + (l & 15)) * ld + kc + 8 * (l >> 4);
    v8bf r;
    for (int j = 0; j < 8; ++j) r[j] = __builtin_bit_cast(__bf16, p[j]);
    return r;
}

struct Dev {
    const float *params;  // f32 master copy
    const float *norm;    // mean[6], std[6]
    const float *x;       // samples x 6
    const float *y;       // targets (training)
    const uint32_t *idx;  // sample order (training), or null
    uint64_t first, count;
    float inv_b;          // 1 / minibatch size
    float *grad;          // gradient accumulator (training)
    float *loss;          // sum of squared errors (training)
    float *out;           // predictions (inference)
};

__device__ void load_shared(Shared &sh, const Dev &d)
{
    for (int i = threadIdx.x; i < 16 * kDH; i += blockDim.x) {
        const int r = i / kDH, c = i % kDH;
        sh.w1[r][c] = r < kDIn ? f2bf(d.params[oW1 + r * kDH + c]) : (uint16_t)0;
    }
    for (int i = threadIdx.x; i < kDH * kDH; i += blockDim.x) sh.w2[i / kDH][i % kDH] = f2bf(d.params[oW2 + i]);
    for (int i = threadIdx.x; i < kDH; i += blockDim.x) {
        sh.w3[i] = bfr(d.params[oW3 + i]);
        sh.b1[i] = d.params[oB1 + i];
        sh.b2[i] = d.params[oB2 + i];
    }
    if (threadIdx.x < kDIn) {
        sh.mean[threadIdx.x] = d.norm[threadIdx.x];
        sh.sd[threadIdx.x] = d.norm[kDIn + threadIdx.x];
    }
    if (threadIdx.x == 0) sh.b3 = d.params[oB3];
}

// forward pass of one tile; rows >= valid are zero inputs
__device__ void forward_tile(const Shared &sh, Wave &w, const Dev &d, uint64_t base, int valid, int l, bool training)
{
    for (int i = l; i < 16 * 16; i += 64) {
        const int s = i / 16, k = i % 16;
        float v = 0.0f;
        if (s < valid && k < kDIn) {
            const uint64_t row = d.idx ? d.idx[base + s] : base + s;
            v = bfr(__fdiv_rn(__fsub_rn(d.x[row * kDIn + k], sh.mean[k]), sh.sd[k]));
        }
        w.xs[s][k] = v;
    }
    if (l < 16) {
        float tv = 0.0f;
        if (training && l < valid) tv = d.y[d.idx ? d.idx[base + l] : base + l];
        w.t[l] = tv;
    }
    __syncthreads();
    for (int nb = 0; nb < 4; ++nb) {
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mfma(opA(&w.xs[0][0], 16, 0, l), opBw(&sh.w1[0][0], kDH, 0, 16 * nb, l), acc);
        for (int j = 0; j < 4; ++j) {
            const int s = 4 * (l >> 4) + j, n = 16 * nb + (l & 15);
            const float z = acc[j] + sh.b1[n];
            w.h1[s][n] = bfr(gelu(z));
            w.g1[s][n] = gelu_grad(z);
        }
    }
    __syncthreads();
    for (int nb = 0; nb < 4; ++nb) {
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int kc = 0; kc < kDH; kc += 32)
            acc = mfma32(opA8(&w.h1[0][0], kDH, kc, l), opBw8(&sh.w2[0][0], kDH, kc, 16 * nb, l), acc);
        for (int j = 0; j < 4; ++j) {
            const int s = 4 * (l >> 4) + j, n = 16 * nb + (l & 15);
            const float z = acc[j] + sh.b2[n];
            w.h2[s][n] = bfr(gelu(z));
            w.g2[s][n] = gelu_grad(z);
        }
    }
    __syncthreads();
    // layer 3: lane (s, q) sums 16 features, quarters combined with xor shuffles
    {
        const int s = l & 15, q = l >> 4;
        float part = 0.0f;
        for (int k = 16 * q; k < 16 * q + 16; ++k) part += w.h2[s][k] * sh.w3[k];
        part += __shfl_xor(part, 16);
        part += __shfl_xor(part, 32);
        if (q == 0) w.y[s] = part + sh.b3;
    }
    __syncthreads();
}

// Inference only, activations kept in registers.  Both layers are computed
// transposed, z^T = W^T x^T, so the outputs of the layer-1 MFMAs (D[n][s]: lane l
// holds hidden units 16 mb + 4 g + j of sample l % 16, g = l / 16) are the B operand
// of layer 2 with no LDS round trip: its K = 32 chunk c takes the tiles mb = 2c and
// 2c + 1, i.e. register j of lane l holds unit k = 32 c + 16 (j / 4) + 4 g + j % 4,
// and W2^T is gathered with the same k.  Layer 3 is a dot product reduced over the
// four lane groups of each sample.  Same arithmetic as the training forward pass
// (bf16 operands, f32 accumulation, bf16-rounded activations); the K order inside the
// layer-2 MFMAs and the layer-3 sum order differ.
constexpr int kInferBlock = 256;

__global__ __launch_bounds__(kInferBlock) void k_infer(Dev d)
{
    const int l = threadIdx.x & 63, g = l >> 4, sl = l & 15;
    // constant operands: A1[mb] = W1^T rows 16 mb .. (input dim padded to 16),
    // A2[c][nb] = W2^T (K chunk c, k as above), per-lane biases / output weights of
    // the units it holds
    v4s a1[4];
    v8bf a2[2][4];
    float b1[4][4], b2[4][4], w3[4][4];
    for (int mb = 0; mb < 4; ++mb) {
        for (int j = 0; j < 4; ++j) {
            const int k = 4 * g + j;  // input feature
            a1[mb][j] = (short)(k < kDIn ? f2bf(d.params[oW1 + k * kDH + 16 * mb + sl]) : 0);
            const int n = 16 * mb + 4 * g + j;
            b1[mb][j] = d.params[oB1 + n];
            b2[mb][j] = d.params[oB2 + n];
            w3[mb][j] = bfr(d.params[oW3 + n]);
        }
    }
    for (int c = 0; c < 2; ++c)
        for (int nb = 0; nb < 4; ++nb)
            for (int j = 0; j < 8; ++j) {
                const int k = 32 * c + 16 * (j >> 2) + 4 * g + (j & 3);
                a2[c][nb][j] = (__bf16)d.params[oW2 + k * kDH + 16 * nb + sl];
            }
    float mean[4], sd[4];
    for (int j = 0; j < 4; ++j) {
        const int k = 4 * g + j;
        mean[j] = k < kDIn ? d.norm[k] : 0.0f;
        sd[j] = k < kDIn ? d.norm[kDIn + k] : 1.0f;
    }
    const float b3 = d.params[oB3];
    const uint64_t tiles = (d.count + 15) / 16;
    const uint64_t waves = (uint64_t)gridDim.x * (kInferBlock / 64);
    for (uint64_t tile = (uint64_t)blockIdx.x * (kInferBlock / 64) + threadIdx.x / 64; tile < tiles;
         tile += waves) {
        const uint64_t row = d.first + tile * 16 + sl;
        const bool valid = row < d.first + d.count;
        v4s xb;  // B1[k][s] = x[s][k]
        for (int j = 0; j < 4; ++j) {
            const int k = 4 * g + j;
            float v = 0.0f;
            if (valid && k < kDIn) v = bfr(__fdiv_rn(__fsub_rn(d.x[row * kDIn + k], mean[j]), sd[j]));
            xb[j] = (short)f2bf(v);
        }
        v8bf h1[2];
        for (int mb = 0; mb < 4; ++mb) {
            const v4f z = mfma(a1[mb], xb, v4f{0.0f, 0.0f, 0.0f, 0.0f});
            for (int j = 0; j < 4; ++j) h1[mb >> 1][4 * (mb & 1) + j] = (__bf16)gelu(z[j] + b1[mb][j]);
        }
        float part = 0.0f;
        for (int nb = 0; nb < 4; ++nb) {
            v4f z = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int c = 0; c < 2; ++c) z = mfma32(a2[c][nb], h1[c], z);
            for (int j = 0; j < 4; ++j) part += bfr(gelu(z[j] + b2[nb][j])) * w3[nb][j];
        }
        part += __shfl_xor(part, 16);
        part += __shfl_xor(part, 32);
        if (g == 0 && valid) d.out[row] = part + b3;
    }
}

__global__ __launch_bounds__(kBlock) void k_train(Dev d)
{
    __shared__ Shared sh;
    __shared__ Wave wv[kWaves];
    load_shared(sh, d);
    __syncthreads();
    const int l = threadIdx.x & 63, wi = threadIdx.x / 64;
    Wave &w = wv[wi];
    v4f gw2[4][4], gw1[4];
    for (int a = 0; a < 4; ++a) {
        gw1[a] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
        for (int b = 0; b < 4; ++b) gw2[a][b] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    }
    float gb1 = 0.0f, gb2 = 0.0f, gw3 = 0.0f, gb3 = 0.0f, sse = 0.0f;
    const uint64_t tiles = (d.count + 15) / 16;
    const uint64_t rounds = (tiles + (uint64_t)gridDim.x * kWaves - 1) / ((uint64_t)gridDim.x * kWaves);
    for (uint64_t r = 0; r < rounds; ++r) {
        const uint64_t tile = (r * gridDim.x + blockIdx.x) * kWaves + wi;
        const uint64_t base = d.first + tile * 16;
        const int valid = tile < tiles ? (int)min<uint64_t>(16, d.first + d.count - base) : 0;
        forward_tile(sh, w, d, base, valid, l, true);
        if (l < 16) {
            const float e = l < valid ? w.y[l] - w.t[l] : 0.0f;
            w.dy[l] = 2.0f * e * d.inv_b;
            sse += e * e;
        }
        __syncthreads();
        // layer 3 gradients; dz2 = dy w3 gelu'(z2)
        for (int s = 0; s < 16; ++s) gw3 += w.dy[s] * w.h2[s][l];
        if (l == 0)
            for (int s = 0; s < 16; ++s) gb3 += w.dy[s];
        for (int s = 0; s < 16; ++s) {
            const float dz = w.dy[s] * sh.w3[l] * w.g2[s][l];
            w.d[s][l] = dz;
            gb2 += dz;
        }
        __syncthreads();
        // gW2 += a1^T bf16(dz2); dh1 = bf16(dz2) W2^T; dz1 = dh1 gelu'(z1) -> g2
        for (int mb = 0; mb < 4; ++mb) {
            const v4s a = opAT(&w.h1[0][0], kDH, 16 * mb, l);
            for (int nb = 0; nb < 4; ++nb) gw2[mb][nb] = mfma(a, opBf(&w.d[0][0], kDH, 16 * nb, l), gw2[mb][nb]);
        }
        for (int nb = 0; nb < 4; ++nb) {
            v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int kc = 0; kc < kDH; kc += 32)
                acc = mfma32(opA8(&w.d[0][0], kDH, kc, l), opBwT8(&sh.w2[0][0], kDH, kc, 16 * nb, l), acc);
            for (int j = 0; j < 4; ++j) {
                const int s = 4 * (l >> 4) + j, n = 16 * nb + (l & 15);
                w.g2[s][n] = acc[j] * w.g1[s][n];
            }
        }
        __syncthreads();
        for (int s = 0; s < 16; ++s) gb1 += w.g2[s][l];
        for (int nb = 0; nb < 4; ++nb) gw1[nb] = mfma(opAT(&w.xs[0][0], 16, 0, l), opBf(&w.g2[0][0], kDH, 16 * nb, l), gw1[nb]);
        __syncthreads();
    }
    // flush this wavefront's partial gradients
    for (int mb = 0; mb < 4; ++mb)
        for (int nb = 0; nb < 4; ++nb)
            for (int j = 0; j < 4; ++j) {
                const int row = 16 * mb + 4 * (l >> 4) + j, col = 16 * nb + (l & 15);
                atomicAdd(d.grad + oW2 + row * kDH + col, gw2[mb][nb][j]);
            }
    for (int nb = 0; nb < 4; ++nb)
        for (int j = 0; j < 4; ++j) {
            const int row = 4 * (l >> 4) + j, col = 16 * nb + (l & 15);
            if (row < kDIn) atomicAdd(d.grad + oW1 + row * kDH + col, gw1[nb][j]);
        }
    atomicAdd(d.grad + oB1 + l, gb1);
    atomicAdd(d.grad + oB2 + l, gb2);
    atomicAdd(d.grad + oW3 + l, gw3);
    if (l == 0) atomicAdd(d.grad + oB3, gb3);
    for (int s = 1; s < 64; s <<= 1) sse += __shfl_xor(sse, s);
    if (l == 0) atomicAdd(d.loss, sse);
}

// Adam (Kingma & Ba 2015) on the f32 master parameters
__global__ void k_adam(float *w, float *m, float *v, const float *g, int n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

__global__ void k_zero(float *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// surrogate inputs of every node: degree, log(1 + degree), log N_1..N_4, with N_t
// the estimate after iteration t
__global__ void k_features(const uint32_t *off, const float *snaps, uint32_t n, float *x)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    const float deg = (float)(off[u + 1] - off[u]);
    float *r = x + (uint64_t)u * kDIn;
    r[0] = deg;
    r[1] = log1pf(deg);
    for (int t = 0; t < 4; ++t) r[2 + t] = logf(snaps[(uint64_t)t * n + u]);
}

// per-graph sums over nodes with a neighbour, per block: part[13 * block + k] with
// k < 6: sum x_k, 6 <= k < 12: sum x_k^2, 12: count (f64; summed over the blocks
// in a fixed order by sum_columns)
constexpr unsigned kSumBlocks = 2048;
__global__ __launch_bounds__(256) void k_feature_sums(const uint32_t *off, const float *x, uint32_t n, double *part)
{
    __shared__ double red[13][4];
    double a[13];
    for (int k = 0; k < 13; ++k) a[k] = 0.0;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        if (off[u + 1] == off[u]) continue;
        const float *r = x + (uint64_t)u * kDIn;
        for (int k = 0; k < kDIn; ++k) {
            const double v = r[k];
            a[k] += v;
            a[6 + k] += v * v;
        }
        a[12] += 1.0;
    }
    for (int k = 0; k < 13; ++k)
        for (int o = 1; o < 64; o <<= 1) a[k] += __shfl_xor(a[k], o);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 13; ++k) red[k][w] = a[k];
    __syncthreads();
    if (threadIdx.x < 13)
        part[13 * blockIdx.x + threadIdx.x] =
            red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
}

// standardise every input by the graph's mean and (population) standard deviation
// over nodes with a neighbour; a constant input maps to 0
__global__ void k_standardize(const double *s, uint32_t n, float *x)
{
    // each column's mean and standard deviation, once per block (the same f64 operations
    // for every block, so every node sees the same values)
    __shared__ double s_mean[kDIn], s_sd[kDIn];
    if (threadIdx.x < kDIn) {
        const int k = (int)threadIdx.x;
        const double cnt = s[12] > 0.0 ? s[12] : 1.0;
        const double mean = s[k] / cnt;
        const double var = s[6 + k] / cnt - mean * mean;
        s_mean[k] = mean;
        s_sd[k] = var > 0.0 ? sqrt(var) : 0.0;
    }
    __syncthreads();
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    float *r = x + (uint64_t)u * kDIn;
    for (int k = 0; k < kDIn; ++k)
        r[k] = s_sd[k] > 0.0 ? (float)(((double)r[k] - s_mean[k]) / s_sd[k]) : 0.0f;
}

// nodes without neighbours have harmonic centrality 0: rank them last
__global__ void k_isolated_last(const uint32_t *off, uint32_t n, float *out)
{
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < n && off[u + 1] == off[u]) out[u] = -3.0e38f;
}

}  // namespace mlp
}  // namespace hb

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {

int mfail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    hb::set_error(buf);
    return code;
}

#define HB_MTRY(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return mfail(e_ == hipErrorOutOfMemory ? HB_ENOMEM : HB_EHIP, "%s failed: %s", #call, \
                         hipGetErrorString(e_));                                               \
    } while (0)

}  // namespace

struct hb_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr, *norm = nullptr, *loss = nullptr;
    float *x = nullptr, *y = nullptr, *out = nullptr;
    uint32_t *idx = nullptr;
    double *stats = nullptr;  // per-graph input sums of the surrogate: 13 totals, then 13 per block
    size_t cap = 0;
    uint64_t step = 0;
    ~hb_mlp()
    {
        void *ptrs[] = {w, m, v, g, norm, loss, x, y, out, idx, stats};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    hipError_t ensure(size_t n)
    {
        if (n <= cap) return hipSuccess;
        for (void **p : {(void **)&x, (void **)&y, (void **)&out, (void **)&idx})
            if (*p) (void)hipFree(*p), *p = nullptr;
        hipError_t e;
        if ((e = hipMalloc(&x, n * 6 * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&y, n * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&out, n * 4)) != hipSuccess) return e;
        if ((e = hipMalloc(&idx, n * 4)) != hipSuccess) return e;
        cap = n;
        return hipSuccess;
    }
};

static unsigned infer_grid(uint64_t count)
{
    const uint64_t tiles = (count + 15) / 16, blocks = (tiles + 3) / 4;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(blocks, 8192));
}

static unsigned mlp_grid(uint64_t count, unsigned cap)
{
    const uint64_t tiles = (count + 15) / 16, blocks = (tiles + hb::mlp::kWaves - 1) / hb::mlp::kWaves;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(blocks, cap));
}

extern "C" {

int hb_mlp_create(hb_mlp **out, int device_ordinal, uint64_t seed)
{
    if (!out) return mfail(HB_EINVAL, "hb_mlp_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    HB_MTRY(hipGetDeviceCount(&ndev));
    if (device_ordinal < 0 || device_ordinal >= ndev) return mfail(HB_EINVAL, "hb_mlp_create: bad device");
    HB_MTRY(hipSetDevice(device_ordinal));
    auto n = std::make_unique<hb_mlp>();
    n->device = device_ordinal;
    HB_MTRY(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
    for (float **p : {&n->w, &n->m, &n->v, &n->g}) HB_MTRY(hipMalloc(p, hb::mlp::kNParams * 4));
    HB_MTRY(hipMalloc(&n->norm, 12 * 4));
    HB_MTRY(hipMalloc(&n->loss, 4));
    HB_MTRY(hipMalloc(&n->stats, 13 * (1 + hb::mlp::kSumBlocks) * sizeof(double)));
    // He-uniform weights for the GELU layers, zero biases; identity normalisation
    std::mt19937_64 rng(seed);
    std::vector<float> w(hb::mlp::kNParams, 0.0f);
    auto fill = [&](int off, int fan_in, int count) {
        const double lim = std::sqrt(6.0 / fan_in);
        std::uniform_real_distribution<double> u(-lim, lim);
        for (int i = 0; i < count; ++i) w[off + i] = (float)u(rng);
    };
    fill(hb::mlp::oW1, 6, 384);
    fill(hb::mlp::oW2, 64, 4096);
    fill(hb::mlp::oW3, 64, 64);
    std::vector<float> norm = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1};
    HB_MTRY(hipMemcpy(n->w, w.data(), w.size() * 4, hipMemcpyHostToDevice));
    HB_MTRY(hipMemset(n->m, 0, hb::mlp::kNParams * 4));
    HB_MTRY(hipMemset(n->v, 0, hb::mlp::kNParams * 4));
    HB_MTRY(hipMemcpy(n->norm, norm.data(), 48, hipMemcpyHostToDevice));
    *out = n.release();
    hb::set_error("");
    return HB_OK;
}

void hb_mlp_destroy(hb_mlp *n)
{
    if (!n) return;
    (void)hipSetDevice(n->device);
    delete n;
}

int hb_mlp_set_params(hb_mlp *n, const float *params, const float *norm)
{
    if (!n) return mfail(HB_EINVAL, "hb_mlp_set_params: NULL model");
    HB_MTRY(hipSetDevice(n->device));
    if (params) HB_MTRY(hipMemcpy(n->w, params, hb::mlp::kNParams * 4, hipMemcpyHostToDevice));
    if (norm) HB_MTRY(hipMemcpy(n->norm, norm, 48, hipMemcpyHostToDevice));
    HB_MTRY(hipMemset(n->m, 0, hb::mlp::kNParams * 4));
    HB_MTRY(hipMemset(n->v, 0, hb::mlp::kNParams * 4));
    n->step = 0;
    return HB_OK;
}

int hb_mlp_get_params(hb_mlp *n, float *params, float *norm)
{
    if (!n) return mfail(HB_EINVAL, "hb_mlp_get_params: NULL model");
    HB_MTRY(hipSetDevice(n->device));
    if (params) HB_MTRY(hipMemcpy(params, n->w, hb::mlp::kNParams * 4, hipMemcpyDeviceToHost));
    if (norm) HB_MTRY(hipMemcpy(norm, n->norm, 48, hipMemcpyDeviceToHost));
    return HB_OK;
}

int hb_mlp_gradient(hb_mlp *n, const float *x, const float *y, uint64_t count, float *grad, float *loss)
{
    if (!n || !x || !y || !grad || count == 0) return mfail(HB_EINVAL, "hb_mlp_gradient: bad arguments");
    HB_MTRY(hipSetDevice(n->device));
    HB_MTRY(n->ensure(count));
    HB_MTRY(hipMemcpy(n->x, x, count * 24, hipMemcpyHostToDevice));
    HB_MTRY(hipMemcpy(n->y, y, count * 4, hipMemcpyHostToDevice));
    HB_MTRY(hipMemsetAsync(n->g, 0, hb::mlp::kNParams * 4, n->stream));
    HB_MTRY(hipMemsetAsync(n->loss, 0, 4, n->stream));
    hb::mlp::Dev d{n->w, n->norm, n->x, n->y, nullptr, 0, count, (float)(1.0 / (double)count), n->g, n->loss, nullptr};
    hb::mlp::k_train<<<mlp_grid(count, 1024), hb::mlp::kBlock, 0, n->stream>>>(d);
    HB_MTRY(hipGetLastError());
    float sse = 0.0f;
    HB_MTRY(hipMemcpyAsync(grad, n->g, hb::mlp::kNParams * 4, hipMemcpyDeviceToHost, n->stream));
    HB_MTRY(hipMemcpyAsync(&sse, n->loss, 4, hipMemcpyDeviceToHost, n->stream));
    HB_MTRY(hipStreamSynchronize(n->stream));
    if (loss) *loss = sse / (float)count;
    return HB_OK;
}

int hb_mlp_train(hb_mlp *n, const float *x, const float *y, uint64_t count, int epochs, int batch, float lr,
                 uint64_t seed, float *epoch_loss)
{
    if (!n || !x || !y || count == 0 || epochs <= 0 || batch <= 0 || count >= (1ull << 32))
        return mfail(HB_EINVAL, "hb_mlp_train: bad arguments");
    HB_MTRY(hipSetDevice(n->device));
    HB_MTRY(n->ensure(count));
    HB_MTRY(hipMemcpy(n->x, x, count * 24, hipMemcpyHostToDevice));
    HB_MTRY(hipMemcpy(n->y, y, count * 4, hipMemcpyHostToDevice));
    // input normalisation from the training set (f64 statistics)
    double mean[6] = {0}, sq[6] = {0};
    for (uint64_t i = 0; i < count; ++i)
        for (int k = 0; k < 6; ++k) {
            mean[k] += x[i * 6 + k];
            sq[k] += (double)x[i * 6 + k] * x[i * 6 + k];
        }
    float norm[12];
    for (int k = 0; k < 6; ++k) {
        const double mu = mean[k] / (double)count, var = std::max(sq[k] / (double)count - mu * mu, 0.0);
        norm[k] = (float)mu;
        norm[6 + k] = (float)std::max(std::sqrt(var), 1e-6);
    }
    HB_MTRY(hipMemcpy(n->norm, norm, 48, hipMemcpyHostToDevice));
    std::vector<uint32_t> perm(count);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        for (uint64_t i = 0; i < count; ++i) perm[i] = (uint32_t)i;
        std::mt19937_64 rng(seed + (uint64_t)ep * 0x9E3779B97F4A7C15ULL);
        for (uint64_t i = count - 1; i > 0; --i) {  // Fisher-Yates
            const uint64_t j = rng() % (i + 1);
            std::swap(perm[i], perm[j]);
        }
        HB_MTRY(hipMemcpyAsync(n->idx, perm.data(), count * 4, hipMemcpyHostToDevice, n->stream));
        HB_MTRY(hipMemsetAsync(n->loss, 0, 4, n->stream));
        for (uint64_t first = 0; first < count; first += (uint64_t)batch) {
            const uint64_t cnt = std::min<uint64_t>(batch, count - first);
            hb::mlp::k_zero<<<(hb::mlp::kNParams + 255) / 256, 256, 0, n->stream>>>(n->g, hb::mlp::kNParams);
            hb::mlp::Dev d{n->w, n->norm, n->x, n->y, n->idx, first, cnt, (float)(1.0 / (double)cnt), n->g, n->loss, nullptr};
            hb::mlp::k_train<<<mlp_grid(cnt, 1024), hb::mlp::kBlock, 0, n->stream>>>(d);
            ++n->step;
            const float bc1 = (float)(1.0 - std::pow(0.9, (double)n->step));
            const float bc2 = (float)(1.0 - std::pow(0.999, (double)n->step));
            hb::mlp::k_adam<<<(hb::mlp::kNParams + 255) / 256, 256, 0, n->stream>>>(n->w, n->m, n->v, n->g,
                                                                                   hb::mlp::kNParams, lr, b1, b2,
                                                                                   eps, bc1, bc2);
        }
        HB_MTRY(hipGetLastError());
        float sse = 0.0f;
        HB_MTRY(hipMemcpyAsync(&sse, n->loss, 4, hipMemcpyDeviceToHost, n->stream));
        HB_MTRY(hipStreamSynchronize(n->stream));
        if (epoch_loss) epoch_loss[ep] = sse / (float)count;
    }
    return HB_OK;
}

int hb_mlp_predict(hb_mlp *n, const float *x, uint64_t count, float *y)
{
    if (!n || !x || !y) return mfail(HB_EINVAL, "hb_mlp_predict: bad arguments");
    if (count == 0) return HB_OK;
    HB_MTRY(hipSetDevice(n->device));
    HB_MTRY(n->ensure(count));
    HB_MTRY(hipMemcpy(n->x, x, count * 24, hipMemcpyHostToDevice));
    hb::mlp::Dev d{n->w, n->norm, n->x, nullptr, nullptr, 0, count, 0.0f, nullptr, nullptr, n->out};
    hb::mlp::k_infer<<<infer_grid(count), hb::mlp::kInferBlock, 0, n->stream>>>(d);
    HB_MTRY(hipGetLastError());
    HB_MTRY(hipMemcpyAsync(y, n->out, count * 4, hipMemcpyDeviceToHost, n->stream));
    HB_MTRY(hipStreamSynchronize(n->stream));
    return HB_OK;
}

}  // extern "C"

// device-resident surrogate on a staged graph (declared in sparkx_hb.h)
namespace hb {
hipError_t launch_surrogate_tail(hb_mlp *n, const uint32_t *off, const float *snaps, uint32_t nodes,
                                 uint32_t infer_nodes, bool standardized_x, float *x_dev, float *out_dev, hipStream_t s)
{
    const unsigned g = (nodes + 255) / 256;
    mlp::k_features<<<g, 256, 0, s>>>(off, snaps, nodes, x_dev);
    return launch_surrogate_infer(n, off, nodes, infer_nodes, standardized_x, x_dev, out_dev, s);
}
hipError_t launch_surrogate_infer(hb_mlp *n, const uint32_t *off, uint32_t nodes, uint32_t infer_nodes,
                                  bool standardized_x, float *x_dev, float *out_dev, hipStream_t s)
{
    const unsigned g = (nodes + 255) / 256;
    mlp::k_feature_sums<<<mlp::kSumBlocks, 256, 0, s>>>(off, x_dev, nodes, n->stats + 13);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = sum_columns(n->stats + 13, mlp::kSumBlocks, 13, n->stats, s);
    if (e != hipSuccess) return e;
    // every row when the caller reads the standardised inputs back, else only the rows
    // the network runs on
    const uint32_t rows = standardized_x ? nodes : infer_nodes;
    if (rows) mlp::k_standardize<<<(rows + 255) / 256, 256, 0, s>>>(n->stats, rows, x_dev);
    if (infer_nodes) {
        mlp::Dev d{n->w, n->norm, x_dev, nullptr, nullptr, 0, infer_nodes, 0.0f, nullptr, nullptr, out_dev};
        mlp::k_infer<<<infer_grid(infer_nodes), mlp::kInferBlock, 0, s>>>(d);
    }
    mlp::k_isolated_last<<<g, 256, 0, s>>>(off, nodes, out_dev);
    return hipGetLastError();
}
int mlp_device(const hb_mlp *n) { return n->device; }
}  // namespace hb

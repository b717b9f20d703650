/*
 * The refinement-priority network on the matrix cores: a 4-layer MLP
 * (KP -> 64 -> 64 -> 64 -> 1, ReLU) with BF16 weights and activations and FP32
 * accumulation.  The three 64-wide layers run as v_mfma_f32_16x16x32_bf16 tiles out of
 * LDS; the 64 -> 1 output layer is a per-row dot product on the VALU (an MFMA tile
 * would be 1/16 used).
 *
 * MFMA 16x16x32 operand layout (wave64): lane l holds A[l%16][8(l/16) .. 8(l/16)+7]
 * and B[8(l/16) .. +7][l%16]; the accumulator's element i is D[4(l/16)+i][l%16].
 * Every operand here is therefore read as 8 consecutive BF16 of a row: activations
 * are stored sample-major, forward weights output-major (W^T) and the copies used to
 * back-propagate input-major (W).
 *
 * Rounding, which the CPU restatement of the network follows:
 *   weights and input features: BF16 (round to nearest even) of the FP32 values
 *   z = sum_k bf(x_k) bf(w_ko) in FP32, + bias (FP32)
 *   h = bf(relu(z)); ReLU derivative taken as [h > 0]
 *   output y = (s0 + s1) + (s2 + s3) + b3, s_q = sequential FP32 sum of 16 products
 *   back-propagated deltas: BF16 of the FP32 products / MFMA sums, masked by [h > 0]
 */
#ifndef VMX_REFINER_H
#define VMX_REFINER_H

#include <hip/hip_runtime.h>

#include "features.h"

namespace vmx {
namespace nn {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8_t __attribute__((ext_vector_type(8)));

constexpr int WS = W + 8;  // LDS row stride of 64-wide activations (16-byte aligned rows)

__device__ inline bf16x8_t ld8(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t*>(p));
}
__device__ inline bf16x8_t gather8(const uint16_t* p, int stride) {
    u16x8_t r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r[i] = p[i * stride];
    return __builtin_bit_cast(bf16x8_t, r);
}

/* One 16x16 tile: sum over K of A[16][K] (rows a, stride lda) times B[K][16] given as
   B^T rows (bt, stride ldb). */
__device__ inline f32x4_t mm_tile(const uint16_t* a, int lda, const uint16_t* bt, int ldb, int K, int lane) {
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
    const int r = lane & 15, kq = (lane >> 4) * 8;
    for (int k0 = 0; k0 < K; k0 += 32)
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ld8(a + r * lda + k0 + kq), ld8(bt + r * ldb + k0 + kq), acc, 0, 0,
                                                      0);
    return acc;
}

/* Tile of A^T B over the sample dim: A[S][m] and B[S][n] both sample-major. */
__device__ inline f32x4_t mm_tile_tn(const uint16_t* a, int lda, const uint16_t* b, int ldb, int S, int lane) {
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
    const int r = lane & 15, kq = (lane >> 4) * 8;
    for (int s0 = 0; s0 < S; s0 += 32)
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(gather8(a + (s0 + kq) * lda + r, lda),
                                                      gather8(b + (s0 + kq) * ldb + r, ldb), acc, 0, 0, 0);
    return acc;
}

/* Forward weights in LDS. */
template <int KP>
struct alignas(16) NetW {  // size a multiple of 16 bytes (copied as 16-byte words)
    uint16_t w0t[W][KP + 8];
    uint16_t w1t[W][WS];
    uint16_t w2t[W][WS];
    float b0[W], b1[W], b2[W], w3[W];
    float b3, mu, sigma, pad;
};
/* NetW<KP> from w1t on, with the layout it has there: the part of the forward weights
   a kernel keeps in LDS when it reads the first layer from device memory. */
struct alignas(16) NetTail {
    uint16_t w1t[W][WS];
    uint16_t w2t[W][WS];
    float b0[W], b1[W], b2[W], w3[W];
    float b3, mu, sigma, pad;
};
template <int KP>
constexpr size_t net_tail_offset() {
    return sizeof(uint16_t) * W * (KP + 8);
}

static_assert(sizeof(NetW<288>) == net_tail_offset<288>() + sizeof(NetTail), "NetTail is the tail of NetW");
static_assert(sizeof(NetW<128>) == net_tail_offset<128>() + sizeof(NetTail), "NetTail is the tail of NetW");

/* Input-major copies for back-propagation. */
struct alignas(16) NetWN {
    uint16_t w1n[W][WS];
    uint16_t w2n[W][WS];
};

template <int KP>
__device__ inline void load_net(NetW<KP>& n, const float* __restrict__ p) {
    using L = Layout<KP>;
    for (int i = threadIdx.x; i < KP * W; i += blockDim.x) n.w0t[i % W][i / W] = f2bf(p[L::w0 + i]);
    for (int i = threadIdx.x; i < W * W; i += blockDim.x) {
        n.w1t[i % W][i / W] = f2bf(p[L::w1 + i]);
        n.w2t[i % W][i / W] = f2bf(p[L::w2 + i]);
    }
    for (int i = threadIdx.x; i < W; i += blockDim.x) {
        n.b0[i] = p[L::b0 + i];
        n.b1[i] = p[L::b1 + i];
        n.b2[i] = p[L::b2 + i];
        n.w3[i] = bf2f(f2bf(p[L::w3 + i]));
    }
    if (threadIdx.x == 0) {
        n.b3 = p[L::b3];
        n.mu = p[L::mu];
        n.sigma = p[L::sigma];
    }
}

/* The LDS image of a network (what load_net builds), written once to device memory so
   kernels that only run the forward pass copy it with 16-byte loads. */
template <int KP>
__global__ void k_pack_net(const float* __restrict__ p, NetW<KP>* out) {
    load_net<KP>(*out, p);
}

template <int KP>
__device__ inline void load_net_n(NetWN& n, const float* __restrict__ p) {
    using L = Layout<KP>;
    for (int i = threadIdx.x; i < W * W; i += blockDim.x) {
        n.w1n[i / W][i % W] = f2bf(p[L::w1 + i]);
        n.w2n[i / W][i % W] = f2bf(p[L::w2 + i]);
    }
}

/* out[R][W] = bf(relu(x[R][K] W + b)); R a multiple of 16; all waves of the block. */
__device__ inline void dense_relu(const uint16_t* wt, int ldw, const float* bias, const uint16_t* x, int ldx, int K,
                                  int R, uint16_t* out) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int units = (R / 16) * (W / 16);
    for (int u = wave; u < units; u += nw) {
        const int rt = u / (W / 16), ct = u % (W / 16);
        const f32x4_t acc = mm_tile(x + rt * 16 * ldx, ldx, wt + ct * 16 * ldw, ldw, K, lane);
        const int col = ct * 16 + (lane & 15);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = rt * 16 + (lane >> 4) * 4 + i;
            const float z = acc[i] + bias[col];
            out[row * WS + col] = f2bf(z > 0.f ? z : 0.f);
        }
    }
}

/* y[r] = h[r] . w3 + b3 for r < R (R a multiple of 16); four lanes per row. */
__device__ inline void dense_out(const float* w3, float b3, const uint16_t* h, int R, float* y) {
    for (int q = threadIdx.x; q < R * 4; q += blockDim.x) {
        const int row = q >> 2, part = q & 3;
        const uint16_t* hr = h + row * WS + part * 16;
        float s = 0.f;
        for (int k = 0; k < 16; ++k) s += bf2f(hr[k]) * w3[part * 16 + k];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        if (part == 0) y[row] = s + b3;
    }
}

/* Forward pass of R rows of BF16 features xs[R][KP+8]: hidden activations into h0, h1,
   h2 ([R][WS] each; h2 may alias h0), raw outputs into y.  Contains __syncthreads. */
template <int KP, typename Net>
__device__ inline void forward_rows(const uint16_t* w0t, const Net& n, const uint16_t* xs, int R, uint16_t* h0,
                                    uint16_t* h1, uint16_t* h2, float* y) {
    dense_relu(w0t, KP + 8, n.b0, xs, KP + 8, KP, R, h0);
    __syncthreads();
    dense_relu(&n.w1t[0][0], WS, n.b1, h0, WS, W, R, h1);
    __syncthreads();
    dense_relu(&n.w2t[0][0], WS, n.b2, h1, WS, W, R, h2);
    __syncthreads();
    dense_out(n.w3, n.b3, h2, R, y);
    __syncthreads();
}
template <int KP>
__device__ inline void forward_block(const NetW<KP>& n, const uint16_t* xs, int R, uint16_t* h0, uint16_t* h1,
                                     uint16_t* h2, float* y) {
    forward_rows<KP>(&n.w0t[0][0], n, xs, R, h0, h1, h2, y);
}
/* The same with the first layer's weights w0t ([W][KP+8], BF16) read from device memory
   (the packed image of NetW<KP>) and the rest from LDS. */
template <int KP>
__device__ inline void forward_block_w0dev(const uint16_t* w0t, const NetTail& n, const uint16_t* xs, int R,
                                           uint16_t* h0, uint16_t* h1, uint16_t* h2, float* y) {
    forward_rows<KP>(w0t, n, xs, R, h0, h1, h2, y);
}

/* ---------------------------------------------------------------- inference
   Raw network outputs for n rows of BF16 features X[n][KP]; 64 rows per tile. */
template <int KP>
__global__ void __launch_bounds__(256) k_mlp_forward(const float* __restrict__ params, const uint16_t* __restrict__ X,
                                                     uint64_t n, float* __restrict__ y) {
    __shared__ NetW<KP> net;
    __shared__ __attribute__((aligned(16))) uint16_t xs[64][KP + 8];
    __shared__ __attribute__((aligned(16))) uint16_t h[2][64][WS];
    __shared__ float yo[64];
    load_net<KP>(net, params);
    for (uint64_t base = uint64_t(blockIdx.x) * 64; base < n; base += uint64_t(gridDim.x) * 64) {
        __syncthreads();
        for (int i = threadIdx.x; i < 64 * KP; i += blockDim.x) {
            const int r = i / KP, k = i % KP;
            xs[r][k] = base + r < n ? X[(base + r) * KP + k] : uint16_t(0);
        }
        __syncthreads();
        forward_block<KP>(net, &xs[0][0], 64, &h[0][0][0], &h[1][0][0], &h[0][0][0], yo);
        if (threadIdx.x < 64 && base + threadIdx.x < n) y[base + threadIdx.x] = yo[threadIdx.x];
    }
}

/* ---------------------------------------------------------------- training
   Gradient of 0.5 * sum_r c (y_r - (t_r - mu) / sigma)^2 over one 64-row tile per block
   (rows idx[blk*64 ..], c = inv_batch), written to part[blk][trained]; the block's loss
   sum goes to loss[blk] (want_grad = 0: the loss only).  Deterministic: no atomics. */
template <int KP>
__global__ void __launch_bounds__(256) k_mlp_grad(const float* __restrict__ params, const uint16_t* __restrict__ X,
                                                  const float* __restrict__ tgt, const uint32_t* __restrict__ idx,
                                                  uint32_t count, float inv_batch, float* __restrict__ part,
                                                  float* __restrict__ loss, int want_grad) {
    using L = Layout<KP>;
    __shared__ NetW<KP> net;
    __shared__ NetWN netn;
    __shared__ __attribute__((aligned(16))) uint16_t xs[64][KP + 8];
    __shared__ __attribute__((aligned(16))) uint16_t h[3][64][WS];
    __shared__ __attribute__((aligned(16))) uint16_t g[2][64][WS];
    __shared__ float yo[64], dy[64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, nw = blockDim.x >> 6;
    const uint32_t base = blockIdx.x * 64u;
    float* out = part + size_t(blockIdx.x) * L::trained;
    load_net<KP>(net, params);
    load_net_n<KP>(netn, params);
    for (int i = tid; i < 64 * KP; i += blockDim.x) {
        const int r = i / KP, k = i % KP;
        xs[r][k] = base + r < count ? X[size_t(idx[base + r]) * KP + k] : uint16_t(0);
    }
    __syncthreads();
    forward_block<KP>(net, &xs[0][0], 64, &h[0][0][0], &h[1][0][0], &h[2][0][0], yo);
    if (tid < 64) {
        float d = 0.f;
        if (base + tid < count) {
            const float tn = (tgt[idx[base + tid]] - net.mu) / net.sigma;
            d = yo[tid] - tn;
        }
        dy[tid] = d * inv_batch;
        yo[tid] = 0.5f * d * d * inv_batch;
    }
    __syncthreads();
    if (!want_grad) {  // loss only (validation)
        if (tid == 0) {
            float ls = 0.f;
            for (int r = 0; r < 64; ++r) ls += yo[r];
            loss[blockIdx.x] = ls;
        }
        return;
    }
    // output layer
    if (tid < W) {
        float s = 0.f;
        for (int r = 0; r < 64; ++r) s += bf2f(h[2][r][tid]) * dy[r];
        out[L::w3 + tid] = s;
    } else if (tid == W) {
        float s = 0.f, ls = 0.f;
        for (int r = 0; r < 64; ++r) {
            s += dy[r];
            ls += yo[r];
        }
        out[L::b3] = s;
        loss[blockIdx.x] = ls;
    }
    for (int i = tid; i < 64 * W; i += blockDim.x) {
        const int r = i / W, k = i % W;
        g[0][r][k] = bf2f(h[2][r][k]) > 0.f ? f2bf(dy[r] * net.w3[k]) : uint16_t(0);
    }
    __syncthreads();
    // hidden layers 2 and 1: dW = h_in^T g, db = sum g, g_in = (g W^T) masked
    for (int layer = 2; layer >= 1; --layer) {
        const uint16_t* hin = &h[layer - 1][0][0];
        const uint16_t* gl = &g[layer == 2 ? 0 : 1][0][0];
        uint16_t* gnext = &g[layer == 2 ? 1 : 0][0][0];
        const uint16_t* wn = layer == 2 ? &netn.w2n[0][0] : &netn.w1n[0][0];
        const int ow = layer == 2 ? L::w2 : L::w1, ob = layer == 2 ? L::b2 : L::b1;
        for (int u = wave; u < 16 + 16; u += nw) {
            if (u < 16) {
                const int mt = u / 4, nt = u % 4;
                const f32x4_t acc = mm_tile_tn(hin + mt * 16, WS, gl + nt * 16, WS, 64, lane);
                const int o = nt * 16 + (lane & 15);
#pragma unroll
                for (int i = 0; i < 4; ++i) out[ow + (mt * 16 + (lane >> 4) * 4 + i) * W + o] = acc[i];
            } else {
                const int rt = (u - 16) / 4, ct = (u - 16) % 4;
                const f32x4_t acc = mm_tile(gl + rt * 16 * WS, WS, wn + ct * 16 * WS, WS, W, lane);
                const int k = ct * 16 + (lane & 15);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int r = rt * 16 + (lane >> 4) * 4 + i;
                    gnext[r * WS + k] = bf2f(hin[r * WS + k]) > 0.f ? f2bf(acc[i]) : uint16_t(0);
                }
            }
        }
        if (tid < W) {
            float s = 0.f;
            for (int r = 0; r < 64; ++r) s += bf2f(gl[r * WS + tid]);
            out[ob + tid] = s;
        }
        __syncthreads();
    }
    // input layer: dW0 = x^T g0 (g0 is in g[0] after the loop)
    for (int u = wave; u < (KP / 16) * 4; u += nw) {
        const int mt = u / 4, nt = u % 4;
        const f32x4_t acc = mm_tile_tn(&xs[0][0] + mt * 16, KP + 8, &g[0][0][0] + nt * 16, WS, 64, lane);
        const int o = nt * 16 + (lane & 15);
#pragma unroll
        for (int i = 0; i < 4; ++i) out[L::w0 + (mt * 16 + (lane >> 4) * 4 + i) * W + o] = acc[i];
    }
    if (tid < W) {
        float s = 0.f;
        for (int r = 0; r < 64; ++r) s += bf2f(g[0][r][tid]);
        out[L::b0 + tid] = s;
    }
}

/* 64-bit FNV-1a over the 32-bit words of row i (kp BF16 features, then the target's
   bits): a content key whose sort gives the pairs one order whatever order the device
   wrote them in. */
static __global__ void k_row_hash(const uint16_t* __restrict__ X, const float* __restrict__ tgt, uint64_t n, int kp,
                           uint64_t* __restrict__ key, uint32_t* __restrict__ idx) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(X + i * uint64_t(kp));
    uint64_t h = 0xcbf29ce484222325ull;
    for (int k = 0; k < kp / 2; ++k) {
        h ^= w[k];
        h *= 0x100000001b3ull;
    }
    h ^= __builtin_bit_cast(uint32_t, tgt[i]);
    h *= 0x100000001b3ull;
    key[i] = h;
    idx[i] = uint32_t(i);
}

/* Sum of the per-block partial gradients in block order. */
static __global__ void k_sum_parts(const float* __restrict__ part, uint32_t nblk, uint32_t P, float* __restrict__ grad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    float s = 0.f;
    for (uint32_t b = 0; b < nblk; ++b) s += part[size_t(b) * P + i];
    grad[i] = s;
}

/* Adam (Kingma & Ba 2015) on the trained parameters; the gradient is the block-ordered
   sum of the partials.  c1, c2 are the bias corrections 1/(1-beta^t). */
static __global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ part, uint32_t nblk, uint32_t P, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    float gsum = 0.f;
    for (uint32_t b = 0; b < nblk; ++b) gsum += part[size_t(b) * P + i];
    const float mi = b1 * m[i] + (1.f - b1) * gsum;
    const float vi = b2 * v[i] + (1.f - b2) * gsum * gsum;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

}  // namespace nn
}  // namespace vmx

#endif

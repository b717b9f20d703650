/*
 * Shared device pieces of the per-bin control-variate path (gfx950, wave64): math
 * policies, histogram keys, per-bin list sort, compensated sums, residual sample
 * draws, plain per-bin Monte Carlo, gathers and device integrand evaluation.
 *
 * Region generation replaces the reference's sequential greedy max-heap
 * (src/nested/regions-generator-adaptive-heap.h:18-45) by exact parallel rounds.  Let
 * b(x) be the minimum heuristic error on the path root..x.  The heap pops nodes in
 * non-increasing b, so its first N-1 pops are the N-1 nodes of largest b (ties aside),
 * and b never grows down the tree.  Each round therefore expands every new node whose
 * b reaches a lower bound of the (N-1)-th largest b among the nodes known so far; a
 * node rejected once is never needed later because that bound only rises.  The exact
 * selection of the N-1 split nodes happens once, after the rounds converge
 * (pipeline.h).
 */
#ifndef VMX_KERNELS_H
#define VMX_KERNELS_H

#include <hip/hip_runtime.h>

#include "philox.h"
#include "region_math.h"

namespace vmx {

/* ---- math policies for integrand evaluation ---- */
/* Each policy also fixes how the integrands round: `fma(a, b, c)` is a*b + c, and
   `cos2pi(p)` is cos(2 pi p). */
struct PreciseMath {  // the host's arithmetic: correctly rounded libm, two roundings per a*b + c
    __device__ static float exp(float x) { return expf(x); }
    __device__ static double exp(double x) { return ::exp(x); }
    __device__ static float sin(float x) { return sinf(x); }
    __device__ static double sin(double x) { return ::sin(x); }
    __device__ static float cos(float x) { return cosf(x); }
    __device__ static double cos(double x) { return ::cos(x); }
    template <typename T>
    __device__ static T fma(T a, T b, T c) { return a * b + c; }
    template <typename T>
    __device__ static T cos2pi(T p) { return cos(T(6.283185307179586476925286766559) * p); }
};
/* FP32 production: the hardware transcendentals (v_exp_f32, v_sin_f32, v_cos_f32; a few
   ulp, arguments here stay within a few hundred radians) and fused multiply-adds.
   FP64 is unchanged. */
struct FastMath {
    __device__ static float exp(float x) { return __expf(x); }
    __device__ static double exp(double x) { return ::exp(x); }
    __device__ static float sin(float x) { return __sinf(x); }
    __device__ static double sin(double x) { return ::sin(x); }
    __device__ static float cos(float x) { return __cosf(x); }
    __device__ static double cos(double x) { return ::cos(x); }
    __device__ static float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
    __device__ static double fma(double a, double b, double c) { return a * b + c; }
    // exact reduction to [-1/2, 1/2] revolutions, then the hardware cosine (revolutions)
    __device__ static float cos2pi(float p) { return __builtin_amdgcn_cosf(p - rintf(p)); }
    __device__ static double cos2pi(double p) { return ::cos(6.283185307179586476925286766559 * p); }
};

/* 16-bit histogram key of a non-negative error: its top 16 bits (order preserving). */
__device__ inline uint32_t key16(float b) { return __float_as_uint(b) >> 16; }
__device__ inline uint32_t key16(double b) { return uint32_t(uint64_t(__double_as_longlong(b)) >> 48); }
__device__ inline void key16_floor(uint32_t k, float& out) { out = __uint_as_float(k << 16); }
__device__ inline void key16_floor(uint32_t k, double& out) { out = __longlong_as_double((long long)(uint64_t(k) << 48)); }

template <typename T>
struct KahanSum {
    T s = 0, c = 0;
    __device__ void add(T x) {
        const T y = x - c;
        const T t2 = s + y;
        c = (t2 - s) - y;
        s = t2;
    }
};

template <typename T>
struct BinArgs {
    T lo[VMX_MAX_DIM], hi[VMX_MAX_DIM];
    T drange[3];
    uint64_t res[3];
    uint64_t nbins;
    uint64_t spp;
    uint32_t seed_lo, seed_hi;
    uint32_t stream;  // integral id in the Philox counter
    int db, mb, form;
    double alpha;
};

/* Uniform draws of one residual sample: region index, then D coordinates. */
template <typename T, int D>
struct SampleDraws {
    static constexpr int NU = 1 + D * (sizeof(T) == 8 ? 2 : 1);
    static constexpr int NC = (NU + 3) / 4;
    uint32_t u[NC * 4];
    __device__ SampleDraws(uint32_t s, uint64_t bin, uint32_t stream, uint32_t k0, uint32_t k1) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const U4 r = philox4x32_10(U4{s, uint32_t(bin), uint32_t(bin >> 32) ^ stream, uint32_t(c)}, k0, k1);
            u[4 * c] = r.x;
            u[4 * c + 1] = r.y;
            u[4 * c + 2] = r.z;
            u[4 * c + 3] = r.w;
        }
    }
    __device__ T coord(int d) const {
        if constexpr (sizeof(T) == 8)
            return u01d(u[1 + 2 * d], u[2 + 2 * d]);
        else
            return u01f(u[1 + d]);
    }
};

/* Uniform draws of a group of 4 consecutive residual samples of a bin: one Philox call
   per word type gives that word to all 4 samples -- the region word, the coordinates
   (one word each in FP32, two in FP64), then XW extra words (the roulette between
   sampling strategies).  The region and coordinate words do not depend on XW, so a job
   draws the same samples in every kernel instantiation.  Every output word is used: at
   D = 2 in FP32 with XW = 0, 3 calls serve 4 samples.  Counter (group, bin,
   stream ^ bin_hi, type). */
template <typename T, int D, int XW = 0>
struct GroupDraws {
    static constexpr int WPC = sizeof(T) == 8 ? 2 : 1;
    static constexpr int NT = 1 + XW + D * WPC;
    uint32_t w[NT][4];
    __device__ GroupDraws(uint32_t group, uint64_t bin, uint32_t stream, uint32_t k0, uint32_t k1) {
#pragma unroll
        for (int c = 0; c < NT; ++c) {
            const U4 r = philox4x32_10(U4{group, uint32_t(bin), uint32_t(bin >> 32) ^ stream, uint32_t(c)}, k0, k1);
            w[c][0] = r.x;
            w[c][1] = r.y;
            w[c][2] = r.z;
            w[c][3] = r.w;
        }
    }
    __device__ uint32_t region_word(int k) const { return w[0][k]; }
    __device__ uint32_t extra_word(int k, int e) const { return w[1 + D * WPC + e][k]; }
    __device__ T coord(int k, int d) const {
        if constexpr (sizeof(T) == 8)
            return u01d(w[1 + 2 * d][k], w[2 + 2 * d][k]);
        else
            return u01f(w[1 + d][k]);
    }
    /* drop sample 0's words: sample k + 1 becomes sample k (for a rolled group loop) */
    __device__ void shift() {
#pragma unroll
        for (int c = 0; c < NT; ++c) {
            w[c][0] = w[c][1];
            w[c][1] = w[c][2];
            w[c][2] = w[c][3];
        }
    }
};

/* Plain per-bin Monte Carlo, the reference's monte_carlo_per_bin
   (src/monte-carlo/monte-carlo-per-bin.h:40-65): `count` uniform samples per bin from
   sample index `first`, accumulated into FP64 sum and sum of squares.  With
   `stratified`, sample s of a bin is jittered inside cell s % (g*g) of a g x g grid over
   the first two bin dims (g*g divides the sample count).  Chunked launches keep every
   kernel short. */
template <typename T, int D, typename F, typename M>
__global__ void k_mc(BinArgs<T> a, const F* __restrict__ f, uint64_t first, uint64_t count, uint32_t grid,
                     double* __restrict__ sum, double* __restrict__ sum2) {
    const uint64_t b = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= a.nbins) return;
    const int db = a.db;
    uint32_t p[3] = {0, 0, 0};
    {
        uint64_t rem = b;
        for (int i = db - 1; i >= 0; --i) {
            p[i] = uint32_t(rem % a.res[i]);
            rem /= a.res[i];
        }
    }
    T blo[D], bhi[D];
    for (int d = 0; d < D; ++d) {
        blo[d] = a.lo[d];
        bhi[d] = a.hi[d];
    }
    for (int i = 0; i < db; ++i) {
        blo[i] = a.lo[i] + T(p[i]) * a.drange[i];
        bhi[i] = a.lo[i] + T(p[i] + 1) * a.drange[i];
    }
    double s1 = 0, s2 = 0;
    for (uint64_t s = first; s < first + count; ++s) {
        const SampleDraws<T, D> u(uint32_t(s), b, a.stream ^ 0x5bd1e995u, a.seed_lo, a.seed_hi);
        T x[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            T c = u.coord(d);
            if (grid > 1 && d < 2 && d < db) {
                const uint32_t cell = uint32_t(s % (uint64_t(grid) * grid));
                const uint32_t ci = d == 0 ? cell % grid : cell / grid;
                c = (T(ci) + c) / T(grid);
            }
            x[d] = blo[d] + c * (bhi[d] - blo[d]);
        }
        const double y = double(f->template eval<M, D>(x));
        s1 += y;
        s2 += y * y;
    }
    sum[b] += s1;
    sum2[b] += s2;
}

/* dst[i*stride + k] = src[order[i]*stride + k] */
template <typename T>
__global__ void k_gather(const T* __restrict__ src, const int32_t* __restrict__ order, uint32_t n, int stride,
                         T* __restrict__ dst) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= uint64_t(n) * stride) return;
    const uint64_t r = i / stride, k = i % stride;
    dst[i] = src[uint64_t(order[r]) * stride + k];
}

/* Integrand at n points (device evaluation check). */
template <typename T, int D, typename F, typename M>
__global__ void k_eval(const F* __restrict__ f, uint64_t n, const double* __restrict__ x, double* __restrict__ out) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T p[D];
    for (int d = 0; d < D; ++d) p[d] = T(x[i * D + d]);
    out[i] = double(f->template eval<M, D>(p));
}

}  // namespace vmx

#endif

/*
 * HIP kernels of the per-bin control-variate path (gfx950, wave64).
 *
 * Region generation replaces the reference's sequential greedy max-heap
 * (src/nested/regions-generator-adaptive-heap.h:18-45) by exact parallel rounds.  Let
 * b(x) be the minimum heuristic error on the path root..x.  The heap pops nodes in
 * non-increasing b, so its first N-1 pops are the N-1 nodes of largest b (ties aside),
 * and b never grows down the tree.  Each round therefore expands every new node whose
 * b reaches a lower bound of the (N-1)-th largest b among the nodes known so far; a
 * node rejected once is never needed later because that bound only rises.  The exact
 * selection of the N-1 split nodes happens once, after the rounds converge.
 *
 * Node arrays are structure-of-arrays in HBM; the 3^D node values of a region sit
 * contiguously (dim 0 fastest) so a region's tile is one coalesced segment.
 */
#ifndef VMX_KERNELS_H
#define VMX_KERNELS_H

#include <hip/hip_runtime.h>

#include "philox.h"
#include "region_math.h"

namespace vmx {

/* ---- math policies for integrand evaluation ---- */
struct PreciseMath {
    __device__ static float exp(float x) { return expf(x); }
    __device__ static double exp(double x) { return ::exp(x); }
    __device__ static float sin(float x) { return sinf(x); }
    __device__ static double sin(double x) { return ::sin(x); }
    __device__ static float cos(float x) { return cosf(x); }
    __device__ static double cos(double x) { return ::cos(x); }
};
/* FP32 production: hardware exp2-based exponential; sin/cos keep full range reduction. */
struct FastMath {
    __device__ static float exp(float x) { return __expf(x); }
    __device__ static double exp(double x) { return ::exp(x); }
    __device__ static float sin(float x) { return sinf(x); }
    __device__ static double sin(double x) { return ::sin(x); }
    __device__ static float cos(float x) { return cosf(x); }
    __device__ static double cos(double x) { return ::cos(x); }
};

enum : int { FLAG_CAPACITY = 1, FLAG_DEPTH = 2, FLAG_EMPTY_BIN = 4 };

/* 16-bit histogram key of a non-negative error: its top 16 bits (order preserving). */
__device__ inline uint32_t key16(float b) { return __float_as_uint(b) >> 16; }
__device__ inline uint32_t key16(double b) { return uint32_t(uint64_t(__double_as_longlong(b)) >> 48); }
__device__ inline void key16_floor(uint32_t k, float& out) { out = __uint_as_float(k << 16); }
__device__ inline void key16_floor(uint32_t k, double& out) { out = __longlong_as_double((long long)(uint64_t(k) << 48)); }

template <typename T>
struct Tree {
    T* lo;          // [cap*D]
    T* hi;          // [cap*D]
    T* v;           // [cap*K]
    T* err;         // [cap]
    T* b;           // [cap] path minimum of err
    T* b1;          // [cap] second entry of the path's suffix-minimum stack (tie-break)
    int32_t* dim;   // [cap]
    int32_t* parent;
    int32_t* child;  // first child, -1 for a leaf
    uint32_t* code;  // [cap*D] dyadic interval code per dim (root 1, children 2c, 2c+1)
    uint32_t* count; // node counter
    uint32_t* hist;  // [65536] counts of key16(b) over all nodes
    int32_t* flags;
    uint32_t cap;
};

/* ---------------------------------------------------------------- root */
template <typename T, int D, typename F, typename M>
__global__ void k_root(Tree<T> t, const F* __restrict__ f, const T* __restrict__ lo, const T* __restrict__ hi,
                       int heur) {
    constexpr int K = pow3(D);
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    T v[K];
    for (int idx = 0; idx < K; ++idx) {
        T x[D];
        for (int d = 0; d < D; ++d) x[d] = node_coord<T>(double(digit3(idx, d)) / 2.0, lo[d], hi[d]);
        v[idx] = f->template eval<M, D>(x);
    }
    const T vol = box_volume<T, D>(lo, hi);
    T e;
    int dm;
    heuristic<T, D>(v, lo, hi, vol, heur, e, dm);
    for (int k = 0; k < K; ++k) t.v[k] = v[k];
    for (int d = 0; d < D; ++d) {
        t.lo[d] = lo[d];
        t.hi[d] = hi[d];
        t.code[d] = 1u;
    }
    t.err[0] = e;
    t.dim[0] = dm;
    t.b[0] = e;
    t.b1[0] = T(INFINITY);
    t.parent[0] = -1;
    t.child[0] = -1;
    *t.count = 1;
    atomicAdd(&t.hist[key16(e)], 1u);
}

/* ---------------------------------------------------------------- round threshold
   theta = smallest value of the 16-bit bucket holding the need-th largest b (a lower
   bound of the exact value), or -inf while fewer than `need` nodes exist.  One block of
   256 threads, each owning 256 consecutive buckets. */
template <typename T>
__global__ void k_theta(Tree<T> t, uint32_t need, T* theta) {
    __shared__ uint32_t part[256];
    const int tid = threadIdx.x;
    uint32_t s = 0;
    for (int i = 0; i < 256; ++i) s += t.hist[tid * 256 + i];
    part[tid] = s;
    __syncthreads();
    if (tid == 0) {
        const uint32_t total = *t.count;
        if (total < need) {
            *theta = T(-INFINITY);
            return;
        }
        uint32_t above = 0;
        int owner = 255;
        for (int i = 255; i >= 0; --i) {
            if (above + part[i] >= need) {
                owner = i;
                break;
            }
            above += part[i];
        }
        uint32_t k = uint32_t(owner) * 256;
        for (int i = 255; i >= 0; --i) {
            above += t.hist[owner * 256 + i];
            if (above >= need) {
                k = uint32_t(owner * 256 + i);
                break;
            }
        }
        T th;
        key16_floor(k, th);
        *theta = th;
    }
}

/* ---------------------------------------------------------------- expansion
   One thread per candidate: if b >= theta, split along its heuristic dim (2*3^(D-1)
   integrand evaluations), score both children and append them as next candidates. */
template <typename T, int D, typename F, typename M>
__global__ void k_expand(Tree<T> t, const F* __restrict__ f, const int32_t* __restrict__ cand_in, uint32_t n_in,
                         int32_t* __restrict__ cand_out, uint32_t* __restrict__ n_out, const T* __restrict__ theta,
                         int heur) {
    constexpr int K = pow3(D);
    constexpr int KM = pow3(D - 1);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_in) return;
    const int32_t node = cand_in[i];
    const T bp = t.b[node];
    if (!(bp >= *theta)) return;
    const uint32_t slot = atomicAdd(t.count, 2u);
    if (slot + 2 > t.cap) {
        atomicOr(t.flags, FLAG_CAPACITY);
        return;
    }
    const int dim = t.dim[node];
    T plo[D], phi[D], lo0[D], hi0[D], lo1[D], hi1[D];
    for (int d = 0; d < D; ++d) {
        plo[d] = t.lo[size_t(node) * D + d];
        phi[d] = t.hi[size_t(node) * D + d];
    }
    split_boxes<T, D>(plo, phi, dim, lo0, hi0, lo1, hi1);
    T pv[K], v0[K], v1[K];
    for (int k = 0; k < K; ++k) pv[k] = t.v[size_t(node) * K + k];
    split_inherit<T, D>(pv, dim, v0, v1);
    for (int c = 0; c < 2; ++c) {
        for (int j = 0; j < KM; ++j) {
            T x[D];
            const int idx = split_new_node<T, D>(plo, phi, dim, c, j, x);
            const T y = f->template eval<M, D>(x);
            if (c == 0)
                v0[idx] = y;
            else
                v1[idx] = y;
        }
    }
    const T b1p = t.b1[node];
    const uint32_t pcode = t.code[size_t(node) * D + dim];
    if (pcode >= 0x80000000u) atomicOr(t.flags, FLAG_DEPTH);
    for (int c = 0; c < 2; ++c) {
        const T* clo = c == 0 ? lo0 : lo1;
        const T* chi = c == 0 ? hi0 : hi1;
        const T* cv = c == 0 ? v0 : v1;
        const T vol = box_volume<T, D>(clo, chi);
        T e;
        int dm;
        heuristic<T, D>(cv, clo, chi, vol, heur, e, dm);
        const size_t s = slot + c;
        for (int d = 0; d < D; ++d) {
            t.lo[s * D + d] = clo[d];
            t.hi[s * D + d] = chi[d];
            t.code[s * D + d] = (d == dim) ? (pcode * 2u + uint32_t(c)) : t.code[size_t(node) * D + d];
        }
        for (int k = 0; k < K; ++k) t.v[s * K + k] = cv[k];
        t.err[s] = e;
        t.dim[s] = dm;
        // suffix-minimum stack of the path, first two entries
        T nb, nb1;
        if (e < bp) {
            nb = e;
            nb1 = T(INFINITY);
        } else if (e < b1p) {
            nb = bp;
            nb1 = e;
        } else {
            nb = bp;
            nb1 = b1p;
        }
        t.b[s] = nb;
        t.b1[s] = nb1;
        t.parent[s] = node;
        t.child[s] = -1;
        atomicAdd(&t.hist[key16(nb)], 1u);
    }
    t.child[node] = int32_t(slot);
    const uint32_t o = atomicAdd(n_out, 2u);
    cand_out[o] = int32_t(slot);
    cand_out[o + 1] = int32_t(slot + 1);
}

/* ---------------------------------------------------------------- final regions */
template <typename T>
struct Regs {
    T* lo;            // [n*D]
    T* hi;            // [n*D]
    T* vol;           // [n]
    double* marg;     // [n*MB] bin-dim marginal (FP64)
    T* coef;          // [n*K] monomial coefficients, normalised coordinates
    uint32_t* box;    // [n*6] pixel box start/end per bin dim
    uint32_t n;
};

template <typename T, int D>
__global__ void k_prepare(Tree<T> t, const int32_t* __restrict__ order, Regs<T> r, int db, int mb,
                          const T* __restrict__ lo, const T* __restrict__ hi, const uint64_t* __restrict__ res) {
    constexpr int K = pow3(D);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.n) return;
    const size_t node = size_t(order[i]);
    T rlo[D], rhi[D], v[K];
    for (int d = 0; d < D; ++d) {
        rlo[d] = t.lo[node * D + d];
        rhi[d] = t.hi[node * D + d];
        r.lo[size_t(i) * D + d] = rlo[d];
        r.hi[size_t(i) * D + d] = rhi[d];
    }
    for (int k = 0; k < K; ++k) v[k] = t.v[node * K + k];
    r.vol[i] = box_volume<T, D>(rlo, rhi);
    double m[27];
    bin_marginal<T, double, D>(v, db, m);
    for (int k = 0; k < mb; ++k) r.marg[size_t(i) * mb + k] = m[k];
    T c[K];
    monomials<T, T, D>(v, c);
    for (int k = 0; k < K; ++k) r.coef[size_t(i) * K + k] = c[k];
    for (int d = 0; d < 3; ++d) {
        uint64_t s = 0, e = 1;
        if (d < db) pixel_range<T>(rlo[d], rhi[d], lo[d], hi[d], res[d], s, e);
        r.box[size_t(i) * 6 + 2 * d] = uint32_t(s);
        r.box[size_t(i) * 6 + 2 * d + 1] = uint32_t(e);
    }
}

/* C-order bin index of a pixel position (bins[i0][i1][i2]). */
__device__ inline uint64_t bin_index(const uint32_t* p, const uint64_t* res, int db) {
    uint64_t k = p[0];
    for (int i = 1; i < db; ++i) k = k * res[i] + p[i];
    return k;
}

/* Per region: count (fill == false) or place (fill == true) its index in every bin of
   its pixel box. */
template <bool Fill>
__global__ void k_stratify(const uint32_t* __restrict__ box, uint32_t n, const uint64_t* __restrict__ res, int db,
                           uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off, uint32_t* __restrict__ lst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* bx = box + size_t(i) * 6;
    uint32_t p[3];
    for (p[2] = bx[4]; p[2] < bx[5]; ++p[2])
        for (p[1] = bx[2]; p[1] < bx[3]; ++p[1])
            for (p[0] = bx[0]; p[0] < bx[1]; ++p[0]) {
                const uint64_t bi = bin_index(p, res, db);
                const uint32_t k = atomicAdd(&cnt[bi], 1u);
                if (Fill) lst[off[bi] + k] = i;
            }
}

/* Exclusive scan of n counts into off[0..n] by one block of 1024 threads. */
__global__ void k_scan(const uint32_t* __restrict__ cnt, uint64_t n, uint32_t* __restrict__ off) {
    __shared__ uint32_t part[1024];
    const int tid = threadIdx.x;
    const uint64_t chunk = (n + 1023) / 1024;
    const uint64_t a = tid * chunk, e = (a + chunk < n) ? a + chunk : n;
    uint32_t s = 0;
    for (uint64_t i = a; i < e; ++i) s += cnt[i];
    part[tid] = s;
    __syncthreads();
    for (int w = 1; w < 1024; w *= 2) {
        const uint32_t add = tid >= w ? part[tid - w] : 0u;
        __syncthreads();
        part[tid] += add;
        __syncthreads();
    }
    uint32_t run = part[tid] - s;
    for (uint64_t i = a; i < e; ++i) {
        off[i] = run;
        run += cnt[i];
    }
    if (tid == 1023) off[n] = part[1023];
}

/* Sort each bin's list ascending (list order = final region order). */
__global__ void k_sort_lists(const uint32_t* __restrict__ off, uint64_t nb, uint32_t* __restrict__ lst) {
    const uint64_t b = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const uint32_t a = off[b], e = off[b + 1];
    for (uint32_t i = a + 1; i < e; ++i) {
        const uint32_t x = lst[i];
        uint32_t j = i;
        while (j > a && lst[j - 1] > x) {
            lst[j] = lst[j - 1];
            --j;
        }
        lst[j] = x;
    }
}

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

/* One thread per bin: control-variate integral over the bin's regions (FP64
   assembly, list order), then `spp` residual samples with uniform region roulette and
   uniform in-cell sampling, weighted per the estimator form.
   References: …regions-integrator-parallel-variance-reduction.h:67-103,
   …regions-integrator-parallel-control-variates.h:39-65, weight-strategy.h:7-110. */
template <typename T, int D, typename F, typename M>
__global__ void k_bins(BinArgs<T> a, Regs<T> r, const uint32_t* __restrict__ off, const uint32_t* __restrict__ lst,
                       const F* __restrict__ f, T* __restrict__ bins, int32_t* flags) {
    constexpr int K = pow3(D);
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
    const uint32_t o = off[b], n = off[b + 1] - o;
    const double factor = double(a.nbins);
    double approx = (a.form == VMX_FORM_CV_ALPHA1) ? double(bins[b]) : 0.0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t ri = lst[o + k];
        double aa[3], bb[3];
        bool empty = false;
        for (int d = 0; d < D; ++d) {
            const T rl = r.lo[size_t(ri) * D + d], rh = r.hi[size_t(ri) * D + d];
            const T il = vmax(blo[d], rl);
            const T ih = vmax(il, vmin(bhi[d], rh));
            if (il >= ih) empty = true;
            if (d < db) {
                aa[d] = double(pos_in(il, rl, rh));
                bb[d] = double(pos_in(ih, rl, rh));
            }
        }
        if (empty) continue;
        const double sub = marginal_subrange<double>(r.marg + size_t(ri) * a.mb, db, aa, bb);
        approx += factor * (double(r.vol[ri]) * sub);
    }
    if (a.spp > 0 && n == 0) {
        atomicOr(flags, FLAG_EMPTY_BIN);
        bins[b] = T(approx);
        return;
    }
    if (a.form == VMX_FORM_CV_ALPHA1) {
        double acc = approx;
        for (uint64_t s = 0; s < a.spp; ++s) {
            const SampleDraws<T, D> u(uint32_t(s), b, a.stream, a.seed_lo, a.seed_hi);
            const uint32_t ri = uniform_index(u.u[0], n);
            const uint32_t reg = lst[o + ri];
            T x[D], t[D], ilo[D], ihi[D];
            bool empty = false;
            for (int d = 0; d < D; ++d) {
                const T rl = r.lo[size_t(reg) * D + d], rh = r.hi[size_t(reg) * D + d];
                ilo[d] = vmax(blo[d], rl);
                ihi[d] = vmax(ilo[d], vmin(bhi[d], rh));
                if (ilo[d] >= ihi[d]) empty = true;
                x[d] = ilo[d] + u.coord(d) * (ihi[d] - ilo[d]);
                t[d] = pos_in(x[d], rl, rh);
            }
            if (empty) continue;
            const T fx = f->template eval<M, D>(x);
            const T px = eval_monomials<T, D>(r.coef + size_t(reg) * K, t);
            acc += double(fx - px) * factor * double(box_volume<T, D>(ilo, ihi)) * double(n) / double(a.spp);
        }
        bins[b] = T(acc);
        return;
    }
    // variance-reduction forms
    KahanSum<T> sum_f, sum_app, e_f, e_ap, e_ap2, e_fap;
    T k_f = 0, k_app = 0;
    for (uint64_t s = 0; s < a.spp; ++s) {
        const SampleDraws<T, D> u(uint32_t(s), b, a.stream, a.seed_lo, a.seed_hi);
        const uint32_t ri = uniform_index(u.u[0], n);
        const uint32_t reg = lst[o + ri];
        T x[D], t[D], ilo[D], ihi[D];
        for (int d = 0; d < D; ++d) {
            const T rl = r.lo[size_t(reg) * D + d], rh = r.hi[size_t(reg) * D + d];
            ilo[d] = vmax(blo[d], rl);
            ihi[d] = vmax(ilo[d], vmin(bhi[d], rh));
            x[d] = ilo[d] + u.coord(d) * (ihi[d] - ilo[d]);
            t[d] = pos_in(x[d], rl, rh);
        }
        const double w = factor * double(n) * double(box_volume<T, D>(ilo, ihi));
        const T fs = T(double(f->template eval<M, D>(x)) * w);
        const T as = T(double(eval_monomials<T, D>(r.coef + size_t(reg) * K, t)) * w);
        if (a.form == VMX_FORM_VR_FIXED) {
            sum_f.add(fs);
            sum_app.add(as);
            continue;
        }
        if (s == 0) {
            k_f = vabs(fs);
            k_app = vabs(as);
        }
        const T df = vabs(fs) - k_f, da = vabs(as) - k_app;
        e_f.add(df);
        e_ap.add(da);
        e_ap2.add(da * da);
        e_fap.add(df * da);
        sum_f.add(fs);
        sum_app.add(as);
    }
    double result;
    const double ns = double(a.spp);
    if (a.form == VMX_FORM_VR_FIXED) {
        result = (a.spp == 0) ? approx
                              : (double(sum_f.s) - a.alpha * double(sum_app.s)) / ns + a.alpha * approx;
    } else if (a.spp < 2) {
        result = approx;
    } else {
        const double ef = e_f.s, ea = e_ap.s, ea2 = e_ap2.s, efa = e_fap.s;
        const double cov = (efa - (ef * ea) / ns) / (ns - 1);
        const double var = (ea2 - (ea * ea) / ns) / (ns - 1);
        double alpha = 1.0;
        const double v = var > 0.0 ? var : 0.0;
        if (v > 0.0) {
            const double c = cov > 0.0 ? (cov < v ? cov : v) : 0.0;
            alpha = c / v;
        }
        result = (double(sum_f.s) - alpha * double(sum_app.s)) / ns + alpha * approx;
    }
    bins[b] = T(result);
}

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

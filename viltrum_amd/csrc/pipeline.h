/*
 * Device-resident, batch-capable pipeline (J integrals of one (scalar, D, family) per
 * launch sequence).  Every array is indexed by an absolute slot; job j owns
 *   node slots    [node_base_j, node_base_j + node_cap_j)
 *   regions       [reg_base_j,  reg_base_j + N_j)
 *   bins          [bin_base_j,  bin_base_j + nbins_j)     (lists, offsets)
 *
 * 1. k_roots, then refinement rounds of every job: per round k_theta_round (boundary
 *    bucket of the N-1 largest path-minimum errors, from per-job histograms) and
 *    k_expand_round (teams of lanes split the new candidates that reach it).
 * 2. Exact selection of the N-1 split nodes: everything above the boundary bucket is
 *    split; inside it one workgroup per job radix-selects on the full key
 *    (b, b1, codes), which is unique per node, so the result is deterministic.
 * 3. Leaves (unsplit children of split nodes) in canonical order, then footprints:
 *    leaves that share their bin-dim extent share their pixel box, and their FP64
 *    bin-dim marginals add up (the control-variate integral is linear in the nodes).
 * 4. Per-bin footprint lists (pairs sorted by bin) and the per-bin control variate +
 *    residual kernel.
 */
#ifndef VMX_PIPELINE_H
#define VMX_PIPELINE_H

#include <type_traits>

#include "fubini.h"
#include "importance.h"
#include "kernels.h"
#include "refiner.h"

namespace vmx {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr int MAX_JOBS = 4096;
constexpr int GEN_THREADS = 256;

enum : uint32_t { JF_CAPACITY = 1, JF_DEPTH = 2, JF_ROUNDS = 4, JF_EMPTY_BIN = 16, JF_LEAVES = 32 };

template <typename T>
struct Job {
    T lo[VMX_MAX_DIM], hi[VMX_MAX_DIM], drange[3];
    uint64_t res[3];
    uint64_t nbins, bin_base, spp;
    void* bins;  // T*, device, C order
    double alpha;
    uint32_t seed_lo, seed_hi, stream;
    uint32_t N, node_base, node_cap, reg_base;
    uint32_t tile_base;  // first tile (of 256 bins) of this job in the per-bin kernel's grid
    uint32_t tile_base2;  // first tile of this job in the region-major per-bin kernel's grid
    int32_t db, mb, form, heur;
    int32_t learn;   // refinement priority from the network (refiner.h)
    int32_t lfloor;  // with learn: every dim is first halved lfloor times, widest first
    float lres[VMX_MAX_DIM];  // log2 of the bin count per dim (0 for integrated dims)
    uint32_t defer_m;  // deferred rounds: at most about this many expansions per round
};

struct JobRt {
    uint32_t count, flags, kbound, above, nbucket, nleaves, ties;
    uint32_t row_floor;  // during the rounds: coarse row of the last boundary bucket
    uint32_t row_ceil;   // no key of the job lies in a coarse row (top 8 key bits) above
                         // it, +infinity keys (row INF_ROW) aside
};

template <typename T>
struct TreeArrays {
    T *lo, *hi, *v, *err, *b, *b1;
    int32_t *dim, *parent, *child;
    uint32_t *code, *job;
    uint8_t* popped;
};

/* The coarse row of a +infinity key (key16 of +inf, FP32 and FP64 alike).  It is always
   read, so +infinity keys (a learned job's root) never raise rt.row_ceil. */
constexpr uint32_t INF_ROW = 0x7Fu;

/* Counts one priority key in job j's 16-bit histogram; raise also lifts rt.row_ceil to
   its coarse row.  A node's b is the minimum of its path, so its key is at most its
   parent's: only a root, a child of a +infinity parent, or a key whose row exceeds its
   parent's (a negative or NaN priority) has to raise the ceiling.  Every other node
   lies at or below its parent's row, which the ceiling already covers. */
__device__ inline void hist_add(uint32_t* hist16, JobRt& rt, uint32_t j, uint32_t k, bool raise) {
    atomicAdd(&hist16[size_t(j) * 65536 + k], 1u);
    if (raise) atomicMax(&rt.row_ceil, k >> 8);
}

/* ------------------------------------------------------------------ block helpers */
/* In-place inclusive suffix sum of s[0..255] (s[t] = sum_{i>=t}); threads 0..255 of a
   block of 256 or more, all of which call it.  Wave-level shuffles within each of the
   four 64-entry quarters, then one exchange of the quarter totals. */
__device__ inline void suffix_sum256(uint32_t* s) {
    __shared__ uint32_t quarter[4];
    const int t = threadIdx.x;
    const int lane = t & 63;
    uint32_t v = t < 256 ? s[t] : 0u;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_down(v, off, 64);
        if (lane + off < 64) v += u;
    }
    if (t < 256 && lane == 0) quarter[t >> 6] = v;
    __syncthreads();
    if (t < 256) {
        uint32_t add = 0;
        for (int q = (t >> 6) + 1; q < 4; ++q) add += quarter[q];
        s[t] = v + add;
    }
    __syncthreads();
}

/* pre[0] = 0, pre[j + 1] = pre[j] + count(j) for j < J, in LDS; every thread of the
   block calls it (blocks of 64..1024 threads).  Each thread sums a run of consecutive
   jobs, a shuffle scan and one exchange of the wave totals give the runs' offsets. */
template <typename Count>
__device__ inline void block_prefix(uint32_t* pre, uint32_t J, Count count) {
    __shared__ uint32_t wave_tot[16];
    const uint32_t per = (J + blockDim.x - 1) / blockDim.x;
    const uint32_t j0 = threadIdx.x * per;
    uint32_t local = 0;
    for (uint32_t k = 0; k < per && j0 + k < J; ++k) {
        const uint32_t c = count(j0 + k);
        pre[j0 + k + 1] = c;
        local += c;
    }
    const int lane = int(threadIdx.x & 63), wave = int(threadIdx.x >> 6);
    uint32_t incl = local;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wave_tot[wave] = incl;
    if (threadIdx.x == 0) pre[0] = 0;
    __syncthreads();
    uint32_t run = incl - local;
    for (int w = 0; w < wave; ++w) run += wave_tot[w];
    for (uint32_t k = 0; k < per && j0 + k < J; ++k) {
        run += pre[j0 + k + 1];
        pre[j0 + k + 1] = run;
    }
    __syncthreads();
}

/* Bucket (16-bit key of b) holding the need-th largest b of one job, and the number of
   nodes in strictly higher buckets.  Requires total >= need >= 1.  NT threads (a
   multiple of 256), all of which call it.  The coarse counts (top 8 bits of the key)
   are the row sums of the fine histogram, formed here (NT/256 threads per row) rather
   than kept by atomics: nodes of a round cluster in few coarse buckets. */
template <int NT>
__device__ inline void boundary_bucket(const uint32_t* h16, uint32_t need, uint32_t* sh, uint32_t& k_out,
                                       uint32_t& above_out, uint32_t row_floor = 0, uint32_t row_ceil = 255) {
    static_assert(NT >= 256 && NT % 256 == 0, "whole rows per thread group");
    constexpr int PER = NT / 256;
    const int t = threadIdx.x;
    __shared__ uint32_t part[NT];
    __shared__ uint32_t sel_c, sel_above;
    {
        // rows below row_floor cannot hold the boundary (the caller knows it lies at or
        // above that row), rows above row_ceil hold no key but +infinity ones (INF_ROW):
        // they count as empty and are not read
        const int row = t / PER, q = t % PER;
        uint32_t acc = 0;
        if (uint32_t(row) >= row_floor && (uint32_t(row) <= row_ceil || uint32_t(row) == INF_ROW)) {
            const uint4* p = reinterpret_cast<const uint4*>(h16 + size_t(row) * 256) + q * (64 / PER);
            for (int i = 0; i < 64 / PER; ++i) {
                const uint4 v = p[i];
                acc += v.x + v.y + v.z + v.w;
            }
        }
        part[t] = acc;
    }
    __syncthreads();
    if (t < 256) {
        uint32_t a = 0;
        for (int i = 0; i < PER; ++i) a += part[t * PER + i];
        sh[t] = a;
    }
    __syncthreads();
    suffix_sum256(sh);
    if (t < 256 && sh[t] >= need && (t == 255 || sh[t + 1] < need)) {
        sel_c = uint32_t(t);
        sel_above = (t == 255) ? 0u : sh[t + 1];
    }
    __syncthreads();
    const uint32_t c = sel_c, above = sel_above;
    __syncthreads();
    if (t < 256) sh[t] = h16[c * 256 + t];
    __syncthreads();
    suffix_sum256(sh);
    const uint32_t need2 = need - above;
    if (t < 256 && sh[t] >= need2 && (t == 255 || sh[t + 1] < need2)) {
        sel_c = c * 256 + uint32_t(t);
        sel_above = above + ((t == 255) ? 0u : sh[t + 1]);
    }
    __syncthreads();
    k_out = sel_c;
    above_out = sel_above;
    __syncthreads();
}

/* ------------------------------------------------------------------ teams
   A team of TEAM lanes (a sub-wave or a whole wave) owns one region split: lanes share
   the new-plane evaluations, the per-line nested errors and the Simpson folds through
   LDS.  Each fold output is computed exactly as the sequential fold computes it, so
   the results are bit-identical. */
template <int D>
struct TeamShape {
    // D = 6: one team per block (its 3^6-value tensors take 32 KB of LDS in FP32)
    static constexpr int TEAM = D <= 2 ? 8 : (D == 3 ? 32 : (D <= 5 ? 64 : 256));
};

template <typename T, int D>
struct TeamLds {
    static constexpr int K = pow3(D), KM = pow3(D - 1);
    T pv[K];
    T cv[2][K];
    T e[2][2 * D * KM];  // double buffer of line errors / folds
    T box[2][2][D];      // child c: lo, hi
    T pbox[2][D];
    int32_t node, slot, dim, job, go;
    int32_t deep[2];  // children that qualify in the same round (slot, or -1)
};

/* Line errors and folds for `nc` regions held in L.cv[c] with boxes L.box[c]; lanes
   `lane` of `team`; returns (err, dim) of region c in lane c (c < nc).  Every lane of
   the block must call it (it contains __syncthreads). */
template <typename T, int D, int TEAM>
__device__ inline void team_heuristic(TeamLds<T, D>& L, int nc, int heur, int lane, bool active, T& err_out,
                                      int& dim_out) {
    constexpr int KM = pow3(D - 1);
    // line (c, d, jl) of region c along dim d lands at (c * D + d) * KM + jl; the loops
    // over c and d are unrolled so every 3^d below is a compile-time constant
    if (active) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (c >= nc) break;
            const T* v = L.cv[c];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int below = pow3(d);
                for (int jl = lane; jl < KM; jl += TEAM) {
                    const int full = (jl % below) + (jl / below) * below * 3;
                    const T p0 = v[full], p1 = v[full + below], p2 = v[full + 2 * below];
                    const T sm = simpson(p0, p1, p2), tz = trapezoid(p0, p2);
                    L.e[0][(c * D + d) * KM + jl] = heur == 0 ? metric_relative(sm, tz) : metric_absolute(sm, tz);
                }
            }
        }
    }
    __syncthreads();
    int buf = 0;
#pragma unroll
    for (int lev = 0; lev < D - 1; ++lev) {
        const int n = KM / pow3(lev + 1);  // a constant in each unrolled level
        if (active)
            for (int q = lane; q < nc * D * n; q += TEAM) {
                const int g = q / n, k = q % n;
                const T* in = L.e[buf] + g * KM;
                L.e[buf ^ 1][g * KM + k] = simpson(in[3 * k], in[3 * k + 1], in[3 * k + 2]);
            }
        buf ^= 1;
        __syncthreads();
    }
    if (active && lane < nc) {
        const int c = lane;
        const T* lo = L.box[c][0];
        const T* hi = L.box[c][1];
        const T vol = box_volume<T, D>(lo, hi);
        const T* E = L.e[buf] + c * D * KM;  // E[d * KM] = folded error of dim d
        if (heur == 0) {
            const double size_weight = 1.e-5, min_size = 1.e-37;
            T max_err = vol * E[0];
            const T ext0 = hi[0] - lo[0];
            if (ext0 < min_size || visnan(ext0))
                max_err = 0;
            else
                max_err += size_weight * vabs(ext0);
            int max_dim = 0;
            for (int d = 1; d < D; ++d) {
                const T ext = hi[d] - lo[d];
                T err = vol * E[d * KM] + size_weight * vabs(ext);
                if (ext < min_size) err = 0;
                if (err >= max_err) {
                    max_err = err;
                    max_dim = d;
                }
            }
            err_out = max_err;
            dim_out = max_dim;
        } else {
            T max_err = 0;
            int max_dim = 0;
            for (int d = 0; d < D; ++d) {
                const T err = vol * E[d * KM];
                if (err > max_err) {
                    max_err = err;
                    max_dim = d;
                }
            }
            err_out = max_err;
            dim_out = max_dim;
        }
    }
}

/* ------------------------------------------------------------------ roots
   One block per job: 256 lanes evaluate the 3^D nodes, one team scores the root. */
template <typename T, int D, typename F, typename M>
__global__ void __launch_bounds__(256) k_roots(const Job<T>* __restrict__ jobs, JobRt* rt, const F* __restrict__ fs,
                                               TreeArrays<T> t, uint32_t* hist16, uint32_t* cstart0,
                                               uint32_t* ncand0) {
    constexpr int K = pow3(D);
    constexpr int TEAM = TeamShape<D>::TEAM;
    __shared__ TeamLds<T, D> L;
    const uint32_t j = blockIdx.x;
    const Job<T>& jb = jobs[j];
    const F& f = fs[j];
    const size_t s = jb.node_base;
    const int tid = threadIdx.x;
    // the job's histogram starts empty and its node slots unused (job NONE)
    {
        uint4* h = reinterpret_cast<uint4*>(hist16 + size_t(j) * 65536);
        for (int i = tid; i < 65536 / 4; i += blockDim.x) h[i] = make_uint4(0u, 0u, 0u, 0u);
        for (uint32_t i = tid; i < jb.node_cap; i += blockDim.x) t.job[s + i] = NONE;
    }
    if (tid < D) {
        L.box[0][0][tid] = jb.lo[tid];
        L.box[0][1][tid] = jb.hi[tid];
    }
    __syncthreads();
    for (int idx = tid; idx < K; idx += blockDim.x) {
        T x[D];
        for (int d = 0; d < D; ++d) x[d] = node_coord<T>(double(digit3(idx, d)) / 2.0, L.box[0][0][d], L.box[0][1][d]);
        const T y = f.template eval<M, D>(x);
        L.cv[0][idx] = y;
        t.v[s * K + idx] = y;
    }
    __syncthreads();
    T e = 0;
    int dm = 0;
    team_heuristic<T, D, TEAM>(L, 1, jb.heur, tid, tid < TEAM, e, dm);
    if (tid == 0) {
        for (int d = 0; d < D; ++d) {
            t.lo[s * D + d] = L.box[0][0][d];
            t.hi[s * D + d] = L.box[0][1][d];
            t.code[s * D + d] = 1u;
        }
        // with a learned priority the root goes first whatever its score (it is the
        // only region), so its children are ranked by their own priorities
        const T pe = jb.learn ? T(INFINITY) : e;
        t.err[s] = pe;
        t.dim[s] = jb.learn && jb.lfloor > 0 ? 0 : dm;
        t.b[s] = pe;
        t.b1[s] = T(INFINITY);
        t.parent[s] = -1;
        t.child[s] = -1;
        t.job[s] = j;
        rt[j].count = jb.node_base + 1;
        hist_add(hist16, rt[j], j, key16(pe), !(pe == T(INFINITY)));
        cstart0[j] = uint32_t(s);
        ncand0[j] = jb.N > 1 ? 1u : 0u;
    }
}

/* A deferred candidate: its slot and b (a deferred node is never split while it waits,
   so the filter reads neither from the tree). */
template <typename T>
struct DefEnt {
    uint32_t slot;
    T b;
};

template <typename T>
struct GenArgs {
    const Job<T>* jobs;
    JobRt* rt;
    TreeArrays<T> t;
    uint32_t* hist16;
    uint32_t* cstart;  // [2*J] first slot of the round's candidates (they are contiguous)
    uint32_t* ncand;   // [2*J] number of candidates
    T* theta;         // [J]
    uint32_t* rounds; // rounds that had candidates
    uint32_t* depth;  // [D] most splits along each dim over all nodes (block maxima)
    uint32_t J;
    uint32_t max_rounds;
    const void* net;   // the priority network's LDS image (nn::NetW<KP>, packed once per network)
    int32_t deepen;    // 1: a split's children with b >= theta are split in the same round
    // deferred rounds (defer = 1 or 2): a round expands its candidates with b >=
    // max(theta, tau), tau about the defer_m-th largest candidate b; those in [theta,
    // tau) wait for a later round in a list, those below theta can never be split.  The
    // candidates are sorted by k_filter_round (1) or by the job's theta block (2)
    int32_t defer;
    uint32_t cap_total;
    T* theta_lo;       // [J] theta itself (theta then holds max(theta, tau))
    uint32_t* ndef;    // [2*J] deferred candidates
    DefEnt<T>* def;    // [2][cap_total] their slots and b, job j's from node_base_j
    uint32_t* nexp;    // [J] the round's expansions
    uint32_t* exp;     // [cap_total] their slots, job j's from node_base_j
    uint32_t* chist;   // [2][J][CHIST] the next round's candidates by the top 12 bits of key16(b)
};
constexpr int CHIST = 4096;

/* Per-team slot requests of one expansion step, combined into one atomic per block and
   job (teams of a step hold consecutive candidates, so a job's teams are one run). */
template <int TPB>
struct SlotLds {
    uint32_t job[TPB];
    int32_t slot[TPB];
    uint8_t want[TPB];
    uint32_t depth[VMX_MAX_DIM];  // the block's deepest new code per dim
};

/* LDS of the fused priority network: the block's 2*TPB new regions form the rows of one
   forward pass per expansion step. */
template <int D>
struct LearnLds {
    static constexpr int KP = nn::Dims<D>::KP;
    static constexpr int TPB = GEN_THREADS / TeamShape<D>::TEAM;
    static constexpr int R = (2 * TPB + 15) / 16 * 16;
    // D = 2: the first layer's weights (5 KB) live in LDS, off the L2 round trip of
    // every round's forward pass; the room comes from the features, which are dead once
    // layer 0 has run and share their space with layer 1's output (same block size).
    // D >= 3: the first layer (9 KB .. 37 KB) would cost a block per CU and is read
    // from device memory (L2)
    static constexpr bool W0_DEV = D != 2;
    static constexpr bool ALIAS = D == 2;
    typename std::conditional<W0_DEV, nn::NetTail, nn::NetW<KP>>::type net;
    union alignas(16) {
        uint16_t xs[R][KP + 8];
        uint16_t h1a[ALIAS ? R : 1][nn::WS];
    };
    alignas(16) uint16_t h[ALIAS ? 1 : 2][R][nn::WS];
    float y[R];
    __device__ uint16_t* h0() { return &h[0][0][0]; }
    __device__ uint16_t* h1() { return ALIAS ? &h1a[0][0] : &h[ALIAS ? 0 : 1][0][0]; }
};

/* nn::region_features of the two children of one expansion (boxes L.box, nodes L.cv,
   the parent's codes pcode, split along dim), computed by the TEAM lanes of a team
   (lane 0..TEAM-1) and written as BF16 to rows[0..1]; jb == nullptr writes zero rows.
   The values are those of nn::region_features bit for bit: s is a max, and each fold
   adds its line terms in the same order.  Lane c < 2 gets child c's s and total
   depth.  Every lane of the block calls it (it contains __syncthreads). */
template <typename T, int D, int TEAM>
__device__ inline void team_child_features(const TeamLds<T, D>& L, const Job<T>* jb, const uint32_t* pcode, int dim,
                                           int lane, uint16_t (*rows)[nn::Dims<D>::KP + 8], float* terms,
                                           float& s_lane, int& depth_lane) {
    constexpr int K = pow3(D), KM = pow3(D - 1), KP = nn::Dims<D>::KP;
    __syncthreads();  // terms overwrite the heuristic's line errors
    // s of both children: lane-strided maxima, then a butterfly inside the team's lanes
    float m0 = 0.f, m1 = 0.f;
    if (jb)
        for (int k = lane; k < K; k += TEAM) {
            m0 = vmax(m0, vabs(float(L.cv[0][k])));
            m1 = vmax(m1, vabs(float(L.cv[1][k])));
        }
#pragma unroll
    for (int o = TEAM / 2; o > 0; o >>= 1) {
        m0 = vmax(m0, __shfl_xor(m0, o));
        m1 = vmax(m1, __shfl_xor(m1, o));
    }
    const float inv0 = 1.f / vmax(m0, 0x1p-100f), inv1 = 1.f / vmax(m1, 0x1p-100f);
    for (int q = lane; q < 2 * K; q += TEAM) {
        const int c = q / K, k = q % K;
        rows[c][3 * D + k] = jb ? nn::f2bf(float(L.cv[c][k]) * (c ? inv1 : inv0)) : uint16_t(0);
    }
    if (jb) {
        // term (c, d, jl) at (c * D + d) * KM + jl; c and d unrolled (constant 3^d)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int below = pow3(d);
                for (int jl = lane; jl < KM; jl += TEAM) {
                    const int full = (jl % below) + (jl / below) * below * 3;
                    const float p0 = float(L.cv[c][full]), p1 = float(L.cv[c][full + below]),
                                p2 = float(L.cv[c][full + 2 * below]);
                    const float diff = vabs(simpson(p0, p1, p2) - trapezoid(p0, p2));
                    float w = 1.f;
                    int rr = jl;
                    for (int e = 0; e < D - 1; ++e) {
                        const int dg = rr % 3;
                        rr /= 3;
                        w *= (dg == 1) ? (4.f / 6.f) : (1.f / 6.f);
                    }
                    terms[(c * D + d) * KM + jl] = w * diff;
                }
            }
    }
    __syncthreads();
    if (lane < 2 * D) {
        const int c = lane / D, d = lane % D;
        float acc = 0.f;
        if (jb)
            for (int jl = 0; jl < KM; ++jl) acc += terms[(c * D + d) * KM + jl];
        rows[c][3 * D + K + d] = jb ? nn::f2bf(acc * (c ? inv1 : inv0)) : uint16_t(0);
    }
    for (int q = lane; q < 2 * (KP - 4 * D - K - 1); q += TEAM) {
        const int c = q / (KP - 4 * D - K - 1), i = 4 * D + K + 1 + q % (KP - 4 * D - K - 1);
        rows[c][i] = 0;
    }
    s_lane = 0.f;
    depth_lane = 0;
    if (lane < 2) {
        const int c = lane;
        uint16_t* row = rows[c];
        if (jb) {
            int depth = 0;
            for (int d = 0; d < D; ++d) {
                const T ctr = T((T(0.5) * (L.box[c][0][d] + L.box[c][1][d]) - jb->lo[d]) / (jb->hi[d] - jb->lo[d]));
                row[d] = nn::f2bf(float(ctr));
                const uint32_t code = d == dim ? pcode[d] * 2u + uint32_t(c) : pcode[d];
                const int n = nn::code_depth(code);
                depth += n;
                row[D + d] = nn::f2bf(-float(n) * (1.f / 16.f));
                row[2 * D + d] = nn::f2bf((jb->lres[d] - float(n)) * (1.f / 16.f));
            }
            row[4 * D + K] = nn::f2bf(float(depth) * (1.f / 32.f));
            s_lane = c ? m1 : m0;
            depth_lane = depth;
        } else {
            for (int i = 0; i < 3 * D; ++i) row[i] = 0;
            row[4 * D + K] = 0;
        }
    }
}

/* Expansion of one candidate per team; all lanes of the block call it the same number
   of times (block-uniform loop), teams without a candidate idle through the syncs. */
template <typename T, int D, typename F, typename M, bool LEARN>
__device__ inline void team_expand(const GenArgs<T>& g, const F* __restrict__ fs, TeamLds<T, D>& L, int par,
                                   const uint32_t* pre, uint64_t w, bool has, int32_t direct, bool may_deepen,
                                   LearnLds<D>* ll, int team, SlotLds<GEN_THREADS / TeamShape<D>::TEAM>& sl) {
    constexpr int K = pow3(D), KM = pow3(D - 1);
    constexpr int TEAM = TeamShape<D>::TEAM;
    constexpr int TPB = GEN_THREADS / TEAM;
    const int lane = threadIdx.x % TEAM;
    const uint32_t J = g.J;
    TreeArrays<T> t = g.t;
    if (lane == 0) {
        L.go = 0;
        L.deep[0] = L.deep[1] = -1;
        sl.want[team] = 0;
        if (has) {
            uint32_t j;
            int32_t node;
            if (direct >= 0) {  // a child split in the round that created it
                node = direct;
                j = t.job[node];
            } else {  // candidate w of the round
                uint32_t lo = 0, hi = J;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) / 2;
                    if (pre[mid] <= w)
                        lo = mid;
                    else
                        hi = mid;
                }
                j = lo;
                node = g.defer ? int32_t(g.exp[g.jobs[j].node_base + (uint32_t(w) - pre[j])])
                               : int32_t(g.cstart[par * J + j] + (uint32_t(w) - pre[j]));
            }
            L.node = node;
            sl.job[team] = j;
            // a candidate already split (in the round that created it) is not split again
            sl.want[team] = t.b[node] >= g.theta[j] && t.child[node] < 0 ? 1 : 0;
        }
    }
    __syncthreads();
    // two child slots per expanding team: one atomic per run of teams of the same job
    if (threadIdx.x == 0) {
        for (int a = 0; a < TPB;) {
            if (!sl.want[a]) {
                sl.slot[a] = -1;
                ++a;
                continue;
            }
            const uint32_t j = sl.job[a];
            int b = a, cnt = 0;
            for (; b < TPB && (!sl.want[b] || sl.job[b] == j); ++b) cnt += sl.want[b];
            const Job<T>& jb = g.jobs[j];
            const uint32_t base = atomicAdd(&g.rt[j].count, 2u * uint32_t(cnt));
            const bool fits = base + 2u * uint32_t(cnt) <= jb.node_base + jb.node_cap;
            if (fits)
                atomicAdd(&g.ncand[(par ^ 1) * J + j], 2u * uint32_t(cnt));
            else
                atomicOr(&g.rt[j].flags, JF_CAPACITY);
            int k = 0;
            for (int c = a; c < b; ++c) {
                sl.slot[c] = (sl.want[c] && fits) ? int32_t(base + 2u * uint32_t(k)) : -1;
                k += sl.want[c];
            }
            a = b;
        }
    }
    __syncthreads();
    if (lane == 0 && sl.want[team] && sl.slot[team] >= 0) {
        const uint32_t j = sl.job[team];
        const int32_t node = L.node;
        L.go = 1;
        L.slot = sl.slot[team];
        L.job = int32_t(j);
        const int dim = t.dim[node];
        L.dim = dim;
        T plo[D], phi[D];
        for (int d = 0; d < D; ++d) {
            plo[d] = t.lo[size_t(node) * D + d];
            phi[d] = t.hi[size_t(node) * D + d];
            L.pbox[0][d] = plo[d];
            L.pbox[1][d] = phi[d];
        }
        split_boxes<T, D>(plo, phi, dim, L.box[0][0], L.box[0][1], L.box[1][0], L.box[1][1]);
    }
    __syncthreads();
    const bool go = L.go != 0;
    if (go)
        for (int k = lane; k < K; k += TEAM) L.pv[k] = t.v[size_t(L.node) * K + k];
    __syncthreads();
    if (go) {
        const int dim = L.dim;
        const int stride = pow3(dim);
        const F& f = fs[L.job];
        for (int idx = lane; idx < K; idx += TEAM) {
            const int jd = digit3(idx, dim);
            if (jd == 1) continue;
            const int base = idx - jd * stride;
            L.cv[0][idx] = L.pv[jd == 0 ? base : base + stride];
            L.cv[1][idx] = L.pv[jd == 0 ? base + stride : base + 2 * stride];
        }
        for (int q = lane; q < 2 * KM; q += TEAM) {
            const int c = q / KM, jj = q % KM;
            T x[D];
            const int idx = split_new_node<T, D>(L.pbox[0], L.pbox[1], dim, c, jj, x);
            L.cv[c][idx] = f.template eval<M, D>(x);
        }
    }
    __syncthreads();
    if (go) {
        const size_t s0 = size_t(L.slot);
        for (int q = lane; q < 2 * K; q += TEAM) t.v[(s0 + q / K) * K + q % K] = L.cv[q / K][q % K];
    }
    const int heur = go ? g.jobs[L.job].heur : 0;
    T e = 0;
    int dm = 0;
    team_heuristic<T, D, TEAM>(L, 2, heur, lane, go, e, dm);
    if constexpr (LEARN) {
        // learned priority of both children (lane c scores child c); the heuristic's
        // split dim is kept, and its error stands in where the network output is NaN
        constexpr int KP = nn::Dims<D>::KP;
        const bool lrn = go && g.jobs[L.job].learn != 0;
        float fsc = 0.f;
        int fdepth = 0;
        uint32_t pcode[D];
        for (int d = 0; d < D; ++d) pcode[d] = lrn ? t.code[size_t(L.node) * D + d] : 0u;
        static_assert(sizeof(L.e) >= sizeof(float) * 2 * D * pow3(D - 1), "fold terms fit the line errors");
        team_child_features<T, D, TEAM>(L, lrn ? &g.jobs[L.job] : nullptr, pcode, lrn ? L.dim : 0, lane,
                                        &ll->xs[team * 2], reinterpret_cast<float*>(&L.e[0][0]), fsc, fdepth);
        __syncthreads();
        if constexpr (LearnLds<D>::W0_DEV)
            nn::forward_block_w0dev<KP>(&static_cast<const nn::NetW<KP>*>(g.net)->w0t[0][0], ll->net,
                                        &ll->xs[0][0], LearnLds<D>::R, ll->h0(), ll->h1(), ll->h0(), ll->y);
        else
            nn::forward_block<KP>(ll->net, &ll->xs[0][0], LearnLds<D>::R, ll->h0(), ll->h1(), ll->h0(), ll->y);
        if (lrn && lane < 2) {
            // lp = log2 of the predicted energy.  Stored as 2^(lp/4): the same order, and
            // the whole span of lp (about -450 for regions of zeros .. +50) stays inside
            // the FP32 range, so no clamp ties distinct predictions.  A child is held
            // below 2^(-1/8) of its parent's predicted energy (2^(-1/32) of its stored
            // priority; the root's is +inf): priorities fall strictly along every path,
            // so b is the node's own priority and a subtree the network overrates cannot
            // tie with its ancestor and grow round after round.
            const float lp = ll->y[team * 2 + lane] * ll->net.sigma + ll->net.mu +
                             2.f * log2f(vmax(fsc, 0x1p-100f)) - float(fdepth);
            if (lp == lp) {
                const float cap = log2f(float(t.err[L.node])) - 0.03125f;
                e = T(exp2f(vmin(vmax(vmin(0.25f * lp, cap), -126.f), 60.f)));
            }
            // size floor: a child still wider than 2^-lfloor along some dim goes before
            // every network-ranked region (2^(126 - depth/4) > 2^60, falling along
            // paths) and is split along its widest dim.  Large regions are rare in the
            // training pairs, so the network's ranking of them is the least reliable.
            const Job<T>& jb = g.jobs[L.job];
            int nmin = 1 << 30, dmin = 0;
            for (int d = 0; d < D; ++d) {
                const uint32_t code = d == L.dim ? pcode[d] * 2u + uint32_t(lane) : pcode[d];
                const int n = nn::code_depth(code);
                if (n < nmin) {
                    nmin = n;
                    dmin = d;
                }
            }
            if (nmin < jb.lfloor) {
                e = T(exp2f(126.f - 0.25f * float(fdepth)));
                dm = dmin;
            }
        }
    }
    if (go && lane < 2) {
        const int c = lane;
        const uint32_t j = uint32_t(L.job);
        const int32_t node = L.node;
        const int dim = L.dim;
        const size_t s = size_t(L.slot) + c;
        const uint32_t pcode = t.code[size_t(node) * D + dim];
        if (pcode >= 0x80000000u) atomicOr(&g.rt[j].flags, JF_DEPTH);
        atomicMax(&sl.depth[dim], 31u - uint32_t(__clz(pcode * 2u + uint32_t(c))));
        for (int d = 0; d < D; ++d) {
            t.lo[s * D + d] = L.box[c][0][d];
            t.hi[s * D + d] = L.box[c][1][d];
            t.code[s * D + d] = (d == dim) ? (pcode * 2u + uint32_t(c)) : t.code[size_t(node) * D + d];
        }
        t.err[s] = e;
        t.dim[s] = dm;
        const T bp = t.b[node], b1p = t.b1[node];
        T nb, nb1;  // first two entries of the path's suffix-minimum stack
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
        L.deep[c] = may_deepen && nb >= g.theta[j] ? int32_t(s) : -1;
        t.parent[s] = node;
        t.child[s] = -1;
        t.job[s] = j;
        const uint32_t kn = key16(nb);
        hist_add(g.hist16, g.rt[j], j, kn, bp == T(INFINITY) || (kn >> 8) > (key16(bp) >> 8));
        if (g.defer) atomicAdd(&g.chist[(size_t(par ^ 1) * J + j) * CHIST + (kn >> 4)], 1u);
        if (c == 0) t.child[node] = int32_t(s);
    }
    __syncthreads();
}

/* Round kernels.  A round is k_theta_round (one block per job: the boundary bucket of
   the job's N-1 largest b among its nodes so far, as a lower bound theta) followed by
   k_expand_round (teams expand the job's new candidates with b >= theta).  Kernel
   boundaries order the rounds, so no grid-wide barrier or cache flush is needed inside
   a kernel.  A round with no candidates exits at once; the host launches rounds in
   chunks and checks for remaining candidates between chunks. */
constexpr int THETA_THREADS = 1024;  // 4 threads per histogram row: 16 vector loads each
template <typename T>
__global__ void __launch_bounds__(THETA_THREADS) k_theta_round(GenArgs<T> g, int par) {
    __shared__ uint32_t sh[256];
    const uint32_t j = blockIdx.x;
    const uint32_t J = g.J;
    const Job<T>& jb = g.jobs[j];
    const uint32_t ncand = g.ncand[par * J + j] + (g.defer ? g.ndef[par * J + j] : 0u);
    // this round's children are appended from the job's current count on; with deferred
    // rounds, the next round's list and candidate histogram start empty
    if (threadIdx.x == 0) {
        g.cstart[(par ^ 1) * J + j] = g.rt[j].count;
        g.ncand[(par ^ 1) * J + j] = 0;
        if (g.defer) {
            g.ndef[(par ^ 1) * J + j] = 0;
            g.nexp[j] = 0;
        }
    }
    if (g.defer) {
        uint4* h = reinterpret_cast<uint4*>(g.chist + (size_t(par ^ 1) * J + j) * CHIST);
        for (int i = threadIdx.x; i < CHIST / 4; i += blockDim.x) h[i] = make_uint4(0u, 0u, 0u, 0u);
    }
    if (ncand == 0) return;
    const uint32_t total = g.rt[j].count - jb.node_base;
    const uint32_t need = jb.N - 1;
    T th;
    if (need == 0) {
        th = T(INFINITY);
    } else if (total < need) {
        th = T(-INFINITY);
    } else {
        // nodes are only added, so the need-th largest b -- and its bucket -- only rise
        // from round to round: the rows below last round's boundary row are skipped
        uint32_t k, above;
        boundary_bucket<THETA_THREADS>(g.hist16 + size_t(j) * 65536, need, sh, k, above, g.rt[j].row_floor,
                                       g.rt[j].row_ceil);
        key16_floor(k, th);
        if (threadIdx.x == 0) g.rt[j].row_floor = k >> 8;
    }
    T eff = th;
    if (g.defer && jb.defer_m > 0 && ncand > jb.defer_m && need > 0) {
        // tau: the lower edge of the 12-bit key bucket holding the defer_m-th largest
        // candidate b (any tau keeps the result exact; this one bounds a round's work)
        const uint32_t* ch = g.chist + (size_t(par) * J + j) * CHIST;
        __shared__ uint32_t sel_row, sel_above;
        if (threadIdx.x < 256) {
            uint32_t a = 0;
            for (int i = 0; i < 16; ++i) a += ch[threadIdx.x * 16 + i];
            sh[threadIdx.x] = a;
        }
        __syncthreads();
        suffix_sum256(sh);
        const uint32_t m = jb.defer_m;
        if (threadIdx.x < 256 && sh[threadIdx.x] >= m && (threadIdx.x == 255 || sh[threadIdx.x + 1] < m)) {
            sel_row = threadIdx.x;
            sel_above = threadIdx.x == 255 ? 0u : sh[threadIdx.x + 1];
        }
        __syncthreads();
        if (sh[0] >= m) {  // block-uniform: the histogram's total
            const uint32_t row = sel_row;
            uint32_t acc = sel_above, b12 = row * 16;
            for (int i = 15; i >= 0; --i) {
                acc += ch[row * 16 + i];
                if (acc >= m) {
                    b12 = row * 16 + uint32_t(i);
                    break;
                }
            }
            T tau;
            key16_floor(b12 << 4, tau);
            eff = tau > th ? tau : th;
        }
    }
    if (threadIdx.x == 0) {
        g.theta[j] = eff;
        if (g.defer) g.theta_lo[j] = th;
    }
    if (g.defer == 2) {
        // jobs of up to 32768 regions: this block sorts its own job's candidates, as
        // k_filter_round does for larger ones (one launch less per round)
        __shared__ uint32_t s_nexp, s_ndef;
        if (threadIdx.x == 0) {
            s_nexp = 0;
            s_ndef = 0;
        }
        __syncthreads();  // also orders the candidate histogram's clearing before its counts
        const uint32_t nc = g.ncand[par * J + j], cs = g.cstart[par * J + j];
        for (uint32_t i = threadIdx.x; i < ncand; i += blockDim.x) {
            uint32_t slot;
            T b;
            if (i < nc) {  // a child of the last round (split already if a deepened pass took it)
                slot = cs + i;
                if (g.t.child[slot] >= 0) continue;
                b = g.t.b[slot];
            } else {
                const DefEnt<T> e = g.def[size_t(par) * g.cap_total + jb.node_base + (i - nc)];
                slot = e.slot;
                b = e.b;
            }
            if (b >= eff) {
                g.exp[jb.node_base + atomicAdd(&s_nexp, 1u)] = slot;
            } else if (b >= th) {
                g.def[size_t(par ^ 1) * g.cap_total + jb.node_base + atomicAdd(&s_ndef, 1u)] = DefEnt<T>{slot, b};
                atomicAdd(&g.chist[(size_t(par ^ 1) * J + j) * CHIST + (key16(b) >> 4)], 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            g.nexp[j] = s_nexp;
            g.ndef[(par ^ 1) * J + j] = s_ndef;
        }
    }
}

/* Deferred rounds: sorts the round's candidates (the last round's children and the
   deferred list) into the expansions (b >= max(theta, tau)), the next round's deferred
   list (theta <= b < tau) and nothing (b < theta: never split, as theta only rises).
   A grid-stride loop over every job's candidates; appends are wave-aggregated. */
template <typename T>
__global__ void __launch_bounds__(256) k_filter_round(GenArgs<T> g, int par) {
    extern __shared__ uint32_t pre[];  // J + 1 entries
    const uint32_t J = g.J;
    block_prefix(pre, J, [&](uint32_t j) { return g.ncand[par * J + j] + g.ndef[par * J + j]; });
    const uint32_t total = pre[J];
    const uint32_t stride = gridDim.x * blockDim.x;
    // the loop bound is uniform across the block: waves leave it together
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += stride) {
        const uint32_t w = base + threadIdx.x;
        uint32_t j = 0, slot = 0, kind = 0;  // kind 1: expand, 2: defer
        T b = T(0);
        if (w < total) {
            uint32_t lo = 0, hi = J;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (pre[mid] <= w)
                    lo = mid;
                else
                    hi = mid;
            }
            j = lo;
            const uint32_t i = w - pre[j], nc = g.ncand[par * J + j];
            const Job<T>& jb = g.jobs[j];
            bool open = true;
            if (i < nc) {  // a child of the last round (split already if a deepened pass took it)
                slot = g.cstart[par * J + j] + i;
                open = g.t.child[slot] < 0;
                b = g.t.b[slot];
            } else {
                const DefEnt<T> e = g.def[size_t(par) * g.cap_total + jb.node_base + (i - nc)];
                slot = e.slot;
                b = e.b;
            }
            if (open) kind = b >= g.theta[j] ? 1u : (b >= g.theta_lo[j] ? 2u : 0u);
        }
        for (uint32_t want = 1; want <= 2; ++want) {
            const bool me = kind == want;
            uint64_t act = __ballot(me);
            while (act) {
                const int leader = __ffsll((unsigned long long)act) - 1;
                const uint32_t jl = __shfl(j, leader);
                const uint64_t same = __ballot(me && j == jl);
                uint32_t pos0 = 0;
                const int lane = int(threadIdx.x % 64);
                if (lane == leader)
                    pos0 = atomicAdd(want == 1 ? &g.nexp[jl] : &g.ndef[(par ^ 1) * J + jl], uint32_t(__popcll(same)));
                pos0 = __shfl(pos0, leader);
                if (me && j == jl) {
                    const uint32_t pos = pos0 + uint32_t(__popcll(same & ((1ull << lane) - 1ull)));
                    const uint32_t nb = g.jobs[j].node_base;
                    if (want == 1) {
                        g.exp[nb + pos] = slot;
                    } else {
                        g.def[size_t(par ^ 1) * g.cap_total + nb + pos] = DefEnt<T>{slot, b};
                        atomicAdd(&g.chist[(size_t(par ^ 1) * J + j) * CHIST + (key16(b) >> 4)], 1u);
                    }
                }
                act &= ~same;
            }
        }
    }
}

template <typename T, int D, typename F, typename M, bool LEARN>
__global__ void __launch_bounds__(GEN_THREADS) k_expand_round(GenArgs<T> g, const F* __restrict__ fs, int par) {
    constexpr int TEAM = TeamShape<D>::TEAM;
    constexpr int TPB = GEN_THREADS / TEAM;  // teams per block
    extern __shared__ uint32_t pre[];        // J + 1 entries: the launch's dynamic LDS
    __shared__ TeamLds<T, D> lds[TPB];
    __shared__ SlotLds<TPB> slots;
    __shared__ typename std::conditional<LEARN, LearnLds<D>, char>::type learn_lds;
    LearnLds<D>* ll = nullptr;
    const uint32_t J = g.J;
    const int team = threadIdx.x / TEAM;
    block_prefix(pre, J, [&](uint32_t j) { return g.defer ? g.nexp[j] : g.ncand[par * J + j]; });
    const uint32_t total = pre[J];
    if (blockIdx.x == 0 && threadIdx.x == 0 && total > 0) atomicAdd(g.rounds, 1u);
    if (uint64_t(blockIdx.x) * TPB >= total) return;  // no candidate for this block (uniform)
    if (threadIdx.x < D) slots.depth[threadIdx.x] = 0;  // ordered before use by team_expand's barriers
    if constexpr (LEARN) {
        ll = &learn_lds;
        constexpr int KP = LearnLds<D>::KP;
        const size_t skip = LearnLds<D>::W0_DEV ? nn::net_tail_offset<KP>() : 0;  // bytes left in device memory
        const uint4* src = reinterpret_cast<const uint4*>(static_cast<const char*>(g.net) + skip);
        uint4* dst = reinterpret_cast<uint4*>(&ll->net);
        for (int i = threadIdx.x; i < int(sizeof(ll->net) / 16); i += blockDim.x) dst[i] = src[i];
        for (int i = threadIdx.x + 2 * TPB * KP; i < LearnLds<D>::R * KP; i += blockDim.x)
            ll->xs[i / KP][i % KP] = 0;  // padding rows stay zero
        __syncthreads();
    }
    // pass 0 splits candidate w; with deepen, passes 1 and 2 split its children that
    // already qualify (b >= theta of this round), one level deeper: every split they
    // add is one a later round would consider, so the selection is unchanged and
    // rounds advance two levels.  Pass predicates are block-wide (team_expand syncs).
    const int passes = g.deepen ? 3 : 1;
    for (uint64_t base = uint64_t(blockIdx.x) * TPB; base < total; base += uint64_t(gridDim.x) * TPB) {
        const uint64_t w = base + team;
        int32_t deep0 = -1, deep1 = -1;
        for (int pass = 0; pass < passes; ++pass) {
            const int32_t direct = pass == 0 ? -1 : (pass == 1 ? deep0 : deep1);
            const bool has = pass == 0 ? w < total : direct >= 0;
            if (pass > 0 && !__syncthreads_or(has)) continue;
            team_expand<T, D, F, M, LEARN>(g, fs, lds[team], par, pre, w, has, direct, pass == 0 && passes > 1, ll,
                                           team, slots);
            if (pass == 0) {
                deep0 = lds[team].deep[0];
                deep1 = lds[team].deep[1];
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < D && slots.depth[threadIdx.x] != 0) atomicMax(&g.depth[threadIdx.x], slots.depth[threadIdx.x]);
}

/* ------------------------------------------------------------------ training pairs
   Features and FP64 target (nn::TargetKind) of every stride-th node the rounds
   created (the regions the method itself generates), one thread per node; job j's
   nodes land at rows out_base[j] + (slot - node_base_j) / stride. */
template <typename T, int D, typename F64>
__global__ void k_train_pairs(const Job<T>* __restrict__ jobs, TreeArrays<T> t, uint32_t cap_total,
                              const uint64_t* __restrict__ out_base, const F64* __restrict__ fs64, int kind,
                              uint32_t stride, uint16_t* __restrict__ X, float* __restrict__ tgt) {
    constexpr int K = pow3(D), KM = pow3(D - 1), KP = nn::Dims<D>::KP;
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= cap_total) return;
    const uint32_t j = t.job[s];
    if (j == NONE) return;
    const Job<T>& jb = jobs[j];
    const uint32_t local = s - jb.node_base;
    if (local % stride != 0) return;
    const uint64_t row = out_base[j] + local / stride;
    const T* lo = t.lo + size_t(s) * D;
    const T* hi = t.hi + size_t(s) * D;
    const T* v = t.v + size_t(s) * K;
    uint16_t* xr = X + row * KP;
    float sc = 0.f;
    int depth = 0;
    nn::region_features<T, D, KP>(lo, hi, jb.lo, jb.hi, v, t.code + size_t(s) * D, jb.lres,
                                  [xr](int i, float x) { xr[i] = nn::f2bf(x); }, sc, depth);
    double vol = 1.0;
    for (int d = 0; d < D; ++d) vol *= double(hi[d]) - double(lo[d]);
    const F64& f = fs64[j];
    double e;
    if (kind == nn::TARGET_ENERGY) {
        e = nn::true_error<T, D, F64, PreciseMath>(f, lo, hi, v);
    } else {
        nn::BinGrid g{};
        g.db = jb.db;
        for (int d = 0; d < jb.db; ++d) {
            g.glo[d] = double(jb.lo[d]);
            g.step[d] = (double(jb.hi[d]) - double(jb.lo[d])) / double(jb.res[d]);
            g.res[d] = jb.res[d];
        }
        double dlo[D], dhi[D], vp[K];
        for (int d = 0; d < D; ++d) {
            dlo[d] = double(lo[d]);
            dhi[d] = double(hi[d]);
        }
        for (int k = 0; k < K; ++k) vp[k] = double(v[k]);
        const int dim = t.dim[s];
        double Vp = 0.0, Vc = 0.0;
        if (kind == nn::TARGET_BINVAR) {
            nn::bin_variances<D, F64, PreciseMath>(f, g, dlo, dhi, vp, dim, nullptr, nullptr, Vp, Vc);
            e = Vp;
        } else {
            // children of the heuristic's split, new planes evaluated in FP64
            double v0[K], v1[K];
            split_inherit<double, D>(vp, dim, v0, v1);
            for (int c = 0; c < 2; ++c)
                for (int jj = 0; jj < KM; ++jj) {
                    double x[D];
                    const int idx = split_new_node<double, D>(dlo, dhi, dim, c, jj, x);
                    (c == 0 ? v0 : v1)[idx] = f.template eval<PreciseMath, D>(x);
                }
            nn::bin_variances<D, F64, PreciseMath>(f, g, dlo, dhi, vp, dim, v0, v1, Vp, Vc);
            e = Vp - Vc;
        }
    }
    tgt[row] = nn::error_target(e, sc, vol);
}

/* ------------------------------------------------------------------ selection */
template <typename T>
__global__ void k_sel_bound(const Job<T>* __restrict__ jobs, JobRt* rt, const uint32_t* hist16, int32_t* leaf_slot) {
    __shared__ uint32_t sh[256];
    const uint32_t j = blockIdx.x;
    const uint32_t need = jobs[j].N - 1;
    // the job's leaf slots start unfilled (k_leaves fills them)
    for (uint32_t i = threadIdx.x; i <= need; i += blockDim.x) leaf_slot[jobs[j].reg_base + i] = -1;
    uint32_t k = NONE, above = 0;
    if (need > 0) boundary_bucket<256>(hist16 + size_t(j) * 65536, need, sh, k, above, 0u, rt[j].row_ceil);
    if (threadIdx.x == 0) {
        rt[j].kbound = k;
        rt[j].above = above;
        rt[j].nbucket = 0;
        rt[j].nleaves = 0;
        rt[j].ties = 0;
    }
}

/* Nodes above the boundary bucket are split; nodes inside it are gathered. */
template <typename T>
__global__ void k_sel_gather(const Job<T>* __restrict__ jobs, JobRt* rt, TreeArrays<T> t, uint32_t cap_total,
                             int32_t* bucket) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= cap_total) return;
    const uint32_t j = t.job[s];
    uint8_t p = 0;
    if (j != NONE && s < rt[j].count && rt[j].kbound != NONE) {
        const uint32_t k = key16(t.b[s]);
        if (k > rt[j].kbound) {
            p = 1;
        } else if (k == rt[j].kbound) {
            const uint32_t pos = atomicAdd(&rt[j].nbucket, 1u);
            bucket[jobs[j].node_base + pos] = int32_t(s);
        }
    }
    t.popped[s] = p;
}

/* Digit `pass` (8 bits, most significant first) of the selection key of node s:
   b below its top 16 bits, b1, then the bitwise complement of every dim's code (so a
   larger key means split earlier; smaller codes win ties). */
template <typename T, int D>
__device__ inline uint32_t sel_digit(const TreeArrays<T>& t, int32_t s, int pass) {
    constexpr int BB = int(sizeof(T));  // bytes of b and b1
    constexpr int nb_b = BB - 2;        // low bytes of b (top 2 bytes fixed by the bucket)
    auto byte_of = [](uint64_t x, int i, int nbytes) { return uint32_t((x >> (8 * (nbytes - 1 - i))) & 0xFFu); };
    if (pass < nb_b) {
        const uint64_t x = (BB == 4) ? uint64_t(__float_as_uint(float(t.b[s])))
                                     : uint64_t(__double_as_longlong(double(t.b[s])));
        return byte_of(x, pass + 2, BB);
    }
    pass -= nb_b;
    if (pass < BB) {
        const uint64_t x = (BB == 4) ? uint64_t(__float_as_uint(float(t.b1[s])))
                                     : uint64_t(__double_as_longlong(double(t.b1[s])));
        return byte_of(x, pass, BB);
    }
    pass -= BB;
    const int d = pass / 4;
    return byte_of(uint64_t(~t.code[size_t(s) * D + d]), pass % 4, 4);
}

/* One workgroup of 256 per job: exact radix select inside the boundary bucket.  bstate
   per bucket item: 0 undecided, 1 split, 2 kept.  Each pass keeps the items whose digit
   is the one holding the rank-th largest; it stops once a single item is left (keys
   are unique per node, so that item is the rank-th largest). */
template <typename T, int D>
__global__ void __launch_bounds__(256) k_sel_resolve(const Job<T>* __restrict__ jobs, JobRt* rt, TreeArrays<T> t,
                                                     const int32_t* __restrict__ bucket, uint8_t* bstate) {
    constexpr int PASSES = (int(sizeof(T)) - 2) + int(sizeof(T)) + 4 * D;
    __shared__ uint32_t hist[256];
    __shared__ uint32_t sel_digit_s, sel_above_s, sel_left_s;
    const uint32_t j = blockIdx.x;
    const Job<T>& jb = jobs[j];
    if (rt[j].kbound == NONE) return;
    const uint32_t nb = rt[j].nbucket;
    const int32_t* lst = bucket + jb.node_base;
    uint8_t* st = bstate + jb.node_base;
    uint32_t rank = (jb.N - 1) - rt[j].above;  // split this many bucket items, largest first
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) st[i] = 0;
    __syncthreads();
    for (int pass = 0; pass < PASSES; ++pass) {
        hist[threadIdx.x] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
            if (st[i] == 0) atomicAdd(&hist[sel_digit<T, D>(t, lst[i], pass)], 1u);
        __syncthreads();
        // the digit holding the rank-th largest: suffix counts, then the one place where
        // they cross rank (1 <= rank <= undecided items)
        const uint32_t own = hist[threadIdx.x];
        __syncthreads();
        suffix_sum256(hist);
        if (hist[threadIdx.x] >= rank && (threadIdx.x == 255 || hist[threadIdx.x + 1] < rank)) {
            sel_digit_s = threadIdx.x;
            sel_above_s = hist[threadIdx.x] - own;
            sel_left_s = own;
        }
        __syncthreads();
        const uint32_t dsel = sel_digit_s;
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
            if (st[i] == 0) {
                const uint32_t dg = sel_digit<T, D>(t, lst[i], pass);
                if (dg > dsel)
                    st[i] = 1;
                else if (dg < dsel)
                    st[i] = 2;
            }
        rank -= sel_above_s;
        const bool last = sel_left_s == 1;
        __syncthreads();
        if (last) break;
    }
    // exactly one undecided item is left (the key is unique per node): the rank-th
    // largest, split too
    __shared__ int32_t kstar;
    if (threadIdx.x == 0) kstar = -1;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
        if (st[i] == 0) kstar = lst[i];
    __syncthreads();
    if (kstar < 0) {
        if (threadIdx.x == 0) atomicOr(&rt[j].flags, JF_LEAVES);
        return;
    }
    const T bk = t.b[kstar], b1k = t.b1[kstar];
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) {
        const int32_t s = lst[i];
        const bool split = st[i] != 2;
        t.popped[s] = split ? 1 : 0;
        if (!split && t.b[s] == bk && t.b1[s] == b1k) atomicAdd(&rt[j].ties, 1u);
    }
}

/* Unsplit children of split nodes (or the root when N = 1), appended per job. */
template <typename T>
__global__ void k_leaves(const Job<T>* __restrict__ jobs, JobRt* rt, TreeArrays<T> t, uint32_t cap_total,
                         int32_t* leaf_slot) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    bool leaf = false;
    uint32_t j = NONE;
    if (s < cap_total) {
        j = t.job[s];
        if (j != NONE && s < rt[j].count && !t.popped[s]) {
            const int32_t p = t.parent[s];
            leaf = (p >= 0) ? (t.popped[p] != 0) : (jobs[j].N == 1);
        }
    }
    // wave-aggregated append: lanes of the leader's job share one atomic
    const uint64_t act = __ballot(leaf);
    if (act == 0) return;
    const int leader = __ffsll((unsigned long long)act) - 1;
    const uint32_t jl = __shfl(j, leader);
    const uint64_t same = __ballot(leaf && j == jl);
    uint32_t base = 0;
    if (int(threadIdx.x % 64) == leader) base = atomicAdd(&rt[jl].nleaves, uint32_t(__popcll(same)));
    base = __shfl(base, leader);
    if (!leaf) return;
    uint32_t pos;
    if (j == jl) {
        const uint64_t below = same & ((1ull << (threadIdx.x % 64)) - 1ull);
        pos = base + uint32_t(__popcll(below));
    } else {
        pos = atomicAdd(&rt[j].nleaves, 1u);
    }
    const Job<T>& jb = jobs[j];
    if (pos >= jb.N) {
        atomicOr(&rt[j].flags, JF_LEAVES);
        return;
    }
    leaf_slot[jb.reg_base + pos] = int32_t(s);
}

/* Field widths (bits) of the codes in a 64-bit leaf key */
template <int D>
struct Widths {
    int w[D];
};

/* Sort keys of one LSD pass over the leaves (canonical order: job, code_0, ...,
   code_{D-1}).  which = -1: job; which = d: code_d (32 bits). */
template <int D>
__global__ void k_leaf_keys(const int32_t* __restrict__ leaf_slot, const uint32_t* __restrict__ code,
                            const uint32_t* __restrict__ job, uint32_t n, int which, uint32_t* keys) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int32_t sl = leaf_slot[r];
    if (sl < 0) {  // unfilled (the host rejects the batch before any slot is used)
        keys[r] = NONE;
        return;
    }
    const size_t s = size_t(sl);
    keys[r] = which < 0 ? job[s] : code[s * D + which];
}

/* One 64-bit key per leaf, (job, code_0, ..., code_{D-1}) in fields of `width` bits
   (code_d < 2^width_d), so one sort gives the lexicographic order of the LSD passes. */
template <int D>
__global__ void k_leaf_keys64(const int32_t* __restrict__ leaf_slot, const uint32_t* __restrict__ code,
                              const uint32_t* __restrict__ job, uint32_t n, Widths<D> wd, uint64_t* keys,
                              int32_t* vals) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int32_t sl = leaf_slot[r];
    vals[r] = sl;
    if (sl < 0) {
        keys[r] = ~uint64_t(0);
        return;
    }
    const size_t s = size_t(sl);
    uint64_t k = job[s];
    for (int d = 0; d < D; ++d) k = (k << wd.w[d]) | code[s * D + d];
    keys[r] = k;
}

/* ------------------------------------------------------------------ regions */
/* Per-region record the residual samples read (one gather per sample): box lo[D],
   hi[D], the inverse extents 1/(hi - lo) (0 for an empty extent), then the 3^D monomial
   coefficients; padded to a multiple of 4 values so a record starts 16-byte aligned
   (64 bytes at D = 2 in FP32: one cache line). */
template <int D>
VMX_HD constexpr int samp_stride() {
    return (3 * D + pow3(D) + 3) & ~3;
}

template <typename T>
struct RegionArrays {
    T *lo, *hi, *vol;
    T* samp;  // sampling records, samp_stride<D>() values per region
    T* v;    // node values (3^D per region) for importance sampling, else null
    T* err;  // Region::error() per region for the rr_error_region roulette, else null
    double* marg;
    uint32_t* job;
    uint32_t* fpflag;  // [n+1]
    uint32_t* fpid;    // [n+1] exclusive scan of fpflag
};

template <typename T>
struct FootArrays {
    T *lo, *hi;          // [F*3] bin dims
    double* scale;       // [F]
    double* marg;        // [F*mb]
    uint32_t *first, *count, *box, *job;  // box [F*6]
    uint32_t* npairs;    // [F+1] bins in the pixel box (exclusive-scanned into pair offsets)
};

/* the residual estimator's parts per form: region roulette (0 uniform, 1 rr_integral_region,
   2 rr_error_region, 3 rr_pdf_region), point sampler, fixed weight */
VMX_HD constexpr int form_roulette(int f) {
    return (f == VMX_FORM_VR_OPTIMIZE_RR_INTEGRAL || f == VMX_FORM_VR_FIXED_RR_INTEGRAL ||
            f == VMX_FORM_VR_OPTIMIZE_RR_INTEGRAL_IMPORTANCE || f == VMX_FORM_VR_FIXED_RR_INTEGRAL_IMPORTANCE)
               ? 1
           : (f == VMX_FORM_VR_OPTIMIZE_RR_ERROR || f == VMX_FORM_VR_FIXED_RR_ERROR ||
              f == VMX_FORM_VR_OPTIMIZE_RR_ERROR_IMPORTANCE || f == VMX_FORM_VR_FIXED_RR_ERROR_IMPORTANCE)
               ? 2
           : (f == VMX_FORM_VR_OPTIMIZE_RR_PDF || f == VMX_FORM_VR_FIXED_RR_PDF ||
              f == VMX_FORM_VR_OPTIMIZE_RR_PDF_IMPORTANCE || f == VMX_FORM_VR_FIXED_RR_PDF_IMPORTANCE)
               ? 3
               : 0;
}
/* the point sampler draws from the region polynomial's pdf (always, or by the roulette
   between strategies of region_sampling_russian_roulette) */
VMX_HD constexpr bool form_sampling_rr(int f) {
    return f == VMX_FORM_VR_OPTIMIZE_SAMPLING_RR || f == VMX_FORM_VR_FIXED_SAMPLING_RR;
}
VMX_HD constexpr bool form_importance(int f) {
    return f == VMX_FORM_VR_OPTIMIZE_IMPORTANCE || f == VMX_FORM_VR_FIXED_IMPORTANCE ||
           (f >= VMX_FORM_VR_OPTIMIZE_RR_INTEGRAL_IMPORTANCE && f <= VMX_FORM_VR_FIXED_RR_PDF_IMPORTANCE) ||
           form_sampling_rr(f);
}
VMX_HD constexpr bool form_fixed(int f) {
    return f == VMX_FORM_VR_FIXED || f == VMX_FORM_VR_FIXED_IMPORTANCE || f == VMX_FORM_VR_FIXED_RR_INTEGRAL ||
           f == VMX_FORM_VR_FIXED_RR_ERROR || f == VMX_FORM_VR_FIXED_RR_PDF ||
           f == VMX_FORM_VR_FIXED_RR_INTEGRAL_IMPORTANCE || f == VMX_FORM_VR_FIXED_RR_ERROR_IMPORTANCE ||
           f == VMX_FORM_VR_FIXED_RR_PDF_IMPORTANCE || f == VMX_FORM_VR_FIXED_SAMPLING_RR ||
           f == VMX_FORM_VR_FIXED_STRATIFIED;
}
/* region_stratification_uniform in place of the uniform region roulette */
VMX_HD constexpr bool form_stratified(int f) {
    return f == VMX_FORM_VR_OPTIMIZE_STRATIFIED || f == VMX_FORM_VR_FIXED_STRATIFIED;
}

template <typename T, int D>
__global__ void k_prepare2(const Job<T>* __restrict__ jobs, const int32_t* __restrict__ leaf_slot, TreeArrays<T> t,
                           RegionArrays<T> r, uint32_t n, int mb, int merge, uint32_t* npairs,
                           unsigned long long* pairs64) {
    constexpr int K = pow3(D);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    // the footprints' pair counts (n + 1 entries) and their 64-bit total start at zero
    if (i < n) npairs[i] = 0;
    if (i == 0) {
        npairs[n] = 0;
        *pairs64 = 0;
    }
    if (i >= n) return;
    if (leaf_slot[i] < 0) {  // selection did not close (the host rejects the result): an empty region
        for (int d = 0; d < D; ++d) r.lo[size_t(i) * D + d] = r.hi[size_t(i) * D + d] = T(0);
        r.vol[i] = T(0);
        for (int k = 0; k < mb; ++k) r.marg[size_t(i) * mb + k] = 0.0;
        for (int k = 0; k < samp_stride<D>(); ++k) r.samp[size_t(i) * samp_stride<D>() + k] = T(0);
        r.job[i] = 0;
        r.fpflag[i] = 1;
        if (i == 0) r.fpflag[n] = 0;
        return;
    }
    const size_t s = size_t(leaf_slot[i]);
    const uint32_t j = t.job[s];
    const Job<T>& jb = jobs[j];
    T lo[D], hi[D], v[K];
    for (int d = 0; d < D; ++d) {
        lo[d] = t.lo[s * D + d];
        hi[d] = t.hi[s * D + d];
        r.lo[size_t(i) * D + d] = lo[d];
        r.hi[size_t(i) * D + d] = hi[d];
    }
    for (int k = 0; k < K; ++k) v[k] = t.v[s * K + k];
    if (r.v)
        for (int k = 0; k < K; ++k) r.v[size_t(i) * K + k] = v[k];
    r.vol[i] = box_volume<T, D>(lo, hi);
    if (r.err) {  // Region::error() (src/newton-cotes/region.h:421-424): dim 0 folded first
        T s[K], tz[K];
        for (int k = 0; k < K; ++k) s[k] = tz[k] = v[k];
        for (int m = K; m > 1; m /= 3)
            for (int q = 0; q < m / 3; ++q) {
                s[q] = simpson(s[3 * q], s[3 * q + 1], s[3 * q + 2]);
                tz[q] = trapezoid(tz[3 * q], tz[3 * q + 2]);
            }
        r.err[i] = r.vol[i] * (s[0] - tz[0]);
    }
    double m[27];
    bin_marginal<T, double, D>(v, jb.db, m);
    for (int k = 0; k < mb; ++k) r.marg[size_t(i) * mb + k] = m[k];
    T c[K];
    monomials<T, T, D>(v, c);
    T* rec = r.samp + size_t(i) * samp_stride<D>();
    for (int d = 0; d < D; ++d) {
        rec[d] = lo[d];
        rec[D + d] = hi[d];
        rec[2 * D + d] = hi[d] > lo[d] ? T(1) / (hi[d] - lo[d]) : T(0);
    }
    for (int k = 0; k < K; ++k) rec[3 * D + k] = c[k];
    for (int k = 3 * D + K; k < samp_stride<D>(); ++k) rec[k] = T(0);
    r.job[i] = j;
    uint32_t flag = 1;
    // the region roulettes weigh each region of a bin: their jobs keep one region per
    // footprint, the batch's other jobs merge as they would alone
    if (merge && form_roulette(jb.form) == 0 && i != jb.reg_base && leaf_slot[i - 1] >= 0) {
        const size_t sp = size_t(leaf_slot[i - 1]);
        flag = 0;
        for (int d = 0; d < jb.db; ++d)
            if (t.code[sp * D + d] != t.code[s * D + d]) flag = 1;
    }
    r.fpflag[i] = flag;
    if (i == 0) r.fpflag[n] = 0;
}

/* k_prepare2 for regions of 4..6 dims: one wave per region, its 3^D node values and
   every intermediate array in LDS (a thread per region spills them to scratch).  Each
   pass of a fold or of the monomial transform runs the independent lines of one dim
   across the lanes, with the operations k_prepare2 applies to each line: the results
   are the same bit for bit. */
template <int D>
constexpr int prep_wpb() {  // waves (regions) per block
    return D >= 6 ? 2 : 4;
}
template <typename T, int D>
struct PrepWave {
    static constexpr int K = pow3(D);
    static constexpr int WPB = prep_wpb<D>();
    T v[K], s[2][K], tz[2][K], c[K];
    double w[K];
};

template <typename T, int D>
__global__ void __launch_bounds__(64 * prep_wpb<D>())
    k_prepare_wave(const Job<T>* __restrict__ jobs, const int32_t* __restrict__ leaf_slot, TreeArrays<T> t,
                   RegionArrays<T> r, uint32_t n, int mb, int merge, uint32_t* npairs,
                   unsigned long long* pairs64) {
    using PW = PrepWave<T, D>;
    constexpr int K = PW::K, WPB = PW::WPB;
    __shared__ PW sw[WPB];
    const int wave = int(threadIdx.x >> 6), lane = int(threadIdx.x & 63);
    const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x;
    // the footprints' pair counts (n + 1 entries) and their 64-bit total start at zero
    if (gt < n) npairs[gt] = 0;
    if (gt == 0) {
        npairs[n] = 0;
        *pairs64 = 0;
    }
    const uint32_t i = blockIdx.x * WPB + uint32_t(wave);
    PW& P = sw[wave];
    // every wave runs every pass (the passes end in block barriers); only regions that
    // exist write results
    const bool real = i < n;
    const bool valid = real && leaf_slot[i] >= 0;
    if (real && !valid) {  // selection did not close (the host rejects the result): an empty region
        for (int k = lane; k < D; k += 64) r.lo[size_t(i) * D + k] = r.hi[size_t(i) * D + k] = T(0);
        for (int k = lane; k < mb; k += 64) r.marg[size_t(i) * mb + k] = 0.0;
        for (int k = lane; k < samp_stride<D>(); k += 64) r.samp[size_t(i) * samp_stride<D>() + k] = T(0);
        if (lane == 0) {
            r.vol[i] = T(0);
            r.job[i] = 0;
            r.fpflag[i] = 1;
            if (i == 0) r.fpflag[n] = 0;
        }
    }
    const size_t s = valid ? size_t(leaf_slot[i]) : 0;
    for (int k = lane; k < K; k += 64) {
        const T x = valid ? t.v[s * K + k] : T(0);
        P.v[k] = x;
        P.s[0][k] = P.tz[0][k] = P.c[k] = x;
        P.w[k] = double(x);
        if (valid && r.v) r.v[size_t(i) * K + k] = x;
    }
    __syncthreads();
    // Region::error() (src/newton-cotes/region.h:421-424): dim 0 folded first
    int cur = 0;
    if (r.err) {
        for (int m = K; m > 1; m /= 3) {
            for (int q = lane; q < m / 3; q += 64) {
                P.s[cur ^ 1][q] = simpson(P.s[cur][3 * q], P.s[cur][3 * q + 1], P.s[cur][3 * q + 2]);
                P.tz[cur ^ 1][q] = trapezoid(P.tz[cur][3 * q], P.tz[cur][3 * q + 2]);
            }
            cur ^= 1;
            __syncthreads();
        }
    }
    // FP64 bin-dim marginal (bin_marginal): the non-bin dims folded, the last first
    int db = 0;
    for (int q = mb; q > 1; q /= 3) ++db;
    int nn = K;
    for (int d = D - 1; d >= db; --d) {
        nn /= 3;
        for (int q = lane; q < nn; q += 64) P.w[q] = simpson_subrange(0.0, 1.0, P.w[q], P.w[q + nn], P.w[q + 2 * nn]);
        __syncthreads();
    }
    // monomial coefficients (monomials): dim 0 first, one line per lane
    int stride = 1;
    for (int d = 0; d < D; ++d) {
        for (int q = lane; q < K / 3; q += 64) {
            const int idx = (q / stride) * stride * 3 + q % stride;
            const T p0 = P.c[idx], p1 = P.c[idx + stride], p2 = P.c[idx + 2 * stride];
            P.c[idx] = p0;
            P.c[idx + stride] = -3 * p0 + 4 * p1 - p2;
            P.c[idx + 2 * stride] = 2 * p0 - 4 * p1 + 2 * p2;
        }
        stride *= 3;
        __syncthreads();
    }
    if (!valid) return;
    const uint32_t j = t.job[s];
    const Job<T>& jb = jobs[j];
    T* rec = r.samp + size_t(i) * samp_stride<D>();
    for (int k = lane; k < mb; k += 64) r.marg[size_t(i) * mb + k] = P.w[k];
    for (int k = lane; k < K; k += 64) rec[3 * D + k] = P.c[k];
    for (int k = 3 * D + K + lane; k < samp_stride<D>(); k += 64) rec[k] = T(0);
    if (lane < D) {
        const T lo = t.lo[s * D + lane], hi = t.hi[s * D + lane];
        r.lo[size_t(i) * D + lane] = lo;
        r.hi[size_t(i) * D + lane] = hi;
        rec[lane] = lo;
        rec[D + lane] = hi;
        rec[2 * D + lane] = hi > lo ? T(1) / (hi - lo) : T(0);
    }
    if (lane == 0) {
        T lo[D], hi[D];
        for (int d = 0; d < D; ++d) {
            lo[d] = t.lo[s * D + d];
            hi[d] = t.hi[s * D + d];
        }
        const T vol = box_volume<T, D>(lo, hi);
        r.vol[i] = vol;
        if (r.err) r.err[i] = vol * (P.s[cur][0] - P.tz[cur][0]);
        r.job[i] = j;
        uint32_t flag = 1;
        // the region roulettes weigh each region of a bin: their jobs keep one region per
        // footprint, the batch's other jobs merge as they would alone
        if (merge && form_roulette(jb.form) == 0 && i != jb.reg_base && leaf_slot[i - 1] >= 0) {
            const size_t sp = size_t(leaf_slot[i - 1]);
            flag = 0;
            for (int d = 0; d < jb.db; ++d)
                if (t.code[sp * D + d] != t.code[s * D + d]) flag = 1;
        }
        r.fpflag[i] = flag;
        if (i == 0) r.fpflag[n] = 0;
    }
}

/* The merged FP64 marginal of a footprint's regions [i, e): sum of vol_r * marginal_r
   in region order, MB = 3^db entries held in registers. */
template <typename T, int MB>
__device__ inline void merge_marginals(const RegionArrays<T>& r, const FootArrays<T>& f, uint32_t i, uint32_t e,
                                       uint32_t id) {
    double acc[MB];
#pragma unroll
    for (int k = 0; k < MB; ++k) acc[k] = 0.0;
    for (uint32_t q = i; q < e; ++q) {
        const double vol = double(r.vol[q]);
#pragma unroll
        for (int k = 0; k < MB; ++k) acc[k] += vol * r.marg[size_t(q) * MB + k];
    }
#pragma unroll
    for (int k = 0; k < MB; ++k) f.marg[size_t(id) * MB + k] = acc[k];
}

/* One thread per footprint: bin-dim box, pixel box, and merged FP64 marginal
   (sum of vol_r * marginal_r over its regions, in region order) or, unmerged, the
   single region's (vol, marginal) so per-region sums keep the reference's rounding. */
template <typename T, int D>
__global__ void k_footprints(const Job<T>* __restrict__ jobs, RegionArrays<T> r, FootArrays<T> f, uint32_t n,
                             int mb, int merge) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !r.fpflag[i]) return;
    const uint32_t id = r.fpid[i];
    uint32_t e = i + 1;
    while (e < n && !r.fpflag[e]) ++e;
    const uint32_t j = r.job[i];
    const Job<T>& jb = jobs[j];
    f.first[id] = i;
    f.count[id] = e - i;
    f.job[id] = j;
    for (int d = 0; d < 3; ++d) {
        uint64_t a = 0, z = 1;
        if (d < jb.db) {
            const T lo = r.lo[size_t(i) * D + d], hi = r.hi[size_t(i) * D + d];
            f.lo[size_t(id) * 3 + d] = lo;
            f.hi[size_t(id) * 3 + d] = hi;
            pixel_range<T>(lo, hi, jb.lo[d], jb.hi[d], jb.res[d], a, z);
        }
        f.box[size_t(id) * 6 + 2 * d] = uint32_t(a);
        f.box[size_t(id) * 6 + 2 * d + 1] = uint32_t(z);
    }
    {
        const uint32_t* bx = f.box + size_t(id) * 6;
        f.npairs[id] = (bx[1] - bx[0]) * (bx[3] - bx[2]) * (bx[5] - bx[4]);
    }
    if (merge && form_roulette(jb.form) == 0) {
        if (mb == 3)
            merge_marginals<T, 3>(r, f, i, e, id);
        else if (mb == 9)
            merge_marginals<T, 9>(r, f, i, e, id);
        else
            merge_marginals<T, 27>(r, f, i, e, id);
        f.scale[id] = 1.0;
    } else {
        for (int k = 0; k < mb; ++k) f.marg[size_t(id) * mb + k] = r.marg[size_t(i) * mb + k];
        f.scale[id] = double(r.vol[i]);
    }
}

/* *total += sum of x[0..n) in 64 bits.  The per-bin list offsets are 32-bit scans of
   these counts, so the host checks this total before it sizes the lists. */
static __global__ void __launch_bounds__(256) k_sum_u64(const uint32_t* __restrict__ x, uint32_t n,
                                                 unsigned long long* __restrict__ total) {
    uint64_t s = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += x[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    __shared__ uint64_t part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(total, (unsigned long long)(part[0] + part[1] + part[2] + part[3]));
}

/* The stratification totals into the rounds' counter words, read back with the job
   states in one copy: [2..3] the 64-bit pair total, [4] the 32-bit pair scan's end,
   [5] the footprint count. */
static __global__ void k_collect_totals(const unsigned long long* __restrict__ pairs64,
                                        const uint32_t* __restrict__ pofs_end, const uint32_t* __restrict__ nfp,
                                        uint32_t* small) {
    if (threadIdx.x == 0) {
        const unsigned long long p = *pairs64;
        small[2] = uint32_t(p);
        small[3] = uint32_t(p >> 32);
        small[4] = *pofs_end;
        small[5] = *nfp;
    }
}

/* q = c / n for the box walks below: a float estimate, then exact integer fix-up. */
__device__ inline uint32_t div_fix(uint32_t c, uint32_t n, float inv) {
    uint32_t q = uint32_t(float(c) * inv);
    int64_t r = int64_t(c) - int64_t(q) * n;
    while (r < 0) {
        --q;
        r += n;
    }
    while (r >= int64_t(n)) {
        ++q;
        r -= n;
    }
    return q;
}

/* Per-bin footprint lists by counting sort, one wave per footprint walking its pixel
   box with the last bin dim fastest (consecutive lanes, consecutive bins).  Pair c of
   footprint w is pair pofs[w] + c of the call.
   PLACE = false: the pair's rank in its bin, rank = cnt[bin]++ (arrival order).
   PLACE = true (off = the counts' exclusive scan): each footprint id goes to
   off[bin] + rank, so every list is filled, in arrival order, with no atomics;
   k_sort_lists then puts each list in ascending footprint id (region order), which
   fixes the per-bin summation order. */
template <typename T, bool PLACE>
__global__ void k_bin_pass(const Job<T>* __restrict__ jobs, FootArrays<T> f, const uint32_t* __restrict__ nfp_ptr,
                           uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off, uint32_t* __restrict__ lst,
                           const uint32_t* __restrict__ pofs, uint32_t* __restrict__ rank) {
    const uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint32_t lane = threadIdx.x % 64;
    if (w >= *nfp_ptr) return;
    const Job<T>& jb = jobs[f.job[w]];
    const uint32_t* bx = f.box + size_t(w) * 6;
    const uint32_t n0 = bx[1] - bx[0], n1 = bx[3] - bx[2], n2 = bx[5] - bx[4];
    const uint32_t total = f.npairs[w];
    const uint64_t r1 = jb.db > 1 ? jb.res[1] : 1, r2 = jb.db > 2 ? jb.res[2] : 1;
    const float i2 = 1.0f / float(n2), i12 = 1.0f / float(n1);
    uint32_t* const rk = rank + pofs[w];
    for (uint32_t c = lane; c < total; c += 64) {
        const uint32_t q2 = div_fix(c, n2, i2);  // c = (a0 * n1 + a1) * n2 + a2
        const uint32_t a2 = c - q2 * n2;
        const uint32_t a0 = div_fix(q2, n1, i12);
        const uint32_t a1 = q2 - a0 * n1;
        const uint64_t local = (uint64_t(bx[0] + a0) * r1 + (bx[2] + a1)) * r2 + (bx[4] + a2);
        const uint64_t bin = jb.bin_base + local;
        if constexpr (PLACE) {
            lst[off[bin] + rk[c]] = w;
        } else {
            rk[c] = atomicAdd(&cnt[bin], 1u);
        }
    }
}

constexpr uint32_t SORT_SHORT = 16;  // longer lists are sorted by a workgroup each
constexpr int LANE_MAXF = 16;  // footprints per bin the per-bin kernel caches in LDS
constexpr int RM_MAXF = 6;     // footprints per bin the region-major per-bin kernel caches in LDS
constexpr int PREFIX_MIN = 6;  // lists longer than this store their prefix counts
static_assert(LANE_MAXF >= PREFIX_MIN && RM_MAXF >= PREFIX_MIN,
              "a list a kernel does not cache whole in LDS has its prefix counts stored");

/* For lists longer than PREFIX_MIN the per-bin kernels find a sample's region beyond
   their LDS cache by binary search over the list's exclusive prefix region counts
   `pcum` and first regions `pfirst`, stored beside the list. */
struct ListPrefix {
    uint32_t* pcum;
    uint32_t* pfirst;
};

/* Ascending footprint ids in every bin's list: insertion sort per lane for short lists,
   staged in the lane's own LDS column (no barriers); longer ones are queued for
   k_sort_long. */
template <typename T>
__global__ void __launch_bounds__(256) k_sort_lists(const uint32_t* __restrict__ off, uint64_t total_bins,
                                                    uint32_t* __restrict__ lst, uint32_t* __restrict__ qlong,
                                                    FootArrays<T> fa, ListPrefix lp) {
    __shared__ uint32_t sv[SORT_SHORT][256];
    const int tid = threadIdx.x;
    const uint64_t b = uint64_t(blockIdx.x) * blockDim.x + tid;
    if (b >= total_bins) return;
    const uint32_t o = off[b], n = off[b + 1] - o;
    if (n <= 1) return;
    if (n > SORT_SHORT) {
        const uint32_t k = atomicAdd(&qlong[0], 1u);
        qlong[1 + k] = uint32_t(b);
        return;
    }
    for (uint32_t i = 0; i < n; ++i) sv[i][tid] = lst[o + i];
    bool moved = false;
    for (uint32_t i = 1; i < n; ++i) {
        const uint32_t v = sv[i][tid];
        uint32_t j = i;
        while (j > 0 && sv[j - 1][tid] > v) {
            sv[j][tid] = sv[j - 1][tid];
            --j;
        }
        sv[j][tid] = v;
        moved |= j != i;
    }
    if (moved)
        for (uint32_t i = 0; i < n; ++i) lst[o + i] = sv[i][tid];
    if (n > uint32_t(PREFIX_MIN)) {
        // four entries' loads ahead of their stores: the stores could alias the
        // footprint arrays for all the compiler knows, so loads behind them would
        // each wait for memory in turn
        uint32_t c = 0;
        for (uint32_t i0 = 0; i0 < n; i0 += 4) {
            uint32_t fv[4], cv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t fi = sv[i0 + u < n ? i0 + u : 0u][tid];  // past the list: entry 0 again
                fv[u] = fa.first[fi];
                cv[u] = fa.count[fi];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (i0 + u < n) {
                    lp.pcum[o + i0 + u] = c;
                    lp.pfirst[o + i0 + u] = fv[u];
                    c += cv[u];
                }
        }
    }
}

/* One workgroup per queued long list: every entry goes to its rank among the list's
   entries (ids within a list are distinct), read from a copy in LDS, or for very long
   lists from a copy in `tmp` read back with device-scope loads.  Blocks stride over the
   queue; the loop bounds are block-uniform. */
constexpr uint32_t SORT_LDS = 8192;
static __global__ void __launch_bounds__(256) k_sort_long(const uint32_t* __restrict__ off, const uint32_t* __restrict__ qlong,
                                                   uint32_t* __restrict__ lst, uint32_t* __restrict__ tmp) {
    __shared__ uint32_t sb[SORT_LDS];
    const uint32_t nq = qlong[0];
    for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
        const uint32_t b = qlong[1 + q];
        const uint32_t o = off[b], n = off[b + 1] - o;
        if (n <= SORT_LDS) {
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) sb[i] = lst[o + i];
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                const uint32_t v = sb[i];
                uint32_t r = 0;
                for (uint32_t k = 0; k < n; ++k) r += sb[k] < v ? 1u : 0u;
                lst[o + r] = v;
            }
        } else {
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) tmp[o + i] = lst[o + i];
            __threadfence();
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                const uint32_t v = __hip_atomic_load(&tmp[o + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                uint32_t r = 0;
                for (uint32_t k = 0; k < n; ++k)
                    r += __hip_atomic_load(&tmp[o + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v ? 1u : 0u;
                lst[o + r] = v;
            }
        }
        __syncthreads();
    }
}

/* Prefix region counts of the queued long lists (sorted by now), one workgroup per list:
   a block-wide scan per 256-entry chunk. */
template <typename T>
__global__ void __launch_bounds__(256) k_list_prefix(const uint32_t* __restrict__ off, const uint32_t* __restrict__ qlong,
                                                     const uint32_t* __restrict__ lst, FootArrays<T> fa, ListPrefix lp) {
    __shared__ uint32_t sc[256];
    const uint32_t nq = qlong[0];
    const int t = threadIdx.x;
    for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
        const uint32_t b = qlong[1 + q];
        const uint32_t o = off[b], n = off[b + 1] - o;
        uint32_t carry = 0;
        for (uint32_t c0 = 0; c0 < n; c0 += 256) {
            const uint32_t i = c0 + uint32_t(t);
            uint32_t fi = 0, cnt = 0;
            if (i < n) {
                fi = lst[o + i];
                cnt = fa.count[fi];
            }
            sc[t] = cnt;
            __syncthreads();
            for (int step = 1; step < 256; step <<= 1) {  // inclusive scan
                const uint32_t add = t >= step ? sc[t - step] : 0u;
                __syncthreads();
                sc[t] += add;
                __syncthreads();
            }
            if (i < n) {
                lp.pcum[o + i] = carry + sc[t] - cnt;
                lp.pfirst[o + i] = fa.first[fi];
            }
            carry += sc[255];
            __syncthreads();
        }
    }
}

template <typename T>
__device__ inline uint32_t job_of_bin(const Job<T>* jobs, uint32_t J, uint64_t b) {
    uint32_t lo = 0, hi = J;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (jobs[mid].bin_base <= b)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

/* Workgroups are dispatched to the 8 XCDs round-robin (workgroup w runs on XCD w % 8).
   This maps them to tiles so that each XCD takes one contiguous run of tiles: neighbouring
   bins share regions, so each XCD's L2 holds one neighbourhood instead of all of them.
   A bijection of [0, G). */
__device__ inline uint32_t xcd_tile(uint32_t w, uint32_t G) {
    constexpr uint32_t NX = 8;
    const uint32_t q = G / NX, r = G % NX, x = w % NX, i = w / NX;
    return x * q + (x < r ? x : r) + i;
}


/* A residual sample's position t in its region: FP64, pos_in (the reference's
   quotient); FP32, (x - lo) times the region's stored inverse extent (one multiply; only
   the polynomial's value at the sample depends on it, while the control-variate
   integral keeps pos_in). */

/* One lane per bin (all jobs of the batch), one block per tile of 256 consecutive bins
   of one job.  The lane sums the control variate over the
   bin's footprints in list order (FP64), then draws the spp residual samples -- a region
   uniform among the bin's regions (binary search over footprint counts cached in LDS),
   a point uniform in bin x region -- and forms the estimator of the job's form
   (…regions-integrator-parallel-variance-reduction.h:67-103,
   …regions-integrator-parallel-control-variates.h:39-65, weight-strategy.h:7-110).
   Moments use the reference's shift by the bin's first sample; each group of 4 samples
   is summed plainly and the group sums are Kahan-compensated in the evaluation type,
   then finished in FP64.  Summation depends on
   the job alone, so a batched integral gives exactly what it gives alone. */
/* The region-major per-bin kernel (bins_rm.h) and the jobs it takes. */
constexpr int RM_SPB = 8;              // samples per bin per pass
constexpr int RM_NS = 256 * RM_SPB;    // samples per pass of a block
constexpr int RM_SLOT_BITS = 11;       // log2(RM_NS)
constexpr uint32_t RM_MAX_REGIONS = 1u << (32 - RM_SLOT_BITS);  // region ids of a job in a key
constexpr uint32_t RM_NONE = 0xFFFFFFFFu;

/* Whether a job runs through k_bins_rm (the host launches it for FP32 regions of 4..6
   dims; k_bins1 skips these jobs then). */
VMX_HD inline bool rm_job(int form, uint32_t N) {
    return (form == VMX_FORM_VR_OPTIMIZE || form == VMX_FORM_VR_FIXED || form == VMX_FORM_CV_ALPHA1) &&
           N < RM_MAX_REGIONS;  // a key never equals RM_NONE
}

/* Tiles of the region-major kernel for a job: 16 x 16 bins with two bin dims, else runs
   of 256 bins. */
inline uint32_t rm_tiles(int db, const uint64_t* res, uint64_t nbins) {
    if (db == 2) return uint32_t(((res[0] + 15) / 16) * ((res[1] + 15) / 16));
    return uint32_t((nbins + 255) / 256);
}

template <typename T, int D, int DB, typename F, typename M, int MODE = 0>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((MODE & 1) == 0 && D <= 3 && sizeof(T) == 4 ? 4 : 1))) k_bins1(const Job<T>* __restrict__ jobs, uint32_t J, uint64_t total_bins,
                                               RegionArrays<T> r, FootArrays<T> fa,
                                               const uint32_t* __restrict__ off, const uint32_t* __restrict__ lst,
                                               const F* __restrict__ fs, JobRt* rt, ListPrefix lp, int xcd,
                                               double* __restrict__ rrw, int rm) {
    // MODE bit 0: importance sampling in the box can occur; bit 1: the region roulettes
    // (rr_integral_region, rr_error_region, rr_pdf_region; rrw: per-list cumulative
    // weights) can; 0: uniform region and uniform point only
    constexpr bool IMP = (MODE & 1) != 0;
    constexpr bool RRI = (MODE & 2) != 0;
    // the stratified split of the samples over the regions: the Fubini form over an
    // infinite range only (the reference builds the optimised integrator for no other)
    constexpr bool STRAT = IsFubiniInf<F>::value;
    constexpr bool FP32 = sizeof(T) == 4;
    constexpr int K = pow3(D);
    // per lane: first region and running region count of its first LANE_MAXF footprints
    __shared__ uint32_t s_first[LANE_MAXF][256];
    __shared__ uint32_t s_cum[LANE_MAXF + 1][256];
    // the integrand's parameters in LDS: every evaluation reads them as LDS broadcasts
    __shared__ alignas(16) unsigned char f_raw[sizeof(F)];
    const int tid = threadIdx.x;
    // the block's job: every job's bins start a fresh tile of blockDim.x bins (tile_base),
    // so a block holds one job's bins and the job's fields are wave-uniform (scalar)
    const uint32_t tile = xcd ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    uint32_t j = 0;
    {
        uint32_t lo = 0, hi = J;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (jobs[mid].tile_base <= tile)
                lo = mid;
            else
                hi = mid;
        }
        j = __builtin_amdgcn_readfirstlane(lo);
    }
    const Job<T>& jb = jobs[j];
    if (rm && rm_job(jb.form, jb.N)) return;  // block-uniform: k_bins_rm runs this job
    {
        const unsigned char* src = reinterpret_cast<const unsigned char*>(&fs[j]);
        for (int i = tid; i < int(sizeof(F) / 4); i += blockDim.x)
            reinterpret_cast<uint32_t*>(f_raw)[i] = reinterpret_cast<const uint32_t*>(src)[i];
        if (tid == 0)
            for (int i = int(sizeof(F) / 4) * 4; i < int(sizeof(F)); ++i) f_raw[i] = src[i];
    }
    __syncthreads();
    const F& f_lds = *reinterpret_cast<const F*>(f_raw);
    {
        const uint64_t b = uint64_t(tile - jb.tile_base) * blockDim.x + tid;  // bin within the job
        if (b >= jb.nbins) return;
        const uint64_t gb = jb.bin_base + b;
        const uint32_t o = off[gb], nf = off[gb + 1] - o;
        uint32_t p[3] = {0, 0, 0};
        {
            uint32_t rem = uint32_t(b);  // a job holds at most 2^31 bins
#pragma unroll
            for (int i = DB - 1; i >= 0; --i) {
                const uint32_t ri = uint32_t(jb.res[i]);
                p[i] = i > 0 ? rem % ri : rem;
                rem /= ri;
            }
        }
        T blo[D], bhi[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            blo[d] = jb.lo[d];
            bhi[d] = jb.hi[d];
        }
#pragma unroll
        for (int i = 0; i < DB; ++i) {
            blo[i] = jb.lo[i] + T(p[i]) * jb.drange[i];
            bhi[i] = jb.lo[i] + T(p[i] + 1) * jb.drange[i];
        }
        const double factor = double(jb.nbins);
        T* bins = static_cast<T*>(jb.bins);
        // control variate: footprint terms in list order
        double approx = (jb.form == VMX_FORM_CV_ALPHA1) ? double(bins[b]) : 0.0;
        const uint32_t nc = nf < uint32_t(LANE_MAXF) ? nf : uint32_t(LANE_MAXF);
        uint32_t n = 0, first0 = 0;
        s_cum[0][tid] = 0;
        for (uint32_t k = 0; k < nf; ++k) {
            const uint32_t fi = lst[o + k];
            n += fa.count[fi];
            if (k < nc) {
                const uint32_t f0 = fa.first[fi];
                if (k == 0) first0 = f0;
                s_first[k][tid] = f0;
                s_cum[k + 1][tid] = n;
            }
            double aa[DB], bb[DB];
            bool empty = false;
#pragma unroll
            for (int d = 0; d < DB; ++d) {
                const T rl = fa.lo[size_t(fi) * 3 + d], rh = fa.hi[size_t(fi) * 3 + d];
                const T il = vmax(blo[d], rl);
                const T ih = vmax(il, vmin(bhi[d], rh));
                if (il >= ih) empty = true;
                aa[d] = double(pos_in(il, rl, rh));
                bb[d] = double(pos_in(ih, rl, rh));
            }
            // FP64 bins keep the reference's divisions by 3; the FP32 path multiplies by 1/3
            const double integ = empty ? 0.0
                                       : fa.scale[fi] * marginal_subrange_n<double, DB, !FP32>(
                                                            fa.marg + size_t(fi) * jb.mb, aa, bb);
            approx += empty ? 0.0 : factor * integ;
            if constexpr (RRI) {  // one region per footprint here
                const int rk = form_roulette(jb.form);
                if (rk == 2) {
                    double share = empty ? 0.0 : 1.0;  // the bin's share of the region
#pragma unroll
                    for (int d = 0; d < DB; ++d) share *= bb[d] - aa[d];
                    rrw[o + k] = vabs(double(r.err[fa.first[fi]])) * share;
                } else if (rk == 3) {
                    // the pdf's integral over the bin's box; dims past the bins span the region
                    const uint32_t rg = fa.first[fi];
                    T an[D], bn[D];
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        an[d] = d < DB ? T(aa[d]) : T(0);
                        bn[d] = d < DB ? T(bb[d]) : T(1);
                    }
                    rrw[o + k] = empty ? 0.0 : double(r.vol[rg] * imp::pdf_fold<T, D, 0>(r.v + size_t(rg) * K, 0, an, bn));
                } else {
                    rrw[o + k] = vabs(integ);
                }
            }
        }
        const uint64_t spp = jb.spp;
        const bool sampling = spp > 0 && n > 0;
        if (spp > 0 && n == 0) atomicOr(&rt[j].flags, JF_EMPTY_BIN);
        // region index -> region: one footprint is one contiguous run of regions; else a
        // binary search over the prefix counts, cached in LDS for the first LANE_MAXF
        // footprints and stored beside the list beyond
        auto region_of = [&](uint32_t idx) -> uint32_t {
            if (nf == 1) return first0 + idx;
            if (idx < s_cum[nc][tid] || nc == nf) {
                uint32_t lo = 0, hi = nc;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) / 2;
                    if (s_cum[mid][tid] <= idx)
                        lo = mid;
                    else
                        hi = mid;
                }
                return s_first[lo][tid] + (idx - s_cum[lo][tid]);
            }
            uint32_t lo = nc, hi = nf;  // lp.pcum[o + nc] = s_cum[nc] <= idx
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (lp.pcum[o + mid] <= idx)
                    lo = mid;
                else
                    hi = mid;
            }
            return lp.pfirst[o + lo] + (idx - lp.pcum[o + lo]);
        };
        const T wn = T(factor * double(n));  // FP32 sample weight factor B * n
        const bool rr_form = RRI && form_roulette(jb.form) != 0;
        double rr_total = 0.0;
        if (rr_form && spp > 0 && n > 0) {  // floor at 1% of the mean weight, then cumulate
            // rr_pdf_region keeps its factor 0.01 in a float
            const double fl = form_roulette(jb.form) == 3 ? double(0.01f) : 0.01;
            double sum = 0.0;
            for (uint32_t k = 0; k < n; ++k) sum += rrw[o + k];
            double cum = 0.0;
            for (uint32_t k = 0; k < n; ++k) {
                const double w = sum <= 0.0 ? 1.0 : vmax(rrw[o + k], fl * sum / double(n));
                cum += w;
                rrw[o + k] = cum;
            }
            rr_total = cum;
        }
        // region_stratification_uniform (region-stratification.h:11-24): sample s of the
        // bin goes to region (start + s) mod n, so region i gets spp / n samples plus one
        // when (i - start) mod n < spp % n -- the reference's counts -- and weight
        // spp / count; start uniform in [0, n) from the bin's own Philox word
        const bool strat = STRAT && form_stratified(jb.form);
        uint32_t st_start = 0;
        uint64_t st_q = 0, st_rem = 0;
        if (strat && sampling) {
            const U4 r0 = philox4x32_10(U4{0xFFFFFFFFu, uint32_t(b), uint32_t(b >> 32) ^ jb.stream, 0x53545254u},
                                        jb.seed_lo, jb.seed_hi);
            st_start = uniform_index(r0.x, n);
            st_q = spp / n;
            st_rem = spp % n;
        }
        const bool imp_form = form_importance(jb.form);
        KahanSum<T> sf, sa, mf, ma, maa, mfa;
        double acc1 = 0.0;  // alpha = 1 form
        const double a1w = spp > 0 ? factor * double(n) / double(spp) : 0.0;
        T k_f = 0, k_a = 0;
        const uint64_t ns_run = sampling ? spp : 0;
        constexpr int SR = samp_stride<D>();
        // samples in groups of 4: one draw per word type serves the group, and the
        // group's partial sums are compensated once
        for (uint64_t s0 = 0; s0 < ns_run; s0 += 4) {
            GroupDraws<T, D, IMP ? 1 : 0> u(uint32_t(s0 >> 2), b, jb.stream, jb.seed_lo, jb.seed_hi);
            T gf = 0, ga = 0, gmf = 0, gma = 0, gmaa = 0, gmfa = 0;
            // sample s0 + k of the bin, drawing from word column kw of the group
            const uint64_t left = ns_run - s0;  // samples of this group that exist: min(left, 4)
            auto sample = [&](int k, int kw) {
                if (uint64_t(k) >= left) return;
                uint32_t pick = 0;
                double rrfactor = double(n);
                const uint32_t uw = u.region_word(kw);
                if (rr_form) {  // first cumulative weight above u * total
                    const double target = double(uw) * 0x1p-32 * rr_total;
                    uint32_t lo = 0, hi = n - 1;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) / 2;
                        if (rrw[o + mid] > target)
                            hi = mid;
                        else
                            lo = mid + 1;
                    }
                    pick = lo;
                    const double wk = rrw[o + pick] - (pick > 0 ? rrw[o + pick - 1] : 0.0);
                    rrfactor = rr_total / wk;
                } else if (strat) {
                    const uint64_t s = s0 + uint64_t(k);
                    const uint64_t sm = s % n;
                    pick = uint32_t((uint64_t(st_start) + sm) % n);
                    rrfactor = double(spp) / double(st_q + (sm < st_rem ? 1u : 0u));
                } else {
                    pick = uniform_index(uw, n);
                }
                const uint32_t reg = region_of(pick);
                const T* rec = r.samp + size_t(reg) * SR;  // lo[D], hi[D], 1/extent[D], coefficients
                T x[D], t[D], ilo[D], ihi[D];
                bool empty = false;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const T rl = rec[d], rh = rec[D + d];
                    if constexpr (FP32) {
                        ilo[d] = __builtin_fmaxf(blo[d], rl);
                        ihi[d] = __builtin_fmaxf(ilo[d], __builtin_fminf(bhi[d], rh));
                    } else {
                        ilo[d] = vmax(blo[d], rl);
                        ihi[d] = vmax(ilo[d], vmin(bhi[d], rh));
                    }
                    if (ilo[d] >= ihi[d]) empty = true;
                    x[d] = mul_add<FP32>(u.coord(kw, d), ihi[d] - ilo[d], ilo[d]);
                    if constexpr (FP32)
                        t[d] = (x[d] - rl) * rec[2 * D + d];
                    else
                        t[d] = pos_in(x[d], rl, rh);
                }
                T vol = box_volume<T, D>(ilo, ihi);
                if constexpr (IMP && D <= imp::MAX_DIM && FP32) {
                    // region_sampling_importance: boxes of volume >= 1e-5 are sampled from
                    // the region polynomial's pdf with weight 1/pdf, smaller ones uniformly;
                    // region_sampling_russian_roulette: every box, importance or uniform by
                    // a roulette on the pdf's integral against the volume
                    const bool srr = form_sampling_rr(jb.form);
                    if (imp_form && !empty && (srr || vol >= T(1.e-5))) {
                        T rlo[D], rhi[D], xi[D], s01[D], wgt;
#pragma unroll
                        for (int d = 0; d < D; ++d) {
                            rlo[d] = rec[d];
                            rhi[d] = rec[D + d];
                            s01[d] = u.coord(kw, d);
                        }
                        if (imp::sample<T, D>(r.v + size_t(reg) * K, rlo, rhi, r.vol[reg], ilo, ihi, s01, xi, wgt,
                                              srr, u01f(u.extra_word(kw, 0)), vol)) {
#pragma unroll
                            for (int d = 0; d < D; ++d) {
                                x[d] = xi[d];
                                t[d] = (x[d] - rlo[d]) * rec[2 * D + d];
                            }
                            vol = wgt;
                        }
                    }
                }
                const T fx = residual_eval<M, D>(f_lds, x, strat);
                const T px = eval_monomials<T, D, FP32>(rec + 3 * D, t);
                if (jb.form == VMX_FORM_CV_ALPHA1) {
                    if (!empty) acc1 += double(fx - px) * (a1w * double(vol));
                    return;
                }
                T fsv, asv;
                if constexpr (FP32) {  // weights in FP32: B * n * vol is O(n), far from the range limits
                    const T w = (rr_form || strat ? T(factor * rrfactor) : wn) * vol;
                    fsv = fx * w;
                    asv = px * w;
                } else {
                    const double w = factor * rrfactor * double(vol);
                    fsv = T(double(fx) * w);
                    asv = T(double(px) * w);
                }
                if (k == 0 && s0 == 0) {  // the bin's first sample shifts the moments
                    k_f = vabs(fsv);
                    k_a = vabs(asv);
                }
                gf += fsv;
                ga += asv;
                const T df = vabs(fsv) - k_f, da = vabs(asv) - k_a;
                gmf += df;
                gma += da;
                gmaa += da * da;
                gmfa += df * da;
            };
            if constexpr (IMP) {  // the importance sampler is large: one copy, words shifted in
#pragma unroll 1
                for (int k = 0; k < 4; ++k) {
                    sample(k, 0);
                    u.shift();
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) sample(k, k);
            }
            sf.add(gf);
            sa.add(ga);
            mf.add(gmf);
            ma.add(gma);
            maa.add(gmaa);
            mfa.add(gmfa);
        }
        if (jb.form == VMX_FORM_CV_ALPHA1) {
            bins[b] = T(approx + acc1);
            return;
        }
        double result;
        const double nsd = double(spp);
        const bool fixed_w = form_fixed(jb.form);
        if (!sampling || (!fixed_w && spp < 2)) {
            result = approx;
        } else {
            const double Sf = double(sf.s) - double(sf.c);
            const double Sa = double(sa.s) - double(sa.c);
            if (fixed_w) {
                result = (Sf - jb.alpha * Sa) / nsd + jb.alpha * approx;
            } else {
                const double Ef = double(mf.s) - double(mf.c);
                const double Ea = double(ma.s) - double(ma.c);
                const double Eaa = double(maa.s) - double(maa.c);
                const double Efa = double(mfa.s) - double(mfa.c);
                const double cov = (Efa - (Ef * Ea) / nsd) / (nsd - 1);
                const double var = (Eaa - (Ea * Ea) / nsd) / (nsd - 1);
                double alpha = 1.0;
                const double v = var > 0.0 ? var : 0.0;
                if (v > 0.0) {
                    const double c = cov > 0.0 ? (cov < v ? cov : v) : 0.0;
                    alpha = c / v;
                }
                result = (Sf - alpha * Sa) / nsd + alpha * approx;
            }
        }
        bins[b] = T(result);
    }
}

}  // namespace vmx

#include "bins_rm.h"

#endif

/*
 * Host orchestration and C ABI of the MI355X per-bin control-variate integrator.
 *
 * Pipeline of a batch of integrals of one (scalar, D, family) (run_batch; one
 * integral is a batch of one):
 *   1. k_roots, then per refinement round k_theta_round + k_expand_round (launched in
 *      chunks; the host checks for remaining candidates between chunks).
 *   2. exact selection of the N-1 split nodes per integral on the device (canonical
 *      order), or, for the reference's heap-array order, libstdc++ push_heap/pop_heap
 *      replayed on the host over the GPU-computed errors
 *      (src/nested/regions-generator-adaptive-heap.h:21-41).
 *   3. k_prepare2 (boxes, volumes, FP64 bin-dim marginals, monomial coefficients),
 *      footprints, per-bin footprint lists (pairs emitted per footprint at scanned
 *      offsets, stable radix sort by bin, boundary offsets).
 *   4. k_bins1: per-bin control variate + residual Monte Carlo, one lane per bin.
 * Host synchronisation: after generation (capacity check) and once to size the bin
 * lists.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <array>
#include <functional>
#include <map>
#include <random>
#include <limits>
#include <type_traits>
#include <typeinfo>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include <cstring>
#include <rocprim/rocprim.hpp>

#include "../../include/viltrum_amd.h"
#include "exact.h"
#include "pipeline.h"

struct vmx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // device memory pool: one arena per purpose, grown on demand
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
    };
    Buf tree, regs, lists, misc, params, bins, bins2;
    Buf nn[VMX_MAX_DIM + 1];                       // FP32 priority-network parameters per dim
    Buf nn_lds[VMX_MAX_DIM + 1];                   // ... and their LDS image (nn::NetW<KP>)
    bool nn_ok[VMX_MAX_DIM + 1] = {};
    Buf tr_x, tr_t, tr_idx, tr_part, tr_loss, tr_state;  // training work space
    hipEvent_t ev[8] = {};
    void* pinned = nullptr;  // host staging for small device-to-host readbacks
    size_t pinned_bytes = 0;
    // one chunk of refinement rounds captured as a graph, reused while its launch
    // arguments (buffers, shapes, stream) stay the same
    struct RoundGraph {
        std::vector<unsigned char> key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    } rounds;
    // launch shapes: the autotuner's choices per workload class, and the shape a tuning
    // trial forces (keys of the classes the trial ran are collected in tuned_keys)
    struct Shape {
        int gen_per_cu = 4, round_chunk = 8, bins_xcd = 0, deepen = 0;
    };
    std::map<std::array<uint64_t, 6>, Shape> tune;
    // node capacity per region (times N) the last call of a workload class needed, so a
    // class whose rounds overflow 8N does not pay a failed attempt on every call
    std::map<std::array<uint64_t, 6>, uint64_t> cap_hint;
    const Shape* trial = nullptr;
    std::vector<std::array<uint64_t, 6>> tuned_keys;
    // last result (for vmx_export_regions): gathered from device memory on demand
    int last_D = 0;
    uint64_t last_n = 0;
    std::function<void(double*, double*, double*, double*, int32_t*)> last_gather;
};

namespace {

struct VmxError : std::runtime_error {
    int code;
    VmxError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(x)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess)                                                                          \
            throw VmxError(VMX_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));            \
    } while (0)

void* ensure(vmx_ctx::Buf& b, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (b.bytes < bytes) {
        if (b.p) HIP_CHECK(hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
        HIP_CHECK(hipMalloc(&b.p, bytes));
        b.bytes = bytes;
    }
    return b.p;
}

void* ensure_pinned(vmx_ctx* ctx, size_t bytes) {
    if (ctx->pinned_bytes < bytes) {
        if (ctx->pinned) HIP_CHECK(hipHostFree(ctx->pinned));
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
        const size_t want = std::max<size_t>(bytes, 1 << 16);
        HIP_CHECK(hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault));
        ctx->pinned_bytes = want;
    }
    return ctx->pinned;
}

/* Carves aligned sub-arrays out of one allocation. */
inline int ilog2(uint64_t x) {
    int r = 0;
    while (x > 1) {
        x >>= 1;
        ++r;
    }
    return r;
}

/* Sub-allocations of one buffer, 256-byte aligned; taking more than the buffer holds
   is a bug in the caller's size computation and throws. */
struct Carve {
    char* base;
    size_t used = 0, cap;
    explicit Carve(const vmx_ctx::Buf& b) : base(static_cast<char*>(b.p)), cap(b.bytes) {}
    template <typename X>
    X* take(size_t n) {
        used = (used + 255) & ~size_t(255);
        if (used + n * sizeof(X) > cap) throw VmxError(VMX_ERR_CAPACITY, "work buffer carved past its size");
        X* r = reinterpret_cast<X*>(base + used);
        used += n * sizeof(X);
        return r;
    }
};
struct Measure {
    size_t used = 0;
    template <typename X>
    void take(size_t n) {
        used = (used + 255) & ~size_t(255);
        used += n * sizeof(X);
    }
};

template <typename P>
struct ParamOf;
template <typename P>
struct F64Of;
template <typename T>
struct F64Of<vmx::Poly<T>> {
    using type = vmx::Poly<double>;
};
template <typename T>
struct F64Of<vmx::Blobs<T>> {
    using type = vmx::Blobs<double>;
};
template <typename T>
struct F64Of<vmx::Lobes5<T>> {
    using type = vmx::Lobes5<double>;
};
template <typename T>
struct ParamOf<vmx::Poly<T>> {
    using type = vmx_poly_params;
};
template <typename T>
struct ParamOf<vmx::Blobs<T>> {
    using type = vmx_blobs_params;
};
template <typename T>
struct ParamOf<vmx::Lobes5<T>> {
    using type = vmx_lobes5_params;
};
template <typename T, typename Inner, int DF, int RD>
struct ParamOf<vmx::Fubini<T, Inner, DF, RD>> {
    using type = vmx_fubini_params;
};

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

/* One (scalar, D, family) batch through the device pipeline (pipeline.h). */
/* A caller-given region list (host memory, the layout vmx_export_regions writes):
   boxes n x D and node values n x 3^D, dim 0 fastest. */
struct GivenRegions {
    uint64_t n;
    const double *lo, *hi, *nodes;
};

struct JobIn {
    const vmx_problem* p;
    const void* params;
    void* bins;  // device pointer
    vmx_stats* stats;
    const GivenRegions* given = nullptr;  // set: skip generation, integrate over these regions
};

/* Training-pair mode of run_batch: stop after generation, report the node count and,
   if alloc hands back device arrays for that many rows, fill them. */
struct PairsReq {
    uint64_t count = 0;
    int kind = vmx::nn::TARGET_ENERGY;
    uint32_t stride = 1;
    std::function<void(uint64_t, uint16_t**, float**)> alloc;
};

constexpr int LEARN_MAX_DIM = 5;  // the fused network fits the expansion kernel's LDS up to here

template <typename T, int D, typename P, typename Math>
void launch_expand(bool learn, const vmx::GenArgs<T>& g, const P* d_f, int par, uint32_t blocks, hipStream_t st) {
    const size_t pre_bytes = size_t(g.J + 1) * sizeof(uint32_t);  // the per-job candidate prefix
    if constexpr (std::is_same<T, float>::value && D <= LEARN_MAX_DIM) {
        if (learn) {
            vmx::k_expand_round<T, D, P, Math, true><<<blocks, vmx::GEN_THREADS, pre_bytes, st>>>(g, d_f, par);
            return;
        }
    }
    vmx::k_expand_round<T, D, P, Math, false><<<blocks, vmx::GEN_THREADS, pre_bytes, st>>>(g, d_f, par);
}

template <typename T>
struct RocprimTemp {
    size_t sort_bytes = 0, scan_bytes = 0;
};

template <typename T, int D, typename P, typename Math>
int run_batch(vmx_ctx* ctx, const std::vector<JobIn>& in, hipStream_t st, PairsReq* pairs_req = nullptr) {
    constexpr int K = vmx::pow3(D);
    const auto t_start = std::chrono::steady_clock::now();
    const uint32_t J = uint32_t(in.size());
    if (J == 0) return VMX_OK;
    if (J > uint32_t(vmx::MAX_JOBS)) throw VmxError(VMX_ERR_INVALID, "at most 4096 integrals per batch");
    const int order = in[0].p->order;
    for (const JobIn& x : in)
        if (x.p->order != order) throw VmxError(VMX_ERR_INVALID, "one region order per batch");
    // a given region list keeps its own order (like the reference order): no merging of
    // regions into shared footprints, so per-bin sums follow the list
    const bool replay = in[0].given != nullptr;
    for (const JobIn& x : in)
        if ((x.given != nullptr) != replay) throw VmxError(VMX_ERR_INVALID, "given regions for all or none");
    if (replay && pairs_req) throw VmxError(VMX_ERR_INVALID, "training pairs need generation");
    // the region roulettes weigh each region of a bin: one footprint per region
    bool any_rr = false, any_rr_err = false;
    for (const JobIn& x : in) {
        any_rr |= vmx::form_roulette(x.p->form) != 0;
        any_rr_err |= vmx::form_roulette(x.p->form) == 2;
    }
    const bool merge = order == VMX_ORDER_CANONICAL && !replay && !any_rr;

    std::vector<vmx::Job<T>> jobs(J);
    std::vector<P> fparams(J);
    uint64_t total_bins = 0, n_reg = 0;
    for (uint32_t j = 0; j < J; ++j) {
        const vmx_problem& p = *in[j].p;
        vmx::Job<T>& jb = jobs[j];
        std::memset(&jb, 0, sizeof(jb));
        for (int d = 0; d < D; ++d) {
            jb.lo[d] = T(p.lo[d]);
            jb.hi[d] = T(p.hi[d]);
        }
        jb.db = p.bin_dims;
        jb.mb = 1;
        jb.nbins = 1;
        for (int i = 0; i < 3; ++i) {
            jb.res[i] = i < jb.db ? p.res[i] : 1;
            jb.drange[i] = i < jb.db ? (jb.hi[i] - jb.lo[i]) / T(p.res[i]) : T(0);
            if (i < jb.db) {
                jb.mb *= 3;
                jb.nbins *= p.res[i];
            }
        }
        jb.bin_base = total_bins;
        total_bins += jb.nbins;
        jb.spp = p.spp;
        jb.bins = in[j].bins;
        jb.alpha = p.alpha;
        jb.seed_lo = uint32_t(p.seed);
        jb.seed_hi = uint32_t(p.seed >> 32);
        jb.stream = 0;  // random streams depend on the seed only, not on the batch position
        jb.N = uint32_t(replay ? in[j].given->n : p.regions);
        jb.reg_base = uint32_t(n_reg);
        n_reg += jb.N;
        jb.form = p.form;
        jb.heur = p.heuristic;
        // training pairs follow each problem's own refiner too: the regions a learned
        // priority generates are where its network must be right (on-policy pairs)
        jb.learn = p.refiner == VMX_REFINER_LEARNED && !replay ? 1 : 0;
        // the size floor of the learned priority costs at most 1/64 of the budget
        jb.lfloor = jb.learn ? std::max(0, (ilog2(p.regions) - 6) / D) : 0;
        for (int d = 0; d < VMX_MAX_DIM; ++d) jb.lres[d] = d < jb.db ? float(std::log2(double(p.res[d]))) : 0.f;
        fparams[j] = P::from(*static_cast<const typename ParamOf<P>::type*>(in[j].params));
    }
    bool learn = false;
    for (const auto& jb : jobs) learn |= jb.learn != 0;
    if (learn) {
        if (vmx::IsFubini<P>::value) throw VmxError(VMX_ERR_UNSUPPORTED, "learned refiner: not for the Fubini form");
        if (!std::is_same<T, float>::value || D > LEARN_MAX_DIM)
            throw VmxError(VMX_ERR_UNSUPPORTED, "learned refiner: FP32 problems of dimension 1..5 only");
        if (!ctx->nn_ok[D]) throw VmxError(VMX_ERR_UNSUPPORTED, "learned refiner: no network for this dimension");
        if (order != VMX_ORDER_CANONICAL)
            throw VmxError(VMX_ERR_INVALID, "learned refiner: canonical region order only");
    }
    if (n_reg > 0x7FFFFFF0ull || total_bins > 0xFFFFFFF0ull) throw VmxError(VMX_ERR_CAPACITY, "batch too large");
    int mb_max = 1;
    for (const auto& jb : jobs) mb_max = std::max(mb_max, int(jb.mb));

    // launch shapes: a tuning trial's, else VMX_* environment overrides, else the
    // autotuner's choice for this workload class, else the defaults
    int dev = ctx->device, n_cu = 0;
    HIP_CHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    const std::array<uint64_t, 6> tune_key = {uint64_t(typeid(P).hash_code()), uint64_t(D), sizeof(T),
                                              uint64_t(learn), uint64_t(ilog2(J)), uint64_t(ilog2(total_bins))};
    vmx_ctx::Shape shape;
    if (ctx->trial) {
        shape = *ctx->trial;
        ctx->tuned_keys.push_back(tune_key);
    } else {
        const auto it = ctx->tune.find(tune_key);
        if (it != ctx->tune.end()) shape = it->second;
        if (const char* e = std::getenv("VMX_GEN_PER_CU")) shape.gen_per_cu = std::max(1, std::atoi(e));
        if (const char* e = std::getenv("VMX_ROUND_CHUNK")) shape.round_chunk = std::max(2, std::atoi(e) & ~1);
        if (const char* e = std::getenv("VMX_BINS_XCD")) shape.bins_xcd = std::atoi(e) != 0;
        if (const char* e = std::getenv("VMX_DEEPEN")) shape.deepen = std::atoi(e) != 0;
    }
    const uint32_t gen_blocks = uint32_t(n_cu) * uint32_t(shape.gen_per_cu);
    const int round_chunk = shape.round_chunk;

    uint32_t rounds = 0;
    uint32_t max_depth[VMX_MAX_DIM] = {};
    uint64_t cap_mult = 8;
    if (const auto h = ctx->cap_hint.find(tune_key); h != ctx->cap_hint.end()) cap_mult = h->second;
    vmx::TreeArrays<T> t{};
    vmx::Job<T>* d_jobs = nullptr;
    vmx::JobRt* d_rt = nullptr;
    P* d_f = nullptr;
    uint32_t *d_hist16 = nullptr, *d_ncand = nullptr, *d_rounds = nullptr;
    T* d_theta = nullptr;
    uint32_t* d_cstart = nullptr;
    uint32_t cap_total = 0;
    std::vector<vmx::JobRt> rt(J);
    // ---------------- 1. generation (retried with more node capacity if needed)
    for (int attempt = 0;; ++attempt) {
        uint64_t cap64 = 0;
        for (auto& jb : jobs) {
            jb.node_base = uint32_t(cap64);
            jb.node_cap = replay ? jb.N : uint32_t(cap_mult * jb.N + 64);
            cap64 += jb.node_cap;
        }
        if (cap64 > 0x7FFFFFF0ull) throw VmxError(VMX_ERR_CAPACITY, "region budget too large for the node arena");
        cap_total = uint32_t(cap64);
        Measure m;
        m.take<vmx::Job<T>>(J);
        m.take<vmx::JobRt>(J);
        m.take<P>(J);
        m.take<T>(size_t(cap_total) * D);
        m.take<T>(size_t(cap_total) * D);
        m.take<T>(size_t(cap_total) * K);
        for (int i = 0; i < 3; ++i) m.take<T>(cap_total);
        for (int i = 0; i < 3; ++i) m.take<int32_t>(cap_total);
        m.take<uint32_t>(size_t(cap_total) * D);
        m.take<uint32_t>(cap_total);
        m.take<uint8_t>(cap_total);
        m.take<uint32_t>(size_t(J) * 65536);
        m.take<uint32_t>(2 * J);
        m.take<uint32_t>(2 * J);
        m.take<T>(J);
        m.take<uint32_t>(16);
        ensure(ctx->tree, m.used);
        Carve c(ctx->tree);
        d_jobs = c.take<vmx::Job<T>>(J);
        d_rt = c.take<vmx::JobRt>(J);
        d_f = c.take<P>(J);
        t.lo = c.take<T>(size_t(cap_total) * D);
        t.hi = c.take<T>(size_t(cap_total) * D);
        t.v = c.take<T>(size_t(cap_total) * K);
        t.err = c.take<T>(cap_total);
        t.b = c.take<T>(cap_total);
        t.b1 = c.take<T>(cap_total);
        t.dim = c.take<int32_t>(cap_total);
        t.parent = c.take<int32_t>(cap_total);
        t.child = c.take<int32_t>(cap_total);
        t.code = c.take<uint32_t>(size_t(cap_total) * D);
        t.job = c.take<uint32_t>(cap_total);
        t.popped = c.take<uint8_t>(cap_total);
        d_hist16 = c.take<uint32_t>(size_t(J) * 65536);
        d_cstart = c.take<uint32_t>(2 * J);
        d_ncand = c.take<uint32_t>(2 * J);
        d_theta = c.take<T>(J);
        uint32_t* small = c.take<uint32_t>(16);  // [0] rounds, [8 + d] deepest split count along dim d
        d_rounds = small;
        HIP_CHECK(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(vmx::Job<T>) * J, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(d_f, fparams.data(), sizeof(P) * J, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemsetAsync(d_rt, 0, sizeof(vmx::JobRt) * J, st));
        HIP_CHECK(hipMemsetAsync(t.job, 0xFF, sizeof(uint32_t) * cap_total, st));
        HIP_CHECK(hipMemsetAsync(d_hist16, 0, sizeof(uint32_t) * J * 65536, st));
        HIP_CHECK(hipMemsetAsync(d_ncand, 0, sizeof(uint32_t) * 2 * J, st));
        HIP_CHECK(hipMemsetAsync(small, 0, sizeof(uint32_t) * 16, st));
        if (replay) {  // the given regions become the nodes, in list order
            for (uint32_t j = 0; j < J; ++j) {
                const GivenRegions& gr = *in[j].given;
                const vmx::Job<T>& jb = jobs[j];
                const size_t n = gr.n;
                std::vector<T> lo(n * D), hi(n * D), v(n * K), err(n, T(0));
                std::vector<uint32_t> job(n, j);
                std::vector<int32_t> dim(n, 0);
                for (size_t i = 0; i < n * D; ++i) {
                    const int d = int(i % D);
                    lo[i] = T(gr.lo[i]);
                    hi[i] = T(gr.hi[i]);
                    if (!(lo[i] < hi[i]) || lo[i] < jb.lo[d] || hi[i] > jb.hi[d])
                        throw VmxError(VMX_ERR_INVALID, "given region outside the range or empty");
                }
                for (size_t i = 0; i < n * K; ++i) {
                    v[i] = T(gr.nodes[i]);
                    if (!std::isfinite(double(v[i]))) throw VmxError(VMX_ERR_INVALID, "non-finite node value");
                }
                const size_t b = jb.node_base;
                HIP_CHECK(hipMemcpyAsync(t.lo + b * D, lo.data(), sizeof(T) * n * D, hipMemcpyHostToDevice, st));
                HIP_CHECK(hipMemcpyAsync(t.hi + b * D, hi.data(), sizeof(T) * n * D, hipMemcpyHostToDevice, st));
                HIP_CHECK(hipMemcpyAsync(t.v + b * K, v.data(), sizeof(T) * n * K, hipMemcpyHostToDevice, st));
                HIP_CHECK(hipMemcpyAsync(t.err + b, err.data(), sizeof(T) * n, hipMemcpyHostToDevice, st));
                HIP_CHECK(hipMemcpyAsync(t.dim + b, dim.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
                HIP_CHECK(hipMemcpyAsync(t.job + b, job.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, st));
                HIP_CHECK(hipStreamSynchronize(st));  // the host vectors go out of scope
                rt[j] = vmx::JobRt{};
                rt[j].count = uint32_t(b + n);
            }
            HIP_CHECK(hipMemcpyAsync(d_rt, rt.data(), sizeof(vmx::JobRt) * J, hipMemcpyHostToDevice, st));
            HIP_CHECK(hipStreamSynchronize(st));
            break;
        }
        vmx::k_roots<T, D, P, Math><<<J, 256, 0, st>>>(d_jobs, d_rt, d_f, t, d_hist16, d_cstart, d_ncand);
        HIP_CHECK(hipGetLastError());
        vmx::GenArgs<T> g{};
        g.jobs = d_jobs;
        g.rt = d_rt;
        g.t = t;
        g.hist16 = d_hist16;
        g.cstart = d_cstart;
        g.ncand = d_ncand;
        g.theta = d_theta;
        g.rounds = d_rounds;
        g.depth = small + 8;
        g.J = J;
        g.net = learn ? ctx->nn_lds[D].p : nullptr;
        g.deepen = shape.deepen;
        // rounds in chunks of round_chunk (even, so each chunk starts at parity 0); after
        // each chunk the candidate counts, job states and round counter are copied to
        // pinned memory.  The host checks chunk k while chunk k+1 runs, so the device is
        // never idle waiting for the host; the chunk after the last busy one finds no
        // candidates and its kernels exit at once.
        const size_t slot_bytes = ((sizeof(uint32_t) * 2 * J + sizeof(vmx::JobRt) * J + 64) + 255) & ~size_t(255);
        char* pin = static_cast<char*>(ensure_pinned(ctx, 2 * slot_bytes));
        auto slot = [&](int k) { return pin + size_t(k & 1) * slot_bytes; };
        // the chunk's 2 * round_chunk kernels as a graph: one launch instead of many
        // small ones (the late rounds are launch-bound); recaptured only when an argument
        // changes.  VMX_NO_GRAPH=1 launches them one by one (e.g. under a kernel tracer
        // that does not follow graph launches).
        static const bool use_graph = !(std::getenv("VMX_NO_GRAPH") && std::atoi(std::getenv("VMX_NO_GRAPH")));
        if (use_graph) {
            std::vector<unsigned char> key(sizeof(g) + 64);
            std::memcpy(key.data(), &g, sizeof(g));
            const void* fn = reinterpret_cast<const void*>(&launch_expand<T, D, P, Math>);
            const void* parts[4] = {fn, d_f, st, reinterpret_cast<const void*>(uintptr_t(learn))};
            std::memcpy(key.data() + sizeof(g), parts, sizeof(parts));
            const uint32_t shape[2] = {gen_blocks, uint32_t(round_chunk)};
            std::memcpy(key.data() + sizeof(g) + sizeof(parts), shape, sizeof(shape));
            if (ctx->rounds.key != key || !ctx->rounds.exec) {
                if (ctx->rounds.exec) HIP_CHECK(hipGraphExecDestroy(ctx->rounds.exec));
                if (ctx->rounds.graph) HIP_CHECK(hipGraphDestroy(ctx->rounds.graph));
                ctx->rounds.exec = nullptr;
                ctx->rounds.graph = nullptr;
                ctx->rounds.key.clear();
                HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
                for (int r = 0; r < round_chunk; ++r) {
                    vmx::k_theta_round<T><<<J, vmx::THETA_THREADS, 0, st>>>(g, r & 1);
                    launch_expand<T, D, P, Math>(learn, g, d_f, r & 1, gen_blocks, st);
                }
                hipGraph_t graph = nullptr;
                const hipError_t e1 = hipStreamEndCapture(st, &graph);
                HIP_CHECK(e1);
                HIP_CHECK(hipGetLastError());
                HIP_CHECK(hipGraphInstantiate(&ctx->rounds.exec, graph, nullptr, nullptr, 0));
                ctx->rounds.graph = graph;
                ctx->rounds.key = key;
            }
        }
        auto launch_chunk = [&](int k) {
            if (use_graph) {
                HIP_CHECK(hipGraphLaunch(ctx->rounds.exec, st));
            } else {
                for (int r = 0; r < round_chunk; ++r) {
                    vmx::k_theta_round<T><<<J, vmx::THETA_THREADS, 0, st>>>(g, r & 1);
                    launch_expand<T, D, P, Math>(learn, g, d_f, r & 1, gen_blocks, st);
                }
                HIP_CHECK(hipGetLastError());
            }
            char* sl = slot(k);
            HIP_CHECK(hipMemcpyAsync(sl, d_ncand, sizeof(uint32_t) * 2 * J, hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipMemcpyAsync(sl + sizeof(uint32_t) * 2 * J, d_rt, sizeof(vmx::JobRt) * J,
                                     hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipMemcpyAsync(sl + sizeof(uint32_t) * 2 * J + sizeof(vmx::JobRt) * J, small, 64,
                                     hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipEventRecord(ctx->ev[2 + (k & 1)], st));
        };
        launch_chunk(0);
        int k = 0;
        for (;; ++k) {
            if ((k + 2) * round_chunk > 4096) {
                HIP_CHECK(hipStreamSynchronize(st));
                throw VmxError(VMX_ERR_CAPACITY, "more than 4096 refinement rounds");
            }
            launch_chunk(k + 1);
            HIP_CHECK(hipEventSynchronize(ctx->ev[2 + (k & 1)]));
            const uint32_t* nc = reinterpret_cast<const uint32_t*>(slot(k));
            bool pending = false;
            for (uint32_t j = 0; j < J; ++j) pending |= nc[j] != 0;
            if (!pending) break;
        }
        std::memcpy(rt.data(), slot(k) + sizeof(uint32_t) * 2 * J, sizeof(vmx::JobRt) * J);
        uint32_t small_h[16];
        std::memcpy(small_h, slot(k) + sizeof(uint32_t) * 2 * J + sizeof(vmx::JobRt) * J, 64);
        rounds = small_h[0];
        for (int d = 0; d < D; ++d) max_depth[d] = small_h[8 + d];
        bool cap_fail = false, too_deep = false;
        for (const auto& r : rt) {
            too_deep |= (r.flags & vmx::JF_DEPTH) != 0;
            cap_fail |= (r.flags & vmx::JF_CAPACITY) != 0;
        }
        if (too_deep || cap_fail) HIP_CHECK(hipStreamSynchronize(st));  // before buffers change hands
        if (too_deep) throw VmxError(VMX_ERR_CAPACITY, "region tree deeper than 31 splits in one dim");
        if (cap_fail) {
            if (attempt >= 3) throw VmxError(VMX_ERR_CAPACITY, "node capacity exceeded");
            cap_mult *= 2;
            continue;
        }
        ctx->cap_hint[tune_key] = cap_mult;
        break;
    }
    const double ms_gen = ms_since(t_start);

    if constexpr (vmx::IsFubini<P>::value) {
        if (pairs_req) throw VmxError(VMX_ERR_UNSUPPORTED, "training pairs: not for the Fubini form");
    } else if (pairs_req) {
        PairsReq* pairs = pairs_req;
        // training pairs: every node of every job, job-major in slot order
        std::vector<uint64_t> base(J);
        uint64_t total = 0;
        const uint32_t stride = std::max<uint32_t>(1, pairs->stride);
        for (uint32_t j = 0; j < J; ++j) {
            base[j] = total;
            total += (rt[j].count - jobs[j].node_base + stride - 1) / stride;
        }
        pairs->count = total;
        uint16_t* X = nullptr;
        float* tg = nullptr;
        HIP_CHECK(hipStreamSynchronize(st));  // the last (idle) round chunk is done
        if (pairs->alloc) pairs->alloc(total, &X, &tg);
        if (X && tg && total > 0) {
            using F64 = typename F64Of<P>::type;
            std::vector<F64> f64(J);
            for (uint32_t j = 0; j < J; ++j)
                f64[j] = F64::from(*static_cast<const typename ParamOf<P>::type*>(in[j].params));
            ensure(ctx->misc, sizeof(uint64_t) * J + sizeof(F64) * J + 1024);
            Carve c(ctx->misc);
            uint64_t* d_base = c.take<uint64_t>(J);
            F64* d_f64 = c.take<F64>(J);
            HIP_CHECK(hipMemcpyAsync(d_base, base.data(), sizeof(uint64_t) * J, hipMemcpyHostToDevice, st));
            HIP_CHECK(hipMemcpyAsync(d_f64, f64.data(), sizeof(F64) * J, hipMemcpyHostToDevice, st));
            vmx::k_train_pairs<T, D, F64><<<(cap_total + 127) / 128, 128, 0, st>>>(d_jobs, t, cap_total, d_base, d_f64,
                                                                                pairs->kind, stride, X, tg);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipStreamSynchronize(st));
        }
        return VMX_OK;
    }

    // ---------------- 2. selection and canonical leaf order
    const auto t_sel = std::chrono::steady_clock::now();
    // canonical order in one sort when (job, code_0, ..., code_{D-1}) fits a 64-bit key:
    // code_d < 2^(max_depth_d + 1); LSD passes over 32-bit keys otherwise
    vmx::Widths<D> widths{};
    int job_bits = 0;
    while (job_bits < 32 && (uint64_t(1) << job_bits) < J) ++job_bits;
    int key_bits64 = job_bits;
    for (int d = 0; d < D; ++d) {
        widths.w[d] = int(max_depth[d]) + 1;
        key_bits64 += widths.w[d];
    }
    const bool one_sort = key_bits64 <= 64;
    size_t sort_bytes = 0, sort64_bytes = 0;
    HIP_CHECK(hipError_t(rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                  (int32_t*)nullptr, (int32_t*)nullptr, size_t(n_reg), 0, 32, st)));
    HIP_CHECK(hipError_t(rocprim::radix_sort_pairs(nullptr, sort64_bytes, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                                  (int32_t*)nullptr, (int32_t*)nullptr, size_t(n_reg), 0,
                                                  std::max(1, std::min(64, key_bits64)), st)));
    sort_bytes = std::max(sort_bytes, sort64_bytes);
    Measure ms;
    ms.take<int32_t>(cap_total);   // bucket
    ms.take<uint8_t>(cap_total);   // bstate
    ms.take<int32_t>(n_reg);       // leaf slots
    ms.take<int32_t>(n_reg);       // sort values tmp
    ms.take<uint64_t>(n_reg);      // sort keys
    ms.take<uint64_t>(n_reg);      // sort keys tmp
    ms.take<char>(sort_bytes + 256);
    ensure(ctx->regs, ms.used);
    Carve cs(ctx->regs);
    int32_t* d_bucket = cs.take<int32_t>(cap_total);
    uint8_t* d_bstate = cs.take<uint8_t>(cap_total);
    int32_t* d_leaf = cs.take<int32_t>(n_reg);
    int32_t* d_leaf_tmp = cs.take<int32_t>(n_reg);
    uint64_t* d_keys64 = cs.take<uint64_t>(n_reg);
    uint64_t* d_keys64_tmp = cs.take<uint64_t>(n_reg);
    uint32_t* d_keys = reinterpret_cast<uint32_t*>(d_keys64);
    uint32_t* d_keys_tmp = reinterpret_cast<uint32_t*>(d_keys64_tmp);
    void* d_sort_tmp = cs.take<char>(sort_bytes + 256);
    const unsigned cap_blocks = (cap_total + 255) / 256;
    const unsigned reg_blocks = unsigned((n_reg + 255) / 256);
    if (replay) {
        std::vector<int32_t> all_leaves(n_reg);
        for (uint32_t j = 0; j < J; ++j)
            for (uint32_t i = 0; i < jobs[j].N; ++i) all_leaves[jobs[j].reg_base + i] = int32_t(jobs[j].node_base + i);
        HIP_CHECK(hipMemcpyAsync(d_leaf, all_leaves.data(), sizeof(int32_t) * n_reg, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipStreamSynchronize(st));
    } else if (order == VMX_ORDER_CANONICAL) {
        HIP_CHECK(hipMemsetAsync(d_leaf, 0xFF, sizeof(int32_t) * n_reg, st));
        vmx::k_sel_bound<T><<<J, 256, 0, st>>>(d_jobs, d_rt, d_hist16);
        vmx::k_sel_gather<T><<<cap_blocks, 256, 0, st>>>(d_jobs, d_rt, t, cap_total, d_bucket);
        vmx::k_sel_resolve<T, D><<<J, 256, 0, st>>>(d_jobs, d_rt, t, d_bucket, d_bstate);
        vmx::k_leaves<T><<<cap_blocks, 256, 0, st>>>(d_jobs, d_rt, t, cap_total, d_leaf);
        HIP_CHECK(hipGetLastError());
        int32_t* vin = d_leaf;
        int32_t* vout = d_leaf_tmp;
        if (one_sort) {
            vmx::k_leaf_keys64<D><<<reg_blocks, 256, 0, st>>>(d_leaf, t.code, t.job, uint32_t(n_reg), widths, d_keys64,
                                                              d_leaf_tmp);
            size_t bytes = sort_bytes;
            HIP_CHECK(hipError_t(rocprim::radix_sort_pairs(d_sort_tmp, bytes, d_keys64, d_keys64_tmp, d_leaf_tmp, d_leaf,
                                                          size_t(n_reg), 0, std::max(1, key_bits64), st)));
        } else {
            // LSD passes: code_{D-1} ... code_0, then job
            for (int which = D - 1; which >= (J > 1 ? -1 : 0); --which) {
                vmx::k_leaf_keys<D><<<reg_blocks, 256, 0, st>>>(vin, t.code, t.job, uint32_t(n_reg), which, d_keys);
                size_t bytes = sort_bytes;
                HIP_CHECK(hipError_t(rocprim::radix_sort_pairs(d_sort_tmp, bytes, d_keys, d_keys_tmp, vin, vout,
                                                              size_t(n_reg), 0, 32, st)));
                std::swap(vin, vout);
            }
        }
        if (vin != d_leaf)
            HIP_CHECK(hipMemcpyAsync(d_leaf, vin, sizeof(int32_t) * n_reg, hipMemcpyDeviceToDevice, st));
        // a job with fewer than N leaves leaves -1 slots behind: k_prepare2 turns them into
        // empty regions and the host rejects the result at its next synchronisation
    } else {
        // the reference's heap-array order: replay push_heap / pop_heap on the errors
        std::vector<int32_t> all_leaves(n_reg);
        for (uint32_t j = 0; j < J; ++j) {
            const auto& jb = jobs[j];
            const uint32_t cnt = rt[j].count - jb.node_base;
            std::vector<T> err(cnt);
            std::vector<int32_t> child(cnt);
            HIP_CHECK(hipMemcpyAsync(err.data(), t.err + jb.node_base, sizeof(T) * cnt, hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipMemcpyAsync(child.data(), t.child + jb.node_base, sizeof(int32_t) * cnt,
                                     hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipStreamSynchronize(st));
            struct E {
                T err;
                int32_t node;
            };
            auto cmp = [](const E& a, const E& b) { return a.err < b.err; };
            std::vector<E> heap;
            heap.reserve(jb.N + 1);
            heap.push_back({err[0], 0});
            for (uint32_t i = 0; i + 1 < jb.N; ++i) {
                const E top = heap.front();
                const int32_t ch = child[top.node];
                if (ch < 0) throw VmxError(VMX_ERR_CAPACITY, "heap replay reached an unexpanded node");
                const int32_t cl = ch - int32_t(jb.node_base);
                std::pop_heap(heap.begin(), heap.end(), cmp);
                heap.pop_back();
                heap.push_back({err[cl], cl});
                std::push_heap(heap.begin(), heap.end(), cmp);
                heap.push_back({err[cl + 1], cl + 1});
                std::push_heap(heap.begin(), heap.end(), cmp);
            }
            for (uint32_t i = 0; i < jb.N; ++i) all_leaves[jb.reg_base + i] = int32_t(jb.node_base) + heap[i].node;
        }
        HIP_CHECK(hipMemcpyAsync(d_leaf, all_leaves.data(), sizeof(int32_t) * n_reg, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipStreamSynchronize(st));
    }
    const double ms_sel = ms_since(t_sel);

    // ---------------- 3. regions, footprints, bin lists
    const auto t_str = std::chrono::steady_clock::now();
    size_t scan_reg_bytes = 0;
    HIP_CHECK(hipError_t(rocprim::exclusive_scan(nullptr, scan_reg_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, 0u,
                                                size_t(n_reg + 1), rocprim::plus<uint32_t>(), st)));

    Measure mr;
    mr.take<T>(n_reg * D);
    mr.take<T>(n_reg * D);
    mr.take<T>(n_reg);
    mr.take<T>(n_reg * K);
    bool any_importance = false, keep_nodes = false;  // node values per region: importance sampling, rr_pdf_region
    for (const auto& jb : jobs) {
        any_importance |= vmx::form_importance(jb.form);
        keep_nodes |= vmx::form_roulette(jb.form) == 3;
    }
    keep_nodes |= any_importance;
    if (keep_nodes) mr.take<T>(n_reg * K);
    if (any_rr_err) mr.take<T>(n_reg);
    mr.take<double>(n_reg * mb_max);
    mr.take<uint32_t>(n_reg);
    mr.take<uint32_t>(n_reg + 1);
    mr.take<uint32_t>(n_reg + 1);
    mr.take<T>(n_reg * 3);
    mr.take<T>(n_reg * 3);
    mr.take<double>(n_reg);
    mr.take<double>(n_reg * mb_max);
    for (int i = 0; i < 3; ++i) mr.take<uint32_t>(n_reg);
    mr.take<uint32_t>(n_reg * 6);
    mr.take<uint32_t>(n_reg + 1);
    mr.take<uint32_t>(n_reg + 1);
    mr.take<uint32_t>(total_bins + 1);
    mr.take<char>(scan_reg_bytes + 256);
    ensure(ctx->lists, mr.used);
    Carve cr(ctx->lists);
    vmx::RegionArrays<T> ra{};
    ra.lo = cr.take<T>(n_reg * D);
    ra.hi = cr.take<T>(n_reg * D);
    ra.vol = cr.take<T>(n_reg);
    ra.coef = cr.take<T>(n_reg * K);
    ra.v = keep_nodes ? cr.take<T>(n_reg * K) : nullptr;
    ra.err = any_rr_err ? cr.take<T>(n_reg) : nullptr;
    ra.marg = cr.take<double>(n_reg * mb_max);
    ra.job = cr.take<uint32_t>(n_reg);
    ra.fpflag = cr.take<uint32_t>(n_reg + 1);
    ra.fpid = cr.take<uint32_t>(n_reg + 1);
    vmx::FootArrays<T> fa{};
    fa.lo = cr.take<T>(n_reg * 3);
    fa.hi = cr.take<T>(n_reg * 3);
    fa.scale = cr.take<double>(n_reg);
    fa.marg = cr.take<double>(n_reg * mb_max);
    fa.first = cr.take<uint32_t>(n_reg);
    fa.count = cr.take<uint32_t>(n_reg);
    fa.job = cr.take<uint32_t>(n_reg);
    fa.box = cr.take<uint32_t>(n_reg * 6);
    fa.npairs = cr.take<uint32_t>(n_reg + 1);
    uint32_t* d_pofs = cr.take<uint32_t>(n_reg + 1);
    uint32_t* d_off = cr.take<uint32_t>(total_bins + 1);
    void* d_scan_tmp = cr.take<char>(scan_reg_bytes + 256);
    // mb differs per job only through db; all jobs of a batch share the footprint stride
    for (const auto& jb : jobs)
        if (jb.mb != mb_max) throw VmxError(VMX_ERR_INVALID, "one bin_dims per batch");
    vmx::k_prepare2<T, D><<<reg_blocks, 256, 0, st>>>(d_jobs, d_leaf, t, ra, uint32_t(n_reg), mb_max, merge);
    HIP_CHECK(hipGetLastError());
    size_t bytes = scan_reg_bytes;
    HIP_CHECK(hipError_t(rocprim::exclusive_scan(d_scan_tmp, bytes, ra.fpflag, ra.fpid, 0u, size_t(n_reg + 1),
                                                rocprim::plus<uint32_t>(), st)));
    HIP_CHECK(hipMemsetAsync(fa.npairs, 0, sizeof(uint32_t) * (n_reg + 1), st));
    vmx::k_footprints<T, D><<<reg_blocks, 256, 0, st>>>(d_jobs, ra, fa, uint32_t(n_reg), mb_max, merge);
    HIP_CHECK(hipGetLastError());
    const uint32_t* d_nfp = ra.fpid + n_reg;
    // per-bin footprint lists: pairs emitted per footprint at scanned offsets, then a
    // stable radix sort by bin (each list ends up in ascending footprint order)
    bytes = scan_reg_bytes;
    HIP_CHECK(hipError_t(rocprim::exclusive_scan(d_scan_tmp, bytes, fa.npairs, d_pofs, 0u, size_t(n_reg + 1),
                                                rocprim::plus<uint32_t>(), st)));
    uint32_t pairs = 0, nfp = 0;
    {
        char* pin = static_cast<char*>(ensure_pinned(ctx, 16 + sizeof(vmx::JobRt) * J));
        HIP_CHECK(hipMemcpyAsync(pin, d_pofs + n_reg, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(pin + 4, d_nfp, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(pin + 16, d_rt, sizeof(vmx::JobRt) * J, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        std::memcpy(&pairs, pin, 4);
        std::memcpy(&nfp, pin + 4, 4);
        std::memcpy(rt.data(), pin + 16, sizeof(vmx::JobRt) * J);
    }
    if (order == VMX_ORDER_CANONICAL && !replay)
        for (uint32_t j = 0; j < J; ++j)
            if ((rt[j].flags & vmx::JF_LEAVES) || rt[j].nleaves != jobs[j].N)
                throw VmxError(VMX_ERR_CAPACITY, "region selection did not close");
    // per-bin footprint lists by counting sort: count, scan, place, then every list in
    // ascending footprint id (region order)
    size_t scan_bins_bytes = 0;
    HIP_CHECK(hipError_t(rocprim::exclusive_scan(nullptr, scan_bins_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, 0u,
                                                size_t(total_bins + 1), rocprim::plus<uint32_t>(), st)));
    Measure mp;
    mp.take<uint32_t>(total_bins + 1);                        // counts
    mp.take<uint32_t>(size_t(pairs) + 1);                     // lists
    mp.take<uint32_t>(size_t(pairs) + 1);                     // copies of very long lists
    mp.take<uint32_t>(size_t(pairs) / vmx::SORT_SHORT + 2);   // queue of long lists
    mp.take<uint32_t>(size_t(pairs) + 1);                     // prefix counts of long lists
    mp.take<uint32_t>(size_t(pairs) + 1);                     // first regions of long lists
    if (any_rr) mp.take<double>(size_t(pairs) + 1);          // roulette weights per pair
    mp.take<char>(scan_bins_bytes + 256);
    ensure(ctx->bins2, mp.used);
    Carve cp(ctx->bins2);
    uint32_t* d_cnt = cp.take<uint32_t>(total_bins + 1);
    uint32_t* d_lst = cp.take<uint32_t>(size_t(pairs) + 1);
    uint32_t* d_lst_tmp = cp.take<uint32_t>(size_t(pairs) + 1);
    uint32_t* d_qlong = cp.take<uint32_t>(size_t(pairs) / vmx::SORT_SHORT + 2);
    vmx::ListPrefix lp{cp.take<uint32_t>(size_t(pairs) + 1), cp.take<uint32_t>(size_t(pairs) + 1)};
    double* d_rrw = any_rr ? cp.take<double>(size_t(pairs) + 1) : nullptr;
    void* d_scan_bins_tmp = cp.take<char>(scan_bins_bytes + 256);
    HIP_CHECK(hipMemsetAsync(d_cnt, 0, sizeof(uint32_t) * (total_bins + 1), st));
    HIP_CHECK(hipMemsetAsync(d_qlong, 0, sizeof(uint32_t), st));
    const unsigned fp_blocks = unsigned((uint64_t(nfp) * 64 + 255) / 256);
    if (pairs > 0)
        vmx::k_bin_pass<T, false><<<fp_blocks, 256, 0, st>>>(d_jobs, fa, d_nfp, d_cnt, nullptr, nullptr);
    bytes = scan_bins_bytes;
    HIP_CHECK(hipError_t(rocprim::exclusive_scan(d_scan_bins_tmp, bytes, d_cnt, d_off, 0u, size_t(total_bins + 1),
                                                rocprim::plus<uint32_t>(), st)));
    if (pairs > 0) {
        vmx::k_bin_pass<T, true><<<fp_blocks, 256, 0, st>>>(d_jobs, fa, d_nfp, d_cnt, d_off, d_lst);
        vmx::k_sort_lists<T><<<unsigned((total_bins + 255) / 256), 256, 0, st>>>(d_off, total_bins, d_lst, d_qlong,
                                                                                 fa, lp);
        vmx::k_sort_long<<<1024, 256, 0, st>>>(d_off, d_qlong, d_lst, d_lst_tmp);
        vmx::k_list_prefix<T><<<1024, 256, 0, st>>>(d_off, d_qlong, d_lst, fa, lp);
    }
    HIP_CHECK(hipGetLastError());
    const double ms_str = ms_since(t_str);

    // ---------------- 4. bins
    // one lane per bin, one block per 256 consecutive bins
    const unsigned bins_grid = unsigned((total_bins + 255) / 256);
    HIP_CHECK(hipEventRecord(ctx->ev[0], st));
    const int db = jobs[0].db;
    // the importance-sampling variant is its own instantiation: its register footprint
    // stays out of the uniform kernel
    auto bins_kernel = [&](auto mode_tag) {
        constexpr int MODE = decltype(mode_tag)::value;
        if (db == 1)
            vmx::k_bins1<T, D, 1, P, Math, MODE><<<bins_grid, 256, 0, st>>>(
                d_jobs, J, total_bins, ra, fa, d_off, d_lst, d_f, d_rt, lp, shape.bins_xcd, d_rrw);
        else if constexpr (D >= 2) {
            if (db == 2)
                vmx::k_bins1<T, D, 2, P, Math, MODE><<<bins_grid, 256, 0, st>>>(
                    d_jobs, J, total_bins, ra, fa, d_off, d_lst, d_f, d_rt, lp, shape.bins_xcd, d_rrw);
            else if constexpr (D >= 3)
                vmx::k_bins1<T, D, 3, P, Math, MODE><<<bins_grid, 256, 0, st>>>(
                    d_jobs, J, total_bins, ra, fa, d_off, d_lst, d_f, d_rt, lp, shape.bins_xcd, d_rrw);
        }
    };
    if constexpr (vmx::IsFubini<P>::value) {
        bins_kernel(std::integral_constant<int, 0>{});
    } else {
        bool done = false;
        if constexpr (std::is_same<T, float>::value && D <= vmx::imp::MAX_DIM) {
            if (any_importance) {
                if (any_rr)
                    bins_kernel(std::integral_constant<int, 3>{});
                else
                    bins_kernel(std::integral_constant<int, 1>{});
                done = true;
            }
        }
        if (!done) {
            if (any_rr)
                bins_kernel(std::integral_constant<int, 2>{});
            else
                bins_kernel(std::integral_constant<int, 0>{});
        }
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(ctx->ev[1], st));
    {
        char* pin = static_cast<char*>(ensure_pinned(ctx, sizeof(vmx::JobRt) * J));
        HIP_CHECK(hipMemcpyAsync(pin, d_rt, sizeof(vmx::JobRt) * J, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        std::memcpy(rt.data(), pin, sizeof(vmx::JobRt) * J);
    }
    float ms_k = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms_k, ctx->ev[0], ctx->ev[1]));
    for (uint32_t j = 0; j < J; ++j)
        if (rt[j].flags & vmx::JF_EMPTY_BIN) throw VmxError(VMX_ERR_NUMERIC, "a bin overlaps no region while spp > 0");

    // region export, gathered lazily from this call's device buffers
    ctx->last_D = D;
    ctx->last_n = n_reg;
    ctx->last_gather = [ctx, t, ra, d_leaf, n_reg](double* olo, double* ohi, double* onodes, double* oerr,
                                                   int32_t* odim) {
        const size_t N = n_reg;
        const size_t total = N * (2 * D + K + 1);
        ensure(ctx->misc, sizeof(T) * total + sizeof(int32_t) * N + 4096);
        Carve c(ctx->misc);
        T* tmp = c.take<T>(total);
        int32_t* tdim = c.take<int32_t>(N);
        const unsigned g1 = unsigned((N * K + 255) / 256);
        vmx::k_gather<T><<<g1, 256, 0, ctx->stream>>>(t.v, d_leaf, uint32_t(N), K, tmp);
        vmx::k_gather<T><<<unsigned((N + 255) / 256), 256, 0, ctx->stream>>>(t.err, d_leaf, uint32_t(N), 1,
                                                                          tmp + N * (2 * D + K));
        vmx::k_gather<int32_t><<<unsigned((N + 255) / 256), 256, 0, ctx->stream>>>(t.dim, d_leaf, uint32_t(N), 1,
                                                                                tdim);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(tmp + N * K, ra.lo, sizeof(T) * N * D, hipMemcpyDeviceToDevice, ctx->stream));
        HIP_CHECK(hipMemcpyAsync(tmp + N * K + N * D, ra.hi, sizeof(T) * N * D, hipMemcpyDeviceToDevice, ctx->stream));
        std::vector<T> h(total);
        std::vector<int32_t> hd(N);
        HIP_CHECK(hipMemcpyAsync(h.data(), tmp, sizeof(T) * total, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipMemcpyAsync(hd.data(), tdim, sizeof(int32_t) * N, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (size_t i = 0; i < N * K; ++i)
            if (onodes) onodes[i] = double(h[i]);
        for (size_t i = 0; i < N * D; ++i) {
            if (olo) olo[i] = double(h[N * K + i]);
            if (ohi) ohi[i] = double(h[N * K + N * D + i]);
        }
        for (size_t i = 0; i < N; ++i) {
            if (oerr) oerr[i] = double(h[N * (2 * D + K) + i]);
            if (odim) odim[i] = hd[i];
        }
    };
    const double ms_total = ms_since(t_start);
    for (uint32_t j = 0; j < J; ++j) {
        vmx_stats* s = in[j].stats;
        if (!s) continue;
        const auto& jb = jobs[j];
        s->regions = jb.N;
        s->nodes = rt[j].count - jb.node_base;
        s->rounds = rounds;
        s->pairs = pairs;
        s->footprints = nfp;
        s->evals_regions = uint64_t(K) + (s->nodes - 1) * uint64_t(vmx::pow3(D - 1));
        if constexpr (vmx::IsFubini<P>::value) s->evals_regions *= fparams[j].mc;  // mc samples per node value
        s->evals_residual = jb.nbins * jb.spp;
        s->ties_unresolved = rt[j].ties;
        s->ms_generate = ms_gen;
        s->ms_select = ms_sel;
        s->ms_stratify = ms_str;
        s->ms_bins = ms_k;
        s->ms_total = ms_total;
    }
    return VMX_OK;
}

template <typename T, typename Math>
int dispatch_T(vmx_ctx* ctx, int family, int D, const std::vector<JobIn>& jobs, hipStream_t st,
               PairsReq* pairs = nullptr) {
    switch (family) {
        case VMX_FAMILY_POLY:
            switch (D) {
                case 1: return run_batch<T, 1, vmx::Poly<T>, Math>(ctx, jobs, st, pairs);
                case 2: return run_batch<T, 2, vmx::Poly<T>, Math>(ctx, jobs, st, pairs);
                case 3: return run_batch<T, 3, vmx::Poly<T>, Math>(ctx, jobs, st, pairs);
                case 4: return run_batch<T, 4, vmx::Poly<T>, Math>(ctx, jobs, st, pairs);
                case 5: return run_batch<T, 5, vmx::Poly<T>, Math>(ctx, jobs, st, pairs);
                default: break;
            }
            break;
        case VMX_FAMILY_BLOBS:
            if (D == 2) return run_batch<T, 2, vmx::Blobs<T>, Math>(ctx, jobs, st, pairs);
            if (D == 3) return run_batch<T, 3, vmx::Blobs<T>, Math>(ctx, jobs, st, pairs);
            break;
        case VMX_FAMILY_LOBES5:
            if (D == 5) return run_batch<T, 5, vmx::Lobes5<T>, Math>(ctx, jobs, st, pairs);
            break;
        case VMX_FAMILY_FUBINI:
            // integrator_crespo2021_infinite is float-only in the reference; built here for
            // (region dims, inner family, inner dims) = (2, blobs, 3), (2, lobes5, 5),
            // (2, poly, 3), (1, poly, 2)
            if constexpr (std::is_same<T, float>::value) {
                const auto* fp = static_cast<const vmx_fubini_params*>(jobs[0].params);
                const int fam = fp->inner_family, dd = fp->inner_dim;
                if (D == 2 && fam == VMX_FAMILY_BLOBS && dd == 3)
                    return run_batch<T, 2, vmx::Fubini<T, vmx::Blobs<T>, 3, 2>, Math>(ctx, jobs, st, pairs);
                if (D == 2 && fam == VMX_FAMILY_LOBES5 && dd == 5)
                    return run_batch<T, 2, vmx::Fubini<T, vmx::Lobes5<T>, 5, 2>, Math>(ctx, jobs, st, pairs);
                if (D == 2 && fam == VMX_FAMILY_POLY && dd == 3)
                    return run_batch<T, 2, vmx::Fubini<T, vmx::Poly<T>, 3, 2>, Math>(ctx, jobs, st, pairs);
                if (D == 1 && fam == VMX_FAMILY_POLY && dd == 2)
                    return run_batch<T, 1, vmx::Fubini<T, vmx::Poly<T>, 2, 1>, Math>(ctx, jobs, st, pairs);
            } else {
                throw VmxError(VMX_ERR_UNSUPPORTED, "Fubini form: FP32 only (as the reference)");
            }
            break;
        default: break;
    }
    throw VmxError(VMX_ERR_UNSUPPORTED, "integrand family / dimension not built");
}

/* Groups a batch by (family, dim) and runs each group through the device pipeline. */
int integrate_groups(vmx_ctx* ctx, uint32_t n, const vmx_problem* const* ps, const vmx_integrand* const* fs,
                     void* const* bins, vmx_stats* const* stats, hipStream_t st,
                     const GivenRegions* const* given = nullptr) {
    const int scalar = ps[0]->scalar;
    std::vector<int> done(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        if (ps[i]->scalar != scalar) throw VmxError(VMX_ERR_INVALID, "one scalar type per batch");
        if (stats[i]) std::memset(stats[i], 0, sizeof(vmx_stats));
    }
    for (uint32_t i = 0; i < n; ++i) {
        if (done[i]) continue;
        std::vector<JobIn> group;
        auto inner = [](const vmx_integrand* f) {
            if (f->family != VMX_FAMILY_FUBINI) return std::pair<int, int>(0, 0);
            const auto* fp = static_cast<const vmx_fubini_params*>(f->params);
            return std::pair<int, int>(fp->inner_family, fp->inner_dim);
        };
        for (uint32_t k = i; k < n; ++k)
            if (!done[k] && fs[k]->family == fs[i]->family && fs[k]->dim == fs[i]->dim && inner(fs[k]) == inner(fs[i])) {
                group.push_back({ps[k], fs[k]->params, bins[k], stats[k], given ? given[k] : nullptr});
                done[k] = 1;
            }
        if (scalar == VMX_F32)
            dispatch_T<float, vmx::FastMath>(ctx, fs[i]->family, fs[i]->dim, group, st);
        else
            dispatch_T<double, vmx::PreciseMath>(ctx, fs[i]->family, fs[i]->dim, group, st);
    }
    return VMX_OK;
}

void validate(const vmx_problem* p, const vmx_integrand* f) {
    if (!p || !f || !f->params) throw VmxError(VMX_ERR_INVALID, "null problem or integrand");
    if (p->scalar != VMX_F32 && p->scalar != VMX_F64) throw VmxError(VMX_ERR_INVALID, "scalar must be 4 or 8");
    if (f->dim < 1 || f->dim > VMX_MAX_DIM) throw VmxError(VMX_ERR_INVALID, "dim must be 1..6");
    if (p->bin_dims < 1 || p->bin_dims > 3 || p->bin_dims > f->dim)
        throw VmxError(VMX_ERR_INVALID, "bin_dims must be 1..min(3, dim)");
    for (int i = 0; i < p->bin_dims; ++i)
        if (p->res[i] < 1 || p->res[i] > (1u << 20)) throw VmxError(VMX_ERR_INVALID, "resolution must be 1..2^20");
    uint64_t nb = 1;
    for (int i = 0; i < p->bin_dims; ++i) nb *= p->res[i];
    if (nb > (1ull << 31)) throw VmxError(VMX_ERR_INVALID, "too many bins");
    if (p->regions < 1 || p->regions > (1ull << 26)) throw VmxError(VMX_ERR_INVALID, "regions must be 1..2^26");
    if (p->heuristic != VMX_HEUR_SIZE_RELATIVE && p->heuristic != VMX_HEUR_DEFAULT_ABSOLUTE)
        throw VmxError(VMX_ERR_INVALID, "unknown heuristic");
    if (p->form < VMX_FORM_VR_OPTIMIZE || p->form > VMX_FORM_VR_FIXED_RR_PDF_IMPORTANCE)
        throw VmxError(VMX_ERR_INVALID, "unknown form");
    if (p->form >= VMX_FORM_VR_OPTIMIZE_RR_INTEGRAL && f->family == VMX_FAMILY_FUBINI)
        throw VmxError(VMX_ERR_UNSUPPORTED, "region roulettes: not for the Fubini form");
    if (vmx::form_importance(p->form) &&
        (p->scalar != VMX_F32 || f->dim > vmx::imp::MAX_DIM || f->family == VMX_FAMILY_FUBINI))
        throw VmxError(VMX_ERR_UNSUPPORTED, "importance sampling: FP32 integrands of dimension 1..3");
    if (p->order != VMX_ORDER_CANONICAL && p->order != VMX_ORDER_REFERENCE)
        throw VmxError(VMX_ERR_INVALID, "unknown region order");
    if (p->refiner != VMX_REFINER_HEURISTIC && p->refiner != VMX_REFINER_LEARNED)
        throw VmxError(VMX_ERR_INVALID, "unknown refiner");
    for (int d = 0; d < f->dim; ++d)
        if (!(p->hi[d] > p->lo[d])) throw VmxError(VMX_ERR_INVALID, "range must have hi > lo");
    if (f->family == VMX_FAMILY_FUBINI) {
        const auto* fp = static_cast<const vmx_fubini_params*>(f->params);
        if (!fp->inner) throw VmxError(VMX_ERR_INVALID, "Fubini form: null inner parameters");
        if (fp->inner_dim <= f->dim || fp->inner_dim > 5 || f->dim > 3)
            throw VmxError(VMX_ERR_INVALID, "Fubini form: 1..3 region dims below the inner integrand's 2..5");
        if (fp->mc_samples < 1 || fp->mc_samples >= 0x80000000u)
            throw VmxError(VMX_ERR_INVALID, "Fubini form: mc_samples must be 1..2^31-1");
        for (int d = f->dim; d < fp->inner_dim; ++d)
            if (!(fp->rest_hi[d] > fp->rest_lo[d])) throw VmxError(VMX_ERR_INVALID, "Fubini form: rest range must have hi > lo");
        if (p->scalar != VMX_F32) throw VmxError(VMX_ERR_UNSUPPORTED, "Fubini form: FP32 only (as the reference)");
        if (p->refiner != VMX_REFINER_HEURISTIC)
            throw VmxError(VMX_ERR_UNSUPPORTED, "Fubini form: heuristic refinement only");
    }
}

int guarded(vmx_ctx* ctx, const std::function<int()>& body) {
    if (!ctx) return VMX_ERR_INVALID;
    try {
        HIP_CHECK(hipSetDevice(ctx->device));
        ctx->err.clear();
        return body();
    } catch (const VmxError& e) {
        ctx->err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        ctx->err = "host allocation failed";
        return VMX_ERR_CAPACITY;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return VMX_ERR_INVALID;
    }
}

namespace {
template <typename T, typename Math>
int eval_T(vmx_ctx* ctx, const vmx_integrand& f, uint64_t n, const double* x, double* out) {
    const int D = f.dim;
    const size_t fbytes = 8192;
    ensure(ctx->misc, fbytes + sizeof(double) * n * (D + 1) + 1024);
    Carve c(ctx->misc);
    void* d_f = c.take<char>(fbytes);
    double* d_x = c.take<double>(n * D);
    double* d_o = c.take<double>(n);
    HIP_CHECK(hipMemcpyAsync(d_x, x, sizeof(double) * n * D, hipMemcpyHostToDevice, ctx->stream));
    const unsigned blocks = unsigned((n + 255) / 256);
    auto launch = [&](auto tag, auto dtag) {
        using P = typename decltype(tag)::type;
        constexpr int DD = decltype(dtag)::value;
        const P fp = P::from(*static_cast<const typename decltype(tag)::param*>(f.params));
        HIP_CHECK(hipMemcpyAsync(d_f, &fp, sizeof(P), hipMemcpyHostToDevice, ctx->stream));
        vmx::k_eval<T, DD, P, Math><<<blocks, 256, 0, ctx->stream>>>(static_cast<const P*>(d_f), n, d_x, d_o);
        HIP_CHECK(hipGetLastError());
    };
    struct PolyTag {
        using type = vmx::Poly<T>;
        using param = vmx_poly_params;
    };
    struct BlobsTag {
        using type = vmx::Blobs<T>;
        using param = vmx_blobs_params;
    };
    struct LobesTag {
        using type = vmx::Lobes5<T>;
        using param = vmx_lobes5_params;
    };
    bool ok = true;
    if (f.family == VMX_FAMILY_POLY) {
        switch (D) {
            case 1: launch(PolyTag{}, std::integral_constant<int, 1>{}); break;
            case 2: launch(PolyTag{}, std::integral_constant<int, 2>{}); break;
            case 3: launch(PolyTag{}, std::integral_constant<int, 3>{}); break;
            case 4: launch(PolyTag{}, std::integral_constant<int, 4>{}); break;
            case 5: launch(PolyTag{}, std::integral_constant<int, 5>{}); break;
            default: ok = false;
        }
    } else if (f.family == VMX_FAMILY_BLOBS && D == 2) {
        launch(BlobsTag{}, std::integral_constant<int, 2>{});
    } else if (f.family == VMX_FAMILY_BLOBS && D == 3) {
        launch(BlobsTag{}, std::integral_constant<int, 3>{});
    } else if (f.family == VMX_FAMILY_LOBES5 && D == 5) {
        launch(LobesTag{}, std::integral_constant<int, 5>{});
    } else {
        ok = false;
    }
    if (!ok) throw VmxError(VMX_ERR_UNSUPPORTED, "integrand family / dimension not built");
    HIP_CHECK(hipMemcpyAsync(out, d_o, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return VMX_OK;
}
}  // namespace

/* ------------------------------------------------------------------ priority network */
int kp_of(int dim) {
    switch (dim) {
        case 1: return vmx::nn::Dims<1>::KP;
        case 2: return vmx::nn::Dims<2>::KP;
        case 3: return vmx::nn::Dims<3>::KP;
        case 4: return vmx::nn::Dims<4>::KP;
        case 5: return vmx::nn::Dims<5>::KP;
        default: throw VmxError(VMX_ERR_UNSUPPORTED, "learned refiner: dimension 1..5");
    }
}
int nf_of(int dim) {
    switch (dim) {
        case 1: return vmx::nn::Dims<1>::F;
        case 2: return vmx::nn::Dims<2>::F;
        case 3: return vmx::nn::Dims<3>::F;
        case 4: return vmx::nn::Dims<4>::F;
        case 5: return vmx::nn::Dims<5>::F;
        default: throw VmxError(VMX_ERR_UNSUPPORTED, "learned refiner: dimension 1..5");
    }
}
uint64_t n_trained(int kp) { return uint64_t(kp) * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1; }
uint64_t n_total(int kp) { return n_trained(kp) + 2; }

template <typename Fn>
void with_kp(int kp, Fn&& fn) {
    if (kp == 32)
        fn(std::integral_constant<int, 32>{});
    else if (kp == 64)
        fn(std::integral_constant<int, 64>{});
    else if (kp == 96)
        fn(std::integral_constant<int, 96>{});
    else if (kp == 128)
        fn(std::integral_constant<int, 128>{});
    else if (kp == 288)
        fn(std::integral_constant<int, 288>{});
    else
        throw VmxError(VMX_ERR_UNSUPPORTED, "learned refiner: unsupported feature width");
}

/* Builds ctx->nn_lds[dim] from the FP32 parameters in ctx->nn[dim]. */
void pack_net(vmx_ctx* ctx, int dim) {
    const int kp = kp_of(dim);
    with_kp(kp, [&](auto K) {
        constexpr int KP = decltype(K)::value;
        auto* out = static_cast<vmx::nn::NetW<KP>*>(ensure(ctx->nn_lds[dim], sizeof(vmx::nn::NetW<KP>)));
        vmx::nn::k_pack_net<KP><<<1, 256, 0, ctx->stream>>>(static_cast<const float*>(ctx->nn[dim].p), out);
    });
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void nn_forward(int kp, const float* params, const uint16_t* X, uint64_t n, float* y, hipStream_t st) {
    with_kp(kp, [&](auto K) {
        constexpr int KP = decltype(K)::value;
        const unsigned blocks = unsigned(std::min<uint64_t>((n + 63) / 64, 8192));
        vmx::nn::k_mlp_forward<KP><<<blocks, 256, 0, st>>>(params, X, n, y);
    });
    HIP_CHECK(hipGetLastError());
}

void nn_grad(int kp, const float* params, const uint16_t* X, const float* tg, const uint32_t* idx, uint32_t count,
             float inv_batch, float* part, float* loss, int want_grad, hipStream_t st) {
    with_kp(kp, [&](auto K) {
        constexpr int KP = decltype(K)::value;
        const unsigned blocks = (count + 63) / 64;
        vmx::nn::k_mlp_grad<KP><<<blocks, 256, 0, st>>>(params, X, tg, idx, count, inv_batch, part, loss, want_grad);
    });
    HIP_CHECK(hipGetLastError());
}

/* Runs generation for n problems of one family and dim and fills training pairs into
   ctx->tr_x / ctx->tr_t (device).  Returns the number of pairs. */
uint64_t make_pairs(vmx_ctx* ctx, uint32_t n, const vmx_problem* problems, const vmx_integrand* integrands,
                    bool fill, const vmx_train_opts& o) {
    if (o.target < 0 || o.target > vmx::nn::TARGET_GAIN)
        throw VmxError(VMX_ERR_INVALID, "unknown training target");
    if (n == 0) throw VmxError(VMX_ERR_INVALID, "no problems");
    if (!problems || !integrands) throw VmxError(VMX_ERR_INVALID, "null argument");
    std::vector<JobIn> group;
    for (uint32_t i = 0; i < n; ++i) {
        validate(&problems[i], &integrands[i]);
        if (integrands[i].family != integrands[0].family || integrands[i].dim != integrands[0].dim ||
            problems[i].scalar != problems[0].scalar)
            throw VmxError(VMX_ERR_INVALID, "training pairs: one family, dim and scalar per call");
        if (problems[i].order != VMX_ORDER_CANONICAL) throw VmxError(VMX_ERR_INVALID, "canonical order only");
        group.push_back({&problems[i], integrands[i].params, nullptr, nullptr});
    }
    const int dim = integrands[0].dim;
    const int kp = kp_of(dim);
    PairsReq req;
    req.kind = o.target;
    req.stride = o.stride ? o.stride : 1;
    if (fill)
        req.alloc = [ctx, kp](uint64_t rows, uint16_t** X, float** tg) {
            *X = static_cast<uint16_t*>(ensure(ctx->tr_x, sizeof(uint16_t) * rows * kp));
            *tg = static_cast<float*>(ensure(ctx->tr_t, sizeof(float) * rows));
        };
    if (problems[0].scalar == VMX_F32)
        dispatch_T<float, vmx::FastMath>(ctx, integrands[0].family, dim, group, ctx->stream, &req);
    else
        dispatch_T<double, vmx::PreciseMath>(ctx, integrands[0].family, dim, group, ctx->stream, &req);
    return req.count;
}

float uniform_pm(std::mt19937_64& rng, float a) { return a * (2.f * float(double(rng() >> 11) * 0x1p-53) - 1.f); }

/* Trains the priority network for `dim` on the `count` pairs in ctx->tr_x / tr_t. */
void train_net(vmx_ctx* ctx, int dim, uint64_t count, const vmx_train_opts& o, vmx_train_report* rep) {
    const int kp = kp_of(dim), F = nf_of(dim);
    const uint64_t P = n_trained(kp), PT = n_total(kp);
    hipStream_t st = ctx->stream;
    const uint32_t epochs = o.epochs ? o.epochs : 8;
    const uint32_t batch = o.batch ? o.batch : 1024;
    if (batch % 64 != 0 || batch > 65536) throw VmxError(VMX_ERR_INVALID, "batch must be a multiple of 64, <= 65536");
    const float lr0 = o.lr > 0 ? o.lr : 2e-3f;
    const double holdout = (o.holdout > 0 && o.holdout < 0.5) ? o.holdout : 0.05;
    if (count < 256 || count > 0xFFFFFFF0ull) throw VmxError(VMX_ERR_INVALID, "need 256 .. 2^32 training pairs");
    const uint16_t* X = static_cast<const uint16_t*>(ctx->tr_x.p);
    const float* tg = static_cast<const float*>(ctx->tr_t.p);

    std::vector<float> th(count);
    HIP_CHECK(hipMemcpyAsync(th.data(), tg, sizeof(float) * count, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    std::mt19937_64 rng(o.seed);
    // the device writes pairs in the order it claimed node slots, which varies between
    // calls; the seeded shuffle starts from the rows sorted by a hash of their content,
    // so the trained network depends on the pair set and the seed only
    std::vector<uint32_t> perm(count);
    {
        size_t sort_bytes = 0;
        HIP_CHECK(hipError_t(rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                                       (uint32_t*)nullptr, (uint32_t*)nullptr, size_t(count), 0, 64,
                                                       st)));
        Measure ms;
        ms.take<uint64_t>(count);
        ms.take<uint64_t>(count);
        ms.take<uint32_t>(count);
        ms.take<uint32_t>(count);
        ms.take<char>(sort_bytes);
        ensure(ctx->tr_idx, ms.used);
        Carve cs(ctx->tr_idx);
        uint64_t* k0 = cs.take<uint64_t>(count);
        uint64_t* k1 = cs.take<uint64_t>(count);
        uint32_t* i0 = cs.take<uint32_t>(count);
        uint32_t* i1 = cs.take<uint32_t>(count);
        void* tmp = cs.take<char>(sort_bytes);
        vmx::nn::k_row_hash<<<unsigned((count + 255) / 256), 256, 0, st>>>(X, tg, count, kp, k0, i0);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipError_t(rocprim::radix_sort_pairs(tmp, sort_bytes, k0, k1, i0, i1, size_t(count), 0, 64, st)));
        HIP_CHECK(hipMemcpyAsync(perm.data(), i1, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
    }
    std::shuffle(perm.begin(), perm.end(), rng);
    const uint64_t n_valid = std::max<uint64_t>(64, uint64_t(double(count) * holdout));
    const uint64_t n_train = count - n_valid;
    double sm = 0, sq = 0;
    for (uint64_t i = 0; i < n_train; ++i) sm += th[perm[i]];
    const double mean = sm / double(n_train);
    for (uint64_t i = 0; i < n_train; ++i) sq += (th[perm[i]] - mean) * (th[perm[i]] - mean);
    const double sd = std::max(1e-3, std::sqrt(sq / double(n_train)));

    // He-uniform initialisation; zero weights on the padding inputs
    using namespace vmx::nn;
    std::vector<float> p(PT, 0.f);
    const uint64_t o_b0 = uint64_t(kp) * 64, o_w1 = o_b0 + 64, o_b1 = o_w1 + 4096, o_w2 = o_b1 + 64,
                   o_b2 = o_w2 + 4096, o_w3 = o_b2 + 64, o_b3 = o_w3 + 64;
    for (int k = 0; k < F; ++k)
        for (int j = 0; j < 64; ++j) p[uint64_t(k) * 64 + j] = uniform_pm(rng, std::sqrt(6.f / float(F)));
    for (int i = 0; i < 4096; ++i) {
        p[o_w1 + i] = uniform_pm(rng, std::sqrt(6.f / 64.f));
        p[o_w2 + i] = uniform_pm(rng, std::sqrt(6.f / 64.f));
    }
    for (int i = 0; i < 64; ++i) p[o_w3 + i] = uniform_pm(rng, std::sqrt(3.f / 64.f));
    (void)o_b3;
    p[P] = float(mean);
    p[P + 1] = float(sd);

    const uint32_t nblk = batch / 64;
    const uint64_t steps = (n_train + batch - 1) / batch;
    const uint64_t vblk = (n_valid + 63) / 64;
    Measure m;
    m.take<float>(PT);             // params
    m.take<float>(P);              // m
    m.take<float>(P);              // v
    m.take<uint32_t>(count);       // idx: train then valid
    ensure(ctx->tr_state, m.used);
    Carve c(ctx->tr_state);
    float* d_p = c.take<float>(PT);
    float* d_m = c.take<float>(P);
    float* d_v = c.take<float>(P);
    uint32_t* d_idx = c.take<uint32_t>(count);
    float* d_part = static_cast<float*>(ensure(ctx->tr_part, sizeof(float) * nblk * P));
    const uint64_t nloss = std::max<uint64_t>(steps * nblk, vblk);
    float* d_loss = static_cast<float*>(ensure(ctx->tr_loss, sizeof(float) * nloss));
    HIP_CHECK(hipMemcpyAsync(d_p, p.data(), sizeof(float) * PT, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemsetAsync(d_m, 0, sizeof(float) * P, st));
    HIP_CHECK(hipMemsetAsync(d_v, 0, sizeof(float) * P, st));
    HIP_CHECK(hipMemcpyAsync(d_idx + n_train, perm.data() + n_train, sizeof(uint32_t) * n_valid,
                             hipMemcpyHostToDevice, st));

    const auto t0 = std::chrono::steady_clock::now();
    const double total_steps = double(steps) * epochs;
    const float b1 = 0.9f, b2 = 0.999f;
    uint64_t t = 0;
    std::vector<float> lh(steps * nblk);
    double first = 0, last = 0;
    std::vector<uint32_t> tr(perm.begin(), perm.begin() + n_train);
    for (uint32_t e = 0; e < epochs; ++e) {
        std::shuffle(tr.begin(), tr.end(), rng);
        HIP_CHECK(hipMemcpyAsync(d_idx, tr.data(), sizeof(uint32_t) * n_train, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemsetAsync(d_loss, 0, sizeof(float) * steps * nblk, st));
        for (uint64_t s = 0; s < steps; ++s, ++t) {
            const uint64_t off = s * batch;
            const uint32_t cnt = uint32_t(std::min<uint64_t>(batch, n_train - off));
            nn_grad(kp, d_p, X, tg, d_idx + off, cnt, 1.f / float(cnt), d_part, d_loss + s * nblk, 1, st);
            const double frac = double(t) / total_steps;
            const float lr = float(lr0 * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.141592653589793 * frac))));
            const float c1 = float(1.0 / (1.0 - std::pow(double(b1), double(t + 1))));
            const float c2 = float(1.0 / (1.0 - std::pow(double(b2), double(t + 1))));
            k_adam<<<unsigned((P + 255) / 256), 256, 0, st>>>(d_p, d_m, d_v, d_part, (cnt + 63) / 64, uint32_t(P), lr,
                                                              b1, b2, 1e-8f, c1, c2);
            HIP_CHECK(hipGetLastError());
        }
        HIP_CHECK(hipMemcpyAsync(lh.data(), d_loss, sizeof(float) * steps * nblk, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        double sum = 0;
        for (float x : lh) sum += x;
        const double mean_loss = sum / double(steps);
        if (e == 0) first = mean_loss;
        last = mean_loss;
    }
    const double ms_train = ms_since(t0);
    // held-out loss
    nn_grad(kp, d_p, X, tg, d_idx + n_train, uint32_t(n_valid), 1.f / float(n_valid), d_part, d_loss, 0, st);
    std::vector<float> vl(vblk);
    HIP_CHECK(hipMemcpyAsync(vl.data(), d_loss, sizeof(float) * vblk, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    double vsum = 0;
    for (float x : vl) vsum += x;
    // keep the parameters
    float* keep = static_cast<float*>(ensure(ctx->nn[dim], sizeof(float) * PT));
    HIP_CHECK(hipMemcpyAsync(keep, d_p, sizeof(float) * PT, hipMemcpyDeviceToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
    pack_net(ctx, dim);
    ctx->nn_ok[dim] = true;
    if (rep) {
        rep->steps = t;
        rep->loss_first = first;
        rep->loss_train = last;
        rep->loss_valid = vsum;
        rep->target_mean = mean;
        rep->target_std = sd;
        rep->ms_train = ms_train;
    }
}

template <typename T, typename Math, int D, typename P>
int mc_run(vmx_ctx* ctx, const vmx_problem& p, const void* fparams, int stratified, void* bins, double* stderr_out) {
    const int db = p.bin_dims;
    uint64_t nb = 1;
    for (int i = 0; i < db; ++i) nb *= p.res[i];
    const P fp = P::from(*static_cast<const typename ParamOf<P>::type*>(fparams));
    vmx::BinArgs<T> ba{};
    double vtot = 1;
    for (int d = 0; d < D; ++d) {
        ba.lo[d] = T(p.lo[d]);
        ba.hi[d] = T(p.hi[d]);
        vtot *= p.hi[d] - p.lo[d];
    }
    for (int i = 0; i < 3; ++i) {
        ba.res[i] = i < db ? p.res[i] : 1;
        ba.drange[i] = i < db ? (ba.hi[i] - ba.lo[i]) / T(p.res[i]) : T(0);
    }
    ba.nbins = nb;
    ba.seed_lo = uint32_t(p.seed);
    ba.seed_hi = uint32_t(p.seed >> 32);
    ba.db = db;
    uint32_t grid = 1;
    if (stratified && db >= 2) {
        grid = uint32_t(std::sqrt(double(p.spp)));
        while (grid > 1 && p.spp % (uint64_t(grid) * grid) != 0) --grid;
    }
    ensure(ctx->misc, sizeof(P) + sizeof(double) * nb * 2 + 4096);
    Carve c(ctx->misc);
    P* d_f = c.take<P>(1);
    double* d_s = c.take<double>(nb);
    double* d_s2 = c.take<double>(nb);
    hipStream_t st = ctx->stream;
    HIP_CHECK(hipMemcpyAsync(d_f, &fp, sizeof(P), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemsetAsync(d_s, 0, sizeof(double) * nb, st));
    HIP_CHECK(hipMemsetAsync(d_s2, 0, sizeof(double) * nb, st));
    const unsigned blocks = unsigned((nb + 255) / 256);
    const uint64_t chunk = 64;
    for (uint64_t s0 = 0; s0 < p.spp; s0 += chunk) {
        const uint64_t cnt = std::min(chunk, p.spp - s0);
        vmx::k_mc<T, D, P, Math><<<blocks, 256, 0, st>>>(ba, d_f, s0, cnt, grid, d_s, d_s2);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(st));
    }
    std::vector<double> s(nb), s2(nb);
    HIP_CHECK(hipMemcpyAsync(s.data(), d_s, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(s2.data(), d_s2, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    const double n = double(p.spp);
    for (uint64_t b = 0; b < nb; ++b) {
        const double mean = s[b] / n;
        const double var = n > 1 ? std::max(0.0, (s2[b] - n * mean * mean) / (n - 1)) : 0.0;
        const double val = vtot * mean;
        if (sizeof(T) == 4)
            static_cast<float*>(bins)[b] = float(val);
        else
            static_cast<double*>(bins)[b] = val;
        if (stderr_out) stderr_out[b] = vtot * std::sqrt(var / n);
    }
    return VMX_OK;
}

template <typename T, typename Math>
int mc_dispatch(vmx_ctx* ctx, const vmx_problem& p, const vmx_integrand& f, int stratified, void* bins, double* se) {
    switch (f.family) {
        case VMX_FAMILY_POLY:
            switch (f.dim) {
                case 1: return mc_run<T, Math, 1, vmx::Poly<T>>(ctx, p, f.params, stratified, bins, se);
                case 2: return mc_run<T, Math, 2, vmx::Poly<T>>(ctx, p, f.params, stratified, bins, se);
                case 3: return mc_run<T, Math, 3, vmx::Poly<T>>(ctx, p, f.params, stratified, bins, se);
                case 4: return mc_run<T, Math, 4, vmx::Poly<T>>(ctx, p, f.params, stratified, bins, se);
                case 5: return mc_run<T, Math, 5, vmx::Poly<T>>(ctx, p, f.params, stratified, bins, se);
                default: break;
            }
            break;
        case VMX_FAMILY_BLOBS:
            if (f.dim == 2) return mc_run<T, Math, 2, vmx::Blobs<T>>(ctx, p, f.params, stratified, bins, se);
            if (f.dim == 3) return mc_run<T, Math, 3, vmx::Blobs<T>>(ctx, p, f.params, stratified, bins, se);
            break;
        case VMX_FAMILY_LOBES5:
            if (f.dim == 5) return mc_run<T, Math, 5, vmx::Lobes5<T>>(ctx, p, f.params, stratified, bins, se);
            break;
        default: break;
    }
    throw VmxError(VMX_ERR_UNSUPPORTED, "integrand family / dimension not built");
}

}  // namespace

namespace {
/* device copies of host network inputs, in ctx->misc */
struct NetIn {
    float* p;
    uint16_t* X;
    float* t;
    float* y;
    uint32_t* idx;
};
NetIn upload_net_inputs(vmx_ctx* ctx, int dim, const float* params, uint64_t n, const uint16_t* X, const float* t) {
    const int kp = kp_of(dim);
    const uint64_t PT = n_total(kp);
    Measure m;
    m.take<float>(PT);
    m.take<uint16_t>(n * kp);
    m.take<float>(n);
    m.take<float>(n);
    m.take<uint32_t>(n);
    ensure(ctx->misc, m.used);
    Carve c(ctx->misc);
    NetIn r{};
    r.p = c.take<float>(PT);
    r.X = c.take<uint16_t>(n * kp);
    r.t = c.take<float>(n);
    r.y = c.take<float>(n);
    r.idx = c.take<uint32_t>(n);
    if (params) {
        HIP_CHECK(hipMemcpyAsync(r.p, params, sizeof(float) * PT, hipMemcpyHostToDevice, ctx->stream));
    } else {
        if (!ctx->nn_ok[dim]) throw VmxError(VMX_ERR_INVALID, "no network for this dimension");
        HIP_CHECK(hipMemcpyAsync(r.p, ctx->nn[dim].p, sizeof(float) * PT, hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIP_CHECK(hipMemcpyAsync(r.X, X, sizeof(uint16_t) * n * kp, hipMemcpyHostToDevice, ctx->stream));
    if (t) HIP_CHECK(hipMemcpyAsync(r.t, t, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    std::vector<uint32_t> id(n);
    for (uint64_t i = 0; i < n; ++i) id[i] = uint32_t(i);
    HIP_CHECK(hipMemcpyAsync(r.idx, id.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return r;
}

}  // namespace

extern "C" {

int vmx_integrate_mc(vmx_ctx* ctx, const vmx_problem* p, const vmx_integrand* f, int32_t stratified, void* bins,
                     double* stderr_out) {
    return guarded(ctx, [&]() -> int {
        validate(p, f);
        if (!bins) throw VmxError(VMX_ERR_INVALID, "null bins");
        if (p->spp < 1) throw VmxError(VMX_ERR_INVALID, "spp must be >= 1");
        if (p->scalar == VMX_F32) return mc_dispatch<float, vmx::FastMath>(ctx, *p, *f, stratified, bins, stderr_out);
        return mc_dispatch<double, vmx::PreciseMath>(ctx, *p, *f, stratified, bins, stderr_out);
    });
}


int vmx_ctx_create(int device, vmx_ctx** out) {
    if (!out) return VMX_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return VMX_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return VMX_ERR_DEVICE;
    vmx_ctx* c = new (std::nothrow) vmx_ctx();
    if (!c) return VMX_ERR_CAPACITY;
    c->device = device;
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return VMX_ERR_DEVICE;
    }
    for (hipEvent_t& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return VMX_ERR_DEVICE;
        }
    *out = c;
    return VMX_OK;
}

void vmx_ctx_destroy(vmx_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    for (vmx_ctx::Buf* b : {&ctx->tree, &ctx->regs, &ctx->lists, &ctx->misc, &ctx->params, &ctx->bins, &ctx->bins2,
                            &ctx->tr_x, &ctx->tr_t, &ctx->tr_idx, &ctx->tr_part, &ctx->tr_loss, &ctx->tr_state})
        if (b->p) (void)hipFree(b->p);
    for (vmx_ctx::Buf& b : ctx->nn)
        if (b.p) (void)hipFree(b.p);
    for (vmx_ctx::Buf& b : ctx->nn_lds)
        if (b.p) (void)hipFree(b.p);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->rounds.exec) (void)hipGraphExecDestroy(ctx->rounds.exec);
    if (ctx->rounds.graph) (void)hipGraphDestroy(ctx->rounds.graph);
    for (hipEvent_t e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* vmx_last_error(const vmx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int vmx_integrate_device(vmx_ctx* ctx, const vmx_problem* p, const vmx_integrand* f, void* bins_device, void* stream,
                         vmx_stats* stats) {
    return guarded(ctx, [&]() -> int {
        validate(p, f);
        if (!bins_device) throw VmxError(VMX_ERR_INVALID, "null bins");
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
        return integrate_groups(ctx, 1, &p, &f, &bins_device, &stats, st);
    });
}

int vmx_integrate_batch(vmx_ctx* ctx, uint32_t n, const vmx_problem* problems, const vmx_integrand* integrands,
                        void* const* bins_device, void* stream, vmx_stats* stats) {
    return guarded(ctx, [&]() -> int {
        if (n == 0) return int(VMX_OK);
        if (!problems || !integrands || !bins_device) throw VmxError(VMX_ERR_INVALID, "null argument");
        std::vector<const vmx_problem*> ps(n);
        std::vector<const vmx_integrand*> fs(n);
        std::vector<vmx_stats*> ss(n);
        for (uint32_t i = 0; i < n; ++i) {
            validate(&problems[i], &integrands[i]);
            if (!bins_device[i]) throw VmxError(VMX_ERR_INVALID, "null bins");
            ps[i] = &problems[i];
            fs[i] = &integrands[i];
            ss[i] = stats ? &stats[i] : nullptr;
        }
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
        return integrate_groups(ctx, n, ps.data(), fs.data(), bins_device, ss.data(), st);
    });
}

int vmx_autotune(vmx_ctx* ctx, uint32_t n, const vmx_problem* problems, const vmx_integrand* integrands,
                 void* const* bins_device, void* stream, vmx_tune_report* report) {
    return guarded(ctx, [&]() -> int {
        if (n == 0) return int(VMX_OK);
        if (!problems || !integrands || !bins_device) throw VmxError(VMX_ERR_INVALID, "null argument");
        std::vector<const vmx_problem*> ps(n);
        std::vector<const vmx_integrand*> fs(n);
        std::vector<vmx_stats*> ss(n, nullptr);
        for (uint32_t i = 0; i < n; ++i) {
            validate(&problems[i], &integrands[i]);
            if (!bins_device[i]) throw VmxError(VMX_ERR_INVALID, "null bins");
            ps[i] = &problems[i];
            fs[i] = &integrands[i];
        }
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
        // one call's wall time at a shape: the faster of two after a warm-up (the first
        // call at a new shape also captures its round graph)
        auto time_shape = [&](const vmx_ctx::Shape& sh) {
            ctx->trial = &sh;
            double best = 1e300;
            for (int r = 0; r < 3; ++r) {
                HIP_CHECK(hipStreamSynchronize(st));
                const auto t0 = std::chrono::steady_clock::now();
                integrate_groups(ctx, n, ps.data(), fs.data(), bins_device, ss.data(), st);
                HIP_CHECK(hipStreamSynchronize(st));
                if (r > 0) best = std::min(best, ms_since(t0));
            }
            ctx->trial = nullptr;
            return best;
        };
        struct Reset {
            vmx_ctx* c;
            ~Reset() { c->trial = nullptr; }
        } reset{ctx};
        ctx->tuned_keys.clear();
        const vmx_ctx::Shape def;
        const double ms_def = time_shape(def);
        vmx_ctx::Shape best = def;
        double ms_best = ms_def;
        int trials = 1;
        for (int g : {2, 4, 8})
            for (int rc : {4, 8})
                for (int x : {0, 1})
                    for (int dp : {0, 1}) {
                        vmx_ctx::Shape sh;
                        sh.gen_per_cu = g;
                        sh.round_chunk = rc;
                        sh.bins_xcd = x;
                        sh.deepen = dp;
                        if (g == def.gen_per_cu && rc == def.round_chunk && x == def.bins_xcd && dp == def.deepen)
                            continue;
                        const double ms = time_shape(sh);
                        ++trials;
                        if (ms < ms_best) {
                            ms_best = ms;
                            best = sh;
                        }
                    }
        for (const auto& k : ctx->tuned_keys) ctx->tune[k] = best;
        ctx->tuned_keys.clear();
        // leave the chosen shape's results in the bins
        integrate_groups(ctx, n, ps.data(), fs.data(), bins_device, ss.data(), st);
        HIP_CHECK(hipStreamSynchronize(st));
        if (report) {
            report->gen_blocks_per_cu = best.gen_per_cu;
            report->round_chunk = best.round_chunk;
            report->bins_xcd = best.bins_xcd;
            report->deepen = best.deepen;
            report->trials = trials;
            report->ms_best = ms_best;
            report->ms_default = ms_def;
        }
        return int(VMX_OK);
    });
}

int vmx_integrate(vmx_ctx* ctx, const vmx_problem* p, const vmx_integrand* f, void* bins, vmx_stats* stats) {
    return guarded(ctx, [&]() -> int {
        validate(p, f);
        if (!bins) throw VmxError(VMX_ERR_INVALID, "null bins");
        uint64_t nb = 1;
        for (int i = 0; i < p->bin_dims; ++i) nb *= p->res[i];
        const size_t bytes = nb * size_t(p->scalar);
        void* d = ensure(ctx->bins, bytes);
        HIP_CHECK(hipMemcpyAsync(d, bins, bytes, hipMemcpyHostToDevice, ctx->stream));
        int rc = integrate_groups(ctx, 1, &p, &f, &d, &stats, ctx->stream);
        HIP_CHECK(hipMemcpyAsync(bins, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return rc;
    });
}

int vmx_integrate_regions(vmx_ctx* ctx, const vmx_problem* p, const vmx_integrand* f, uint64_t n, const double* lo,
                          const double* hi, const double* nodes, void* bins, vmx_stats* stats) {
    return guarded(ctx, [&]() -> int {
        if (!p) throw VmxError(VMX_ERR_INVALID, "null problem");
        if (!bins || !lo || !hi || !nodes) throw VmxError(VMX_ERR_INVALID, "null bins or region arrays");
        if (n < 1 || n > (1ull << 26)) throw VmxError(VMX_ERR_INVALID, "1..2^26 regions");
        vmx_problem q = *p;
        q.regions = n;
        q.refiner = VMX_REFINER_HEURISTIC;
        validate(&q, f);
        const GivenRegions gr{n, lo, hi, nodes};
        const GivenRegions* grp = &gr;
        uint64_t nb = 1;
        for (int i = 0; i < q.bin_dims; ++i) nb *= q.res[i];
        const size_t bytes = nb * size_t(q.scalar);
        void* d = ensure(ctx->bins, bytes);
        HIP_CHECK(hipMemcpyAsync(d, bins, bytes, hipMemcpyHostToDevice, ctx->stream));
        const vmx_problem* qp = &q;
        int rc = integrate_groups(ctx, 1, &qp, &f, &d, &stats, ctx->stream, &grp);
        HIP_CHECK(hipMemcpyAsync(bins, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return rc;
    });
}

int vmx_export_regions(vmx_ctx* ctx, uint64_t* n, double* lo, double* hi, double* nodes, double* err,
                       int32_t* dim) {
    return guarded(ctx, [&]() -> int {
        if (!n) throw VmxError(VMX_ERR_INVALID, "null n");
        const uint64_t N = ctx->last_n;
        *n = N;
        if ((lo || hi || nodes || err || dim) && ctx->last_gather) ctx->last_gather(lo, hi, nodes, err, dim);
        return VMX_OK;
    });
}


int vmx_refiner_size(int32_t dim, uint64_t* n_params, int32_t* n_features, int32_t* row_width) {
    try {
        const int kp = kp_of(dim);
        if (n_params) *n_params = n_total(kp);
        if (n_features) *n_features = nf_of(dim);
        if (row_width) *row_width = kp;
        return VMX_OK;
    } catch (const VmxError& e) {
        return e.code;
    }
}

int vmx_refiner_pairs(vmx_ctx* ctx, uint32_t n, const vmx_problem* problems, const vmx_integrand* integrands,
                      const vmx_train_opts* opts, uint64_t cap, uint16_t* features, float* targets, uint64_t* count) {
    return guarded(ctx, [&]() -> int {
        if (!count) throw VmxError(VMX_ERR_INVALID, "null count");
        vmx_train_opts o{};
        if (opts) o = *opts;
        const bool fill = features && targets;
        const uint64_t rows = make_pairs(ctx, n, problems, integrands, fill, o);
        *count = rows;
        if (!fill) return int(VMX_OK);
        const uint64_t out = rows < cap ? rows : cap;
        const int kp = kp_of(integrands[0].dim);
        HIP_CHECK(hipMemcpyAsync(features, ctx->tr_x.p, sizeof(uint16_t) * out * kp, hipMemcpyDeviceToHost,
                                 ctx->stream));
        HIP_CHECK(hipMemcpyAsync(targets, ctx->tr_t.p, sizeof(float) * out, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return int(VMX_OK);
    });
}

int vmx_refiner_train(vmx_ctx* ctx, uint32_t n, const vmx_problem* problems, const vmx_integrand* integrands,
                      const vmx_train_opts* opts, vmx_train_report* report) {
    return guarded(ctx, [&]() -> int {
        vmx_train_opts o{};
        if (opts) o = *opts;
        if (report) std::memset(report, 0, sizeof(*report));
        const auto t0 = std::chrono::steady_clock::now();
        const uint64_t rows = make_pairs(ctx, n, problems, integrands, true, o);
        const double ms_pairs = ms_since(t0);
        train_net(ctx, integrands[0].dim, rows, o, report);
        if (report) {
            report->pairs = rows;
            report->ms_pairs = ms_pairs;
        }
        return int(VMX_OK);
    });
}

int vmx_refiner_get(vmx_ctx* ctx, int32_t dim, float* params) {
    return guarded(ctx, [&]() -> int {
        const int kp = kp_of(dim);
        if (!params) throw VmxError(VMX_ERR_INVALID, "null params");
        if (!ctx->nn_ok[dim]) throw VmxError(VMX_ERR_INVALID, "no network for this dimension");
        HIP_CHECK(hipMemcpy(params, ctx->nn[dim].p, sizeof(float) * n_total(kp), hipMemcpyDeviceToHost));
        return int(VMX_OK);
    });
}

int vmx_refiner_set(vmx_ctx* ctx, int32_t dim, const float* params) {
    return guarded(ctx, [&]() -> int {
        const int kp = kp_of(dim);
        if (!params) throw VmxError(VMX_ERR_INVALID, "null params");
        const uint64_t PT = n_total(kp);
        for (uint64_t i = 0; i < PT; ++i)
            if (!std::isfinite(params[i])) throw VmxError(VMX_ERR_INVALID, "non-finite parameter");
        if (!(params[PT - 1] > 0.f)) throw VmxError(VMX_ERR_INVALID, "target scale must be > 0");
        float* d = static_cast<float*>(ensure(ctx->nn[dim], sizeof(float) * PT));
        HIP_CHECK(hipMemcpy(d, params, sizeof(float) * PT, hipMemcpyHostToDevice));
        pack_net(ctx, dim);
        ctx->nn_ok[dim] = true;
        return int(VMX_OK);
    });
}

int vmx_refiner_forward(vmx_ctx* ctx, int32_t dim, const float* params, uint64_t n, const uint16_t* features,
                        float* y) {
    return guarded(ctx, [&]() -> int {
        if (!features || !y) throw VmxError(VMX_ERR_INVALID, "null argument");
        if (n == 0) return int(VMX_OK);
        if (n > (1ull << 26)) throw VmxError(VMX_ERR_INVALID, "at most 2^26 rows");
        const NetIn in = upload_net_inputs(ctx, dim, params, n, features, nullptr);
        nn_forward(kp_of(dim), in.p, in.X, n, in.y, ctx->stream);
        HIP_CHECK(hipMemcpyAsync(y, in.y, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return int(VMX_OK);
    });
}

int vmx_refiner_grad(vmx_ctx* ctx, int32_t dim, const float* params, uint64_t n, const uint16_t* features,
                     const float* targets, float* grad, double* loss) {
    return guarded(ctx, [&]() -> int {
        if (!params || !features || !targets || !grad) throw VmxError(VMX_ERR_INVALID, "null argument");
        if (n == 0 || n > (1ull << 20)) throw VmxError(VMX_ERR_INVALID, "1 .. 2^20 rows");
        const int kp = kp_of(dim);
        const uint64_t P = n_trained(kp);
        if (!(params[P + 1] > 0.f)) throw VmxError(VMX_ERR_INVALID, "target scale must be > 0");
        const NetIn in = upload_net_inputs(ctx, dim, params, n, features, targets);
        const uint32_t nblk = uint32_t((n + 63) / 64);
        float* d_part = static_cast<float*>(ensure(ctx->tr_part, sizeof(float) * nblk * P));
        float* d_loss = static_cast<float*>(ensure(ctx->tr_loss, sizeof(float) * nblk));
        float* d_grad = static_cast<float*>(ensure(ctx->params, sizeof(float) * P));
        nn_grad(kp, in.p, in.X, in.t, in.idx, uint32_t(n), 1.f / float(n), d_part, d_loss, 1, ctx->stream);
        vmx::nn::k_sum_parts<<<unsigned((P + 255) / 256), 256, 0, ctx->stream>>>(d_part, nblk, uint32_t(P), d_grad);
        HIP_CHECK(hipGetLastError());
        std::vector<float> lh(nblk);
        HIP_CHECK(hipMemcpyAsync(grad, d_grad, sizeof(float) * P, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipMemcpyAsync(lh.data(), d_loss, sizeof(float) * nblk, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (loss) {
            double sm = 0;
            for (float x : lh) sm += x;
            *loss = sm;
        }
        return int(VMX_OK);
    });
}

int vmx_exact_bins(vmx_ctx* ctx, const vmx_problem* p, const vmx_integrand* f, double* bins) {
    return guarded(ctx, [&]() -> int {
        validate(p, f);
        if (!bins) throw VmxError(VMX_ERR_INVALID, "null bins");
        if (f->family != VMX_FAMILY_BLOBS || f->dim != 2 || p->bin_dims != 2)
            throw VmxError(VMX_ERR_UNSUPPORTED, "closed form: 2-D blob integrands with 2 bin dims only");
        const auto& q = *static_cast<const vmx_blobs_params*>(f->params);
        if (q.step_mode == 2) throw VmxError(VMX_ERR_UNSUPPORTED, "closed form: additive step only");
        vmx::ExactBlobs2D e{};
        e.nblobs = q.nblobs;
        e.step_mode = q.step_mode;
        for (int k = 0; k < q.nblobs; ++k) {
            if (q.blob[k][2] != 0.0 || q.blob[k][3] != 0.0)
                throw VmxError(VMX_ERR_UNSUPPORTED, "closed form: static blobs only");
            for (int i = 0; i < 7; ++i) e.blob[k][i] = q.blob[k][i];
        }
        for (int i = 0; i < 5; ++i) e.step[i] = q.step[i];
        for (int i = 0; i < 3; ++i) e.wave[i] = q.wave[i];
        const uint64_t nb = p->res[0] * p->res[1];
        ensure(ctx->misc, sizeof(e) + sizeof(double) * nb + 1024);
        Carve c(ctx->misc);
        vmx::ExactBlobs2D* d_e = c.take<vmx::ExactBlobs2D>(1);
        double* d_o = c.take<double>(nb);
        hipStream_t st = ctx->stream;
        HIP_CHECK(hipMemcpyAsync(d_e, &e, sizeof(e), hipMemcpyHostToDevice, st));
        const double dx = (p->hi[0] - p->lo[0]) / double(p->res[0]), dy = (p->hi[1] - p->lo[1]) / double(p->res[1]);
        vmx::k_exact_bins<<<unsigned((nb + 255) / 256), 256, 0, st>>>(d_e, p->lo[0], p->lo[1], dx, dy, p->res[0],
                                                                       p->res[1], d_o);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(bins, d_o, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        return int(VMX_OK);
    });
}

int vmx_eval(vmx_ctx* ctx, const vmx_integrand* f, int32_t scalar, uint64_t n, const double* x, double* out) {
    return guarded(ctx, [&]() -> int {
        if (!f || !f->params || !x || !out) throw VmxError(VMX_ERR_INVALID, "null argument");
        if (n == 0) return int(VMX_OK);
        if (scalar == VMX_F32) return eval_T<float, vmx::FastMath>(ctx, *f, n, x, out);
        if (scalar == VMX_F64) return eval_T<double, vmx::PreciseMath>(ctx, *f, n, x, out);
        throw VmxError(VMX_ERR_INVALID, "scalar must be 4 or 8");
    });
}

}  // extern "C"

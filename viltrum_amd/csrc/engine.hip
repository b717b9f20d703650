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
 *      footprints, per-bin footprint lists (k_stratify2, device scan, k_sort_lists).
 *   4. k_bins2: per-bin control variate + residual Monte Carlo.
 * Host synchronisation: after generation (capacity check) and once to size the bin
 * lists.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <type_traits>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include <cstring>
#include <rocprim/rocprim.hpp>

#include "../../include/viltrum_amd.h"
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
    hipEvent_t ev[8] = {};
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

/* Carves aligned sub-arrays out of one allocation. */
struct Carve {
    char* base;
    size_t used = 0;
    explicit Carve(void* p) : base(static_cast<char*>(p)) {}
    template <typename X>
    X* take(size_t n) {
        used = (used + 255) & ~size_t(255);
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

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

/* One (scalar, D, family) batch through the device pipeline (pipeline.h). */
struct JobIn {
    const vmx_problem* p;
    const void* params;
    void* bins;  // device pointer
    vmx_stats* stats;
};

template <typename T>
struct RocprimTemp {
    size_t sort_bytes = 0, scan_bytes = 0;
};

template <typename T, int D, typename P, typename Math>
int run_batch(vmx_ctx* ctx, const std::vector<JobIn>& in, hipStream_t st) {
    constexpr int K = vmx::pow3(D);
    const auto t_start = std::chrono::steady_clock::now();
    const uint32_t J = uint32_t(in.size());
    if (J == 0) return VMX_OK;
    if (J > uint32_t(vmx::MAX_JOBS)) throw VmxError(VMX_ERR_INVALID, "at most 4096 integrals per batch");
    const int order = in[0].p->order;
    for (const JobIn& x : in)
        if (x.p->order != order) throw VmxError(VMX_ERR_INVALID, "one region order per batch");
    const bool merge = order == VMX_ORDER_CANONICAL;

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
        jb.N = uint32_t(p.regions);
        jb.reg_base = uint32_t(n_reg);
        n_reg += p.regions;
        jb.form = p.form;
        jb.heur = p.heuristic;
        fparams[j] = P::from(*static_cast<const typename ParamOf<P>::type*>(in[j].params));
    }
    if (n_reg > 0x7FFFFFF0ull || total_bins > 0xFFFFFFF0ull) throw VmxError(VMX_ERR_CAPACITY, "batch too large");
    int mb_max = 1;
    for (const auto& jb : jobs) mb_max = std::max(mb_max, int(jb.mb));

    // launch shapes: expansion grid and lanes per bin (VMX_* overrides for tuning)
    int dev = ctx->device, n_cu = 0;
    HIP_CHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    uint32_t gen_blocks = uint32_t(n_cu) * 2;
    if (const char* e = std::getenv("VMX_GEN_BLOCKS")) gen_blocks = std::max(1, std::atoi(e));
    int round_chunk = 8;
    if (const char* e = std::getenv("VMX_ROUND_CHUNK")) round_chunk = std::max(1, std::atoi(e));

    uint32_t rounds = 0;
    uint64_t cap_mult = 8;
    vmx::TreeArrays<T> t{};
    vmx::Job<T>* d_jobs = nullptr;
    vmx::JobRt* d_rt = nullptr;
    P* d_f = nullptr;
    uint32_t *d_hist16 = nullptr, *d_hist8 = nullptr, *d_ncand = nullptr, *d_rounds = nullptr;
    T* d_theta = nullptr;
    int32_t* d_cand[2] = {nullptr, nullptr};
    uint32_t cap_total = 0;
    std::vector<vmx::JobRt> rt(J);
    // ---------------- 1. generation (retried with more node capacity if needed)
    for (int attempt = 0;; ++attempt) {
        uint64_t cap64 = 0;
        for (auto& jb : jobs) {
            jb.node_base = uint32_t(cap64);
            jb.node_cap = uint32_t(cap_mult * jb.N + 64);
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
        m.take<uint32_t>(size_t(J) * 256);
        m.take<int32_t>(cap_total);
        m.take<int32_t>(cap_total);
        m.take<uint32_t>(2 * J);
        m.take<T>(J);
        m.take<uint32_t>(4);
        Carve c(ensure(ctx->tree, m.used));
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
        d_hist8 = c.take<uint32_t>(size_t(J) * 256);
        d_cand[0] = c.take<int32_t>(cap_total);
        d_cand[1] = c.take<int32_t>(cap_total);
        d_ncand = c.take<uint32_t>(2 * J);
        d_theta = c.take<T>(J);
        uint32_t* small = c.take<uint32_t>(4);
        d_rounds = small;
        HIP_CHECK(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(vmx::Job<T>) * J, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(d_f, fparams.data(), sizeof(P) * J, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemsetAsync(d_rt, 0, sizeof(vmx::JobRt) * J, st));
        HIP_CHECK(hipMemsetAsync(t.job, 0xFF, sizeof(uint32_t) * cap_total, st));
        HIP_CHECK(hipMemsetAsync(d_hist16, 0, sizeof(uint32_t) * J * 65536, st));
        HIP_CHECK(hipMemsetAsync(d_hist8, 0, sizeof(uint32_t) * J * 256, st));
        HIP_CHECK(hipMemsetAsync(d_ncand, 0, sizeof(uint32_t) * 2 * J, st));
        HIP_CHECK(hipMemsetAsync(small, 0, sizeof(uint32_t) * 4, st));
        vmx::k_roots<T, D, P, Math><<<J, 256, 0, st>>>(d_jobs, d_rt, d_f, t, d_hist16, d_hist8, d_cand[0], d_ncand);
        HIP_CHECK(hipGetLastError());
        vmx::GenArgs<T> g{};
        g.jobs = d_jobs;
        g.rt = d_rt;
        g.t = t;
        g.hist16 = d_hist16;
        g.hist8 = d_hist8;
        g.cand[0] = d_cand[0];
        g.cand[1] = d_cand[1];
        g.ncand = d_ncand;
        g.theta = d_theta;
        g.rounds = d_rounds;
        g.J = J;
        std::vector<uint32_t> ncand_h(2 * J);
        int par = 0;
        for (int launched = 0;;) {
            for (int k = 0; k < round_chunk; ++k, ++launched, par ^= 1) {
                vmx::k_theta_round<T><<<J, 256, 0, st>>>(g, par);
                vmx::k_expand_round<T, D, P, Math><<<gen_blocks, vmx::GEN_THREADS, 0, st>>>(g, d_f, par);
            }
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipMemcpyAsync(ncand_h.data(), d_ncand, sizeof(uint32_t) * 2 * J, hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipStreamSynchronize(st));
            bool pending = false;
            for (uint32_t j = 0; j < J; ++j) pending |= ncand_h[par * J + j] != 0;
            if (!pending) break;
            if (launched >= 4096) throw VmxError(VMX_ERR_CAPACITY, "more than 4096 refinement rounds");
        }
        uint32_t small_h[4];
        HIP_CHECK(hipMemcpyAsync(rt.data(), d_rt, sizeof(vmx::JobRt) * J, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(small_h, small, sizeof(small_h), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        rounds = small_h[0];
        bool cap_fail = false;
        for (const auto& r : rt) {
            if (r.flags & vmx::JF_DEPTH) throw VmxError(VMX_ERR_CAPACITY, "region tree deeper than 31 splits in one dim");
            if (r.flags & vmx::JF_CAPACITY) cap_fail = true;
        }
        if (cap_fail) {
            if (attempt >= 3) throw VmxError(VMX_ERR_CAPACITY, "node capacity exceeded");
            cap_mult *= 2;
            continue;
        }
        break;
    }
    const double ms_gen = ms_since(t_start);

    // ---------------- 2. selection and canonical leaf order
    const auto t_sel = std::chrono::steady_clock::now();
    Measure ms;
    ms.take<int32_t>(cap_total);   // bucket
    ms.take<uint8_t>(cap_total);   // bstate
    ms.take<int32_t>(n_reg);       // leaf slots
    ms.take<int32_t>(n_reg);       // sort values tmp
    ms.take<uint32_t>(n_reg);      // sort keys
    ms.take<uint32_t>(n_reg);      // sort keys tmp
    size_t sort_bytes = 0;
    HIP_CHECK(hipError_t(rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                  (int32_t*)nullptr, (int32_t*)nullptr, size_t(n_reg), 0, 32, st)));
    ms.take<char>(sort_bytes + 256);
    Carve cs(ensure(ctx->regs, ms.used));
    int32_t* d_bucket = cs.take<int32_t>(cap_total);
    uint8_t* d_bstate = cs.take<uint8_t>(cap_total);
    int32_t* d_leaf = cs.take<int32_t>(n_reg);
    int32_t* d_leaf_tmp = cs.take<int32_t>(n_reg);
    uint32_t* d_keys = cs.take<uint32_t>(n_reg);
    uint32_t* d_keys_tmp = cs.take<uint32_t>(n_reg);
    void* d_sort_tmp = cs.take<char>(sort_bytes + 256);
    const unsigned cap_blocks = (cap_total + 255) / 256;
    const unsigned reg_blocks = unsigned((n_reg + 255) / 256);
    if (order == VMX_ORDER_CANONICAL) {
        HIP_CHECK(hipMemsetAsync(d_leaf, 0xFF, sizeof(int32_t) * n_reg, st));
        vmx::k_sel_bound<T><<<J, 256, 0, st>>>(d_jobs, d_rt, d_hist16, d_hist8);
        vmx::k_sel_gather<T><<<cap_blocks, 256, 0, st>>>(d_jobs, d_rt, t, cap_total, d_bucket);
        vmx::k_sel_resolve<T, D><<<J, 256, 0, st>>>(d_jobs, d_rt, t, d_bucket, d_bstate);
        vmx::k_leaves<T><<<cap_blocks, 256, 0, st>>>(d_jobs, d_rt, t, cap_total, d_leaf);
        HIP_CHECK(hipGetLastError());
        // LSD passes: code_{D-1} ... code_0, then job
        int32_t* vin = d_leaf;
        int32_t* vout = d_leaf_tmp;
        for (int which = D - 1; which >= (J > 1 ? -1 : 0); --which) {
            vmx::k_leaf_keys<D><<<reg_blocks, 256, 0, st>>>(vin, t.code, t.job, uint32_t(n_reg), which, d_keys);
            size_t bytes = sort_bytes;
            HIP_CHECK(hipError_t(rocprim::radix_sort_pairs(d_sort_tmp, bytes, d_keys, d_keys_tmp, vin, vout,
                                                          size_t(n_reg), 0, 32, st)));
            std::swap(vin, vout);
        }
        if (vin != d_leaf)
            HIP_CHECK(hipMemcpyAsync(d_leaf, vin, sizeof(int32_t) * n_reg, hipMemcpyDeviceToDevice, st));
        // every job must have exactly N leaves before any kernel indexes by them
        HIP_CHECK(hipMemcpyAsync(rt.data(), d_rt, sizeof(vmx::JobRt) * J, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        for (uint32_t j = 0; j < J; ++j)
            if ((rt[j].flags & vmx::JF_LEAVES) || rt[j].nleaves != jobs[j].N)
                throw VmxError(VMX_ERR_CAPACITY, "region selection did not close");
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
    size_t scan_reg_bytes = 0, scan_bin_bytes = 0;
    HIP_CHECK(hipError_t(rocprim::exclusive_scan(nullptr, scan_reg_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, 0u,
                                                size_t(n_reg + 1), rocprim::plus<uint32_t>(), st)));
    HIP_CHECK(hipError_t(rocprim::exclusive_scan(nullptr, scan_bin_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                0u, size_t(total_bins + 1), rocprim::plus<uint32_t>(), st)));
    Measure mr;
    mr.take<T>(n_reg * D);
    mr.take<T>(n_reg * D);
    mr.take<T>(n_reg);
    mr.take<T>(n_reg * K);
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
    mr.take<uint32_t>(total_bins + 1);
    mr.take<uint32_t>(total_bins + 1);
    mr.take<char>(std::max(scan_reg_bytes, scan_bin_bytes) + 256);
    Carve cr(ensure(ctx->lists, mr.used));
    vmx::RegionArrays<T> ra{};
    ra.lo = cr.take<T>(n_reg * D);
    ra.hi = cr.take<T>(n_reg * D);
    ra.vol = cr.take<T>(n_reg);
    ra.coef = cr.take<T>(n_reg * K);
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
    uint32_t* d_cnt = cr.take<uint32_t>(total_bins + 1);
    uint32_t* d_off = cr.take<uint32_t>(total_bins + 1);
    void* d_scan_tmp = cr.take<char>(std::max(scan_reg_bytes, scan_bin_bytes) + 256);
    // mb differs per job only through db; all jobs of a batch share the footprint stride
    for (const auto& jb : jobs)
        if (jb.mb != mb_max) throw VmxError(VMX_ERR_INVALID, "one bin_dims per batch");
    vmx::k_prepare2<T, D><<<reg_blocks, 256, 0, st>>>(d_jobs, d_leaf, t, ra, uint32_t(n_reg), mb_max, merge);
    HIP_CHECK(hipGetLastError());
    size_t bytes = scan_reg_bytes;
    HIP_CHECK(hipError_t(rocprim::exclusive_scan(d_scan_tmp, bytes, ra.fpflag, ra.fpid, 0u, size_t(n_reg + 1),
                                                rocprim::plus<uint32_t>(), st)));
    vmx::k_footprints<T, D><<<reg_blocks, 256, 0, st>>>(d_jobs, ra, fa, uint32_t(n_reg), mb_max, merge);
    HIP_CHECK(hipGetLastError());
    const uint32_t* d_nfp = ra.fpid + n_reg;
    const unsigned strat_blocks = unsigned((n_reg * 64 + 255) / 256);
    HIP_CHECK(hipMemsetAsync(d_cnt, 0, sizeof(uint32_t) * (total_bins + 1), st));
    vmx::k_stratify2<T, false><<<strat_blocks, 256, 0, st>>>(d_jobs, fa, d_nfp, d_cnt, nullptr, nullptr);
    HIP_CHECK(hipGetLastError());
    bytes = scan_bin_bytes;
    HIP_CHECK(hipError_t(rocprim::exclusive_scan(d_scan_tmp, bytes, d_cnt, d_off, 0u, size_t(total_bins + 1),
                                                rocprim::plus<uint32_t>(), st)));
    uint32_t pairs = 0, nfp = 0;
    HIP_CHECK(hipMemcpyAsync(&pairs, d_off + total_bins, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(&nfp, d_nfp, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(rt.data(), d_rt, sizeof(vmx::JobRt) * J, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    uint32_t* d_lst = static_cast<uint32_t*>(ensure(ctx->bins2, sizeof(uint32_t) * (size_t(pairs) + 1)));
    HIP_CHECK(hipMemsetAsync(d_cnt, 0, sizeof(uint32_t) * (total_bins + 1), st));
    vmx::k_stratify2<T, true><<<strat_blocks, 256, 0, st>>>(d_jobs, fa, d_nfp, d_cnt, d_off, d_lst);
    HIP_CHECK(hipGetLastError());
    const unsigned bin_blocks = unsigned((total_bins + 255) / 256);
    vmx::k_sort_lists<<<bin_blocks, 256, 0, st>>>(d_off, total_bins, d_lst);
    HIP_CHECK(hipGetLastError());
    const double ms_str = ms_since(t_str);

    // ---------------- 4. bins
    // lanes per bin: one residual sample per lane up to a wave
    uint64_t spp_max = 0;
    for (const auto& jb : jobs) spp_max = std::max(spp_max, jb.spp);
    int L = 8;
    while (L < 64 && uint64_t(L) * 8 < spp_max) L *= 2;
    if (const char* e = std::getenv("VMX_BIN_LANES")) L = std::max(1, std::min(64, std::atoi(e)));
    if (L & (L - 1)) throw VmxError(VMX_ERR_INVALID, "lanes per bin must be a power of two");
    const int teams = 256 / L;
    const size_t smem = sizeof(vmx::BinLds<T>) * teams;
    const uint64_t want_blocks = (total_bins + teams - 1) / teams;
    const unsigned bins_grid = unsigned(std::min<uint64_t>(want_blocks, uint64_t(n_cu) * 64));
    HIP_CHECK(hipEventRecord(ctx->ev[0], st));
    const int db = jobs[0].db;
    if (db == 1)
        vmx::k_bins3<T, D, 1, P, Math><<<bins_grid, 256, smem, st>>>(d_jobs, J, total_bins, L, ra, fa, d_off, d_lst,
                                                                     d_f, d_rt);
    else if constexpr (D >= 2) {
        if (db == 2)
            vmx::k_bins3<T, D, 2, P, Math><<<bins_grid, 256, smem, st>>>(d_jobs, J, total_bins, L, ra, fa, d_off,
                                                                         d_lst, d_f, d_rt);
        else if constexpr (D >= 3)
            vmx::k_bins3<T, D, 3, P, Math><<<bins_grid, 256, smem, st>>>(d_jobs, J, total_bins, L, ra, fa, d_off,
                                                                         d_lst, d_f, d_rt);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(ctx->ev[1], st));
    HIP_CHECK(hipMemcpyAsync(rt.data(), d_rt, sizeof(vmx::JobRt) * J, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
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
        Carve c(ensure(ctx->misc, sizeof(T) * total + sizeof(int32_t) * N + 4096));
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
int dispatch_T(vmx_ctx* ctx, int family, int D, const std::vector<JobIn>& jobs, hipStream_t st) {
    switch (family) {
        case VMX_FAMILY_POLY:
            switch (D) {
                case 1: return run_batch<T, 1, vmx::Poly<T>, Math>(ctx, jobs, st);
                case 2: return run_batch<T, 2, vmx::Poly<T>, Math>(ctx, jobs, st);
                case 3: return run_batch<T, 3, vmx::Poly<T>, Math>(ctx, jobs, st);
                case 4: return run_batch<T, 4, vmx::Poly<T>, Math>(ctx, jobs, st);
                case 5: return run_batch<T, 5, vmx::Poly<T>, Math>(ctx, jobs, st);
                default: break;
            }
            break;
        case VMX_FAMILY_BLOBS:
            if (D == 2) return run_batch<T, 2, vmx::Blobs<T>, Math>(ctx, jobs, st);
            if (D == 3) return run_batch<T, 3, vmx::Blobs<T>, Math>(ctx, jobs, st);
            break;
        case VMX_FAMILY_LOBES5:
            if (D == 5) return run_batch<T, 5, vmx::Lobes5<T>, Math>(ctx, jobs, st);
            break;
        default: break;
    }
    throw VmxError(VMX_ERR_UNSUPPORTED, "integrand family / dimension not built");
}

/* Groups a batch by (family, dim) and runs each group through the device pipeline. */
int integrate_groups(vmx_ctx* ctx, uint32_t n, const vmx_problem* const* ps, const vmx_integrand* const* fs,
                     void* const* bins, vmx_stats* const* stats, hipStream_t st) {
    const int scalar = ps[0]->scalar;
    std::vector<int> done(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        if (ps[i]->scalar != scalar) throw VmxError(VMX_ERR_INVALID, "one scalar type per batch");
        if (stats[i]) std::memset(stats[i], 0, sizeof(vmx_stats));
    }
    for (uint32_t i = 0; i < n; ++i) {
        if (done[i]) continue;
        std::vector<JobIn> group;
        for (uint32_t k = i; k < n; ++k)
            if (!done[k] && fs[k]->family == fs[i]->family && fs[k]->dim == fs[i]->dim) {
                group.push_back({ps[k], fs[k]->params, bins[k], stats[k]});
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
    if (p->form < VMX_FORM_VR_OPTIMIZE || p->form > VMX_FORM_CV_ALPHA1) throw VmxError(VMX_ERR_INVALID, "unknown form");
    if (p->order != VMX_ORDER_CANONICAL && p->order != VMX_ORDER_REFERENCE)
        throw VmxError(VMX_ERR_INVALID, "unknown region order");
    if (p->refiner != VMX_REFINER_HEURISTIC) throw VmxError(VMX_ERR_UNSUPPORTED, "learned refiner not loaded");
    for (int d = 0; d < f->dim; ++d)
        if (!(p->hi[d] > p->lo[d])) throw VmxError(VMX_ERR_INVALID, "range must have hi > lo");
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
    Carve c(ensure(ctx->misc, fbytes + sizeof(double) * n * (D + 1) + 1024));
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
    Carve c(ensure(ctx->misc, sizeof(P) + sizeof(double) * nb * 2 + 4096));
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
    for (vmx_ctx::Buf* b : {&ctx->tree, &ctx->regs, &ctx->lists, &ctx->misc, &ctx->params, &ctx->bins, &ctx->bins2})
        if (b->p) (void)hipFree(b->p);
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

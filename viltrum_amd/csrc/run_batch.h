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
#ifndef VMX_RUN_BATCH_H
#define VMX_RUN_BATCH_H

#include "engine_core.h"

namespace vmxe {

template <typename T, int D, typename P, typename Math>
int run_batch(vmx_ctx* ctx, const std::vector<JobIn>& in, hipStream_t st, PairsReq* pairs_req) {
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
    // the region roulettes weigh each region of a bin: one footprint per region for
    // their jobs (k_prepare2 and k_footprints decide per job)
    bool any_rr = false, any_rr_err = false;
    for (const JobIn& x : in) {
        any_rr |= vmx::form_roulette(x.p->form) != 0;
        any_rr_err |= vmx::form_roulette(x.p->form) == 2;
    }
    const bool merge = order == VMX_ORDER_CANONICAL && !replay;

    std::vector<vmx::Job<T>> jobs(J);
    std::vector<P> fparams(J);
    uint64_t total_bins = 0, n_reg = 0, total_tiles = 0, total_tiles2 = 0;
    bool any_rm = false;  // a job the region-major per-bin kernel takes
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
        jb.tile_base = uint32_t(total_tiles);
        total_tiles += (jb.nbins + 255) / 256;
        jb.tile_base2 = uint32_t(total_tiles2);
        total_tiles2 += vmx::rm_tiles(jb.db, jb.res, jb.nbins);
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
    // FP32 regions of 4..6 dims with the uniform estimator go through the region-major
    // per-bin kernel (bins_rm.h); k_bins1 takes the batch's other jobs
    constexpr bool RM_KIND = std::is_same<T, float>::value && D >= 4 && !vmx::IsFubini<P>::value;
    // VMX_NO_RM=1 sends these jobs through k_bins1 too (the same estimator, other draws)
    static const bool rm_env = !(std::getenv("VMX_NO_RM") && std::atoi(std::getenv("VMX_NO_RM")));
    bool all_rm = RM_KIND && rm_env;
    if (RM_KIND && rm_env) {
        for (const auto& jb : jobs) {
            const bool r = vmx::rm_job(jb.form, jb.N);
            any_rm |= r;
            all_rm &= r;
        }
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
    // deferred rounds: a round expands about N / k of a job's candidates, largest b
    // first (at least 256).  Exact for any k; they pay where rounds hold much work (a
    // call of 65536 regions or more), and cost an extra kernel per round elsewhere.
    // VMX_DEFER = k sets k for every call (0: off).  Training pairs (every node the
    // rounds create) keep the undeferred rounds' nodes
    int defer_k = n_reg >= 65536 ? DEFER_K : 0;
    if (const char* e = std::getenv("VMX_DEFER")) defer_k = std::max(0, std::atoi(e));
    const bool defer = defer_k > 0 && !replay && !pairs_req;
    if (defer)
        for (auto& jb : jobs) jb.defer_m = std::max<uint32_t>(256u, jb.N / uint32_t(defer_k));
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
    // the round-count hint's class: the workload class and the largest budget's octave
    uint32_t n_max = 1;
    for (const auto& jb : jobs) n_max = std::max(n_max, jb.N);
    std::array<uint64_t, 7> hint_key{};
    std::copy(tune_key.begin(), tune_key.end(), hint_key.begin());
    hint_key[6] = uint64_t(ilog2(n_max));
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
    // the rounds' shared state, one block so that one copy reads it back: candidate
    // counts (2 x J), job states (J), 16 words of counters ([0] rounds, [2..5] the pair
    // totals after stratification, [8 + d] deepest split count along dim d) and deferred
    // candidate counts (2 x J)
    const size_t state_bytes = sizeof(uint32_t) * 2 * J + sizeof(vmx::JobRt) * J + 64 + sizeof(uint32_t) * 2 * J;
    const size_t slot_bytes = (state_bytes + 255) & ~size_t(255);
    char* d_state = nullptr;
    uint32_t* small = nullptr;
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
        m.take<P>(J);
        m.take<char>(state_bytes);
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
        m.take<T>(J);
        if (defer) {
            m.take<T>(J);
            m.take<uint32_t>(J);
            m.take<vmx::DefEnt<T>>(size_t(cap_total) * 2);
            m.take<uint32_t>(cap_total);
            m.take<uint32_t>(size_t(J) * 2 * vmx::CHIST);
        }
        ensure(ctx->tree, m.used);
        Carve c(ctx->tree);
        // jobs, parameters and state first: one contiguous image the host uploads at once
        d_jobs = c.take<vmx::Job<T>>(J);
        d_f = c.take<P>(J);
        d_state = c.take<char>(state_bytes);
        d_ncand = reinterpret_cast<uint32_t*>(d_state);
        d_rt = reinterpret_cast<vmx::JobRt*>(d_state + sizeof(uint32_t) * 2 * J);
        small = reinterpret_cast<uint32_t*>(d_state + sizeof(uint32_t) * 2 * J + sizeof(vmx::JobRt) * J);
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
        d_theta = c.take<T>(J);
        d_rounds = small;
        T* d_theta_lo = nullptr;
        uint32_t *d_nexp = nullptr, *d_exp = nullptr, *d_chist = nullptr;
        vmx::DefEnt<T>* d_def = nullptr;
        uint32_t* d_ndef = small + 16;
        if (defer) {
            d_theta_lo = c.take<T>(J);
            d_nexp = c.take<uint32_t>(J);
            d_def = c.take<vmx::DefEnt<T>>(size_t(cap_total) * 2);
            d_exp = c.take<uint32_t>(cap_total);
            d_chist = c.take<uint32_t>(size_t(J) * 2 * vmx::CHIST);
        }
        {
            // jobs, parameters and zeroed state in one copy from pinned staging (after the
            // round loop's two readback slots) where it fits; k_roots clears each job's
            // histogram and node slots
            char* base = reinterpret_cast<char*>(d_jobs);
            const size_t img = size_t(d_state + state_bytes - base);
            const size_t need = 2 * slot_bytes + img;
            if (!ctx->noalloc || need <= ctx->pinned_bytes) {
                char* pin = static_cast<char*>(ensure_pinned(ctx, need)) + 2 * slot_bytes;
                std::memcpy(pin, jobs.data(), sizeof(vmx::Job<T>) * J);
                std::memcpy(pin + (reinterpret_cast<char*>(d_f) - base), fparams.data(), sizeof(P) * J);
                std::memset(pin + (d_state - base), 0, state_bytes);
                HIP_CHECK(hipMemcpyAsync(base, pin, img, hipMemcpyHostToDevice, st));
            } else {
                HIP_CHECK(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(vmx::Job<T>) * J, hipMemcpyHostToDevice, st));
                HIP_CHECK(hipMemcpyAsync(d_f, fparams.data(), sizeof(P) * J, hipMemcpyHostToDevice, st));
                HIP_CHECK(hipMemsetAsync(d_state, 0, state_bytes, st));
            }
        }
        if (replay) {
            HIP_CHECK(hipMemsetAsync(t.job, 0xFF, sizeof(uint32_t) * cap_total, st));  // the given regions become the nodes, in list order
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
        g.defer = defer ? (n_max <= 32768 ? 2 : 1) : 0;
        g.cap_total = cap_total;
        g.theta_lo = d_theta_lo;
        g.ndef = d_ndef;
        g.def = d_def;
        g.nexp = d_nexp;
        g.exp = d_exp;
        g.chist = d_chist;
        // rounds in chunks of round_chunk (even, so each chunk starts at parity 0); after
        // each chunk the candidate counts, job states and round counter are copied to
        // pinned memory.  The host checks chunk k while chunk k+1 runs, so the device is
        // never idle waiting for the host; the chunk after the last busy one finds no
        // candidates and its kernels exit at once.
        char* pin = static_cast<char*>(ensure_pinned(ctx, 2 * slot_bytes));
        auto slot = [&](int k) { return pin + size_t(k & 1) * slot_bytes; };
        // the chunk's 2 * round_chunk kernels as a graph: one launch instead of many
        // small ones (the late rounds are launch-bound); recaptured only when an argument
        // changes.  VMX_NO_GRAPH=1 launches them one by one (e.g. under a kernel tracer
        // that does not follow graph launches).
        // (nor under the no-alloc entry point: instantiating a graph allocates)
        static const bool graph_env = !(std::getenv("VMX_NO_GRAPH") && std::atoi(std::getenv("VMX_NO_GRAPH")));
        const bool use_graph = graph_env && !ctx->noalloc;
        // chunk sizes: round_chunk, then 2 once the rounds near the count the last call of
        // this workload class needed (fewer idle rounds after the last busy one)
        const auto hint_it = ctx->round_hint.find(hint_key);
        const uint32_t hint = hint_it == ctx->round_hint.end() ? 0u : hint_it->second;
        const int tail_chunk = std::min(2, round_chunk);
        const size_t pre_bytes = size_t(J + 1) * sizeof(uint32_t);
        auto round_kernels = [&](int par) {
            vmx::k_theta_round<T><<<J, vmx::THETA_THREADS, 0, st>>>(g, par);
            if (g.defer == 1) vmx::k_filter_round<T><<<gen_blocks, 256, pre_bytes, st>>>(g, par);
            launch_expand<T, D, P, Math>(learn, g, d_f, par, gen_blocks, st);
        };
        auto ensure_graph = [&](vmx_ctx::RoundGraph& rg, int size) {
            std::vector<unsigned char> key(sizeof(g) + 64);
            std::memcpy(key.data(), &g, sizeof(g));
            const void* fn = reinterpret_cast<const void*>(&launch_expand<T, D, P, Math>);
            const void* parts[4] = {fn, d_f, st, reinterpret_cast<const void*>(uintptr_t(learn))};
            std::memcpy(key.data() + sizeof(g), parts, sizeof(parts));
            const uint32_t shape[2] = {gen_blocks, uint32_t(size)};
            std::memcpy(key.data() + sizeof(g) + sizeof(parts), shape, sizeof(shape));
            if (rg.key == key && rg.exec) return;
            if (rg.exec) HIP_CHECK(hipGraphExecDestroy(rg.exec));
            if (rg.graph) HIP_CHECK(hipGraphDestroy(rg.graph));
            rg.exec = nullptr;
            rg.graph = nullptr;
            rg.key.clear();
            HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
            for (int r = 0; r < size; ++r) round_kernels(r & 1);
            hipGraph_t graph = nullptr;
            const hipError_t e1 = hipStreamEndCapture(st, &graph);
            HIP_CHECK(e1);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipGraphInstantiate(&rg.exec, graph, nullptr, nullptr, 0));
            rg.graph = graph;
            rg.key = key;
        };
        if (use_graph) {
            ensure_graph(ctx->rounds, round_chunk);
            if (hint > 0 && tail_chunk < round_chunk) ensure_graph(ctx->rounds_tail, tail_chunk);
        }
        uint32_t launched = 0;
        // sizes are even, so every chunk starts at parity 0
        auto launch_chunk = [&](int k) {
            const bool tail = hint > 0 && launched + uint32_t(round_chunk) + 2 > hint && launched < hint + 8;
            const int size = tail ? tail_chunk : round_chunk;
            if (use_graph) {
                HIP_CHECK(hipGraphLaunch(tail && tail_chunk < round_chunk ? ctx->rounds_tail.exec : ctx->rounds.exec,
                                         st));
            } else {
                for (int r = 0; r < size; ++r) round_kernels(r & 1);
                HIP_CHECK(hipGetLastError());
            }
            launched += uint32_t(size);
            HIP_CHECK(hipMemcpyAsync(slot(k), d_state, state_bytes, hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipEventRecord(ctx->ev[2 + (k & 1)], st));
        };
        launch_chunk(0);
        int k = 0;
        for (;; ++k) {
            if (launched + uint32_t(round_chunk) > 4096) {
                HIP_CHECK(hipStreamSynchronize(st));
                throw VmxError(VMX_ERR_CAPACITY, "more than 4096 refinement rounds");
            }
            launch_chunk(k + 1);
            HIP_CHECK(hipEventSynchronize(ctx->ev[2 + (k & 1)]));
            const uint32_t* nc = reinterpret_cast<const uint32_t*>(slot(k));
            const uint32_t* nd = reinterpret_cast<const uint32_t*>(slot(k) + sizeof(uint32_t) * 2 * J +
                                                                   sizeof(vmx::JobRt) * J + 64);
            bool pending = false;
            for (uint32_t j = 0; j < J; ++j) pending |= nc[j] != 0 || nd[j] != 0;
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
        ctx->round_hint[hint_key] = rounds;
        break;
    }
    const double ms_gen = ms_since(t_start);

    if constexpr (vmx::IsFubini<P>::value || D > LEARN_MAX_DIM) {
        if (pairs_req) throw VmxError(VMX_ERR_UNSUPPORTED, "training pairs: not for the Fubini form or D > 5");
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
        vmx::k_sel_bound<T><<<J, 256, 0, st>>>(d_jobs, d_rt, d_hist16, d_leaf);
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
    mr.take<T>(n_reg * vmx::samp_stride<D>());
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
    mr.take<unsigned long long>(1);
    mr.take<uint32_t>(n_reg + 1);
    mr.take<uint32_t>(total_bins + 1);
    mr.take<char>(scan_reg_bytes + 256);
    ensure(ctx->lists, mr.used);
    Carve cr(ctx->lists);
    vmx::RegionArrays<T> ra{};
    ra.lo = cr.take<T>(n_reg * D);
    ra.hi = cr.take<T>(n_reg * D);
    ra.vol = cr.take<T>(n_reg);
    ra.samp = cr.take<T>(n_reg * vmx::samp_stride<D>());
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
    unsigned long long* d_pairs64 = cr.take<unsigned long long>(1);
    uint32_t* d_pofs = cr.take<uint32_t>(n_reg + 1);
    uint32_t* d_off = cr.take<uint32_t>(total_bins + 1);
    void* d_scan_tmp = cr.take<char>(scan_reg_bytes + 256);
    // mb differs per job only through db; all jobs of a batch share the footprint stride
    for (const auto& jb : jobs)
        if (jb.mb != mb_max) throw VmxError(VMX_ERR_INVALID, "one bin_dims per batch");
    if constexpr (D >= 4) {
        constexpr unsigned WPB = vmx::PrepWave<T, D>::WPB;
        vmx::k_prepare_wave<T, D><<<unsigned((n_reg + WPB - 1) / WPB), 64 * WPB, 0, st>>>(
            d_jobs, d_leaf, t, ra, uint32_t(n_reg), mb_max, merge, fa.npairs, d_pairs64);
    } else {
        vmx::k_prepare2<T, D><<<reg_blocks, 256, 0, st>>>(d_jobs, d_leaf, t, ra, uint32_t(n_reg), mb_max, merge,
                                                          fa.npairs, d_pairs64);
    }
    HIP_CHECK(hipGetLastError());
    size_t bytes = scan_reg_bytes;
    HIP_CHECK(hipError_t(rocprim::exclusive_scan(d_scan_tmp, bytes, ra.fpflag, ra.fpid, 0u, size_t(n_reg + 1),
                                                rocprim::plus<uint32_t>(), st)));
    vmx::k_footprints<T, D><<<reg_blocks, 256, 0, st>>>(d_jobs, ra, fa, uint32_t(n_reg), mb_max, merge);
    HIP_CHECK(hipGetLastError());
    const uint32_t* d_nfp = ra.fpid + n_reg;
    // per-bin footprint lists: pairs emitted per footprint at scanned offsets, then a
    // stable radix sort by bin (each list ends up in ascending footprint order)
    bytes = scan_reg_bytes;
    HIP_CHECK(hipError_t(rocprim::exclusive_scan(d_scan_tmp, bytes, fa.npairs, d_pofs, 0u, size_t(n_reg + 1),
                                                rocprim::plus<uint32_t>(), st)));
    // the pair total in 64 bits: the 32-bit scan above (and the per-bin offsets below)
    // only hold when it fits
    vmx::k_sum_u64<<<std::min(reg_blocks, 1024u), 256, 0, st>>>(fa.npairs, uint32_t(n_reg), d_pairs64);
    vmx::k_collect_totals<<<1, 64, 0, st>>>(d_pairs64, d_pofs + n_reg, d_nfp, small);
    HIP_CHECK(hipGetLastError());
    uint32_t pairs = 0, nfp = 0;
    uint64_t pairs64 = 0;
    {
        char* pin = static_cast<char*>(ensure_pinned(ctx, state_bytes));
        HIP_CHECK(hipMemcpyAsync(pin, d_state, state_bytes, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        uint32_t w[4];
        std::memcpy(w, pin + sizeof(uint32_t) * 2 * J + sizeof(vmx::JobRt) * J + 2 * sizeof(uint32_t), sizeof(w));
        pairs64 = uint64_t(w[0]) | (uint64_t(w[1]) << 32);
        pairs = w[2];
        nfp = w[3];
        std::memcpy(rt.data(), pin + sizeof(uint32_t) * 2 * J, sizeof(vmx::JobRt) * J);
    }
    check_pair_total(pairs64);
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
    const size_t qlong_len = size_t(pairs) / vmx::SORT_SHORT + 2;
    mp.take<uint32_t>(total_bins + 1 + qlong_len);            // counts, then the queue of long lists
    mp.take<uint32_t>(size_t(pairs) + 1);                     // lists
    mp.take<uint32_t>(size_t(pairs) + 1);                     // copies of very long lists
    mp.take<uint32_t>(size_t(pairs) + 1);                     // prefix counts of long lists
    mp.take<uint32_t>(size_t(pairs) + 1);                     // first regions of long lists
    if (any_rr) mp.take<double>(size_t(pairs) + 1);          // roulette weights per pair
    mp.take<char>(scan_bins_bytes + 256);
    ensure(ctx->bins2, mp.used);
    Carve cp(ctx->bins2);
    // the per-bin counts, then the queue of long lists (its word 0 counts the entries
    // that follow): one fill clears the counts and the queue's counter
    uint32_t* d_cnt = cp.take<uint32_t>(total_bins + 1 + qlong_len);
    uint32_t* d_qlong = d_cnt + total_bins + 1;
    uint32_t* d_lst = cp.take<uint32_t>(size_t(pairs) + 1);
    uint32_t* d_lst_tmp = cp.take<uint32_t>(size_t(pairs) + 1);
    vmx::ListPrefix lp{cp.take<uint32_t>(size_t(pairs) + 1), cp.take<uint32_t>(size_t(pairs) + 1)};
    double* d_rrw = any_rr ? cp.take<double>(size_t(pairs) + 1) : nullptr;
    void* d_scan_bins_tmp = cp.take<char>(scan_bins_bytes + 256);
    HIP_CHECK(hipMemsetAsync(d_cnt, 0, sizeof(uint32_t) * (total_bins + 2), st));
    const unsigned fp_blocks = unsigned((uint64_t(nfp) * 64 + 255) / 256);
    if (pairs > 0)
        vmx::k_bin_pass<T, false><<<fp_blocks, 256, 0, st>>>(d_jobs, fa, d_nfp, d_cnt, nullptr, nullptr, d_pofs,
                                                             d_lst_tmp);
    bytes = scan_bins_bytes;
    HIP_CHECK(hipError_t(rocprim::exclusive_scan(d_scan_bins_tmp, bytes, d_cnt, d_off, 0u, size_t(total_bins + 1),
                                                rocprim::plus<uint32_t>(), st)));
    if (pairs > 0) {
        // the pairs' ranks live in d_lst_tmp until k_sort_long reuses it
        vmx::k_bin_pass<T, true><<<fp_blocks, 256, 0, st>>>(d_jobs, fa, d_nfp, d_cnt, d_off, d_lst, d_pofs,
                                                            d_lst_tmp);
        vmx::k_sort_lists<T><<<unsigned((total_bins + 255) / 256), 256, 0, st>>>(d_off, total_bins, d_lst, d_qlong,
                                                                                 fa, lp);
        vmx::k_sort_long<<<1024, 256, 0, st>>>(d_off, d_qlong, d_lst, d_lst_tmp);
        vmx::k_list_prefix<T><<<1024, 256, 0, st>>>(d_off, d_qlong, d_lst, fa, lp);
    }
    HIP_CHECK(hipGetLastError());
    const double ms_str = ms_since(t_str);

    // ---------------- 4. bins
    // one lane per bin, one block per tile of 256 consecutive bins of one job
    const unsigned bins_grid = unsigned(total_tiles);  // every job starts a fresh tile
    HIP_CHECK(hipEventRecord(ctx->ev[0], st));
    const int db = jobs[0].db;
    // the importance-sampling variant is its own instantiation: its register footprint
    // stays out of the uniform kernel
    const int rm = any_rm ? 1 : 0;
    auto bins_kernel = [&](auto mode_tag) {
        constexpr int MODE = decltype(mode_tag)::value;
        if constexpr (RM_KIND) {
            if (any_rm) {
                const unsigned grid2 = unsigned(total_tiles2);
                if (db == 1)
                    vmx::k_bins_rm<T, D, 1, P, Math><<<grid2, 256, 0, st>>>(d_jobs, J, ra, fa, d_off, d_lst, d_f, d_rt,
                                                                             lp, shape.bins_xcd);
                else if (db == 2)
                    vmx::k_bins_rm<T, D, 2, P, Math><<<grid2, 256, 0, st>>>(d_jobs, J, ra, fa, d_off, d_lst, d_f, d_rt,
                                                                             lp, shape.bins_xcd);
                else
                    vmx::k_bins_rm<T, D, 3, P, Math><<<grid2, 256, 0, st>>>(d_jobs, J, ra, fa, d_off, d_lst, d_f, d_rt,
                                                                             lp, shape.bins_xcd);
            }
        }
        if (all_rm) return;
        if (db == 1)
            vmx::k_bins1<T, D, 1, P, Math, MODE><<<bins_grid, 256, 0, st>>>(
                d_jobs, J, total_bins, ra, fa, d_off, d_lst, d_f, d_rt, lp, shape.bins_xcd, d_rrw, rm);
        else if constexpr (D >= 2) {
            if (db == 2)
                vmx::k_bins1<T, D, 2, P, Math, MODE><<<bins_grid, 256, 0, st>>>(
                    d_jobs, J, total_bins, ra, fa, d_off, d_lst, d_f, d_rt, lp, shape.bins_xcd, d_rrw, rm);
            else if constexpr (D >= 3)
                vmx::k_bins1<T, D, 3, P, Math, MODE><<<bins_grid, 256, 0, st>>>(
                    d_jobs, J, total_bins, ra, fa, d_off, d_lst, d_f, d_rt, lp, shape.bins_xcd, d_rrw, rm);
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

}  // namespace vmxe

/* explicit instantiation of one (scalar, D, family, math policy) pipeline */
#define VMX_INSTANTIATE_BATCH(T, D, P, M) \
    template int vmxe::run_batch<T, D, P, M>(vmx_ctx*, const std::vector<vmxe::JobIn>&, hipStream_t, vmxe::PairsReq*);

#endif

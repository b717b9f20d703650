/*
 * Host orchestration and C ABI of the MI355X per-bin control-variate integrator.
 *
 * Pipeline of one integral (vmx_integrate*):
 *   1. k_root, then rounds of k_theta + k_expand until no new node reaches the round
 *      threshold (region generation, exact parallel rounds).
 *   2. exact selection of the N-1 split nodes; the final N regions are the unsplit
 *      children, ordered canonically (per-dim dyadic codes) or in the reference's
 *      heap-array order (libstdc++ push_heap/pop_heap replayed over the GPU-computed
 *      errors, src/nested/regions-generator-adaptive-heap.h:21-41).
 *   3. k_prepare (boxes, volumes, FP64 bin-dim marginals, monomial coefficients, pixel
 *      boxes), bin lists (k_stratify count, k_scan, k_stratify fill, k_sort_lists).
 *   4. k_bins: per-bin control variate + residual Monte Carlo.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <type_traits>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/viltrum_amd.h"
#include "kernels.h"

struct vmx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // device memory pool: one arena per purpose, grown on demand
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
    };
    Buf tree, regs, lists, misc, params, bins;
    hipEvent_t ev[8] = {};
    // last result (for vmx_export_regions): gathered from device memory on demand
    int last_D = 0;
    uint64_t last_n = 0;
    std::vector<double> last_err;
    std::vector<int32_t> last_dim;
    std::function<void(double*, double*, double*)> last_gather;
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

template <typename T>
struct HostNodes {
    std::vector<T> b, b1, err;
    std::vector<int32_t> parent, child, dim;
    std::vector<uint32_t> code;
};

/* Final region order: popped set by (b desc, b1 desc, codes asc) or by heap replay. */
template <typename T>
std::vector<int32_t> select_regions(const HostNodes<T>& h, uint32_t M, int D, uint64_t N, int order,
                                    uint64_t* ties_unresolved) {
    std::vector<int32_t> out;
    *ties_unresolved = 0;
    if (N == 1) {
        out.push_back(0);
        return out;
    }
    const uint64_t need = N - 1;
    auto code_less = [&](int32_t x, int32_t y) {
        for (int d = 0; d < D; ++d) {
            const uint32_t a = h.code[size_t(x) * D + d], c = h.code[size_t(y) * D + d];
            if (a != c) return a < c;
        }
        return x < y;
    };
    std::vector<char> popped(M, 0);
    if (order == VMX_ORDER_REFERENCE) {
        // replay the reference's heap on the scalar errors
        struct E {
            T err;
            int32_t node;
        };
        std::vector<E> heap;
        heap.reserve(N + 1);
        auto cmp = [](const E& a, const E& b) { return a.err < b.err; };
        heap.push_back({h.err[0], 0});
        for (uint64_t i = 0; i < need; ++i) {
            const E top = heap.front();
            const int32_t c = h.child[top.node];
            if (c < 0) throw VmxError(VMX_ERR_CAPACITY, "heap replay reached an unexpanded node");
            std::pop_heap(heap.begin(), heap.end(), cmp);
            heap.pop_back();
            heap.push_back({h.err[c], c});
            std::push_heap(heap.begin(), heap.end(), cmp);
            heap.push_back({h.err[c + 1], c + 1});
            std::push_heap(heap.begin(), heap.end(), cmp);
        }
        for (const E& e : heap) out.push_back(e.node);
        return out;
    }
    std::vector<int32_t> idx(M);
    for (uint32_t i = 0; i < M; ++i) idx[i] = int32_t(i);
    auto key_greater = [&](int32_t x, int32_t y) {
        if (h.b[x] != h.b[y]) return h.b[x] > h.b[y];
        if (h.b1[x] != h.b1[y]) return h.b1[x] > h.b1[y];
        return code_less(x, y);
    };
    if (need > M) throw VmxError(VMX_ERR_CAPACITY, "fewer nodes than splits");
    std::nth_element(idx.begin(), idx.begin() + (need - 1), idx.end(), key_greater);
    const int32_t last = idx[need - 1];
    for (uint64_t i = 0; i < need; ++i) popped[idx[i]] = 1;
    uint64_t unpopped_eq = 0;
    for (uint32_t i = 0; i < M; ++i)
        if (!popped[i] && h.b[i] == h.b[last] && h.b1[i] == h.b1[last]) ++unpopped_eq;
    *ties_unresolved = unpopped_eq;
    for (uint32_t i = 0; i < M; ++i) {
        if (popped[i]) {
            if (h.child[i] < 0) throw VmxError(VMX_ERR_CAPACITY, "selected node was not expanded");
            continue;
        }
        if (h.parent[i] >= 0 && popped[h.parent[i]]) out.push_back(int32_t(i));
    }
    if (out.size() != N) throw VmxError(VMX_ERR_CAPACITY, "region selection did not close");
    std::sort(out.begin(), out.end(), code_less);
    return out;
}

template <typename T, int D, typename P, typename Math>
int run(vmx_ctx* ctx, const vmx_problem& p, const void* fparams, T* bins_dev, hipStream_t st, vmx_stats* stats) {
    constexpr int K = vmx::pow3(D);
    const auto t_start = std::chrono::steady_clock::now();
    const int db = p.bin_dims;
    const uint64_t N = p.regions;
    uint64_t nb = 1;
    for (int i = 0; i < db; ++i) nb *= p.res[i];
    int mb = 1;
    for (int i = 0; i < db; ++i) mb *= 3;

    // integrand parameters in the evaluation type
    const P fp = P::from(*static_cast<const typename ParamOf<P>::type*>(fparams));
    T lo[D], hi[D];
    for (int d = 0; d < D; ++d) {
        lo[d] = T(p.lo[d]);
        hi[d] = T(p.hi[d]);
    }
    Measure pm;
    pm.take<P>(1);
    pm.take<T>(D);
    pm.take<T>(D);
    pm.take<uint64_t>(3);
    pm.take<T>(1);
    pm.take<uint32_t>(2);
    pm.take<int32_t>(1);
    Carve pc(ensure(ctx->params, pm.used));
    P* d_f = pc.take<P>(1);
    T* d_lo = pc.take<T>(D);
    T* d_hi = pc.take<T>(D);
    uint64_t* d_res = pc.take<uint64_t>(3);
    T* d_theta = pc.take<T>(1);
    uint32_t* d_ncand = pc.take<uint32_t>(2);
    int32_t* d_flags = pc.take<int32_t>(1);
    uint64_t res3[3] = {p.res[0], db > 1 ? p.res[1] : 1, db > 2 ? p.res[2] : 1};
    HIP_CHECK(hipMemcpyAsync(d_f, &fp, sizeof(P), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(d_lo, lo, sizeof(T) * D, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(d_hi, hi, sizeof(T) * D, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(d_res, res3, sizeof(res3), hipMemcpyHostToDevice, st));

    // ---------------- 1. region generation
    uint64_t cap64 = 8 * N + 64;
    uint32_t nnodes = 0, rounds = 0;
    vmx::Tree<T> tr{};
    int32_t* cand[2] = {nullptr, nullptr};
    for (int attempt = 0;; ++attempt) {
        if (cap64 > 0x7FFFFFF0ull) throw VmxError(VMX_ERR_CAPACITY, "region budget too large");
        const uint32_t cap = uint32_t(cap64);
        Measure tm;
        tm.take<T>(size_t(cap) * D);
        tm.take<T>(size_t(cap) * D);
        tm.take<T>(size_t(cap) * K);
        tm.take<T>(cap);
        tm.take<T>(cap);
        tm.take<T>(cap);
        tm.take<int32_t>(cap);
        tm.take<int32_t>(cap);
        tm.take<int32_t>(cap);
        tm.take<uint32_t>(size_t(cap) * D);
        tm.take<uint32_t>(1);
        tm.take<uint32_t>(65536);
        tm.take<int32_t>(cap);
        tm.take<int32_t>(cap);
        Carve tc(ensure(ctx->tree, tm.used));
        tr.lo = tc.take<T>(size_t(cap) * D);
        tr.hi = tc.take<T>(size_t(cap) * D);
        tr.v = tc.take<T>(size_t(cap) * K);
        tr.err = tc.take<T>(cap);
        tr.b = tc.take<T>(cap);
        tr.b1 = tc.take<T>(cap);
        tr.dim = tc.take<int32_t>(cap);
        tr.parent = tc.take<int32_t>(cap);
        tr.child = tc.take<int32_t>(cap);
        tr.code = tc.take<uint32_t>(size_t(cap) * D);
        tr.count = tc.take<uint32_t>(1);
        tr.hist = tc.take<uint32_t>(65536);
        tr.flags = d_flags;
        tr.cap = cap;
        cand[0] = tc.take<int32_t>(cap);
        cand[1] = tc.take<int32_t>(cap);
        HIP_CHECK(hipMemsetAsync(tr.hist, 0, 65536 * sizeof(uint32_t), st));
        HIP_CHECK(hipMemsetAsync(d_flags, 0, sizeof(int32_t), st));
        vmx::k_root<T, D, P, Math><<<1, 64, 0, st>>>(tr, d_f, d_lo, d_hi, p.heuristic);
        HIP_CHECK(hipGetLastError());
        const int32_t root = 0;
        HIP_CHECK(hipMemcpyAsync(cand[0], &root, sizeof(int32_t), hipMemcpyHostToDevice, st));
        uint32_t n_in = 1;
        int cur = 0;
        rounds = 0;
        int32_t flags = 0;
        while (N > 1 && n_in > 0) {
            vmx::k_theta<T><<<1, 256, 0, st>>>(tr, uint32_t(N - 1), d_theta);
            HIP_CHECK(hipMemsetAsync(d_ncand, 0, sizeof(uint32_t), st));
            const unsigned blocks = (n_in + 255) / 256;
            vmx::k_expand<T, D, P, Math><<<blocks, 256, 0, st>>>(tr, d_f, cand[cur], n_in, cand[cur ^ 1], d_ncand,
                                                               d_theta, p.heuristic);
            HIP_CHECK(hipGetLastError());
            uint32_t n_out = 0;
            HIP_CHECK(hipMemcpyAsync(&n_out, d_ncand, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipMemcpyAsync(&flags, d_flags, sizeof(int32_t), hipMemcpyDeviceToHost, st));
            HIP_CHECK(hipStreamSynchronize(st));
            ++rounds;
            if (flags) break;
            n_in = n_out;
            cur ^= 1;
        }
        HIP_CHECK(hipMemcpyAsync(&flags, d_flags, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipMemcpyAsync(&nnodes, tr.count, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (flags & vmx::FLAG_DEPTH) throw VmxError(VMX_ERR_CAPACITY, "region tree deeper than 31 splits in one dim");
        if (flags & vmx::FLAG_CAPACITY) {
            if (attempt >= 4) throw VmxError(VMX_ERR_CAPACITY, "node capacity exceeded");
            cap64 *= 2;
            continue;
        }
        break;
    }
    const double ms_gen = ms_since(t_start);

    // ---------------- 2. selection
    const auto t_sel = std::chrono::steady_clock::now();
    HostNodes<T> hn;
    hn.b.resize(nnodes);
    hn.b1.resize(nnodes);
    hn.err.resize(nnodes);
    hn.parent.resize(nnodes);
    hn.child.resize(nnodes);
    hn.dim.resize(nnodes);
    hn.code.resize(size_t(nnodes) * D);
    HIP_CHECK(hipMemcpyAsync(hn.b.data(), tr.b, sizeof(T) * nnodes, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(hn.b1.data(), tr.b1, sizeof(T) * nnodes, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(hn.err.data(), tr.err, sizeof(T) * nnodes, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(hn.parent.data(), tr.parent, sizeof(int32_t) * nnodes, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(hn.child.data(), tr.child, sizeof(int32_t) * nnodes, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(hn.dim.data(), tr.dim, sizeof(int32_t) * nnodes, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(hn.code.data(), tr.code, sizeof(uint32_t) * nnodes * D, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    uint64_t ties = 0;
    std::vector<int32_t> order = select_regions<T>(hn, nnodes, D, N, p.order, &ties);
    const double ms_sel = ms_since(t_sel);

    // ---------------- 3. regions and bin lists
    const auto t_str = std::chrono::steady_clock::now();
    Measure rm;
    rm.take<int32_t>(N);
    rm.take<T>(N * D);
    rm.take<T>(N * D);
    rm.take<T>(N);
    rm.take<double>(N * mb);
    rm.take<T>(N * K);
    rm.take<uint32_t>(N * 6);
    rm.take<uint32_t>(nb + 1);
    rm.take<uint32_t>(nb + 1);
    Carve rc(ensure(ctx->regs, rm.used));
    int32_t* d_order = rc.take<int32_t>(N);
    vmx::Regs<T> rg{};
    rg.lo = rc.take<T>(N * D);
    rg.hi = rc.take<T>(N * D);
    rg.vol = rc.take<T>(N);
    rg.marg = rc.take<double>(N * mb);
    rg.coef = rc.take<T>(N * K);
    rg.box = rc.take<uint32_t>(N * 6);
    rg.n = uint32_t(N);
    uint32_t* d_cnt = rc.take<uint32_t>(nb + 1);
    uint32_t* d_off = rc.take<uint32_t>(nb + 1);
    HIP_CHECK(hipMemcpyAsync(d_order, order.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, st));
    const unsigned rblocks = unsigned((N + 255) / 256);
    vmx::k_prepare<T, D><<<rblocks, 256, 0, st>>>(tr, d_order, rg, db, mb, d_lo, d_hi, d_res);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemsetAsync(d_cnt, 0, sizeof(uint32_t) * (nb + 1), st));
    vmx::k_stratify<false><<<rblocks, 256, 0, st>>>(rg.box, rg.n, d_res, db, d_cnt, nullptr, nullptr);
    HIP_CHECK(hipGetLastError());
    vmx::k_scan<<<1, 1024, 0, st>>>(d_cnt, nb, d_off);
    HIP_CHECK(hipGetLastError());
    uint32_t pairs = 0;
    HIP_CHECK(hipMemcpyAsync(&pairs, d_off + nb, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    uint32_t* d_lst = static_cast<uint32_t*>(ensure(ctx->lists, sizeof(uint32_t) * (size_t(pairs) + 1)));
    HIP_CHECK(hipMemsetAsync(d_cnt, 0, sizeof(uint32_t) * (nb + 1), st));
    vmx::k_stratify<true><<<rblocks, 256, 0, st>>>(rg.box, rg.n, d_res, db, d_cnt, d_off, d_lst);
    HIP_CHECK(hipGetLastError());
    const unsigned bblocks = unsigned((nb + 255) / 256);
    vmx::k_sort_lists<<<bblocks, 256, 0, st>>>(d_off, nb, d_lst);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st));
    const double ms_str = ms_since(t_str);

    // ---------------- 4. bins
    const auto t_bins = std::chrono::steady_clock::now();
    vmx::BinArgs<T> ba{};
    for (int d = 0; d < D; ++d) {
        ba.lo[d] = lo[d];
        ba.hi[d] = hi[d];
    }
    for (int i = 0; i < 3; ++i) {
        ba.res[i] = res3[i];
        ba.drange[i] = i < db ? (hi[i] - lo[i]) / T(p.res[i]) : T(0);
    }
    ba.nbins = nb;
    ba.spp = p.spp;
    ba.seed_lo = uint32_t(p.seed);
    ba.seed_hi = uint32_t(p.seed >> 32);
    ba.stream = 0;
    ba.db = db;
    ba.mb = mb;
    ba.form = p.form;
    ba.alpha = p.alpha;
    HIP_CHECK(hipMemsetAsync(d_flags, 0, sizeof(int32_t), st));
    HIP_CHECK(hipEventRecord(ctx->ev[0], st));
    vmx::k_bins<T, D, P, Math><<<bblocks, 256, 0, st>>>(ba, rg, d_off, d_lst, d_f, bins_dev, d_flags);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(ctx->ev[1], st));
    int32_t flags = 0;
    HIP_CHECK(hipMemcpyAsync(&flags, d_flags, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    float ms_kernel_bins = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms_kernel_bins, ctx->ev[0], ctx->ev[1]));
    const double ms_bins = ms_kernel_bins;
    (void)t_bins;
    if (flags & vmx::FLAG_EMPTY_BIN) throw VmxError(VMX_ERR_NUMERIC, "a bin overlaps no region while spp > 0");

    // region list export, gathered lazily from the device buffers of this call
    ctx->last_D = D;
    ctx->last_n = N;
    ctx->last_err.resize(N);
    ctx->last_dim.resize(N);
    for (uint64_t i = 0; i < N; ++i) {
        ctx->last_err[i] = double(hn.err[order[i]]);
        ctx->last_dim[i] = hn.dim[order[i]];
    }
    ctx->last_gather = [ctx, tr, rg, d_order, N](double* olo, double* ohi, double* onodes) {
        const size_t total = size_t(N) * (2 * D + K);
        Carve c(ensure(ctx->misc, sizeof(T) * total + 1024));
        T* tmp = c.take<T>(total);
        const unsigned nb1 = unsigned((N * K + 255) / 256);
        vmx::k_gather<T><<<nb1, 256, 0, ctx->stream>>>(tr.v, d_order, uint32_t(N), K, tmp);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(tmp + N * K, rg.lo, sizeof(T) * N * D, hipMemcpyDeviceToDevice, ctx->stream));
        HIP_CHECK(hipMemcpyAsync(tmp + N * K + N * D, rg.hi, sizeof(T) * N * D, hipMemcpyDeviceToDevice, ctx->stream));
        std::vector<T> h(total);
        HIP_CHECK(hipMemcpyAsync(h.data(), tmp, sizeof(T) * total, hipMemcpyDeviceToHost, ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (size_t i = 0; i < size_t(N) * K; ++i)
            if (onodes) onodes[i] = double(h[i]);
        for (size_t i = 0; i < size_t(N) * D; ++i) {
            if (olo) olo[i] = double(h[N * K + i]);
            if (ohi) ohi[i] = double(h[N * K + N * D + i]);
        }
    };
    if (stats) {
        stats->regions = N;
        stats->nodes = nnodes;
        stats->rounds = rounds;
        stats->pairs = pairs;
        stats->footprints = N;
        stats->evals_regions = uint64_t(K) + uint64_t(nnodes - 1) * vmx::pow3(D - 1);
        stats->evals_residual = nb * p.spp;
        stats->ties_unresolved = ties;
        stats->ms_generate = ms_gen;
        stats->ms_select = ms_sel;
        stats->ms_stratify = ms_str;
        stats->ms_bins = ms_bins;
        stats->ms_total = ms_since(t_start);
    }
    return VMX_OK;
}

template <typename T, typename Math>
int dispatch_T(vmx_ctx* ctx, const vmx_problem& p, const vmx_integrand& f, T* bins, hipStream_t st, vmx_stats* s) {
    const int D = f.dim;
    switch (f.family) {
        case VMX_FAMILY_POLY:
            switch (D) {
                case 1: return run<T, 1, vmx::Poly<T>, Math>(ctx, p, f.params, bins, st, s);
                case 2: return run<T, 2, vmx::Poly<T>, Math>(ctx, p, f.params, bins, st, s);
                case 3: return run<T, 3, vmx::Poly<T>, Math>(ctx, p, f.params, bins, st, s);
                case 4: return run<T, 4, vmx::Poly<T>, Math>(ctx, p, f.params, bins, st, s);
                case 5: return run<T, 5, vmx::Poly<T>, Math>(ctx, p, f.params, bins, st, s);
                default: break;
            }
            break;
        case VMX_FAMILY_BLOBS:
            if (D == 2) return run<T, 2, vmx::Blobs<T>, Math>(ctx, p, f.params, bins, st, s);
            if (D == 3) return run<T, 3, vmx::Blobs<T>, Math>(ctx, p, f.params, bins, st, s);
            break;
        case VMX_FAMILY_LOBES5:
            if (D == 5) return run<T, 5, vmx::Lobes5<T>, Math>(ctx, p, f.params, bins, st, s);
            break;
        default: break;
    }
    throw VmxError(VMX_ERR_UNSUPPORTED, "integrand family / dimension not built");
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
    for (vmx_ctx::Buf* b : {&ctx->tree, &ctx->regs, &ctx->lists, &ctx->misc, &ctx->params, &ctx->bins})
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
        if (stats) std::memset(stats, 0, sizeof(*stats));
        if (p->scalar == VMX_F32)
            return dispatch_T<float, vmx::FastMath>(ctx, *p, *f, static_cast<float*>(bins_device), st, stats);
        return dispatch_T<double, vmx::PreciseMath>(ctx, *p, *f, static_cast<double*>(bins_device), st, stats);
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
        if (stats) std::memset(stats, 0, sizeof(*stats));
        int rc = (p->scalar == VMX_F32)
                     ? dispatch_T<float, vmx::FastMath>(ctx, *p, *f, static_cast<float*>(d), ctx->stream, stats)
                     : dispatch_T<double, vmx::PreciseMath>(ctx, *p, *f, static_cast<double*>(d), ctx->stream, stats);
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
        if (err) std::memcpy(err, ctx->last_err.data(), sizeof(double) * N);
        if (dim) std::memcpy(dim, ctx->last_dim.data(), sizeof(int32_t) * N);
        if ((lo || hi || nodes) && ctx->last_gather) ctx->last_gather(lo, hi, nodes);
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

/*
 * C ABI of the MI355X per-bin control-variate integrator: validation, grouping of
 * batches by (scalar, D, family), plain Monte Carlo, device evaluation, the priority
 * network's training and inference entry points.  The batch pipeline itself is
 * run_batch.h.
 */
#include <mutex>
#include <tuple>

#include "engine_core.h"
#include "exact.h"

namespace vmxe {

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
                case 6: return run_batch<T, 6, vmx::Poly<T>, Math>(ctx, jobs, st, pairs);
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
        case VMX_FAMILY_CUSTOM:
            // a user-defined functor: its pipeline is instantiated in the caller's own
            // translation unit, reached through the descriptor's hook
            return static_cast<const vmx_custom_integrand*>(jobs[0].params)
                ->run(ctx, int32_t(sizeof(T)), &jobs, st, pairs);
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
                // a user-defined inner integrand (of D dims, or of a sequence over an
                // infinite range): the Fubini instantiation its own translation unit made
                if (fam == VMX_FAMILY_CUSTOM) {
                    const auto* c = static_cast<const vmx_custom_integrand*>(fp->inner);
                    if (!c->run_fubini)
                        throw VmxError(VMX_ERR_UNSUPPORTED,
                                       "user integrand not instantiated for the Fubini form (device_integrand_fubini)");
                    if (pairs) throw VmxError(VMX_ERR_UNSUPPORTED, "training pairs: not for the Fubini form");
                    return c->run_fubini(ctx, D, &jobs, st);
                }
                // over an infinite range: the Dyson series, 1..4 region dims
                if (fam == VMX_FAMILY_DYSON && dd == VMX_INFINITE_DIM) {
                    if (D == 1)
                        return run_batch<T, 1, vmx::Fubini<T, vmx::Dyson<T>, VMX_INFINITE_DIM, 1>, Math>(ctx, jobs, st,
                                                                                                       pairs);
                    if (D == 2)
                        return run_batch<T, 2, vmx::Fubini<T, vmx::Dyson<T>, VMX_INFINITE_DIM, 2>, Math>(ctx, jobs, st,
                                                                                                       pairs);
                    if (D == 3)
                        return run_batch<T, 3, vmx::Fubini<T, vmx::Dyson<T>, VMX_INFINITE_DIM, 3>, Math>(ctx, jobs, st,
                                                                                                       pairs);
                    if (D == 4)
                        return run_batch<T, 4, vmx::Fubini<T, vmx::Dyson<T>, VMX_INFINITE_DIM, 4>, Math>(ctx, jobs, st,
                                                                                                       pairs);
                }
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
        // what else must agree within a group: the inner integrand of a Fubini form,
        // the instantiation hook of a user-defined functor
        auto inner = [](const vmx_integrand* f) {
            if (f->family == VMX_FAMILY_CUSTOM)
                return std::make_tuple(0, 0,
                                       reinterpret_cast<const void*>(static_cast<const vmx_custom_integrand*>(f->params)->run));
            if (f->family != VMX_FAMILY_FUBINI) return std::make_tuple(0, 0, static_cast<const void*>(nullptr));
            const auto* fp = static_cast<const vmx_fubini_params*>(f->params);
            const void* hook = nullptr;
            if (fp->inner_family == VMX_FAMILY_CUSTOM)
                hook = reinterpret_cast<const void*>(static_cast<const vmx_custom_integrand*>(fp->inner)->run_fubini);
            return std::make_tuple(fp->inner_family, fp->inner_dim, hook);
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
    if (p->form < VMX_FORM_VR_OPTIMIZE || p->form > VMX_FORM_VR_FIXED_STRATIFIED)
        throw VmxError(VMX_ERR_INVALID, "unknown form");
    if (f->family == VMX_FAMILY_CUSTOM) {
        const auto* c = static_cast<const vmx_custom_integrand*>(f->params);
        if (!c->run || !c->functor) throw VmxError(VMX_ERR_INVALID, "user integrand: null hook or functor");
    }
    if (p->form >= VMX_FORM_VR_OPTIMIZE_RR_INTEGRAL && p->form <= VMX_FORM_VR_FIXED_SAMPLING_RR &&
        f->family == VMX_FAMILY_FUBINI)
        throw VmxError(VMX_ERR_UNSUPPORTED, "region roulettes: not for the Fubini form");
    const bool infinite = f->family == VMX_FAMILY_FUBINI &&
                          static_cast<const vmx_fubini_params*>(f->params)->inner_dim == VMX_INFINITE_DIM;
    if (vmx::form_stratified(p->form) && !infinite)
        throw VmxError(VMX_ERR_UNSUPPORTED,
                       "stratified forms: the Fubini form over an infinite range only (as the reference)");
    if (vmx::form_importance(p->form) &&
        (p->scalar != VMX_F32 || f->dim > vmx::imp::MAX_DIM || f->family == VMX_FAMILY_FUBINI))
        throw VmxError(VMX_ERR_UNSUPPORTED, "importance sampling: FP32 integrands of dimension 1..6");
    if (p->order != VMX_ORDER_CANONICAL && p->order != VMX_ORDER_REFERENCE)
        throw VmxError(VMX_ERR_INVALID, "unknown region order");
    if (p->refiner != VMX_REFINER_HEURISTIC && p->refiner != VMX_REFINER_LEARNED)
        throw VmxError(VMX_ERR_INVALID, "unknown refiner");
    for (int d = 0; d < f->dim; ++d)
        if (!(p->hi[d] > p->lo[d])) throw VmxError(VMX_ERR_INVALID, "range must have hi > lo");
    if (f->family == VMX_FAMILY_FUBINI) {
        const auto* fp = static_cast<const vmx_fubini_params*>(f->params);
        if (!fp->inner) throw VmxError(VMX_ERR_INVALID, "Fubini form: null inner parameters");
        if (fp->inner_family == VMX_FAMILY_CUSTOM && !static_cast<const vmx_custom_integrand*>(fp->inner)->functor)
            throw VmxError(VMX_ERR_INVALID, "Fubini form: user integrand without a functor");
        if (fp->mc_samples < 1 || fp->mc_samples >= 0x80000000u)
            throw VmxError(VMX_ERR_INVALID, "Fubini form: mc_samples must be 1..2^31-1");
        if (infinite) {
            if (fp->inner_family != VMX_FAMILY_DYSON && fp->inner_family != VMX_FAMILY_CUSTOM)
                throw VmxError(VMX_ERR_UNSUPPORTED,
                               "Fubini form over an infinite range: the Dyson series or a user integrand of a sequence");
            if (f->dim > 4) throw VmxError(VMX_ERR_INVALID, "Fubini form over an infinite range: 1..4 region dims");
            if (fp->rest_count < 0 || fp->rest_count > VMX_MAX_DIM - f->dim)
                throw VmxError(VMX_ERR_INVALID, "Fubini form over an infinite range: rest_count must be 0..6 - dim");
            for (int j = 0; j < fp->rest_count; ++j)
                if (!(fp->rest_hi[f->dim + j] > fp->rest_lo[f->dim + j]))
                    throw VmxError(VMX_ERR_INVALID, "Fubini form: rest range must have hi > lo");
            if (fp->inner_family == VMX_FAMILY_DYSON) {
                const double decay = static_cast<const vmx_dyson_params*>(fp->inner)->decay;
                if (!(decay > 0.0 && decay < 1.0))
                    throw VmxError(VMX_ERR_INVALID, "Dyson series: decay must be in (0, 1)");
            }
        } else {
            if (fp->inner_family == VMX_FAMILY_DYSON)
                throw VmxError(VMX_ERR_INVALID, "Dyson series: only over an infinite range (inner_dim VMX_INFINITE_DIM)");
            if (fp->inner_dim <= f->dim || fp->inner_dim > 5 || f->dim > 3)
                throw VmxError(VMX_ERR_INVALID, "Fubini form: 1..3 region dims below the inner integrand's 2..5");
            for (int d = f->dim; d < fp->inner_dim; ++d)
                if (!(fp->rest_hi[d] > fp->rest_lo[d]))
                    throw VmxError(VMX_ERR_INVALID, "Fubini form: rest range must have hi > lo");
        }
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
            case 6: launch(PolyTag{}, std::integral_constant<int, 6>{}); break;
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
    if (!ok)
        throw VmxError(VMX_ERR_UNSUPPORTED, f.family == VMX_FAMILY_CUSTOM
                                                ? "vmx_eval: not for user integrands (call the functor itself)"
                                                : "integrand family / dimension not built");
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
        if (integrands[i].family == VMX_FAMILY_CUSTOM &&
            static_cast<const vmx_custom_integrand*>(integrands[i].params)->run !=
                static_cast<const vmx_custom_integrand*>(integrands[0].params)->run)
            throw VmxError(VMX_ERR_INVALID, "training pairs: one user integrand type per call");
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
                case 6: return mc_run<T, Math, 6, vmx::Poly<T>>(ctx, p, f.params, stratified, bins, se);
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

}  // namespace vmxe

using namespace vmxe;

/* Work memory of vmx_integrate_device_noalloc: every arena of its pipeline is carved
   from this array (one per GPU, held until the process ends), so that entry point
   calls no hipMalloc.  Host code never passes `scratch` itself to a kernel or a copy:
   on the host the name is only a stand-in, and the array's device address on each GPU
   comes from hipGetSymbolAddress (noalloc_scratch below). */
static __device__ __align__(256) unsigned char scratch[VMX_NOALLOC_SCRATCH_BYTES];

namespace {
/* The device address of `scratch` on `device` (the current device), resolved once per
   device and kept. */
hipError_t noalloc_scratch(int device, char** out) {
    static std::mutex mu;
    static std::map<int, char*> addr;
    std::lock_guard<std::mutex> lock(mu);
    const auto it = addr.find(device);
    if (it != addr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return e != hipSuccess ? e : hipErrorInvalidSymbol;
    addr[device] = static_cast<char*>(ptr);
    *out = static_cast<char*>(ptr);
    return hipSuccess;
}

/* For the duration of a no-alloc call: the pipeline's arenas become views into the
   device's scratch array (bump-allocated from its start), and graphs and pinned growth
   are off.  The context's own arenas are put back afterwards. */
struct NoAllocScope {
    vmx_ctx* ctx;
    vmx_ctx::Fixed fixed;
    vmx_ctx::Buf saved[5];
    vmx_ctx::Buf* bufs[5];
    explicit NoAllocScope(vmx_ctx* c) : ctx(c), bufs{&c->tree, &c->regs, &c->lists, &c->misc, &c->bins2} {
        fixed.base = c->scratch;
        fixed.cap = VMX_NOALLOC_SCRATCH_BYTES;
        fixed.used = 0;
        for (int i = 0; i < 5; ++i) {
            saved[i] = *bufs[i];
            *bufs[i] = vmx_ctx::Buf{};
            bufs[i]->fixed = &fixed;
        }
        c->noalloc = true;
    }
    ~NoAllocScope() {
        for (int i = 0; i < 5; ++i) *bufs[i] = saved[i];
        ctx->noalloc = false;
        // the call's regions live in the shared scratch: nothing to export afterwards
        ctx->last_n = 0;
        ctx->last_gather = nullptr;
    }
};
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
            vmx_ctx_destroy(c);
            return VMX_ERR_DEVICE;
        }
    // set-up for vmx_integrate_device_noalloc: the scratch array's address on this device
    // and the pinned staging its readbacks use, so that entry point allocates nothing
    if (noalloc_scratch(device, &c->scratch) != hipSuccess ||
        hipHostMalloc(&c->pinned, 1 << 16, hipHostMallocDefault) != hipSuccess) {
        vmx_ctx_destroy(c);
        return VMX_ERR_DEVICE;
    }
    c->pinned_bytes = 1 << 16;
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
    for (auto* rg : {&ctx->rounds, &ctx->rounds_tail}) {
        if (rg->exec) (void)hipGraphExecDestroy(rg->exec);
        if (rg->graph) (void)hipGraphDestroy(rg->graph);
    }
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

int vmx_integrate_device_noalloc(vmx_ctx* ctx, const vmx_problem* p, const vmx_integrand* f, void* bins_device,
                                 void* stream, vmx_stats* stats) {
    return guarded(ctx, [&]() -> int {
        validate(p, f);
        if (!bins_device) throw VmxError(VMX_ERR_INVALID, "null bins");
        if (!ctx->scratch) throw VmxError(VMX_ERR_DEVICE, "no-alloc scratch not set up");
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
        NoAllocScope scope(ctx);
        try {
            return integrate_groups(ctx, 1, &p, &f, &bins_device, &stats, st);
        } catch (...) {
            // work already queued may still read the scratch: let it finish before the
            // array is handed to the next call
            (void)hipStreamSynchronize(st);
            throw;
        }
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

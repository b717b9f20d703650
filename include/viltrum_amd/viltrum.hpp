/*
 * C++17 drop-in for viltrum's per-bin control-variate path on an MI355X.
 *
 * The call shape of the reference is kept (src/integrate.h:72-173):
 *
 *     viltrum::integrate(integrator, bins, resolution, f, range);   // accessor bins
 *     viltrum::integrate(integrator, bins_vector_of_vectors, f, range);
 *     viltrum::integrate(integrator, bins_vector, f, range);
 *     auto v = viltrum::integrate(integrator, f, range);            // one bin
 *
 * with the integrators of the reference's factories:
 *
 *     integrator_crespo2021(iterations, spp, seed)                      (integrator-crespo2021.h:7-22)
 *     integrator_adaptive_control_variates_parallel(iterations, spp, seed, heuristic)
 *                                                                       (integrator-adaptive-control-variates.h:17-20)
 *     integrator_variance_reduction_fixed(alpha, iterations, spp, seed) (regions-integrator-parallel-variance-reduction.h:127-130)
 *     integrator_variance_reduction_importance[_fixed](...) (region_sampling_importance, region-sampling.h:22-44)
 *     integrator_variance_reduction_sampling_russian_roulette[_fixed](...)
 *                                                (region_sampling_russian_roulette, region-sampling.h:47-84)
 *     integrator_crespo2021_infinite<DIM>(iterations, mc_samples, spp, seed)
 *                                                                       (integrator-crespo2021.h:24-44; float bins)
 *
 * What differs: the integrand is device-evaluable, either a built-in family descriptor
 * (viltrum_amd::blobs, poly, lobes5, dyson) or a user functor / __host__ __device__
 * lambda wrapped by viltrum_amd::device_integrand<D> (viltrum_amd/device.hpp, in a
 * translation unit compiled by hipcc); a plain host lambda, which a GPU cannot call,
 * is a compile-time error, not a silent CPU fallback.  The bin value type picks
 * the evaluation precision (float: FP32 evaluation with FP64 assembly; double: FP64).
 *
 * Include this header instead of <viltrum.h>; `namespace viltrum` is an alias of
 * viltrum_amd unless VILTRUM_AMD_NO_ALIAS is defined (do not include both libraries'
 * headers in one translation unit).  Link with -lviltrum_amd.
 */
#ifndef VILTRUM_AMD_VILTRUM_HPP
#define VILTRUM_AMD_VILTRUM_HPP

#include <array>
#include <cstddef>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../viltrum_amd.h"

namespace viltrum_amd {

/* ---------------------------------------------------------------- ranges */
template <typename T, std::size_t DIM>
class Range {
    std::array<T, DIM> lo_, hi_;

public:
    static constexpr std::size_t dimensions = DIM;
    using value_type = T;
    Range(const std::array<T, DIM>& a, const std::array<T, DIM>& b) : lo_(a), hi_(b) {}
    const std::array<T, DIM>& min() const { return lo_; }
    const std::array<T, DIM>& max() const { return hi_; }
    T min(std::size_t i) const { return lo_[i]; }
    T max(std::size_t i) const { return hi_[i]; }
    T volume() const {
        T v = T(1);
        for (std::size_t i = 0; i < DIM; ++i) v *= (hi_[i] - lo_[i]);
        return v;
    }
};

template <typename T, std::size_t DIM>
Range<T, DIM> range(const std::array<T, DIM>& a, const std::array<T, DIM>& b) {
    return Range<T, DIM>(a, b);
}

template <std::size_t N, typename T = float>
Range<T, N> range_primary() {
    std::array<T, N> a, b;
    a.fill(T(0));
    b.fill(T(1));
    return Range<T, N>(a, b);
}

/* A range of unbounded dimension (src/range-infinite.h:16-87): explicit entries
   [a[i], b[i]], then [0, 1] for every entry after them. */
template <typename T>
class RangeInfinite {
    std::vector<T> lo_, hi_;

public:
    using value_type = T;
    RangeInfinite(const std::vector<T>& a = std::vector<T>(), const std::vector<T>& b = std::vector<T>())
        : lo_(a), hi_(b) {
        if (a.size() != b.size()) throw std::invalid_argument("viltrum_amd: range_infinite entries differ");
    }
    const std::vector<T>& min() const { return lo_; }
    const std::vector<T>& max() const { return hi_; }
    T min(std::size_t i) const { return i < lo_.size() ? lo_[i] : T(0); }
    T max(std::size_t i) const { return i < hi_.size() ? hi_[i] : T(1); }
    T volume() const {
        T v = T(1);
        for (std::size_t i = 0; i < lo_.size(); ++i) v *= (hi_[i] - lo_[i]);
        return v;
    }
};

template <typename T>
RangeInfinite<T> range_infinite(const std::vector<T>& a, const std::vector<T>& b) {
    return RangeInfinite<T>(a, b);
}

template <typename T = float>
RangeInfinite<T> range_primary_infinite() {
    return RangeInfinite<T>();
}

template <typename R>
struct is_range : std::false_type {};
template <typename T, std::size_t DIM>
struct is_range<Range<T, DIM>> : std::true_type {};
template <typename T>
struct is_range<RangeInfinite<T>> : std::true_type {};

/* ---------------------------------------------------------------- integrands */
struct Integrand {
    vmx_integrand desc{};
    std::shared_ptr<void> params;
};

inline Integrand blobs(const vmx_blobs_params& p, int dim) {
    Integrand f;
    auto q = std::make_shared<vmx_blobs_params>(p);
    f.params = q;
    f.desc = vmx_integrand{VMX_FAMILY_BLOBS, dim, q.get()};
    return f;
}
inline Integrand poly(const vmx_poly_params& p, int dim) {
    Integrand f;
    auto q = std::make_shared<vmx_poly_params>(p);
    f.params = q;
    f.desc = vmx_integrand{VMX_FAMILY_POLY, dim, q.get()};
    return f;
}
inline Integrand lobes5(const vmx_lobes5_params& p) {
    Integrand f;
    auto q = std::make_shared<vmx_lobes5_params>(p);
    f.params = q;
    f.desc = vmx_integrand{VMX_FAMILY_LOBES5, 5, q.get()};
    return f;
}

/* The Dyson series of a sequence (the reference's infinite-range example integrand,
   main/compilation-tests/crespo21.cc:6-25); integrated over an infinite range only. */
inline Integrand dyson(double decay = 0.75) {
    Integrand f;
    auto q = std::make_shared<vmx_dyson_params>(vmx_dyson_params{decay});
    f.params = q;
    f.desc = vmx_integrand{VMX_FAMILY_DYSON, VMX_INFINITE_DIM, q.get()};
    return f;
}

/* ---------------------------------------------------------------- context */
class Context {
    vmx_ctx* ctx_ = nullptr;

public:
    explicit Context(int device = 0) {
        if (vmx_ctx_create(device, &ctx_) != VMX_OK) throw std::runtime_error("viltrum_amd: no usable GPU");
    }
    ~Context() { vmx_ctx_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    vmx_ctx* get() const { return ctx_; }
    void check(int rc) const {
        if (rc != VMX_OK) throw std::runtime_error(std::string("viltrum_amd: ") + vmx_last_error(ctx_));
    }
    /* one context per host thread (a context is not shared between threads) */
    static Context& thread_default() {
        thread_local Context c(0);
        return c;
    }
};

class LoggerNull {
public:
    template <typename Data>
    void log(const Data&) {}
};

/* A region list in the layout of vmx_export_regions / vmx_integrate_regions: boxes
   count x dim, node values count x 3^dim (dim 0 fastest), error and split dim. */
struct RegionList {
    std::size_t dim = 0, count = 0;
    std::vector<double> lo, hi, nodes, err;
    std::vector<int32_t> split_dim;
};

/* What an integrator hands to logger.log after generation, in place of the reference's
   region sequence (src/newton-cotes/integrator-region-based.h:19): the list is read
   from the device only if the logger asks for it, and only until the context's next
   call. */
class RegionLog {
    Context* ctx_;
    std::size_t dim_;

public:
    RegionLog(Context& ctx, std::size_t dim) : ctx_(&ctx), dim_(dim) {}
    RegionList fetch() const {
        RegionList r;
        r.dim = dim_;
        uint64_t n = 0;
        ctx_->check(vmx_export_regions(ctx_->get(), &n, nullptr, nullptr, nullptr, nullptr, nullptr));
        std::size_t k = 1;
        for (std::size_t d = 0; d < dim_; ++d) k *= 3;
        r.count = n;
        r.lo.resize(n * dim_);
        r.hi.resize(n * dim_);
        r.nodes.resize(n * k);
        r.err.resize(n);
        r.split_dim.resize(n);
        ctx_->check(vmx_export_regions(ctx_->get(), &n, r.lo.data(), r.hi.data(), r.nodes.data(), r.err.data(),
                                       r.split_dim.data()));
        return r;
    }
};

/* ---------------------------------------------------------------- integrators */
class IntegratorGpu {
    vmx_problem base_{};

public:
    IntegratorGpu(std::size_t iterations, std::size_t spp, std::size_t seed, int heuristic, int form, double alpha,
                  int order = VMX_ORDER_CANONICAL) {
        std::memset(&base_, 0, sizeof(base_));
        base_.heuristic = heuristic;
        base_.form = form;
        base_.order = order;
        base_.refiner = VMX_REFINER_HEURISTIC;
        base_.alpha = alpha;
        base_.regions = iterations + 1;
        base_.spp = spp;
        base_.seed = seed;
    }
    const vmx_problem& base() const { return base_; }

    /* IntegratorRegionBased::integrate (src/newton-cotes/integrator-region-based.h:13-22):
       bins(pos) is assigned (variance-reduction forms) or accumulated (alpha = 1 form),
       pos in dim-0-outer C order. */
    template <typename Bins, std::size_t DB, typename Float, std::size_t DIM, typename Logger>
    void integrate(Bins& bins, const std::array<std::size_t, DB>& res, const Integrand& f,
                   const Range<Float, DIM>& range, Logger& logger, Context& ctx = Context::thread_default()) const {
        static_assert(DB >= 1 && DB <= 3 && DB <= DIM, "1..3 bin dimensions, no more than the integrand's");
        using T = std::decay_t<decltype(bins(std::array<std::size_t, DB>()))>;
        static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                      "bins must hold float (FP32 path) or double (FP64 path)");
        if (f.desc.dim != int(DIM)) throw std::invalid_argument("viltrum_amd: range and integrand dims differ");
        vmx_problem p = base_;
        p.scalar = sizeof(T) == 4 ? VMX_F32 : VMX_F64;
        p.bin_dims = int(DB);
        for (std::size_t d = 0; d < DIM; ++d) {
            p.lo[d] = double(range.min(d));
            p.hi[d] = double(range.max(d));
        }
        std::size_t nb = 1;
        for (std::size_t i = 0; i < DB; ++i) {
            p.res[i] = res[i];
            nb *= res[i];
        }
        std::vector<T> flat(nb, T(0));
        auto for_each_pos = [&](auto&& fn) {
            std::array<std::size_t, DB> pos{};
            for (std::size_t k = 0; k < nb; ++k) {
                std::size_t rem = k;
                for (std::size_t i = DB; i-- > 0;) {
                    pos[i] = rem % res[i];
                    rem /= res[i];
                }
                fn(k, pos);
            }
        };
        if (p.form == VMX_FORM_CV_ALPHA1) for_each_pos([&](std::size_t k, const auto& pos) { flat[k] = bins(pos); });
        vmx_stats st{};
        ctx.check(vmx_integrate(ctx.get(), &p, &f.desc, flat.data(), &st));
        for_each_pos([&](std::size_t k, const auto& pos) { bins(pos) = flat[k]; });
        logger.log(st);
        logger.log(RegionLog(ctx, DIM));
    }

    /* RegionsIntegrator::integrate_regions
       (src/control-variates/regions-integrator-parallel-variance-reduction.h:32-34): this
       integrator's per-bin phase (form, spp, seed) over a given region list, e.g. one a
       RegionLog fetched, replayed at another spp, seed or resolution. */
    template <typename Bins, std::size_t DB, typename Float, std::size_t DIM, typename Logger>
    void integrate_regions(Bins& bins, const std::array<std::size_t, DB>& res, const RegionList& regions,
                           const Integrand& f, const Range<Float, DIM>& range, Logger& logger,
                           Context& ctx = Context::thread_default()) const {
        static_assert(DB >= 1 && DB <= 3 && DB <= DIM, "1..3 bin dimensions, no more than the integrand's");
        using T = std::decay_t<decltype(bins(std::array<std::size_t, DB>()))>;
        static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                      "bins must hold float (FP32 path) or double (FP64 path)");
        if (f.desc.dim != int(DIM) || regions.dim != DIM)
            throw std::invalid_argument("viltrum_amd: range, integrand and region dims differ");
        vmx_problem p = base_;
        p.scalar = sizeof(T) == 4 ? VMX_F32 : VMX_F64;
        p.bin_dims = int(DB);
        for (std::size_t d = 0; d < DIM; ++d) {
            p.lo[d] = double(range.min(d));
            p.hi[d] = double(range.max(d));
        }
        std::size_t nb = 1;
        for (std::size_t i = 0; i < DB; ++i) {
            p.res[i] = res[i];
            nb *= res[i];
        }
        std::vector<T> flat(nb, T(0));
        auto for_each_pos = [&](auto&& fn) {
            std::array<std::size_t, DB> pos{};
            for (std::size_t k = 0; k < nb; ++k) {
                std::size_t rem = k;
                for (std::size_t i = DB; i-- > 0;) {
                    pos[i] = rem % res[i];
                    rem /= res[i];
                }
                fn(k, pos);
            }
        };
        if (p.form == VMX_FORM_CV_ALPHA1) for_each_pos([&](std::size_t k, const auto& pos) { flat[k] = bins(pos); });
        vmx_stats st{};
        ctx.check(vmx_integrate_regions(ctx.get(), &p, &f.desc, regions.count, regions.lo.data(), regions.hi.data(),
                                        regions.nodes.data(), flat.data(), &st));
        for_each_pos([&](std::size_t k, const auto& pos) { bins(pos) = flat[k]; });
        logger.log(st);
    }

    template <typename Bins, std::size_t DB, typename F, typename Float, std::size_t DIM, typename Logger>
    void integrate(Bins&, const std::array<std::size_t, DB>&, const F&, const Range<Float, DIM>&, Logger&) const {
        static_assert(std::is_same<F, Integrand>::value,
                      "viltrum_amd integrates device-evaluable integrands: a built-in family (viltrum_amd::blobs, "
                      "poly, lobes5) or viltrum_amd::device_integrand<D>(functor) from viltrum_amd/device.hpp; a "
                      "host-only lambda cannot run on the GPU");
    }
};

inline IntegratorGpu integrator_crespo2021(std::size_t iterations, std::size_t spp, std::size_t seed = 0) {
    return IntegratorGpu(iterations, spp, seed, VMX_HEUR_SIZE_RELATIVE, VMX_FORM_VR_OPTIMIZE, 1.0);
}
inline IntegratorGpu integrator_adaptive_control_variates_parallel(std::size_t iterations, std::size_t spp,
                                                                   std::size_t seed = 0,
                                                                   int heuristic = VMX_HEUR_DEFAULT_ABSOLUTE) {
    return IntegratorGpu(iterations, spp, seed, heuristic, VMX_FORM_CV_ALPHA1, 1.0);
}
inline IntegratorGpu integrator_variance_reduction_fixed(double alpha, std::size_t iterations, std::size_t spp,
                                                         std::size_t seed = 0,
                                                         int heuristic = VMX_HEUR_SIZE_RELATIVE) {
    return IntegratorGpu(iterations, spp, seed, heuristic, VMX_FORM_VR_FIXED, alpha);
}
/* regions_integrator_parallel_variance_reduction(rr_uniform_region(), cv_optimize_weight(),
   region_sampling_importance(), rng, spp) behind the crespo2021 generator
   (src/control-variates/region-sampling.h:22-44); float bins, dims 1..6 */
inline IntegratorGpu integrator_variance_reduction_importance(std::size_t iterations, std::size_t spp,
                                                              std::size_t seed = 0,
                                                              int heuristic = VMX_HEUR_SIZE_RELATIVE) {
    return IntegratorGpu(iterations, spp, seed, heuristic, VMX_FORM_VR_OPTIMIZE_IMPORTANCE, 1.0);
}
/* regions_integrator_parallel_variance_reduction(rr_integral_region(), cv_optimize_weight(),
   region_sampling_uniform(), rng, spp) (src/control-variates/region-russian-roulette.h:30-67) */
inline IntegratorGpu integrator_variance_reduction_rr_integral(std::size_t iterations, std::size_t spp,
                                                               std::size_t seed = 0,
                                                               int heuristic = VMX_HEUR_SIZE_RELATIVE) {
    return IntegratorGpu(iterations, spp, seed, heuristic, VMX_FORM_VR_OPTIMIZE_RR_INTEGRAL, 1.0);
}
/* ... with cv_fixed_weight(alpha) */
inline IntegratorGpu integrator_variance_reduction_rr_integral_fixed(double alpha, std::size_t iterations,
                                                                     std::size_t spp, std::size_t seed = 0,
                                                                     int heuristic = VMX_HEUR_SIZE_RELATIVE) {
    return IntegratorGpu(iterations, spp, seed, heuristic, VMX_FORM_VR_FIXED_RR_INTEGRAL, alpha);
}
/* ... with rr_error_region() (src/control-variates/region-russian-roulette.h:68-104) */
inline IntegratorGpu integrator_variance_reduction_rr_error(std::size_t iterations, std::size_t spp,
                                                            std::size_t seed = 0,
                                                            int heuristic = VMX_HEUR_SIZE_RELATIVE) {
    return IntegratorGpu(iterations, spp, seed, heuristic, VMX_FORM_VR_OPTIMIZE_RR_ERROR, 1.0);
}
inline IntegratorGpu integrator_variance_reduction_rr_error_fixed(double alpha, std::size_t iterations,
                                                                  std::size_t spp, std::size_t seed = 0,
                                                                  int heuristic = VMX_HEUR_SIZE_RELATIVE) {
    return IntegratorGpu(iterations, spp, seed, heuristic, VMX_FORM_VR_FIXED_RR_ERROR, alpha);
}
/* ... with rr_pdf_region() (src/control-variates/region-russian-roulette.h:106-149) */
inline IntegratorGpu integrator_variance_reduction_rr_pdf(std::size_t iterations, std::size_t spp,
                                                          std::size_t seed = 0,
                                                          int heuristic = VMX_HEUR_SIZE_RELATIVE) {
    return IntegratorGpu(iterations, spp, seed, heuristic, VMX_FORM_VR_OPTIMIZE_RR_PDF, 1.0);
}
inline IntegratorGpu integrator_variance_reduction_rr_pdf_fixed(double alpha, std::size_t iterations,
                                                                std::size_t spp, std::size_t seed = 0,
                                                                int heuristic = VMX_HEUR_SIZE_RELATIVE) {
    return IntegratorGpu(iterations, spp, seed, heuristic, VMX_FORM_VR_FIXED_RR_PDF, alpha);
}
/* ... with rr_pdf_region() and region_sampling_importance() (the region by its pdf mass,
   the point from its pdf; main/compilation-tests/multiple-parameters-2d.cc:101-103).  The
   other roulettes with importance sampling: IntegratorGpu with the forms
   VMX_FORM_VR_{OPTIMIZE,FIXED}_RR_{INTEGRAL,ERROR}_IMPORTANCE. */
inline IntegratorGpu integrator_variance_reduction_rr_pdf_importance(std::size_t iterations, std::size_t spp,
                                                                     std::size_t seed = 0,
                                                                     int heuristic = VMX_HEUR_SIZE_RELATIVE) {
    return IntegratorGpu(iterations, spp, seed, heuristic, VMX_FORM_VR_OPTIMIZE_RR_PDF_IMPORTANCE, 1.0);
}
inline IntegratorGpu integrator_variance_reduction_rr_pdf_importance_fixed(double alpha, std::size_t iterations,
                                                                           std::size_t spp, std::size_t seed = 0,
                                                                           int heuristic = VMX_HEUR_SIZE_RELATIVE) {
    return IntegratorGpu(iterations, spp, seed, heuristic, VMX_FORM_VR_FIXED_RR_PDF_IMPORTANCE, alpha);
}
/* regions_integrator_parallel_variance_reduction(rr_uniform_region(), cv_optimize_weight(),
   region_sampling_russian_roulette(), rng, spp) (src/control-variates/region-sampling.h:
   47-84): importance or uniform sampling per sample by a roulette; float bins, dims 1..6 */
inline IntegratorGpu integrator_variance_reduction_sampling_russian_roulette(std::size_t iterations, std::size_t spp,
                                                                             std::size_t seed = 0,
                                                                             int heuristic = VMX_HEUR_SIZE_RELATIVE) {
    return IntegratorGpu(iterations, spp, seed, heuristic, VMX_FORM_VR_OPTIMIZE_SAMPLING_RR, 1.0);
}
inline IntegratorGpu integrator_variance_reduction_sampling_russian_roulette_fixed(
    double alpha, std::size_t iterations, std::size_t spp, std::size_t seed = 0,
    int heuristic = VMX_HEUR_SIZE_RELATIVE) {
    return IntegratorGpu(iterations, spp, seed, heuristic, VMX_FORM_VR_FIXED_SAMPLING_RR, alpha);
}
/* ... with cv_fixed_weight(alpha) */
inline IntegratorGpu integrator_variance_reduction_importance_fixed(double alpha, std::size_t iterations,
                                                                    std::size_t spp, std::size_t seed = 0,
                                                                    int heuristic = VMX_HEUR_SIZE_RELATIVE) {
    return IntegratorGpu(iterations, spp, seed, heuristic, VMX_FORM_VR_FIXED_IMPORTANCE, alpha);
}

/* integrator_crespo2021_infinite<RD> (src/control-variates/integrator-crespo2021.h:24-44):
   crespo2021 over the first RD dims of the range; the remaining dims of the integrand are
   integrated by Monte Carlo, mc_samples per region node value and one per residual
   sample.  Float bins only, as the reference (its rest-dims lambda takes float arrays). */
template <std::size_t RD>
class IntegratorGpuFubini {
    IntegratorGpu base_;
    std::size_t mc_, seed_;

public:
    IntegratorGpuFubini(std::size_t iterations, std::size_t mc_samples, std::size_t spp, std::size_t seed)
        : base_(iterations, spp, seed, VMX_HEUR_SIZE_RELATIVE, VMX_FORM_VR_OPTIMIZE, 1.0), mc_(mc_samples),
          seed_(seed) {}

    template <typename Bins, std::size_t DB, typename Float, std::size_t DIM, typename Logger>
    void integrate(Bins& bins, const std::array<std::size_t, DB>& res, const Integrand& f,
                   const Range<Float, DIM>& range, Logger& logger, Context& ctx = Context::thread_default()) const {
        static_assert(RD >= 1 && RD < DIM && DB <= RD, "region dims 1..DIM-1, bins over region dims");
        using T = std::decay_t<decltype(bins(std::array<std::size_t, DB>()))>;
        static_assert(std::is_same<T, float>::value, "the Fubini form integrates into float bins");
        if (f.desc.dim != int(DIM)) throw std::invalid_argument("viltrum_amd: range and integrand dims differ");
        vmx_fubini_params fp{};
        fp.inner_family = f.desc.family;
        fp.inner_dim = f.desc.dim;
        fp.mc_samples = uint32_t(mc_);
        fp.seed = uint32_t(2 * seed_ + 1);
        for (std::size_t d = 0; d < DIM; ++d) {
            fp.rest_lo[d] = double(range.min(d));
            fp.rest_hi[d] = double(range.max(d));
        }
        fp.inner = f.desc.params;
        Integrand g;
        g.desc = vmx_integrand{VMX_FAMILY_FUBINI, int(RD), &fp};
        std::array<Float, RD> lo, hi;
        for (std::size_t d = 0; d < RD; ++d) {
            lo[d] = range.min(d);
            hi[d] = range.max(d);
        }
        base_.integrate(bins, res, g, Range<Float, RD>(lo, hi), logger, ctx);
    }

    /* over an infinite range: regions and bins over its first RD entries, the sequence
       after them Monte Carlo (the integrand reads an unbounded sequence) */
    template <typename Bins, std::size_t DB, typename Float, typename Logger>
    void integrate(Bins& bins, const std::array<std::size_t, DB>& res, const Integrand& f,
                   const RangeInfinite<Float>& range, Logger& logger, Context& ctx = Context::thread_default()) const {
        integrate_infinite<RD>(base_, uint32_t(2 * seed_ + 1), mc_, bins, res, f, range, logger, ctx);
    }

    template <typename Bins, std::size_t DB, typename F, typename Float, std::size_t DIM, typename Logger>
    void integrate(Bins&, const std::array<std::size_t, DB>&, const F&, const Range<Float, DIM>&, Logger&) const {
        static_assert(std::is_same<F, Integrand>::value,
                      "viltrum_amd integrates device-evaluable integrands: a built-in family (viltrum_amd::blobs, "
                      "poly, lobes5) or viltrum_amd::device_integrand<D>(functor) from viltrum_amd/device.hpp; a "
                      "host-only lambda cannot run on the GPU");
    }

    template <std::size_t R, typename Bins, std::size_t DB, typename Float, typename Logger>
    static void integrate_infinite(const IntegratorGpu& base, uint32_t mc_seed, std::size_t mc, Bins& bins,
                                   const std::array<std::size_t, DB>& res, const Integrand& f,
                                   const RangeInfinite<Float>& range, Logger& logger, Context& ctx) {
        static_assert(R >= 1 && R <= 4 && DB <= R && DB <= 3, "1..4 region dims, bins over region dims");
        using T = std::decay_t<decltype(bins(std::array<std::size_t, DB>()))>;
        static_assert(std::is_same<T, float>::value, "the Fubini form integrates into float bins");
        if (f.desc.dim != VMX_INFINITE_DIM)
            throw std::invalid_argument("viltrum_amd: an infinite range goes with an integrand of a sequence");
        const std::size_t nrest = range.min().size() > R ? range.min().size() - R : 0;
        if (R + nrest > VMX_MAX_DIM) throw std::invalid_argument("viltrum_amd: too many explicit range entries");
        vmx_fubini_params fp{};
        fp.inner_family = f.desc.family;
        fp.inner_dim = VMX_INFINITE_DIM;
        fp.mc_samples = uint32_t(mc);
        fp.seed = mc_seed;
        for (std::size_t d = 0; d < R + nrest; ++d) {
            fp.rest_lo[d] = double(range.min(d));
            fp.rest_hi[d] = double(range.max(d));
        }
        fp.rest_count = int32_t(nrest);
        fp.inner = f.desc.params;
        Integrand g;
        g.desc = vmx_integrand{VMX_FAMILY_FUBINI, int(R), &fp};
        std::array<Float, R> lo, hi;
        for (std::size_t d = 0; d < R; ++d) {
            lo[d] = range.min(d);
            hi[d] = range.max(d);
        }
        base.integrate(bins, res, g, Range<Float, R>(lo, hi), logger, ctx);
    }
};

template <std::size_t RD>
IntegratorGpuFubini<RD> integrator_crespo2021_infinite(std::size_t iterations, std::size_t mc_samples,
                                                       std::size_t spp, std::size_t seed = 0) {
    return IntegratorGpuFubini<RD>(iterations, mc_samples, spp, seed);
}

/* integrator_adaptive_fubini_variance_reduction_parallel_optimized<RD>(nested(simpson,
   trapezoidal), error_heuristic_size(error_metric_relative(), 1e-5), iterations,
   mc_samples, region_stratification_uniform(), cv_optimize_weight() or
   cv_fixed_weight(alpha), region_sampling_uniform(), spp, seed)
   (src/control-variates/integrator-adaptive-fubini-variance-reduction-optimized.h:17-23):
   node values Monte Carlo from seed, residual samples from seed + 1, spread evenly over
   each bin's regions.  Over infinite ranges only, as the reference builds it. */
template <std::size_t RD>
class IntegratorGpuFubiniOptimized {
    IntegratorGpu base_;
    std::size_t mc_, seed_;

public:
    IntegratorGpuFubiniOptimized(std::size_t iterations, std::size_t mc_samples, std::size_t spp, std::size_t seed,
                                 int form, double alpha)
        : base_(iterations, spp, seed + 1, VMX_HEUR_SIZE_RELATIVE, form, alpha), mc_(mc_samples), seed_(seed) {}

    template <typename Bins, std::size_t DB, typename Float, typename Logger>
    void integrate(Bins& bins, const std::array<std::size_t, DB>& res, const Integrand& f,
                   const RangeInfinite<Float>& range, Logger& logger, Context& ctx = Context::thread_default()) const {
        IntegratorGpuFubini<RD>::template integrate_infinite<RD>(base_, uint32_t(seed_), mc_, bins, res, f, range,
                                                                 logger, ctx);
    }
};

template <std::size_t RD>
IntegratorGpuFubiniOptimized<RD> integrator_adaptive_fubini_variance_reduction_parallel_optimized(
    std::size_t iterations, std::size_t mc_samples, std::size_t spp, std::size_t seed = 0) {
    return IntegratorGpuFubiniOptimized<RD>(iterations, mc_samples, spp, seed, VMX_FORM_VR_OPTIMIZE_STRATIFIED, 1.0);
}

/* the same with cv_fixed_weight(alpha) */
template <std::size_t RD>
IntegratorGpuFubiniOptimized<RD> integrator_adaptive_fubini_variance_reduction_parallel_optimized_fixed(
    double alpha, std::size_t iterations, std::size_t mc_samples, std::size_t spp, std::size_t seed = 0) {
    return IntegratorGpuFubiniOptimized<RD>(iterations, mc_samples, spp, seed, VMX_FORM_VR_FIXED_STRATIFIED, alpha);
}

/* ---------------------------------------------------------------- integrate overloads
   (src/integrate.h:72-79, 93-97, 105-130, 132-148, 169-173) */
template <typename Integrator, typename Bins, std::size_t DB, typename Rg, typename Logger>
void integrate(const Integrator& integrator, Bins& bins, const std::array<std::size_t, DB>& resolution,
               const Integrand& f, const Rg& range, Logger& logger) {
    integrator.integrate(bins, resolution, f, range, logger);
}

template <typename Integrator, typename Bins, std::size_t DB, typename Rg>
void integrate(const Integrator& integrator, Bins& bins, const std::array<std::size_t, DB>& resolution,
               const Integrand& f, const Rg& range) {
    LoggerNull log;
    integrate(integrator, bins, resolution, f, range, log);
}

template <typename Integrator, typename T, typename Rg>
void integrate(const Integrator& integrator, std::vector<std::vector<T>>& bins, const Integrand& f,
               const Rg& range) {
    std::size_t max1 = 0;
    for (const auto& v : bins) max1 = v.size() > max1 ? v.size() : max1;
    for (auto& v : bins) v.resize(max1);
    auto b = [&bins](const std::array<std::size_t, 2>& i) -> T& { return bins[i[0]][i[1]]; };
    const std::array<std::size_t, 2> res{bins.size(), max1};
    integrate(integrator, b, res, f, range);
}

template <typename Integrator, typename T, typename Rg, typename Logger>
void integrate(const Integrator& integrator, std::vector<std::vector<T>>& bins, const Integrand& f,
               const Rg& range, Logger& logger) {
    std::size_t max1 = 0;
    for (const auto& v : bins) max1 = v.size() > max1 ? v.size() : max1;
    for (auto& v : bins) v.resize(max1);
    auto b = [&bins](const std::array<std::size_t, 2>& i) -> T& { return bins[i[0]][i[1]]; };
    const std::array<std::size_t, 2> res{bins.size(), max1};
    integrate(integrator, b, res, f, range, logger);
}

/* src/integrate.h:150-167: three bin dims, inner vectors padded to the longest */
template <typename Integrator, typename T, typename Rg, typename Logger>
void integrate(const Integrator& integrator, std::vector<std::vector<std::vector<T>>>& bins, const Integrand& f,
               const Rg& range, Logger& logger) {
    std::size_t max1 = 0, max2 = 0;
    for (const auto& v : bins) max1 = v.size() > max1 ? v.size() : max1;
    for (auto& v : bins) v.resize(max1);
    for (const auto& vv : bins)
        for (const auto& v : vv) max2 = v.size() > max2 ? v.size() : max2;
    for (auto& vv : bins)
        for (auto& v : vv) v.resize(max2);
    auto b = [&bins](const std::array<std::size_t, 3>& i) -> T& { return bins[i[0]][i[1]][i[2]]; };
    const std::array<std::size_t, 3> res{bins.size(), max1, max2};
    integrate(integrator, b, res, f, range, logger);
}

template <typename Integrator, typename T, typename Rg>
void integrate(const Integrator& integrator, std::vector<std::vector<std::vector<T>>>& bins, const Integrand& f,
               const Rg& range) {
    LoggerNull log;
    integrate(integrator, bins, f, range, log);
}

template <typename Integrator, typename T, typename Rg, typename Logger>
void integrate(const Integrator& integrator, std::vector<T>& bins, const Integrand& f, const Rg& range,
               Logger& logger) {
    auto b = [&bins](const std::array<std::size_t, 1>& i) -> T& { return bins[i[0]]; };
    const std::array<std::size_t, 1> res{bins.size()};
    integrate(integrator, b, res, f, range, logger);
}

template <typename Integrator, typename T, typename Rg>
void integrate(const Integrator& integrator, std::vector<T>& bins, const Integrand& f, const Rg& range) {
    auto b = [&bins](const std::array<std::size_t, 1>& i) -> T& { return bins[i[0]]; };
    const std::array<std::size_t, 1> res{bins.size()};
    integrate(integrator, b, res, f, range);
}

template <typename Integrator, typename Rg, typename = std::enable_if_t<is_range<Rg>::value>>
typename Rg::value_type integrate(const Integrator& integrator, const Integrand& f, const Rg& range) {
    using Float = typename Rg::value_type;
    Float sol(0);
    auto b = [&sol](const std::array<std::size_t, 1>&) -> Float& { return sol; };
    const std::array<std::size_t, 1> res{1};
    integrate(integrator, b, res, f, range);
    return sol;
}

}  // namespace viltrum_amd

#ifndef VILTRUM_AMD_NO_ALIAS
namespace viltrum = viltrum_amd;
#endif

#endif

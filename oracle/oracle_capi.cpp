/*
 * TEST INFRASTRUCTURE ONLY -- C entry points of the oracle (liboracle.so), called from
 * tests/ and from bench.py's cpu_baseline leg through ctypes.  Never linked by the
 * product.  Dispatches (scalar type, D, integrand family) onto oracle.hpp.
 */
#include <chrono>
#include <cstdio>
#include <string>

#include "oracle.hpp"

namespace {

thread_local std::string g_err;

/* Integrand wrapper; optionally logs every evaluation (x, f(x)) in call order. */
template <typename T, int D, typename P>
struct Fn {
    P p;
    std::vector<double>* log = nullptr;
    T operator()(const T* x) const {
        const T y = p.template eval<vmo::HostMath, D>(x);
        if (log) {
            for (int d = 0; d < D; ++d) log->push_back(double(x[d]));
            log->push_back(double(y));
        }
        return y;
    }
};

template <typename T, int D, typename Body>
int with_f(const vmx_integrand* f, Body&& body, std::vector<double>* log = nullptr) {
    if (f->family == VMX_FAMILY_POLY) {
        Fn<T, D, vmx::Poly<T>> fn{vmx::Poly<T>::from(*static_cast<const vmx_poly_params*>(f->params)), log};
        return body(fn);
    }
    if constexpr (D == 2 || D == 3) {
        if (f->family == VMX_FAMILY_BLOBS) {
            Fn<T, D, vmx::Blobs<T>> fn{vmx::Blobs<T>::from(*static_cast<const vmx_blobs_params*>(f->params)), log};
            return body(fn);
        }
    }
    if constexpr (D == 5) {
        if (f->family == VMX_FAMILY_LOBES5) {
            Fn<T, D, vmx::Lobes5<T>> fn{vmx::Lobes5<T>::from(*static_cast<const vmx_lobes5_params*>(f->params)),
                                        log};
            return body(fn);
        }
    }
    g_err = "unsupported integrand family for this dimension";
    return -2;
}

template <template <typename, int> class Job, typename... A>
int dispatch(int scalar_bytes, int D, A&&... a) {
    try {
#define VMO_CASE(DD)                                                              \
    case DD:                                                                      \
        return scalar_bytes == 4 ? Job<float, DD>::run(std::forward<A>(a)...)    \
                                 : Job<double, DD>::run(std::forward<A>(a)...);
        if (scalar_bytes != 4 && scalar_bytes != 8) {
            g_err = "scalar_bytes must be 4 or 8";
            return -1;
        }
        switch (D) {
            VMO_CASE(1)
            VMO_CASE(2)
            VMO_CASE(3)
            VMO_CASE(4)
            VMO_CASE(5)
            VMO_CASE(6)
            default:
                g_err = "D must be in 1..6";
                return -1;
        }
#undef VMO_CASE
    } catch (const std::exception& e) {
        g_err = e.what();
        return -3;
    }
}

template <typename T, int D>
void load_box(const double* lo, const double* hi, T* tlo, T* thi) {
    for (int d = 0; d < D; ++d) {
        tlo[d] = T(lo[d]);
        thi[d] = T(hi[d]);
    }
}

template <typename T, int D>
void store_regions(const std::vector<vmo::Region<T, D>>& regs, double* out_lo, double* out_hi, double* out_nodes,
                   double* out_err, int32_t* out_dim) {
    for (size_t r = 0; r < regs.size(); ++r) {
        for (int d = 0; d < D; ++d) {
            if (out_lo) out_lo[r * D + d] = regs[r].lo[d];
            if (out_hi) out_hi[r * D + d] = regs[r].hi[d];
        }
        if (out_nodes)
            for (int k = 0; k < vmo::Region<T, D>::K; ++k) out_nodes[r * vmo::Region<T, D>::K + k] = regs[r].v[k];
        if (out_err) out_err[r] = regs[r].err;
        if (out_dim) out_dim[r] = regs[r].dim;
    }
}

template <typename T, int D>
struct GenerateJob {
    static int run(const vmx_integrand* f, const double* lo, const double* hi, int heuristic, uint64_t iterations,
                   double* out_lo, double* out_hi, double* out_nodes, double* out_err, int32_t* out_dim,
                   double* out_evals, uint64_t* out_nevals) {
        T tlo[D], thi[D];
        load_box<T, D>(lo, hi, tlo, thi);
        std::vector<double> log;
        int rc = with_f<T, D>(
            f,
            [&](const auto& fn) {
                auto regs = vmo::generate<T, D>(fn, tlo, thi, heuristic, iterations);
                store_regions(regs, out_lo, out_hi, out_nodes, out_err, out_dim);
                return int(regs.size());
            },
            out_evals ? &log : nullptr);
        if (out_evals) {
            std::copy(log.begin(), log.end(), out_evals);
            *out_nevals = log.size() / (D + 1);
        }
        return rc;
    }
};

struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double lap() {
        auto t1 = std::chrono::steady_clock::now();
        double s = std::chrono::duration<double>(t1 - t0).count();
        t0 = t1;
        return s;
    }
};

template <typename T, int D, typename F>
int run_regions(int integrator, double alpha, const F& fn, const std::vector<vmo::Region<T, D>>& regs, const T* tlo,
                const T* thi, int db, const uint64_t* res, uint64_t spp, uint64_t seed, int nthreads,
                double* bins_inout, vmo::Stats* st) {
    uint64_t nb = 1;
    for (int i = 0; i < db; ++i) nb *= res[i];
    std::vector<T> out(nb);
    for (uint64_t b = 0; b < nb; ++b) out[b] = T(bins_inout[b]);
    if (integrator == vmo::INT_VR_OPTIMIZE)
        vmo::integrate_regions_vr<T, D, F, vmo::AccOptimize<T>>(regs, fn, tlo, thi, db, res, spp, seed, alpha,
                                                                  nthreads, out.data(), st);
    else if (integrator == vmo::INT_VR_FIXED)
        vmo::integrate_regions_vr<T, D, F, vmo::AccFixed<T>>(regs, fn, tlo, thi, db, res, spp, seed, alpha, nthreads,
                                                               out.data(), st);
    else if (integrator == vmo::INT_CV_ALPHA1)
        vmo::integrate_regions_cv<T, D, F>(regs, fn, tlo, thi, db, res, spp, seed, out.data(), st);
    else if (integrator == vmo::INT_VR_OPTIMIZE_IMPORTANCE)
        vmo::integrate_regions_vr<T, D, F, vmo::AccOptimize<T>, true>(regs, fn, tlo, thi, db, res, spp, seed, alpha,
                                                                        nthreads, out.data(), st);
    else if (integrator == vmo::INT_VR_FIXED_IMPORTANCE)
        vmo::integrate_regions_vr<T, D, F, vmo::AccFixed<T>, true>(regs, fn, tlo, thi, db, res, spp, seed, alpha,
                                                                     nthreads, out.data(), st);
    else if (integrator == vmo::INT_VR_OPTIMIZE_RR_INTEGRAL)
        vmo::integrate_regions_vr<T, D, F, vmo::AccOptimize<T>, false, 1>(regs, fn, tlo, thi, db, res, spp, seed,
                                                                            alpha, nthreads, out.data(), st);
    else if (integrator == vmo::INT_VR_FIXED_RR_INTEGRAL)
        vmo::integrate_regions_vr<T, D, F, vmo::AccFixed<T>, false, 1>(regs, fn, tlo, thi, db, res, spp, seed,
                                                                         alpha, nthreads, out.data(), st);
    else if (integrator == vmo::INT_VR_OPTIMIZE_RR_ERROR)
        vmo::integrate_regions_vr<T, D, F, vmo::AccOptimize<T>, false, 2>(regs, fn, tlo, thi, db, res, spp, seed,
                                                                            alpha, nthreads, out.data(), st);
    else if (integrator == vmo::INT_VR_FIXED_RR_ERROR)
        vmo::integrate_regions_vr<T, D, F, vmo::AccFixed<T>, false, 2>(regs, fn, tlo, thi, db, res, spp, seed,
                                                                         alpha, nthreads, out.data(), st);
    else if (integrator == vmo::INT_VR_OPTIMIZE_RR_PDF)
        vmo::integrate_regions_vr<T, D, F, vmo::AccOptimize<T>, false, 3>(regs, fn, tlo, thi, db, res, spp, seed,
                                                                            alpha, nthreads, out.data(), st);
    else if (integrator == vmo::INT_VR_FIXED_RR_PDF)
        vmo::integrate_regions_vr<T, D, F, vmo::AccFixed<T>, false, 3>(regs, fn, tlo, thi, db, res, spp, seed,
                                                                         alpha, nthreads, out.data(), st);
    else if (integrator == vmo::INT_VR_OPTIMIZE_SAMPLING_RR)
        vmo::integrate_regions_vr<T, D, F, vmo::AccOptimize<T>, 2>(regs, fn, tlo, thi, db, res, spp, seed, alpha,
                                                                     nthreads, out.data(), st);
    else if (integrator == vmo::INT_VR_FIXED_SAMPLING_RR)
        vmo::integrate_regions_vr<T, D, F, vmo::AccFixed<T>, 2>(regs, fn, tlo, thi, db, res, spp, seed, alpha,
                                                                  nthreads, out.data(), st);
    else if (integrator >= vmo::INT_VR_OPTIMIZE_RR_INTEGRAL_IMPORTANCE &&
             integrator <= vmo::INT_VR_FIXED_RR_PDF_IMPORTANCE) {
        // codes 11..16: roulette (integral, error, pdf) x weight (optimised, fixed), importance sampled
        const int k = integrator - vmo::INT_VR_OPTIMIZE_RR_INTEGRAL_IMPORTANCE;
        auto run = [&](auto acc, auto rr) {
            vmo::integrate_regions_vr<T, D, F, decltype(acc), true, decltype(rr)::value>(
                regs, fn, tlo, thi, db, res, spp, seed, alpha, nthreads, out.data(), st);
        };
        using R1 = std::integral_constant<int, 1>;
        using R2 = std::integral_constant<int, 2>;
        using R3 = std::integral_constant<int, 3>;
        const bool fixed = k % 2 == 1;
        if (k / 2 == 0)
            fixed ? run(vmo::AccFixed<T>{}, R1{}) : run(vmo::AccOptimize<T>{}, R1{});
        else if (k / 2 == 1)
            fixed ? run(vmo::AccFixed<T>{}, R2{}) : run(vmo::AccOptimize<T>{}, R2{});
        else
            fixed ? run(vmo::AccFixed<T>{}, R3{}) : run(vmo::AccOptimize<T>{}, R3{});
    } else {
        g_err = "unknown integrator";
        return -1;
    }
    for (uint64_t b = 0; b < nb; ++b) bins_inout[b] = double(out[b]);
    return 0;
}

template <typename T, int D>
struct IntegrateJob {
    static int run(int integrator, int heuristic, double alpha, const vmx_integrand* f, const double* lo,
                   const double* hi, int db, const uint64_t* res, uint64_t iterations, uint64_t spp, uint64_t seed,
                   int nthreads, double* bins_inout, vmo::Stats* st) {
        if (db < 1 || db > 3 || db > D) {
            g_err = "bin dims must be 1..min(3, D)";
            return -1;
        }
        T tlo[D], thi[D];
        load_box<T, D>(lo, hi, tlo, thi);
        return with_f<T, D>(f, [&](const auto& fn) {
            Timer tm;
            uint64_t ev = 0;
            auto regs = vmo::generate<T, D>(fn, tlo, thi, heuristic, iterations, &ev);
            const double tg = tm.lap();
            int rc = run_regions<T, D>(integrator, alpha, fn, regs, tlo, thi, db, res, spp, seed, nthreads,
                                       bins_inout, st);
            if (st) {
                st->regions = regs.size();
                st->evals_regions = ev;
                st->t_generate = tg;
                st->t_bins = tm.lap();
            }
            return rc;
        });
    }
};

template <typename T, int D>
struct RegionsJob {
    static int run(int integrator, double alpha, const vmx_integrand* f, const double* lo, const double* hi,
                   uint64_t nregions, const double* reg_lo, const double* reg_hi, const double* reg_nodes, int db,
                   const uint64_t* res, uint64_t spp, uint64_t seed, int nthreads, double* bins_inout,
                   vmo::Stats* st) {
        if (db < 1 || db > 3 || db > D) {
            g_err = "bin dims must be 1..min(3, D)";
            return -1;
        }
        T tlo[D], thi[D];
        load_box<T, D>(lo, hi, tlo, thi);
        std::vector<vmo::Region<T, D>> regs(nregions);
        for (uint64_t r = 0; r < nregions; ++r) {
            for (int d = 0; d < D; ++d) {
                regs[r].lo[d] = T(reg_lo[r * D + d]);
                regs[r].hi[d] = T(reg_hi[r * D + d]);
            }
            regs[r].vol = vmo::box_volume<T, D>(regs[r].lo, regs[r].hi);
            for (int k = 0; k < vmo::Region<T, D>::K; ++k)
                regs[r].v[k] = T(reg_nodes[r * vmo::Region<T, D>::K + k]);
            regs[r].err = 0;
            regs[r].dim = 0;
        }
        return with_f<T, D>(f, [&](const auto& fn) {
            return run_regions<T, D>(integrator, alpha, fn, regs, tlo, thi, db, res, spp, seed, nthreads, bins_inout,
                                     st);
        });
    }
};

template <typename T, int D>
struct EvalJob {
    static int run(const vmx_integrand* f, uint64_t n, const double* x, double* out) {
        return with_f<T, D>(f, [&](const auto& fn) {
            for (uint64_t i = 0; i < n; ++i) {
                T p[D];
                for (int d = 0; d < D; ++d) p[d] = T(x[i * D + d]);
                out[i] = double(fn(p));
            }
            return 0;
        });
    }
};

/* Heuristic (err, dim) of one region given its box and nodes. */
template <typename T, int D>
struct HeuristicJob {
    static int run(int heuristic, uint64_t n, const double* reg_lo, const double* reg_hi, const double* reg_nodes,
                   double* out_err, int32_t* out_dim) {
        for (uint64_t r = 0; r < n; ++r) {
            vmo::Region<T, D> g;
            for (int d = 0; d < D; ++d) {
                g.lo[d] = T(reg_lo[r * D + d]);
                g.hi[d] = T(reg_hi[r * D + d]);
            }
            g.vol = vmo::box_volume<T, D>(g.lo, g.hi);
            for (int k = 0; k < vmo::Region<T, D>::K; ++k) g.v[k] = T(reg_nodes[r * vmo::Region<T, D>::K + k]);
            vmo::apply_heuristic(g, heuristic);
            out_err[r] = g.err;
            out_dim[r] = g.dim;
        }
        return 0;
    }
};

}  // namespace

/* integrator_crespo2021_infinite<RD> (src/control-variates/integrator-crespo2021.h:24-44)
   over a float range of DF dims: regions and bins over the first RD, node values and the
   residual integrate the rest by Monte Carlo (see vmo::FubiniNodes / FubiniSamples).
   gen_copies / bin_copies: copy-constructions of the Monte Carlo objects between their
   seeding and use (each reseeds from one draw). */
template <int RD, int DF>
int fubini_run(const vmx_integrand* f, const double* lo, const double* hi, int db, const uint64_t* res,
               uint64_t iterations, uint64_t mc, uint64_t spp, uint64_t seed, int gen_copies, int bin_copies,
               int nthreads, double* bins_inout, vmo::Stats* st) {
    if (db < 1 || db > RD || db > 3) {
        g_err = "bin dims must be 1..min(3, region dims)";
        return -1;
    }
    if (mc < 1) {
        g_err = "mc samples must be >= 1";
        return -1;
    }
    float flo[DF], fhi[DF];
    load_box<float, DF>(lo, hi, flo, fhi);
    return with_f<float, DF>(f, [&](const auto& fn) {
        using FN = std::decay_t<decltype(fn)>;
        float vol = 1.0f;
        for (int d = RD; d < DF; ++d) vol *= (fhi[d] - flo[d]);
        vmo::FubiniNodes<RD, DF, FN> nodes{fn, {}, {}, vol, mc,
                                           vmo::copied_engine(std::mt19937(2 * seed + 1), gen_copies)};
        vmo::FubiniSamples<RD, DF, FN> samples{fn, {}, {}, vol, bin_copies};
        for (int d = 0; d < DF; ++d) {
            nodes.rlo[d] = samples.rlo[d] = flo[d];
            nodes.rhi[d] = samples.rhi[d] = fhi[d];
        }
        Timer tm;
        uint64_t ev = 0;
        auto regs = vmo::generate<float, RD>(nodes, flo, fhi, vmo::HEUR_SIZE_RELATIVE, iterations, &ev);
        const double tg = tm.lap();
        uint64_t nb = 1;
        for (int i = 0; i < db; ++i) nb *= res[i];
        std::vector<float> out(nb);
        for (uint64_t b = 0; b < nb; ++b) out[b] = float(bins_inout[b]);
        vmo::integrate_regions_vr<float, RD, decltype(samples), vmo::AccOptimize<float>>(
            regs, samples, flo, fhi, db, res, spp, seed, 1.0, nthreads, out.data(), st);
        for (uint64_t b = 0; b < nb; ++b) bins_inout[b] = double(out[b]);
        if (st) {
            st->regions = regs.size();
            st->evals_regions = ev * mc;
            st->t_generate = tg;
            st->t_bins = tm.lap();
        }
        return 0;
    });
}

extern "C" {

typedef struct vmo_stats {
    uint64_t regions, pairs, evals_regions, evals_residual;
    double t_generate, t_stratify, t_bins;
} vmo_stats;

static void copy_stats(const vmo::Stats& s, vmo_stats* o) {
    if (!o) return;
    o->regions = s.regions;
    o->pairs = s.pairs;
    o->evals_regions = s.evals_regions;
    o->evals_residual = s.evals_residual;
    o->t_generate = s.t_generate;
    o->t_stratify = s.t_stratify;
    o->t_bins = s.t_bins;
}

const char* vmo_last_error(void) { return g_err.c_str(); }

/* Region list of RegionsGeneratorAdaptiveHeap::generate in heap-array order, and, if
   out_evals is given, every integrand evaluation in call order as rows (x.., f(x)).
   Returns the number of regions. */
int vmo_generate(int scalar_bytes, const vmx_integrand* f, const double* lo, const double* hi, int heuristic,
                 uint64_t iterations, double* out_lo, double* out_hi, double* out_nodes, double* out_err,
                 int32_t* out_dim, double* out_evals, uint64_t* out_nevals) {
    return dispatch<GenerateJob>(scalar_bytes, f->dim, f, lo, hi, heuristic, iterations, out_lo, out_hi, out_nodes,
                                 out_err, out_dim, out_evals, out_nevals);
}

/* Full integrate: generate + per-bin CV + residual.  bins_inout uses C order
   bins[i0][i1][i2]; it is read (the alpha=1 form accumulates) and written. */
int vmo_integrate(int scalar_bytes, int integrator, int heuristic, double alpha, const vmx_integrand* f,
                  const double* lo, const double* hi, int db, const uint64_t* res, uint64_t iterations,
                  uint64_t spp, uint64_t seed, int nthreads, double* bins_inout, vmo_stats* st) {
    vmo::Stats s;
    int rc = dispatch<IntegrateJob>(scalar_bytes, f->dim, integrator, heuristic, alpha, f, lo, hi, db, res,
                                    iterations, spp, seed, nthreads, bins_inout, &s);
    copy_stats(s, st);
    return rc;
}

/* Per-bin phase only, on a caller-given region list (any order). */
int vmo_integrate_regions(int scalar_bytes, int integrator, double alpha, const vmx_integrand* f, const double* lo,
                          const double* hi, uint64_t nregions, const double* reg_lo, const double* reg_hi,
                          const double* reg_nodes, int db, const uint64_t* res, uint64_t spp, uint64_t seed,
                          int nthreads, double* bins_inout, vmo_stats* st) {
    vmo::Stats s;
    int rc = dispatch<RegionsJob>(scalar_bytes, f->dim, integrator, alpha, f, lo, hi, nregions, reg_lo, reg_hi,
                                  reg_nodes, db, res, spp, seed, nthreads, bins_inout, &s);
    copy_stats(s, st);
    return rc;
}

/* The Fubini form above for (rd, f->dim) with 1 <= rd <= 3, rd < f->dim <= 5. */
int vmo_integrate_fubini(int rd, const vmx_integrand* f, const double* lo, const double* hi, int db,
                         const uint64_t* res, uint64_t iterations, uint64_t mc, uint64_t spp, uint64_t seed,
                         int gen_copies, int bin_copies, int nthreads, double* bins_inout, vmo_stats* st) {
    vmo::Stats s;
    int rc = -1;
    try {
#define VMO_FB(R, F)                                                                                          \
    if (rd == R && f->dim == F)                                                                               \
        rc = fubini_run<R, F>(f, lo, hi, db, res, iterations, mc, spp, seed, gen_copies, bin_copies, nthreads, \
                              bins_inout, &s);                                                                \
    else
        VMO_FB(1, 2) VMO_FB(1, 3) VMO_FB(1, 5) VMO_FB(2, 3) VMO_FB(2, 4) VMO_FB(2, 5) VMO_FB(3, 4) VMO_FB(3, 5) {
            g_err = "fubini: region dims 1..3 below the integrand's dims (2..5)";
            rc = -1;
        }
#undef VMO_FB
    } catch (const std::exception& e) {
        g_err = e.what();
        rc = -3;
    }
    copy_stats(s, st);
    return rc;
}

/* integrator_crespo2021_infinite<rd> (form 0) or
   integrator_adaptive_fubini_variance_reduction_parallel_optimized<rd> with
   region_stratification_uniform, region_sampling_uniform and cv_optimize_weight (form 1)
   or cv_fixed_weight(alpha) (form 2), over an infinite range: region dims
   [lo, hi] (rd entries), the rest's first rest_count entries [rest_lo, rest_hi], [0, 1]
   after them; the integrand is the Dyson series of the sequence. */
int vmo_integrate_fubini_inf(int rd, int form, const vmx_dyson_params* dp, const double* lo, const double* hi,
                             int rest_count, const double* rest_lo, const double* rest_hi, int db, const uint64_t* res,
                             uint64_t iterations, uint64_t mc, uint64_t spp, uint64_t seed, double alpha,
                             int gen_copies, int bin_copies, int nthreads, double* bins_inout, vmo_stats* st) {
    vmo::Stats s;
    int rc = -1;
    try {
        if (db < 1 || db > rd || db > 3) throw std::runtime_error("bin dims must be 1..min(3, region dims)");
        if (mc < 1) throw std::runtime_error("mc samples must be >= 1");
        auto run = [&](auto rdc) {
            constexpr int RD = decltype(rdc)::value;
            const vmx::Dyson<float> dy = vmx::Dyson<float>::from(*dp);
            vmo::RestInf rest;
            for (int k = 0; k < rest_count; ++k) {
                rest.lo.push_back(float(rest_lo[k]));
                rest.hi.push_back(float(rest_hi[k]));
            }
            float flo[RD], fhi[RD];
            load_box<float, RD>(lo, hi, flo, fhi);
            const uint64_t gen_seed = form == 0 ? 2 * seed + 1 : seed;
            vmo::FubiniNodesInf<RD, vmx::Dyson<float>> nodes{dy, rest, mc,
                                                              vmo::copied_engine(std::mt19937(gen_seed), gen_copies)};
            Timer tm;
            uint64_t ev = 0;
            auto regs = vmo::generate<float, RD>(nodes, flo, fhi, vmo::HEUR_SIZE_RELATIVE, iterations, &ev);
            const double tg = tm.lap();
            uint64_t nb = 1;
            for (int i = 0; i < db; ++i) nb *= res[i];
            std::vector<float> out(nb);
            for (uint64_t b = 0; b < nb; ++b) out[b] = float(bins_inout[b]);
            if (form == 0) {
                vmo::FubiniSamplesInf<RD, vmx::Dyson<float>> samples{dy, rest, bin_copies};
                vmo::integrate_regions_vr<float, RD, decltype(samples), vmo::AccOptimize<float>>(
                    regs, samples, flo, fhi, db, res, spp, seed, 1.0, nthreads, out.data(), &s);
            } else {
                vmo::RestFunctionInf<RD, vmx::Dyson<float>> restf{dy, rest};
                if (form == 1)
                    vmo::integrate_regions_strat<float, RD, decltype(restf), vmo::AccOptimize<float>>(
                        regs, restf, flo, fhi, db, res, spp, seed + 1, alpha, nthreads, out.data(), &s);
                else
                    vmo::integrate_regions_strat<float, RD, decltype(restf), vmo::AccFixed<float>>(
                        regs, restf, flo, fhi, db, res, spp, seed + 1, alpha, nthreads, out.data(), &s);
            }
            for (uint64_t b = 0; b < nb; ++b) bins_inout[b] = double(out[b]);
            s.regions = regs.size();
            s.evals_regions = ev * mc;
            s.t_generate = tg;
            s.t_bins = tm.lap();
            return 0;
        };
        if (rd == 1)
            rc = run(std::integral_constant<int, 1>{});
        else if (rd == 2)
            rc = run(std::integral_constant<int, 2>{});
        else if (rd == 3)
            rc = run(std::integral_constant<int, 3>{});
        else if (rd == 4)
            rc = run(std::integral_constant<int, 4>{});
        else
            throw std::runtime_error("infinite Fubini form: region dims 1..4");
    } catch (const std::exception& e) {
        g_err = e.what();
        rc = -3;
    }
    copy_stats(s, st);
    return rc;
}

/* Rules::sample_line (src/newton-cotes/rules.h:216-237, float): a point of [t0, t1] from
   the uniform s and the line's node values p[3]. */
float vmo_imp_sample_line(float s, float t0, float t1, const float* p) {
    return vmo::imp::sample_line<float>(s, t0, t1, std::array<float, 3>{p[0], p[1], p[2]});
}

/* region_sampling_importance's draw over the box [slo, shi] of the float region [rlo, rhi]
   with node values v, from given uniforms s[D] (Region::sample_subrange and pdf_subrange,
   src/newton-cotes/region.h:220-330): the point x[D] and its weight 1/pdf (0 where
   pdf < 1e-5). */
int vmo_imp_sample(int D, const float* v, const float* rlo, const float* rhi, const float* slo, const float* shi,
                   const float* s, float* x, float* w) {
    auto run = [&](auto dd) {
        constexpr int DD = decltype(dd)::value;
        vmo::Region<float, DD> r;
        for (int d = 0; d < DD; ++d) {
            r.lo[d] = rlo[d];
            r.hi[d] = rhi[d];
        }
        r.vol = vmo::box_volume<float, DD>(r.lo, r.hi);
        for (int k = 0; k < vmo::Region<float, DD>::K; ++k) r.v[k] = v[k];
        float an[DD], bn[DD], un[DD], pn[DD];
        auto in_range = [&](float p, int d) {
            return (r.lo[d] >= r.hi[d]) ? r.lo[d] : (p - r.lo[d]) / (r.hi[d] - r.lo[d]);
        };
        for (int d = 0; d < DD; ++d) {
            an[d] = in_range(slo[d], d);
            bn[d] = in_range(shi[d], d);
        }
        vmo::imp::sample_subrange<float, DD>(r, s, an, bn, un);
        for (int d = 0; d < DD; ++d) x[d] = un[d] * (r.hi[d] - r.lo[d]) + r.lo[d];
        for (int d = 0; d < DD; ++d) pn[d] = in_range(x[d], d);
        const float pdf = vmo::imp::pdf_subrange<float, DD>(r, pn, an, bn);
        *w = float((pdf < 1.e-5) ? 0.0 : (1.0 / pdf));
        return 1;
    };
    switch (D) {
        case 1: return run(std::integral_constant<int, 1>{});
        case 2: return run(std::integral_constant<int, 2>{});
        case 3: return run(std::integral_constant<int, 3>{});
        case 4: return run(std::integral_constant<int, 4>{});
        case 5: return run(std::integral_constant<int, 5>{});
        case 6: return run(std::integral_constant<int, 6>{});
        default: return -1;
    }
}

int vmo_eval(int scalar_bytes, const vmx_integrand* f, uint64_t n, const double* x, double* out) {
    return dispatch<EvalJob>(scalar_bytes, f->dim, f, n, x, out);
}

int vmo_heuristic(int scalar_bytes, int D, int heuristic, uint64_t n, const double* reg_lo, const double* reg_hi,
                  const double* reg_nodes, double* out_err, int32_t* out_dim) {
    return dispatch<HeuristicJob>(scalar_bytes, D, heuristic, n, reg_lo, reg_hi, reg_nodes, out_err, out_dim);
}

/* FP64 closed-form bins for a 2-D static blob integrand over [lo, hi], C order. */
int vmo_blobs2d_exact_bins(const vmx_blobs_params* p, const double* lo, const double* hi, uint64_t res0,
                           uint64_t res1, double* out) {
    for (int k = 0; k < p->nblobs; ++k)
        if (p->blob[k][2] != 0 || p->blob[k][3] != 0) {
            g_err = "closed form needs static blobs";
            return -1;
        }
    if (p->step_mode == 2) {
        g_err = "closed form needs step_mode 0 or 1";
        return -1;
    }
    const double dx = (hi[0] - lo[0]) / double(res0), dy = (hi[1] - lo[1]) / double(res1);
    for (uint64_t i = 0; i < res0; ++i)
        for (uint64_t j = 0; j < res1; ++j) {
            const double x0 = lo[0] + i * dx, x1 = lo[0] + (i + 1) * dx;
            const double y0 = lo[1] + j * dy, y1 = lo[1] + (j + 1) * dy;
            out[i * res1 + j] = double(res0 * res1) * vmo::blobs2d_exact_rect(*p, x0, x1, y0, y1);
        }
    return 0;
}

/* Exact integral over one rectangle (used to label per-region true errors). */
double vmo_blobs2d_exact_rect(const vmx_blobs_params* p, double x0, double x1, double y0, double y1) {
    return vmo::blobs2d_exact_rect(*p, x0, x1, y0, y1);
}

}  // extern "C"

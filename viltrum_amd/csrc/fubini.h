/*
 * The Fubini form of an integrand, for integrator_crespo2021_infinite<RD>
 * (src/control-variates/integrator-crespo2021.h:24-44): regions and bins live in the
 * first RD dims of an inner integrand of DF dims, and the remaining dims are integrated
 * by Monte Carlo (src/combination/fubini.h:51-54, regions-generator-fubini.h:16-22):
 *   - a region node value is the estimate with mc uniform samples over the rest range,
 *     sum of f(x | xr_k) * (volume / mc)  (MonteCarlo::integrate,
 *     src/monte-carlo/monte-carlo.h:39-60);
 *   - a residual sample uses one, f(x | xr) * volume  (MonteCarloPerBin::integrate,
 *     src/monte-carlo/monte-carlo-per-bin.h:37-62).
 * The rest-dim samples are counter-based: Philox4x32-10 on the bits of the point's
 * region coordinates and the sample index.  A node shared by several regions therefore
 * gets one value, and results do not depend on evaluation order, batching or launch
 * shape (the reference draws them from one sequential engine).
 */
#ifndef VMX_FUBINI_H
#define VMX_FUBINI_H

#include <string.h>

#include <type_traits>

#include "philox.h"

namespace vmx {

template <typename F>
struct FamilyParams;
template <typename T>
struct FamilyParams<Poly<T>> {
    using type = vmx_poly_params;
};
template <typename T>
struct FamilyParams<Blobs<T>> {
    using type = vmx_blobs_params;
};
template <typename T>
struct FamilyParams<Lobes5<T>> {
    using type = vmx_lobes5_params;
};

template <typename T>
VMX_HD inline uint32_t coord_bits(T v) {
    if constexpr (sizeof(T) == 4) {
        uint32_t u;
        memcpy(&u, &v, 4);
        return u;
    } else {
        uint64_t u;
        memcpy(&u, &v, 8);
        return uint32_t(u) ^ uint32_t(u >> 32);
    }
}

template <typename T, typename Inner, int DF, int RD>
struct Fubini {
    static_assert(RD >= 1 && RD <= 3 && RD < DF && DF - RD <= 4, "1..3 region dims, 1..4 rest dims");
    static constexpr uint32_t RESIDUAL = 0x80000000u;  // sample-index tag of residual samples
    static constexpr uint32_t KEY1 = 0x46554249u;
    Inner in;
    T rlo[DF], rhi[DF];  // entries RD..DF-1: the rest range
    T vol;               // product of the rest extents, in T (Range::volume)
    uint32_t mc, seed;

    static Fubini from(const vmx_fubini_params& p) {
        Fubini q{};
        q.in = Inner::from(*static_cast<const typename FamilyParams<Inner>::type*>(p.inner));
        q.vol = T(1);
        for (int d = 0; d < DF; ++d) {
            q.rlo[d] = T(p.rest_lo[d]);
            q.rhi[d] = T(p.rest_hi[d]);
            if (d >= RD) q.vol *= (q.rhi[d] - q.rlo[d]);
        }
        q.mc = p.mc_samples;
        q.seed = p.seed;
        return q;
    }

    /* f at the region point x (RD coords) and rest sample k of its stream */
    template <class M>
    VMX_HD T at_sample(const T* x, uint32_t k) const {
        const U4 r = philox4x32_10(
            U4{coord_bits(x[0]), RD > 1 ? coord_bits(x[RD > 1 ? 1 : 0]) : 0u,
               RD > 2 ? coord_bits(x[RD > 2 ? 2 : 0]) : 0u, k},
            seed, KEY1);
        const uint32_t w[4] = {r.x, r.y, r.z, r.w};
        T xf[DF];
        for (int d = 0; d < RD; ++d) xf[d] = x[d];
        for (int d = RD; d < DF; ++d) xf[d] = M::fma(T(u01f(w[d - RD])), rhi[d] - rlo[d], rlo[d]);
        return in.template eval<M, DF>(xf);
    }

    /* region node value: mc samples, sum of f * (volume / mc) */
    template <class M, int D>
    VMX_HD T eval(const T* x) const {
        static_assert(D == RD, "evaluated over the region dims");
        const T factor = T(double(vol) / double(mc));
        T sol = T(0);
        for (uint32_t k = 0; k < mc; ++k) sol = M::fma(at_sample<M>(x, k), factor, sol);
        return sol;
    }

    /* residual sample: one rest sample, f * volume */
    template <class M, int D>
    VMX_HD T eval_residual(const T* x) const {
        static_assert(D == RD, "evaluated over the region dims");
        return at_sample<M>(x, RESIDUAL) * vol;
    }
};

template <typename F>
struct IsFubini : std::false_type {};
template <typename T, typename Inner, int DF, int RD>
struct IsFubini<Fubini<T, Inner, DF, RD>> : std::true_type {};

/* The integrand of one residual sample: the Fubini form draws one rest sample, every
   other family is its own evaluation. */
template <class M, int D, class F, typename T>
VMX_HD inline T residual_eval(const F& f, const T* x) {
    if constexpr (IsFubini<F>::value)
        return f.template eval_residual<M, D>(x);
    else
        return f.template eval<M, D>(x);
}

}  // namespace vmx

#endif

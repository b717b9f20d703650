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
struct FamilyParams<Dyson<T>> {
    using type = vmx_dyson_params;
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

/* The Fubini form over an infinite range (RangeInfinite, src/range-infinite.h:16-87), for
   integrator_crespo2021_infinite<RD> and the optimised Fubini integrator: the inner
   integrand reads a sequence -- the region point's RD coordinates (concat,
   src/combination/concat.h:8-33), then unbounded rest entries, entry j uniform in
   [rlo[j], rlo[j] + rext[j]] for j < nrest and in [0, 1] after them (RangeInfinite::min(i)
   / max(i)).  A node value is the mean of mc sequences times the rest's volume
   (MonteCarlo::integrate over RangeInfinite, src/monte-carlo/monte-carlo.h:62-84); a
   residual sample reads one sequence, times the volume for crespo2021_infinite
   (monte-carlo-per-bin.h:64-87) and without it for the optimised integrator
   (MonteCarloFunction, regions-integrator-parallel-variance-reduction-optimized.h:36-47).
   Rest entries come 4 per Philox4x32-10 call: counter (the point's coordinate bits, the
   sample index), key (seed, a constant plus the block of 4).  The integrand reads the
   sequence lazily, so a Dyson series that stops early draws nothing more.  Inner is any
   integrand of a sequence with Inner::from(its parameter block) and eval_seq(Seq&): the
   Dyson series or a user functor (custom.h). */
template <typename T, typename Inner, int RD>
struct Fubini<T, Inner, VMX_INFINITE_DIM, RD> {
    static_assert(RD >= 1 && RD <= 4, "1..4 region dims");
    static constexpr uint32_t RESIDUAL = 0x80000000u;  // sample-index tag of residual samples
    static constexpr uint32_t KEY1 = 0x44595349u;
    Inner in;
    T rlo[VMX_MAX_DIM], rext[VMX_MAX_DIM];  // explicit rest entries: lower end, extent
    int32_t nrest;
    T vol;  // product of the explicit extents, in T (RangeInfinite::volume)
    uint32_t mc, seed;

    static Fubini from(const vmx_fubini_params& p) {
        Fubini q{};
        q.in = Inner::from(*static_cast<const typename FamilyParams<Inner>::type*>(p.inner));
        q.nrest = p.rest_count < 0 ? 0 : p.rest_count > VMX_MAX_DIM - RD ? VMX_MAX_DIM - RD : p.rest_count;
        q.vol = T(1);
        for (int j = 0; j < VMX_MAX_DIM; ++j) {
            const bool ex = j < q.nrest;
            q.rlo[j] = ex ? T(p.rest_lo[RD + j]) : T(0);
            q.rext[j] = ex ? T(p.rest_hi[RD + j]) - q.rlo[j] : T(1);
            if (ex) q.vol *= q.rext[j];
        }
        q.mc = p.mc_samples;
        q.seed = p.seed;
        return q;
    }

    /* the sequence of one sample; every index into the point and the words is resolved
       by selects, so nothing here is indexed at run time in registers */
    struct Seq {
        const Fubini* fb;
        T xs[RD];
        uint32_t c0, c1, c2, k;
        int i;
        uint32_t w0, w1, w2, w3;
        T v;
        VMX_HD Seq(const Fubini* f, const T* p, uint32_t sample) : fb(f), k(sample), i(0), w0(0), w1(0), w2(0), w3(0) {
#pragma unroll
            for (int d = 0; d < RD; ++d) xs[d] = p[d];
            c0 = coord_bits(xs[0]);
            c1 = RD > 1 ? coord_bits(xs[RD > 1 ? 1 : 0]) : 0u;
            c2 = RD > 2 ? coord_bits(xs[RD > 2 ? 2 : 0]) : 0u;
            if (RD > 3) c2 ^= coord_bits(xs[RD > 3 ? 3 : 0]) * 0x9E3779B1u;
            v = xs[0];
        }
        VMX_HD T cur() const { return v; }
        VMX_HD void next() {
            ++i;
            if (i < RD) {
                T nv = xs[0];
#pragma unroll
                for (int d = 1; d < RD; ++d)
                    if (i == d) nv = xs[d];
                v = nv;
                return;
            }
            const int j = i - RD, q = j & 3;
            if (q == 0) {
                const U4 r = philox4x32_10(U4{c0, c1, c2, k}, fb->seed, KEY1 + uint32_t(j >> 2) * 0x9E3779B9u);
                w0 = r.x;
                w1 = r.y;
                w2 = r.z;
                w3 = r.w;
            }
            const uint32_t wq = q == 0 ? w0 : q == 1 ? w1 : q == 2 ? w2 : w3;
            const T u = T(u01f(wq));
            v = j < fb->nrest ? u * fb->rext[j] + fb->rlo[j] : u;
        }
    };

    /* the inner integrand on sequence `sample` of the point x */
    VMX_HD T at_sample(const T* x, uint32_t sample) const {
        Seq s(this, x, sample);
        return in.eval_seq(s);
    }

    /* region node value: mc sequences, sum of f * (volume / mc) */
    template <class M, int D>
    VMX_HD T eval(const T* x) const {
        static_assert(D == RD, "evaluated over the region dims");
        const T factor = T(double(vol) / double(mc));
        T sol = T(0);
        for (uint32_t k = 0; k < mc; ++k) sol = M::fma(at_sample(x, k), factor, sol);
        return sol;
    }

    /* residual sample: one sequence, times the volume unless `bare` */
    template <class M, int D>
    VMX_HD T eval_residual(const T* x, bool bare = false) const {
        static_assert(D == RD, "evaluated over the region dims");
        const T fx = at_sample(x, RESIDUAL);
        return bare ? fx : fx * vol;
    }
};

template <typename F>
struct IsFubini : std::false_type {};
template <typename T, typename Inner, int DF, int RD>
struct IsFubini<Fubini<T, Inner, DF, RD>> : std::true_type {};

/* the Fubini form over an infinite range */
template <typename F>
struct IsFubiniInf : std::false_type {};
template <typename T, typename Inner, int RD>
struct IsFubiniInf<Fubini<T, Inner, VMX_INFINITE_DIM, RD>> : std::true_type {};

/* The integrand of one residual sample: the Fubini form draws one rest sample, every
   other family is its own evaluation.  `bare`: the Fubini form over an infinite range
   leaves out the rest's volume (the optimised integrator's residual). */
template <class M, int D, class F, typename T>
VMX_HD inline T residual_eval(const F& f, const T* x, bool bare = false) {
    if constexpr (IsFubiniInf<F>::value)
        return f.template eval_residual<M, D>(x, bare);
    else if constexpr (IsFubini<F>::value)
        return f.template eval_residual<M, D>(x);
    else
        return f.template eval<M, D>(x);
}

}  // namespace vmx

#endif

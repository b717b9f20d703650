/*
 * User-defined integrands: a device functor as an integrand family.
 *
 * The reference integrates any callable f(x0, x1, ...) or f(std::array<Float, D>)
 * (src/integrate.h:11-70, detail::adapt).  Here the callable is a trivially copyable
 * functor whose operator() is __host__ __device__ (or __device__), compiled in the
 * caller's HIP translation unit, which instantiates the pipeline for it
 * (include/viltrum_amd/device.hpp).  The functor's bytes travel like any family's
 * parameter block: copied to device memory per integral and staged in LDS by the
 * per-bin kernels.
 */
#ifndef VMX_CUSTOM_H
#define VMX_CUSTOM_H

#include <array>
#include <cstring>
#include <type_traits>
#include <utility>

#include "engine_core.h"

namespace vmx {

/* The three call forms of the reference's adapter, tried in this order: f(const T*),
   f(std::array<T, D>), f(x0, ..., x_{D-1}). */
template <typename Fn, typename T, int D, typename Seq = std::make_index_sequence<D>>
struct CallForms;
template <typename Fn, typename T, int D, std::size_t... I>
struct CallForms<Fn, T, D, std::index_sequence<I...>> {
    static constexpr bool ptr = std::is_invocable<const Fn&, const T*>::value;
    static constexpr bool arr = std::is_invocable<const Fn&, const std::array<T, D>&>::value;
    static constexpr bool args = std::is_invocable<const Fn&, decltype(T(I))...>::value;
    static constexpr bool ok = ptr || arr || args;

    VMX_HD static T call(const Fn& f, const T* x) {
        if constexpr (ptr) {
            return T(f(x));
        } else if constexpr (arr) {
            const std::array<T, D> a{{x[I]...}};
            return T(f(a));
        } else {
            return T(f(x[I]...));
        }
    }
};

/* A functor with an FP32 form only is called with FP32 coordinates where the pipeline
   asks for an FP64 value (the refiner's FP64 error targets of an FP32 run). */
template <typename Fn, typename T, int D>
struct CustomCall {
    static constexpr bool ok = CallForms<Fn, T, D>::ok;

    VMX_HD static T call(const Fn& f, const T* x) {
        if constexpr (CallForms<Fn, T, D>::ok) {
            return CallForms<Fn, T, D>::call(f, x);
        } else {
            float xf[D];
            for (int d = 0; d < D; ++d) xf[d] = float(x[d]);
            return T(CallForms<Fn, float, D>::call(f, xf));
        }
    }
};

template <typename T, int DIM, typename Fn>
struct Custom {
    // the functor's bytes: a closure type need not be default-constructible, the
    // pipeline's parameter arrays need their element type to be
    alignas(alignof(Fn) > 8 ? alignof(Fn) : 8) unsigned char raw[sizeof(Fn)];

    static Custom from(const vmx_custom_integrand& c) {
        Custom q;
        std::memcpy(q.raw, c.functor, sizeof(Fn));
        return q;
    }

    /* The math policy M of the built-in families does not apply: the functor computes
       its value as its own code says. */
    template <class M, int D>
    VMX_HD T eval(const T* x) const {
        static_assert(D == DIM, "a user integrand is instantiated for its own dimension");
        return CustomCall<Fn, T, DIM>::call(*reinterpret_cast<const Fn*>(raw), x);
    }
};

template <typename T, int DIM, typename Fn>
struct FamilyParams<Custom<T, DIM, Fn>> {
    using type = vmx_custom_integrand;
};

/* The sequence a user integrand of an infinite range reads, in the shape of the
   reference's random sequences (src/monte-carlo/random-sequence-ref-dis.h:10-38):
   begin() gives a forward iterator, *it is the current entry, ++it steps to the next one,
   and end() is never reached.  Gen is the Fubini form's lazy generator (cur(), next()):
   the entries are the point's region coordinates, then the Monte Carlo rest.  The
   sequence is read in one pass: every iterator steps the same generator. */
template <typename T, class Gen>
struct SeqIter {
    Gen* g;
    VMX_HD T operator*() const { return g->cur(); }
    VMX_HD SeqIter& operator++() {
        g->next();
        return *this;
    }
    VMX_HD bool operator==(const SeqIter&) const { return false; }
    VMX_HD bool operator!=(const SeqIter&) const { return true; }
};

template <typename T, class Gen>
struct SeqView {
    using value_type = T;
    using const_iterator = SeqIter<T, Gen>;
    Gen* g;
    VMX_HD const_iterator begin() const { return const_iterator{g}; }
    VMX_HD const_iterator end() const { return const_iterator{g}; }
};

/* A functor of a sequence, f(const Sequence&) (the reference's integrands of an infinite
   range, e.g. main/doc/ranges-infinite.cc:5-27), as the inner integrand of the Fubini
   form over an infinite range. */
template <typename T, typename Fn>
struct CustomSeq {
    alignas(alignof(Fn) > 8 ? alignof(Fn) : 8) unsigned char raw[sizeof(Fn)];

    static CustomSeq from(const vmx_custom_integrand& c) {
        CustomSeq q;
        std::memcpy(q.raw, c.functor, sizeof(Fn));
        return q;
    }

    template <class Gen>
    VMX_HD T eval_seq(Gen& s) const {
        return T((*reinterpret_cast<const Fn*>(raw))(SeqView<T, Gen>{&s}));
    }
};

template <typename T, typename Fn>
struct FamilyParams<CustomSeq<T, Fn>> {
    using type = vmx_custom_integrand;
};

}  // namespace vmx

namespace vmxe {

template <typename T, int DIM, typename Fn>
struct ParamOf<vmx::Custom<T, DIM, Fn>> {
    using type = vmx_custom_integrand;
};
template <typename T, int DIM, typename Fn>
struct F64Of<vmx::Custom<T, DIM, Fn>> {
    using type = vmx::Custom<double, DIM, Fn>;
};

}  // namespace vmxe

#endif

/*
 * User-defined integrands for the C++ drop-in: any device-evaluable functor.
 *
 * The reference integrates any callable, f(x0, x1, ...) or f(std::array<Float, D>)
 * (src/integrate.h:11-70).  A GPU can only call code compiled for it, so a user
 * integrand is a functor whose call operator is __host__ __device__ (or __device__),
 * in a translation unit compiled by hipcc for gfx950.  Wrapping it with
 * viltrum_amd::device_integrand<D>(f) instantiates the whole pipeline (generation,
 * bin lists, per-bin kernels) for that functor in this translation unit; the result
 * is an ordinary viltrum_amd::Integrand, so every integrate overload, integrator and
 * C entry point (vmx_integrate, vmx_integrate_batch, vmx_integrate_device,
 * vmx_integrate_regions, the refiner's training pairs) takes it:
 *
 *     #include <viltrum_amd/device.hpp>
 *     auto f = viltrum::device_integrand<2>([] __host__ __device__ (float x, float y) {
 *         return sinf(3.f * x) * expf(y);
 *     });
 *     viltrum::integrate(viltrum::integrator_crespo2021(1023, 16), bins, f, viltrum::range_primary<2>());
 *
 * The functor is taken as f(const T*), f(const std::array<T, D>&) or f(x0, ..., x_{D-1}),
 * in that order of preference.  It must be trivially copyable and at most
 * VMX_CUSTOM_MAX_BYTES: its bytes are copied to device memory per integral (captured
 * parameters travel with it, pointers in it must be device pointers).
 *
 * The scalar types to instantiate follow the dimension argument: device_integrand<D>(f) builds
 * the FP32 path (float bins), device_integrand<D, double>(f) the FP64 path and
 * device_integrand<D, float, double>(f) both; bins of a type not instantiated are
 * refused with VMX_ERR_UNSUPPORTED.  Each instantiation compiles the pipeline's kernels
 * for the functor (about a minute of hipcc per scalar type).
 *
 * As the inner integrand of the Fubini form (integrator_crespo2021_infinite<RD> over a
 * finite range), wrap it with viltrum_amd::device_integrand_fubini<D[, RD...]>(f); a
 * functor of a sequence, for the same form over an infinite range, with
 * viltrum_amd::device_integrand_infinite<RD...>(f).
 *
 * Compile with hipcc --offload-arch=gfx950 -std=c++17 -I<repo>/include
 * -Xclang -target-feature -Xclang -packed-fp32-ops, and link -lviltrum_amd.
 */
#ifndef VILTRUM_AMD_DEVICE_HPP
#define VILTRUM_AMD_DEVICE_HPP

#include <initializer_list>
#include <memory>
#include <tuple>
#include <type_traits>
#include <vector>

#include "viltrum.hpp"
#include "../../viltrum_amd/csrc/custom.h"
#include "../../viltrum_amd/csrc/run_batch.h"

namespace viltrum_amd {
namespace detail {

template <typename T, typename... Ts>
constexpr bool one_of = (std::is_same<T, Ts>::value || ...);

/* The hook the library calls for a group of integrals of this functor type. */
template <int D, typename Fn, typename... Ts>
int32_t run_device_integrand(vmx_ctx* ctx, int32_t scalar, const void* group, void* stream, void* pairs) {
    const auto& jobs = *static_cast<const std::vector<vmxe::JobIn>*>(group);
    const auto st = static_cast<hipStream_t>(stream);
    auto* req = static_cast<vmxe::PairsReq*>(pairs);
    if (scalar == VMX_F32) {
        if constexpr (one_of<float, Ts...>)
            return vmxe::run_batch<float, D, vmx::Custom<float, D, Fn>, vmx::FastMath>(ctx, jobs, st, req);
    } else if (scalar == VMX_F64) {
        if constexpr (one_of<double, Ts...>)
            return vmxe::run_batch<double, D, vmx::Custom<double, D, Fn>, vmx::PreciseMath>(ctx, jobs, st, req);
    }
    throw vmxe::VmxError(VMX_ERR_UNSUPPORTED,
                         scalar == VMX_F32 ? "user integrand not instantiated for float bins (device_integrand<D, float>)"
                                           : "user integrand not instantiated for double bins (device_integrand<D, double>)");
}

/* The hook for the Fubini form with this functor (FP32, D dims) as the inner integrand:
   region dims RDs... (1..3, below D) are instantiated. */
template <int D, typename Fn, int... RDs>
int32_t run_device_integrand_fubini(vmx_ctx* ctx, int32_t rd, const void* group, void* stream) {
    const auto& jobs = *static_cast<const std::vector<vmxe::JobIn>*>(group);
    const auto st = static_cast<hipStream_t>(stream);
    if (static_cast<const vmx_fubini_params*>(jobs[0].params)->inner_dim != D)
        throw vmxe::VmxError(VMX_ERR_INVALID, "Fubini form: inner_dim differs from the user integrand's dimension");
    using C = vmx::Custom<float, D, Fn>;
    int32_t rc = VMX_ERR_UNSUPPORTED;
    bool done = false;
    (void)std::initializer_list<int>{
        (rd == RDs && !done
             ? (rc = vmxe::run_batch<float, RDs, vmx::Fubini<float, C, D, RDs>, vmx::FastMath>(ctx, jobs, st, nullptr),
                done = true, 0)
             : 0)...};
    if (!done)
        throw vmxe::VmxError(VMX_ERR_UNSUPPORTED,
                             "user integrand not instantiated for the Fubini form at these region dims");
    return rc;
}

/* The hooks of a functor of a sequence: as the inner integrand of the Fubini form over
   an infinite range, region dims RDs... (1..4) instantiated; alone it is refused (the
   library never integrates a sequence integrand without a range to draw it from). */
template <typename Fn, int... RDs>
int32_t run_device_integrand_infinite(vmx_ctx* ctx, int32_t rd, const void* group, void* stream) {
    const auto& jobs = *static_cast<const std::vector<vmxe::JobIn>*>(group);
    const auto st = static_cast<hipStream_t>(stream);
    if (static_cast<const vmx_fubini_params*>(jobs[0].params)->inner_dim != VMX_INFINITE_DIM)
        throw vmxe::VmxError(VMX_ERR_INVALID, "an integrand of a sequence goes with an infinite range only");
    using C = vmx::CustomSeq<float, Fn>;
    int32_t rc = VMX_ERR_UNSUPPORTED;
    bool done = false;
    (void)std::initializer_list<int>{
        (rd == RDs && !done ? (rc = vmxe::run_batch<float, RDs, vmx::Fubini<float, C, VMX_INFINITE_DIM, RDs>,
                                                     vmx::FastMath>(ctx, jobs, st, nullptr),
                               done = true, 0)
                            : 0)...};
    if (!done)
        throw vmxe::VmxError(VMX_ERR_UNSUPPORTED,
                             "user integrand of a sequence not instantiated for these region dims");
    return rc;
}

inline int32_t run_sequence_alone(vmx_ctx*, int32_t, const void*, void*, void*) {
    throw vmxe::VmxError(VMX_ERR_UNSUPPORTED, "an integrand of a sequence goes with an infinite range only");
}

template <typename Fn>
struct DeviceIntegrandHolder {
    vmx_custom_integrand desc;
    Fn fn;
    explicit DeviceIntegrandHolder(const Fn& f) : desc{}, fn(f) {}
};

}  // namespace detail

template <int D, typename Fn, typename... Ts>
Integrand make_device_integrand(const Fn& fn, std::tuple<Ts...>*) {
    static_assert((... && (std::is_same<Ts, float>::value || std::is_same<Ts, double>::value)),
                  "scalar types: float and/or double");
    static_assert((... && vmx::CustomCall<Fn, Ts, D>::ok),
                  "the functor must be callable as f(const T*), f(std::array<T, D>) or f(x0, ..., x_{D-1})");
    auto h = std::make_shared<detail::DeviceIntegrandHolder<Fn>>(fn);
    h->desc.run = &detail::run_device_integrand<D, Fn, Ts...>;
    h->desc.functor = &h->fn;
    Integrand f;
    f.desc = vmx_integrand{VMX_FAMILY_CUSTOM, D, &h->desc};
    f.params = h;
    return f;
}

template <std::size_t D, typename... Ts, typename Fn>
Integrand device_integrand(const Fn& fn) {
    static_assert(D >= 1 && D <= VMX_MAX_DIM, "1..6 dimensions (the reference's region cap)");
    static_assert(std::is_trivially_copyable<Fn>::value,
                  "a user integrand is copied to the GPU byte for byte: it must be trivially copyable");
    static_assert(sizeof(Fn) <= VMX_CUSTOM_MAX_BYTES, "a user integrand is staged in LDS: at most 2 KB");
    using Scalars = std::conditional_t<sizeof...(Ts) == 0, std::tuple<float>, std::tuple<Ts...>>;
    return make_device_integrand<int(D), Fn>(fn, static_cast<Scalars*>(nullptr));
}

/* A functor of D dims for the Fubini form (integrator_crespo2021_infinite<RD> over a
   finite range, src/integrator-crespo2021.h:24-44) as well as the plain integrators:
   FP32, region dims RDs... (each 1..3 and below D; all of them if none is given).  Each
   region-dim count instantiates the pipeline once more. */
template <std::size_t D, int... RDs, typename Fn>
Integrand device_integrand_fubini(const Fn& fn) {
    static_assert(D >= 2 && D <= 5, "the Fubini form takes inner integrands of 2..5 dims");
    static_assert((... && (RDs >= 1 && RDs <= 3 && RDs < int(D))), "region dims 1..3, below the integrand's");
    Integrand f = device_integrand<D, float>(fn);
    auto* desc = const_cast<vmx_custom_integrand*>(static_cast<const vmx_custom_integrand*>(f.desc.params));
    if constexpr (sizeof...(RDs) > 0) {
        desc->run_fubini = &detail::run_device_integrand_fubini<int(D), Fn, RDs...>;
    } else if constexpr (D == 2) {
        desc->run_fubini = &detail::run_device_integrand_fubini<2, Fn, 1>;
    } else if constexpr (D == 3) {
        desc->run_fubini = &detail::run_device_integrand_fubini<3, Fn, 1, 2>;
    } else {
        desc->run_fubini = &detail::run_device_integrand_fubini<int(D), Fn, 1, 2, 3>;
    }
    return f;
}

/* A functor of a sequence for the Fubini form over an infinite range
   (integrator_crespo2021_infinite<RD> and the optimised Fubini integrators over a
   range_infinite / range_primary_infinite, src/integrate.h:80-120): FP32, region dims
   RDs... (each 1..4; all four if none is given).  The functor is called as f(seq) with
   seq a single-pass sequence whose begin() gives an iterator it: *it is the current
   entry, ++it steps on (the point's region coordinates first, then entries drawn
   uniformly over the rest of the range, [0, 1] past its explicit entries), as the
   reference's integrands of an infinite range read theirs:
       struct Series {
           float decay;
           template <class Seq>
           __host__ __device__ float operator()(const Seq& seq) const {
               auto x = seq.begin();
               float sum = 0.f, term = 1.f;
               while (*x < decay) { ++x; term *= 2.f * (*x); ++x; sum += term; }
               return sum;
           }
       };
       auto f = viltrum_amd::device_integrand_infinite<2>(Series{0.5f});
   Each region-dim count instantiates the pipeline once. */
template <int... RDs, typename Fn>
Integrand device_integrand_infinite(const Fn& fn) {
    static_assert((... && (RDs >= 1 && RDs <= 4)), "region dims 1..4");
    static_assert(std::is_trivially_copyable<Fn>::value,
                  "a user integrand is copied to the GPU byte for byte: it must be trivially copyable");
    static_assert(sizeof(Fn) <= VMX_CUSTOM_MAX_BYTES, "a user integrand is staged in LDS: at most 2 KB");
    using Gen = typename vmx::Fubini<float, vmx::CustomSeq<float, Fn>, VMX_INFINITE_DIM, 1>::Seq;
    static_assert(std::is_invocable<const Fn&, const vmx::SeqView<float, Gen>&>::value,
                  "the functor must be callable as f(const Sequence&)");
    auto h = std::make_shared<detail::DeviceIntegrandHolder<Fn>>(fn);
    h->desc.run = &detail::run_sequence_alone;
    h->desc.functor = &h->fn;
    if constexpr (sizeof...(RDs) > 0)
        h->desc.run_fubini = &detail::run_device_integrand_infinite<Fn, RDs...>;
    else
        h->desc.run_fubini = &detail::run_device_integrand_infinite<Fn, 1, 2, 3, 4>;
    Integrand f;
    f.desc = vmx_integrand{VMX_FAMILY_CUSTOM, VMX_INFINITE_DIM, &h->desc};
    f.params = h;
    return f;
}

}  // namespace viltrum_amd

#endif

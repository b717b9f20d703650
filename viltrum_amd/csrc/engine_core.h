/*
 * Shared host declarations of the engine's translation units: the context, device
 * arena helpers, job descriptors and the batch pipeline's entry (run_batch, defined in
 * run_batch.h and instantiated per integrand family in batch_*.hip so the families
 * compile in parallel).
 */
#ifndef VMX_ENGINE_CORE_H
#define VMX_ENGINE_CORE_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <array>
#include <functional>
#include <map>
#include <random>
#include <limits>
#include <type_traits>
#include <typeinfo>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include <cstring>
#include <rocprim/rocprim.hpp>

#include "../../include/viltrum_amd.h"
#include "exact.h"
#include "pipeline.h"


struct vmx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // bump allocator over a fixed device array, for the entry point that allocates
    // nothing: arenas that point at it are carved from it instead of hipMalloc
    struct Fixed {
        char* base = nullptr;
        size_t cap = 0, used = 0;
    };
    // device memory pool: one arena per purpose, grown on demand
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
        Fixed* fixed = nullptr;  // set: a view into a Fixed array, never hipMalloc'd or freed
    };
    Buf tree, regs, lists, misc, params, bins, bins2;
    Buf nn[VMX_MAX_DIM + 1];                       // FP32 priority-network parameters per dim
    Buf nn_lds[VMX_MAX_DIM + 1];                   // ... and their LDS image (nn::NetW<KP>)
    bool nn_ok[VMX_MAX_DIM + 1] = {};
    Buf tr_x, tr_t, tr_idx, tr_part, tr_loss, tr_state;  // training work space
    hipEvent_t ev[8] = {};
    void* pinned = nullptr;  // host staging for small device-to-host readbacks
    size_t pinned_bytes = 0;
    // one chunk of refinement rounds captured as a graph, reused while its launch
    // arguments (buffers, shapes, stream) stay the same
    struct RoundGraph {
        std::vector<unsigned char> key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    } rounds, rounds_tail;  // chunks of the tuned size, and of 2 rounds near the end
    // launch shapes: the autotuner's choices per workload class, and the shape a tuning
    // trial forces (keys of the classes the trial ran are collected in tuned_keys)
    struct Shape {
        int gen_per_cu = 4, round_chunk = 8, bins_xcd = 0, deepen = 0;
    };
    std::map<std::array<uint64_t, 6>, Shape> tune;
    // node capacity per region (times N) the last call of a workload class needed, so a
    // class whose rounds overflow 8N does not pay a failed attempt on every call
    std::map<std::array<uint64_t, 6>, uint64_t> cap_hint;
    // busy rounds the last call of a workload class (and budget octave) needed: where the
    // round chunks shrink
    std::map<std::array<uint64_t, 7>, uint32_t> round_hint;
    const Shape* trial = nullptr;
    std::vector<std::array<uint64_t, 6>> tuned_keys;
    // last result (for vmx_export_regions): gathered from device memory on demand
    int last_D = 0;
    uint64_t last_n = 0;
    std::function<void(double*, double*, double*, double*, int32_t*)> last_gather;
    // the no-alloc entry point: this device's scratch array (its device address, resolved
    // at context creation) and whether the current call runs on it
    char* scratch = nullptr;
    bool noalloc = false;
};


namespace vmxe {


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

inline void* ensure(vmx_ctx::Buf& b, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (b.fixed) {  // carve from the fixed array: grow in place when b is its last carve
        if (b.bytes >= bytes) return b.p;
        vmx_ctx::Fixed& f = *b.fixed;
        const bool last = b.p && static_cast<char*>(b.p) + b.bytes == f.base + f.used;
        const size_t start = last ? size_t(static_cast<char*>(b.p) - f.base) : (f.used + 255) & ~size_t(255);
        if (start + bytes > f.cap)
            throw VmxError(VMX_ERR_CAPACITY, "problem needs more work memory than the no-alloc scratch holds (" +
                                                 std::to_string(f.cap >> 20) + " MiB)");
        b.p = f.base + start;
        b.bytes = bytes;
        f.used = start + bytes;
        return b.p;
    }
    if (b.bytes < bytes) {
        if (b.p) HIP_CHECK(hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
        HIP_CHECK(hipMalloc(&b.p, bytes));
        b.bytes = bytes;
    }
    return b.p;
}

/* A batch whose (bin, footprint) pair total does not fit the 32-bit list offsets. */
inline void check_pair_total(uint64_t pairs) {
    if (pairs > vmx::MAX_PAIRS)
        throw VmxError(VMX_ERR_CAPACITY, "more (bin, region footprint) pairs than 32-bit list offsets hold: " +
                                             std::to_string(pairs));
}

inline void* ensure_pinned(vmx_ctx* ctx, size_t bytes) {
    if (ctx->pinned_bytes < bytes) {
        if (ctx->noalloc)
            throw VmxError(VMX_ERR_CAPACITY, "no-alloc call needs more pinned staging than the context set up");
        if (ctx->pinned) HIP_CHECK(hipHostFree(ctx->pinned));
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
        const size_t want = std::max<size_t>(bytes, 1 << 16);
        HIP_CHECK(hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault));
        ctx->pinned_bytes = want;
    }
    return ctx->pinned;
}

/* Carves aligned sub-arrays out of one allocation. */
inline int ilog2(uint64_t x) {
    int r = 0;
    while (x > 1) {
        x >>= 1;
        ++r;
    }
    return r;
}

/* Sub-allocations of one buffer, 256-byte aligned; taking more than the buffer holds
   is a bug in the caller's size computation and throws. */
struct Carve {
    char* base;
    size_t used = 0, cap;
    explicit Carve(const vmx_ctx::Buf& b) : base(static_cast<char*>(b.p)), cap(b.bytes) {}
    template <typename X>
    X* take(size_t n) {
        used = (used + 255) & ~size_t(255);
        if (used + n * sizeof(X) > cap) throw VmxError(VMX_ERR_CAPACITY, "work buffer carved past its size");
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
template <typename P>
struct F64Of;
template <typename T>
struct F64Of<vmx::Poly<T>> {
    using type = vmx::Poly<double>;
};
template <typename T>
struct F64Of<vmx::Blobs<T>> {
    using type = vmx::Blobs<double>;
};
template <typename T>
struct F64Of<vmx::Lobes5<T>> {
    using type = vmx::Lobes5<double>;
};
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
template <typename T, typename Inner, int DF, int RD>
struct ParamOf<vmx::Fubini<T, Inner, DF, RD>> {
    using type = vmx_fubini_params;
};

inline double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

/* One (scalar, D, family) batch through the device pipeline (pipeline.h). */
/* A caller-given region list (host memory, the layout vmx_export_regions writes):
   boxes n x D and node values n x 3^D, dim 0 fastest. */
struct GivenRegions {
    uint64_t n;
    const double *lo, *hi, *nodes;
};

struct JobIn {
    const vmx_problem* p;
    const void* params;
    void* bins;  // device pointer
    vmx_stats* stats;
    const GivenRegions* given = nullptr;  // set: skip generation, integrate over these regions
};

/* Training-pair mode of run_batch: stop after generation, report the node count and,
   if alloc hands back device arrays for that many rows, fill them. */
struct PairsReq {
    uint64_t count = 0;
    int kind = vmx::nn::TARGET_ENERGY;
    uint32_t stride = 1;
    std::function<void(uint64_t, uint16_t**, float**)> alloc;
};

constexpr int LEARN_MAX_DIM = 5;  // the fused network fits the expansion kernel's LDS up to here
constexpr int DEFER_K = 32;        // deferred rounds: about N / DEFER_K expansions per job and round

template <typename T, int D, typename P, typename Math>
void launch_expand(bool learn, const vmx::GenArgs<T>& g, const P* d_f, int par, uint32_t blocks, hipStream_t st) {
    const size_t pre_bytes = size_t(g.J + 1) * sizeof(uint32_t);  // the per-job candidate prefix
    if constexpr (std::is_same<T, float>::value && D <= LEARN_MAX_DIM) {
        if (learn) {
            vmx::k_expand_round<T, D, P, Math, true><<<blocks, vmx::GEN_THREADS, pre_bytes, st>>>(g, d_f, par);
            return;
        }
    }
    vmx::k_expand_round<T, D, P, Math, false><<<blocks, vmx::GEN_THREADS, pre_bytes, st>>>(g, d_f, par);
}

template <typename T>
struct RocprimTemp {
    size_t sort_bytes = 0, scan_bytes = 0;
};

/* One (scalar, D, family) batch through the device pipeline (pipeline.h). */
template <typename T, int D, typename P, typename Math>
int run_batch(vmx_ctx* ctx, const std::vector<JobIn>& in, hipStream_t st, PairsReq* pairs_req = nullptr);

}  // namespace vmxe

#endif

/*
 * TEST INFRASTRUCTURE: the product's device arithmetic (viltrum_amd/csrc/region_math.h,
 * philox.h, features.h) compiled for the host with g++, exposed through C entry points so tests can
 * check it against the oracle on the CPU.  Built by tests/native/build.py.
 */
#include <cstdint>
#include <cstring>
#include <type_traits>

#include "../../viltrum_amd/csrc/philox.h"
#include "../../viltrum_amd/csrc/region_math.h"
#include "../../viltrum_amd/csrc/features.h"
#include "../../viltrum_amd/csrc/importance.h"

namespace {
template <typename T, int D>
void heur_n(int heur, uint64_t n, const double* lo, const double* hi, const double* nodes, double* err, int32_t* dim) {
    constexpr int K = vmx::pow3(D);
    for (uint64_t r = 0; r < n; ++r) {
        T l[D], h[D], v[K];
        for (int d = 0; d < D; ++d) {
            l[d] = T(lo[r * D + d]);
            h[d] = T(hi[r * D + d]);
        }
        for (int k = 0; k < K; ++k) v[k] = T(nodes[r * K + k]);
        T e;
        int dm;
        vmx::heuristic<T, D>(v, l, h, vmx::box_volume<T, D>(l, h), heur, e, dm);
        err[r] = double(e);
        dim[r] = dm;
    }
}
/* split: child boxes (4*D), new-node coordinates (2*3^(D-1)*D), their node indices, and
   inherited child values (2*3^D, NaN where a new node lands) */
template <typename T, int D>
void split_n(int dim, const double* lo, const double* hi, const double* pv, double* boxes, double* xs, int32_t* idxs,
             double* cv) {
    constexpr int K = vmx::pow3(D), KM = vmx::pow3(D - 1);
    T l[D], h[D], l0[D], h0[D], l1[D], h1[D], p[K], v0[K], v1[K];
    for (int d = 0; d < D; ++d) {
        l[d] = T(lo[d]);
        h[d] = T(hi[d]);
    }
    for (int k = 0; k < K; ++k) {
        p[k] = T(pv[k]);
        v0[k] = v1[k] = T(0) / T(0);
    }
    vmx::split_boxes<T, D>(l, h, dim, l0, h0, l1, h1);
    for (int d = 0; d < D; ++d) {
        boxes[d] = l0[d];
        boxes[D + d] = h0[d];
        boxes[2 * D + d] = l1[d];
        boxes[3 * D + d] = h1[d];
    }
    for (int c = 0; c < 2; ++c)
        for (int j = 0; j < KM; ++j) {
            T x[D];
            idxs[c * KM + j] = vmx::split_new_node<T, D>(l, h, dim, c, j, x);
            for (int d = 0; d < D; ++d) xs[(c * KM + j) * D + d] = x[d];
        }
    vmx::split_inherit<T, D>(p, dim, v0, v1);
    for (int k = 0; k < K; ++k) {
        cv[k] = v0[k];
        cv[K + k] = v1[k];
    }
}
/* vol * marginal_subrange over the bin dims of the (bin, region) intersection */
template <typename T, int D>
double sub_n(int db, const double* rlo, const double* rhi, const double* nodes, const double* slo, const double* shi) {
    constexpr int K = vmx::pow3(D);
    T l[D], h[D], v[K];
    for (int d = 0; d < D; ++d) {
        l[d] = T(rlo[d]);
        h[d] = T(rhi[d]);
    }
    for (int k = 0; k < K; ++k) v[k] = T(nodes[k]);
    double m[27], a[3], b[3];
    vmx::bin_marginal<T, double, D>(v, db, m);
    for (int d = 0; d < db; ++d) {
        a[d] = double(vmx::pos_in(T(slo[d]), l[d], h[d]));
        b[d] = double(vmx::pos_in(T(shi[d]), l[d], h[d]));
    }
    return double(vmx::box_volume<T, D>(l, h)) * vmx::marginal_subrange<double>(m, db, a, b);
}
template <typename T, int D>
double mono_n(const double* rlo, const double* rhi, const double* nodes, const double* x) {
    constexpr int K = vmx::pow3(D);
    T v[K], c[K], t[D];
    for (int k = 0; k < K; ++k) v[k] = T(nodes[k]);
    vmx::monomials<T, T, D>(v, c);
    for (int d = 0; d < D; ++d) t[d] = vmx::pos_in(T(x[d]), T(rlo[d]), T(rhi[d]));
    return double(vmx::eval_monomials<T, D>(c, t));
}
}  // namespace

#define VMH_DISPATCH(FN, ...)                                        \
    switch (D) {                                                     \
        case 1: return scalar == 4 ? FN<float, 1>(__VA_ARGS__) : FN<double, 1>(__VA_ARGS__); \
        case 2: return scalar == 4 ? FN<float, 2>(__VA_ARGS__) : FN<double, 2>(__VA_ARGS__); \
        case 3: return scalar == 4 ? FN<float, 3>(__VA_ARGS__) : FN<double, 3>(__VA_ARGS__); \
        case 4: return scalar == 4 ? FN<float, 4>(__VA_ARGS__) : FN<double, 4>(__VA_ARGS__); \
        case 5: return scalar == 4 ? FN<float, 5>(__VA_ARGS__) : FN<double, 5>(__VA_ARGS__); \
        default: break;                                              \
    }

extern "C" {
void vmh_philox(const uint32_t* c, const uint32_t* k, uint32_t* out) {
    const vmx::U4 r = vmx::philox4x32_10(vmx::U4{c[0], c[1], c[2], c[3]}, k[0], k[1]);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
}
void vmh_heuristic(int scalar, int D, int heur, uint64_t n, const double* lo, const double* hi, const double* nodes,
                   double* err, int32_t* dim) {
    VMH_DISPATCH(heur_n, heur, n, lo, hi, nodes, err, dim)
}
void vmh_split(int scalar, int D, int dim, const double* lo, const double* hi, const double* pv, double* boxes,
               double* xs, int32_t* idxs, double* cv) {
    VMH_DISPATCH(split_n, dim, lo, hi, pv, boxes, xs, idxs, cv)
}
double vmh_subrange(int scalar, int D, int db, const double* rlo, const double* rhi, const double* nodes,
                    const double* slo, const double* shi) {
    VMH_DISPATCH(sub_n, db, rlo, rhi, nodes, slo, shi)
    return 0;
}
double vmh_monomial(int scalar, int D, const double* rlo, const double* rhi, const double* nodes, const double* x) {
    VMH_DISPATCH(mono_n, rlo, rhi, nodes, x)
    return 0;
}
uint64_t vmh_max_pairs(void) { return vmx::MAX_PAIRS; }

/* the importance sampler (importance.h) on one line: a point of [t0, t1] from the
   uniform s and the line's node values p[3] */
float vmh_imp_sample_line(float s, float t0, float t1, const float* p) {
    return vmx::imp::sample_line<float>(s, t0, t1, p);
}

/* ... on a region: point x[D] of the box [slo, shi] from the uniforms s[D], weight 1/pdf;
   returns 0 where the pdf's integral over the box is not positive and finite */
int vmh_imp_sample(int D, const float* v, const float* rlo, const float* rhi, float rvol, const float* slo,
                   const float* shi, const float* s, float* x, float* w) {
    switch (D) {
        case 1: return vmx::imp::sample<float, 1>(v, rlo, rhi, rvol, slo, shi, s, x, *w) ? 1 : 0;
        case 2: return vmx::imp::sample<float, 2>(v, rlo, rhi, rvol, slo, shi, s, x, *w) ? 1 : 0;
        case 3: return vmx::imp::sample<float, 3>(v, rlo, rhi, rvol, slo, shi, s, x, *w) ? 1 : 0;
        case 4: return vmx::imp::sample<float, 4>(v, rlo, rhi, rvol, slo, shi, s, x, *w) ? 1 : 0;
        case 5: return vmx::imp::sample<float, 5>(v, rlo, rhi, rvol, slo, shi, s, x, *w) ? 1 : 0;
        case 6: return vmx::imp::sample<float, 6>(v, rlo, rhi, rvol, slo, shi, s, x, *w) ? 1 : 0;
        default: return -1;
    }
}

void vmh_pixel_range(int scalar, double rlo, double rhi, double lo, double hi, uint64_t res, uint64_t* se) {
    if (scalar == 4)
        vmx::pixel_range<float>(float(rlo), float(rhi), float(lo), float(hi), res, se[0], se[1]);
    else
        vmx::pixel_range<double>(rlo, rhi, lo, hi, res, se[0], se[1]);
}

/* the region features of the priority network (FP32 regions), as the kernels compute them */
int vmh_region_features(int D, const double* lo, const double* hi, const double* rlo, const double* rhi,
                        const double* v, const uint32_t* code, const float* lres, float* x, float* s, int* depth) {
    auto run = [&](auto dd) {
        constexpr int DD = decltype(dd)::value;
        constexpr int K = vmx::pow3(DD), KP = vmx::nn::Dims<DD>::KP;
        float l[DD], h[DD], rl[DD], rh[DD], vv[K];
        for (int d = 0; d < DD; ++d) {
            l[d] = float(lo[d]);
            h[d] = float(hi[d]);
            rl[d] = float(rlo[d]);
            rh[d] = float(rhi[d]);
        }
        for (int k = 0; k < K; ++k) vv[k] = float(v[k]);
        vmx::nn::region_features<float, DD, KP>(l, h, rl, rh, vv, code, lres, [x](int i, float val) { x[i] = val; },
                                                *s, *depth);
        return 0;
    };
    switch (D) {
        case 1: return run(std::integral_constant<int, 1>{});
        case 2: return run(std::integral_constant<int, 2>{});
        case 3: return run(std::integral_constant<int, 3>{});
        case 4: return run(std::integral_constant<int, 4>{});
        default: return -1;
    }
}
uint16_t vmh_f2bf(float x) { return vmx::nn::f2bf(x); }
}

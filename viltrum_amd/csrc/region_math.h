/*
 * Region arithmetic of the nested Simpson control variate, shared by every kernel.
 * Plain __host__ __device__ templates: the same header compiles with g++ for the
 * host-side unit tests and with hipcc for gfx950.
 *
 * Operation order and float/double promotions follow the reference so that the FP64
 * instantiation reproduces it bit for bit (device code is built with
 * -ffp-contract=off):
 *   node map      Float(double(p) * (hi - lo) + lo)          src/newton-cotes/region.h:40-47
 *   Simpson       (p0 + 4.0*p1 + p2) / 6.0, double inside     src/newton-cotes/rules.h:63-65
 *   trapezoid     (p0 + p2) / 2.0 on the outer nodes          src/newton-cotes/rules.h:13-15,
 *                                                             src/nested/nested.h:16-23
 *   at / subrange on monomial coefficients                    src/newton-cotes/rules.h:76-100
 *   error(dim)    fold dim with metric(S, T), fold_all Simpson src/newton-cotes/region.h:387-393
 *   integral_subrange folds last dim first                    src/newton-cotes/region.h:141-169
 */
#ifndef VMX_REGION_MATH_H
#define VMX_REGION_MATH_H

#include <math.h>
#include <stdint.h>

#include "../../include/viltrum_amd/integrands.h"

namespace vmx {

VMX_HD constexpr int pow3(int d) { return d <= 0 ? 1 : 3 * pow3(d - 1); }

/* |x| with a +0 result for -0 (std::abs) */
VMX_HD inline float vabs(float x) { return fabsf(x); }
VMX_HD inline double vabs(double x) { return fabs(x); }
template <typename T>
VMX_HD inline T vmax(T a, T b) {  // std::max(a, b)
    return (a < b) ? b : a;
}
template <typename T>
VMX_HD inline T vmin(T a, T b) {  // std::min(a, b)
    return (b < a) ? b : a;
}
template <typename T>
VMX_HD inline bool visnan(T x) {
    return x != x;
}

/* digit d (base 3, dim 0 fastest) of a node index */
VMX_HD inline int digit3(int idx, int d) {
    for (int i = 0; i < d; ++i) idx /= 3;
    return idx % 3;
}

template <typename T>
VMX_HD inline T node_coord(double p, T lo, T hi) {
    return T(p * (hi - lo) + lo);
}

template <typename T>
VMX_HD inline T simpson(T p0, T p1, T p2) {
    return T((p0 + 4.0 * p1 + p2) / 6.0);
}
template <typename T>
VMX_HD inline T trapezoid(T p0, T p2) {
    return T((p0 + p2) / 2.0);
}
template <typename T>
VMX_HD inline T simpson_at(T t, T p0, T p1, T p2) {
    const T c0 = p0;
    const T c1 = -3 * p0 + 4 * p1 - p2;
    const T c2 = 2 * p0 - 4 * p1 + 2 * p2;
    return (c2 * t + c1) * t + c0;
}
/* Exact integral over [a, b] (normalised) of the quadratic through p at 0, .5, 1. */
template <typename T>
VMX_HD inline T simpson_subrange(T a, T b, T p0, T p1, T p2) {
    const T c0 = p0;
    const T c1 = -3 * p0 + 4 * p1 - p2;
    const T c2 = 2 * p0 - 4 * p1 + 2 * p2;
    return T(((c2 * b / 3.0 + c1 / 2.0) * b + c0) * b - ((c2 * a / 3.0 + c1 / 2.0) * a + c0) * a);
}

/* simpson_subrange with the divisions by 3 replaced by products with 1/3 (within an ulp
   of the quotient): for sums whose result is rounded to FP32 afterwards. */
template <typename T>
VMX_HD inline T simpson_subrange_mul(T a, T b, T p0, T p1, T p2) {
    const T third = T(1) / T(3);
    const T c0 = p0;
    const T c1 = -3 * p0 + 4 * p1 - p2;
    const T c2 = 2 * p0 - 4 * p1 + 2 * p2;
    return ((c2 * b * third + c1 * T(0.5)) * b + c0) * b - ((c2 * a * third + c1 * T(0.5)) * a + c0) * a;
}

template <typename T>
VMX_HD inline T metric_relative(T a, T b) {  // error_metric_relative(1e-37)
    const T ma = vmax(vabs(a), vabs(b));
    if (double(ma) < 1.e-37) return vabs(b - a);
    return vabs(b - a) / ma;
}
template <typename T>
VMX_HD inline T metric_absolute(T a, T b) {
    return vabs(b - a);
}

template <typename T, int D>
VMX_HD inline T box_volume(const T* lo, const T* hi) {
    T v = T(1);
    for (int i = 0; i < D; ++i) v *= (hi[i] - lo[i]);
    return v;
}

/* Iterated Simpson over k dims, dim 0 fastest (fold_all). Overwrites a. */
template <typename T>
VMX_HD inline T fold_all_simpson(T* a, int k) {
    int n = pow3(k);
    while (n > 1) {
        n /= 3;
        for (int j = 0; j < n; ++j) a[j] = simpson(a[3 * j], a[3 * j + 1], a[3 * j + 2]);
    }
    return a[0];
}

/* Region::error(dim, metric): fold `dim` with metric(Simpson, trapezoid) line by line,
   then fold_all Simpson over the remaining dims (in increasing order). */
template <typename T, int D, bool Relative>
VMX_HD inline T region_error(const T* v, T vol, int dim) {
    constexpr int KM = pow3(D - 1);
    T e[KM];
    const int stride = pow3(dim);
    const int below = stride;  // number of index values of the dims below `dim`
    for (int j = 0; j < KM; ++j) {
        // remaining dims keep their relative order: split j at `dim`
        const int full = (j % below) + (j / below) * below * 3;
        const T p0 = v[full], p1 = v[full + stride], p2 = v[full + 2 * stride];
        const T s = simpson(p0, p1, p2), t = trapezoid(p0, p2);
        e[j] = Relative ? metric_relative(s, t) : metric_absolute(s, t);
    }
    return vol * fold_all_simpson(e, D - 1);
}

/* (err, dim) of a region under the chosen heuristic. */
template <typename T, int D>
VMX_HD inline void heuristic(const T* v, const T* lo, const T* hi, T vol, int which, T& err_out, int& dim_out) {
    if (which == 0) {  // error_heuristic_size(error_metric_relative(), 1e-5, 1e-37)
        const double size_weight = 1.e-5, min_size = 1.e-37;
        T max_err = region_error<T, D, true>(v, vol, 0);
        const T ext0 = hi[0] - lo[0];
        if (ext0 < min_size || visnan(ext0))
            max_err = 0;
        else
            max_err += size_weight * vabs(ext0);
        int max_dim = 0;
        for (int d = 1; d < D; ++d) {
            const T ext = hi[d] - lo[d];
            T err = region_error<T, D, true>(v, vol, d) + size_weight * vabs(ext);
            if (ext < min_size) err = 0;
            if (err >= max_err) {
                max_err = err;
                max_dim = d;
            }
        }
        err_out = max_err;
        dim_out = max_dim;
    } else {  // error_heuristic_default(error_metric_absolute())
        T max_err = 0;
        int max_dim = 0;
        for (int d = 0; d < D; ++d) {
            const T err = region_error<T, D, false>(v, vol, d);
            if (err > max_err) {
                max_err = err;
                max_dim = d;
            }
        }
        err_out = max_err;
        dim_out = max_dim;
    }
}

/* Child boxes of a split in two along `dim`. */
template <typename T, int D>
VMX_HD inline void split_boxes(const T* lo, const T* hi, int dim, T* lo0, T* hi0, T* lo1, T* hi1) {
    const T dd = (hi[dim] - lo[dim]) / T(2);
    for (int d = 0; d < D; ++d) {
        lo0[d] = lo1[d] = lo[d];
        hi0[d] = hi1[d] = hi[d];
    }
    hi0[dim] = lo[dim] + dd * T(1);
    lo1[dim] = hi0[dim];
}

/* j with a base-3 digit 0 inserted at position d: (j mod 3^d) + (j div 3^d) * 3^(d+1).
   The switch keeps every divisor a compile-time constant (an integer division by a
   run-time value costs tens of instructions on the GPU). */
VMX_HD inline int insert_digit0(int j, int d) {
    const unsigned u = unsigned(j);
    switch (d) {
        case 0: return int(u * 3u);
        case 1: return int(u % 3u + (u / 3u) * 9u);
        case 2: return int(u % 9u + (u / 9u) * 27u);
        case 3: return int(u % 27u + (u / 27u) * 81u);
        case 4: return int(u % 81u + (u / 81u) * 243u);
        default: return int(u % 243u + (u / 243u) * 729u);
    }
}

/* Coordinates of the j-th new node of a split (j < 3^(D-1)): child c (0 or 1) gets its
   middle plane at parent-relative .25 / .75 along `dim`; the other dims take {0,.5,1}
   through the parent's node map.  Returns the child's node index the value lands on. */
template <typename T, int D>
VMX_HD inline int split_new_node(const T* lo, const T* hi, int dim, int c, int j, T* x) {
    const int idx = insert_digit0(j, dim) + pow3(dim);  // digit_dim = 1
    for (int d = 0; d < D; ++d)
        x[d] = node_coord<T>(d == dim ? (c == 0 ? 0.25 : 0.75) : double(digit3(idx, d)) / 2.0, lo[d], hi[d]);
    return idx;
}

/* Child node values inherited from the parent planes (entries with digit_dim != 1). */
template <typename T, int D>
VMX_HD inline void split_inherit(const T* pv, int dim, T* v0, T* v1) {
    const int stride = pow3(dim);
    for (int idx = 0; idx < pow3(D); ++idx) {
        const int j = digit3(idx, dim);
        const int base = idx - j * stride;
        if (j == 0) {
            v0[idx] = pv[base];
            v1[idx] = pv[base + stride];
        } else if (j == 2) {
            v0[idx] = pv[base + stride];
            v1[idx] = pv[base + 2 * stride];
        }
    }
}

/* Range::pos_in_range */
template <typename T>
VMX_HD inline T pos_in(T x, T lo, T hi) {
    if (lo >= hi) return lo;
    return (x - lo) / (hi - lo);
}

/* Bin-dim marginal of a region: fold the non-bin dims D-1 .. db with subrange(0, 1),
   exactly as integral_subrange does before it reaches the bin dims.  Output has 3^db
   entries, dim 0 fastest, in the accumulation type A. */
template <typename T, typename A, int D>
VMX_HD inline void bin_marginal(const T* v, int db, A* out) {
    A w[pow3(D)];
    for (int i = 0; i < pow3(D); ++i) w[i] = A(v[i]);
    int n = pow3(D);
    for (int d = D - 1; d >= db; --d) {
        n /= 3;
        for (int j = 0; j < n; ++j) w[j] = simpson_subrange(A(0), A(1), w[j], w[j + n], w[j + 2 * n]);
    }
    for (int i = 0; i < n; ++i) out[i] = w[i];
}

/* integral_subrange restricted to the bin dims on a marginal m (3^DB entries): folds
   dim DB-1 first, dim 0 last.  a/b are normalised bounds. */
template <typename A, int DB, bool EXACT = true>
VMX_HD inline A marginal_subrange_n(const A* m, const A* a, const A* b) {
    A w[pow3(DB)];
#pragma unroll
    for (int i = 0; i < pow3(DB); ++i) w[i] = m[i];
#pragma unroll
    for (int d = DB - 1; d >= 0; --d) {
        const int n = pow3(d);
#pragma unroll
        for (int j = 0; j < n; ++j)
            w[j] = EXACT ? simpson_subrange(a[d], b[d], w[j], w[j + n], w[j + 2 * n])
                         : simpson_subrange_mul(a[d], b[d], w[j], w[j + n], w[j + 2 * n]);
    }
    return w[0];
}
template <typename A>
VMX_HD inline A marginal_subrange(const A* m, int db, const A* a, const A* b) {
    if (db == 1) return marginal_subrange_n<A, 1>(m, a, b);
    if (db == 2) return marginal_subrange_n<A, 2>(m, a, b);
    return marginal_subrange_n<A, 3>(m, a, b);
}

/* The per-bin footprint lists are indexed by 32-bit offsets (an exclusive scan of the
   pair counts, one extra slot for the end): a batch's (bin, footprint) pair total must
   not exceed this. */
constexpr uint64_t MAX_PAIRS = 0xFFFFFFFEull;

/* pixels_in_region: bin index box [start, end) along bin dim i. */
template <typename T>
VMX_HD inline void pixel_range(T rlo, T rhi, T lo, T hi, uint64_t res, uint64_t& start, uint64_t& end) {
    start = uint64_t(T(res) * (rlo - lo) / (hi - lo));
    // the reference indexes past the last bin when rounding lands exactly on `res`
    // (undefined behaviour there); clamp so every bin index stays in bounds
    if (start >= res) start = res - 1;
    const uint64_t e = uint64_t(0.99f + (T(res) * (rhi - lo) / (hi - lo)));
    end = e < res ? e : res;
    if (end < start + 1) end = start + 1;
}

/* Monomial coefficients of the tensor quadratic through the 3^D nodes, in normalised
   coordinates: P(t) = sum_k c_k prod_d t_d^(k_d).  Per line: c0 = p0,
   c1 = -3 p0 + 4 p1 - p2, c2 = 2 p0 - 4 p1 + 2 p2 (src/newton-cotes/rules.h:76-83). */
template <typename T, typename C, int D>
VMX_HD inline void monomials(const T* v, C* c) {
    for (int i = 0; i < pow3(D); ++i) c[i] = C(v[i]);
    int stride = 1;
    for (int d = 0; d < D; ++d) {
        for (int idx = 0; idx < pow3(D); ++idx) {
            if (digit3(idx, d) != 0) continue;
            const C p0 = c[idx], p1 = c[idx + stride], p2 = c[idx + 2 * stride];
            c[idx] = p0;
            c[idx + stride] = -3 * p0 + 4 * p1 - p2;
            c[idx + 2 * stride] = 2 * p0 - 4 * p1 + 2 * p2;
        }
        stride *= 3;
    }
}

/* a*b + c, fused (one rounding) when FUSED */
template <bool FUSED, typename C>
VMX_HD inline C mul_add(C a, C b, C c) {
    if constexpr (FUSED)
        return C(__builtin_fma(a, b, c));
    else
        return a * b + c;
}
template <>
VMX_HD inline float mul_add<true, float>(float a, float b, float c) {
    return __builtin_fmaf(a, b, c);
}

/* Evaluate the monomial form at normalised t: nested Horner, dim 0 innermost, written
   as template recursion so every coefficient stays in registers.  FUSED: each Horner
   step is one fused multiply-add. */
template <typename C, int DD, bool FUSED = false>
struct MonoHorner {
    VMX_HD static C eval(const C* c, const C* t) {
        constexpr int S = pow3(DD - 1);
        const C a0 = MonoHorner<C, DD - 1, FUSED>::eval(c, t);
        const C a1 = MonoHorner<C, DD - 1, FUSED>::eval(c + S, t);
        const C a2 = MonoHorner<C, DD - 1, FUSED>::eval(c + 2 * S, t);
        return mul_add<FUSED>(mul_add<FUSED>(a2, t[DD - 1], a1), t[DD - 1], a0);
    }
};
template <typename C, bool FUSED>
struct MonoHorner<C, 0, FUSED> {
    VMX_HD static C eval(const C* c, const C*) { return c[0]; }
};
template <typename C, int D, bool FUSED = false>
VMX_HD inline C eval_monomials(const C* c, const C* t) {
    return MonoHorner<C, D, FUSED>::eval(c, t);
}

}  // namespace vmx

#endif

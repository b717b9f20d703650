/*
 * Learned refinement priority: region features, BF16 rounding, the network's
 * parameter layout and the FP64 training targets.  Plain __host__ __device__ code, so
 * the host unit tests compile the very functions the kernels run.  The reference has
 * no learned component; its error heuristic (src/nested/error-heuristic.h:21-47) stays
 * the split-dimension rule, the fallback and the baseline.
 *
 * Priority of a region r with 3^D node values v and dyadic depth n_d per dim:
 *     s  = max_k |v_k|                      (scale of the integrand on r)
 *     p  = 2^(y * sigma + mu) * s^2 * 2^(-sum_d n_d)
 * where y is the network output on the features below, i.e. the network predicts
 * t = log2(E / (s^2 vol_r)) for one of the FP64 targets E (TargetKind):
 *     ENERGY    integral over r of (f - P_r)^2, P_r the tensor-quadratic control variate
 *               (the default: the per-bin estimator picks one of the n regions meeting a
 *               bin uniformly and weights it by n times its overlap, so a region's
 *               residual mean costs variance as well as its spread)
 *     BINVAR    summed over the bins b meeting r, int_{b^r} (f - P_r)^2 -
 *               (int_{b^r} (f - P_r))^2 / |b^r|: the residual's variance within each bin
 *               only, exact when one region covers each bin
 *     GAIN      BINVAR of r minus BINVAR of its two children (split along the
 *               heuristic's dim), the reduction one split buys
 * Features (F = 4D + 3^D + 1):
 *     [0, D)            centre of r in the integration range, in [0, 1]
 *     [D, 2D)           -n_d / 16      (relative extent is exactly 2^-n_d)
 *     [2D, 3D)          (log2 res_d - n_d) / 16: log2 of the bins across r (0 = one
 *                       bin; integrated dims count one bin)
 *     [3D, 3D+K)        v_k / s
 *     [3D+K, 4D+K)      Simpson fold over the other dims of |S - T| of the lines along d, / s
 *     [4D+K]            sum_d n_d / 32
 * then zero padding to KP (a multiple of 32, the MFMA K step).
 */
#ifndef VMX_FEATURES_H
#define VMX_FEATURES_H

#include "region_math.h"

namespace vmx {
namespace nn {

constexpr int W = 64;  // hidden width of the 4-layer MLP

template <int D>
struct Dims {
    static constexpr int K = pow3(D);
    static constexpr int F = 4 * D + K + 1;
    static constexpr int KP = (F + 31) / 32 * 32;
};

/* FP32 master parameters, one flat array:
   W0[KP][W] b0[W] W1[W][W] b1[W] W2[W][W] b2[W] w3[W] b3 | mu sigma
   (W_l[i][o]: input i to output o).  The last two entries standardise the target and
   are not trained. */
template <int KP>
struct Layout {
    static constexpr int w0 = 0;
    static constexpr int b0 = w0 + KP * W;
    static constexpr int w1 = b0 + W;
    static constexpr int b1 = w1 + W * W;
    static constexpr int w2 = b1 + W;
    static constexpr int b2 = w2 + W * W;
    static constexpr int w3 = b2 + W;
    static constexpr int b3 = w3 + W;
    static constexpr int trained = b3 + 1;
    static constexpr int mu = trained;
    static constexpr int sigma = trained + 1;
    static constexpr int total = trained + 2;
};

/* float -> bfloat16, round to nearest even (NaN stays a quiet NaN). */
VMX_HD inline uint16_t f2bf(float x) {
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
VMX_HD inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, uint32_t(h) << 16); }

/* number of halvings along a dim, from the dyadic code (leading 1 then the path) */
VMX_HD inline int code_depth(uint32_t code) {
    int n = 0;
    while (code > 1u) {
        code >>= 1;
        ++n;
    }
    return n;
}

/* Features of one region, handed to put(i, value) for i = 0..KP-1 (floats before BF16
   rounding), plus s and the total depth, from which the caller forms the priority
   scale 2 log2 s - depth.  lres[d] = log2 of the bin count along dim d (0 for
   integrated dims). */
template <typename T, int D, int KP, typename Put>
VMX_HD inline void region_features(const T* lo, const T* hi, const T* rlo, const T* rhi, const T* v,
                                   const uint32_t* code, const float* lres, Put&& put, float& s_out, int& depth_out) {
    constexpr int K = pow3(D), KM = pow3(D - 1);
    float s = 0.f;
    for (int k = 0; k < K; ++k) s = vmax(s, vabs(float(v[k])));
    const float se = vmax(s, 0x1p-100f);
    const float inv = 1.f / se;
    int depth = 0;
    for (int d = 0; d < D; ++d) {
        const T c = T((T(0.5) * (lo[d] + hi[d]) - rlo[d]) / (rhi[d] - rlo[d]));
        put(d, float(c));
        const int n = code_depth(code[d]);
        depth += n;
        put(D + d, -float(n) * (1.f / 16.f));
        put(2 * D + d, (lres[d] - float(n)) * (1.f / 16.f));
    }
    for (int k = 0; k < K; ++k) put(3 * D + k, float(v[k]) * inv);
    for (int d = 0; d < D; ++d) {
        const int below = pow3(d);
        float acc = 0.f;
        for (int jl = 0; jl < KM; ++jl) {
            const int full = (jl % below) + (jl / below) * below * 3;
            const float p0 = float(v[full]), p1 = float(v[full + below]), p2 = float(v[full + 2 * below]);
            const float diff = vabs(simpson(p0, p1, p2) - trapezoid(p0, p2));
            // Simpson weight of line jl over the other dims: prod (1,4,1)/6 by digit
            float w = 1.f;
            int r = jl;
            for (int e = 0; e < D - 1; ++e) {
                const int dg = r % 3;
                r /= 3;
                w *= (dg == 1) ? (4.f / 6.f) : (1.f / 6.f);
            }
            acc += w * diff;
        }
        put(3 * D + K + d, acc * inv);
    }
    put(4 * D + K, float(depth) * (1.f / 32.f));
    for (int i = 4 * D + K + 1; i < KP; ++i) put(i, 0.f);
    s_out = s;
    depth_out = depth;
}

/* ---------------------------------------------------------------- FP64 true error
   e_r = integral over r of (f - P_r)^2, by a composite 3-point Gauss-Legendre product
   rule with Q cells per dim (exact for r of degree <= 5 per cell and dim). */
VMX_HD inline double lagrange3(int k, double u) {
    return k == 0 ? 2.0 * (u - 0.5) * (u - 1.0) : (k == 1 ? -4.0 * u * (u - 1.0) : 2.0 * u * (u - 0.5));
}

template <typename T, int D>
VMX_HD inline double quadratic_at(const T* v, const double* u) {
    constexpr int K = pow3(D);
    double acc = 0.0;
    for (int k = 0; k < K; ++k) {
        double w = double(v[k]);
        int r = k;
        for (int d = 0; d < D; ++d) {
            w *= lagrange3(r % 3, u[d]);
            r /= 3;
        }
        acc += w;
    }
    return acc;
}

template <int D>
struct TrueErrorRule {
    static constexpr int Q = D <= 2 ? 3 : (D == 3 ? 2 : 1);  // cells per dim
    static constexpr int G = 3 * Q;                         // points per dim
};

VMX_HD inline void gauss3_point(int i, int Q, double& u, double& w) {
    const double h = 1.0 / double(Q);
    const int cell = i / 3, g = i % 3;
    const double off = 0.5 * 0.7745966692414834;  // sqrt(3/5) / 2
    const double t = g == 0 ? 0.5 - off : (g == 1 ? 0.5 : 0.5 + off);
    u = (double(cell) + t) * h;
    w = (g == 1 ? 8.0 / 18.0 : 5.0 / 18.0) * h;
}

/* f64 is the integrand evaluated in double (its own eval) */
template <typename T, int D, typename F64, typename M>
VMX_HD inline double true_error(const F64& f64, const T* lo, const T* hi, const T* v) {
    constexpr int G = TrueErrorRule<D>::G, Q = TrueErrorRule<D>::Q;
    int total = 1;
    for (int d = 0; d < D; ++d) total *= G;
    double vol = 1.0;
    for (int d = 0; d < D; ++d) vol *= double(hi[d]) - double(lo[d]);
    double acc = 0.0;
    for (int q = 0; q < total; ++q) {
        double u[D], x[D], w = 1.0;
        int r = q;
        for (int d = 0; d < D; ++d) {
            double ud, wd;
            gauss3_point(r % G, Q, ud, wd);
            r /= G;
            u[d] = ud;
            w *= wd;
            x[d] = double(lo[d]) + ud * (double(hi[d]) - double(lo[d]));
        }
        const double res = f64.template eval<M, D>(x) - quadratic_at<T, D>(v, u);
        acc += w * res * res;
    }
    return vol * acc;
}

/* ---------------------------------------------------------------- FP64 bin variance */
enum TargetKind { TARGET_ENERGY = 0, TARGET_BINVAR = 1, TARGET_GAIN = 2 };

/* Bins of width step[d] from glo[d] along the bin dims d < db. */
struct BinGrid {
    double glo[3], step[3];
    uint64_t res[3];
    int db;
};

constexpr int BINVAR_MAX_BINS = 4;  // bins sampled per bin dim and region (evenly spaced)

/* 3-point Gauss-Legendre product rule over the box [a, b]: sums of g and g^2 for the
   residuals gp = f - P(vp on [plo, phi]) and gc = f - P(vc on [clo, chi]) at the same
   points (f evaluated once per point). */
template <int D, typename F64, typename M>
VMX_HD inline void piece_sums(const F64& f, const double* a, const double* b, const double* plo, const double* phi,
                              const double* vp, const double* clo, const double* chi, const double* vc,
                              double* sp, double* sc) {
    const double off = 0.5 * 0.7745966692414834;
    for (int q = 0; q < pow3(D); ++q) {
        double x[D], up[D], uc[D], w = 1.0;
        int r = q;
        for (int d = 0; d < D; ++d) {
            const int g = r % 3;
            r /= 3;
            const double t = g == 0 ? 0.5 - off : (g == 1 ? 0.5 : 0.5 + off);
            x[d] = a[d] + t * (b[d] - a[d]);
            w *= (g == 1 ? 8.0 / 18.0 : 5.0 / 18.0) * (b[d] - a[d]);
            up[d] = (x[d] - plo[d]) / (phi[d] - plo[d]);
            if (vc) uc[d] = (x[d] - clo[d]) / (chi[d] - clo[d]);
        }
        const double fx = f.template eval<M, D>(x);
        const double gp = fx - quadratic_at<double, D>(vp, up);
        sp[0] += w * gp;
        sp[1] += w * gp * gp;
        if (vc) {
            const double gc = fx - quadratic_at<double, D>(vc, uc);
            sc[0] += w * gc;
            sc[1] += w * gc * gc;
        }
    }
}

/* Within-bin residual variance of region [lo, hi] (nodes vp), and, when v0/v1 are given,
   of its two children split at the middle of `dim` (nodes v0, v1).  A region meeting more
   than BINVAR_MAX_BINS bins along a bin dim is measured on that many evenly spaced bins
   and scaled by the count. */
template <int D, typename F64, typename M>
VMX_HD inline void bin_variances(const F64& f, const BinGrid& g, const double* lo, const double* hi,
                                 const double* vp, int dim, const double* v0, const double* v1, double& Vp,
                                 double& Vc) {
    double clo[2][D], chi[2][D];
    for (int d = 0; d < D; ++d) {
        clo[0][d] = clo[1][d] = lo[d];
        chi[0][d] = chi[1][d] = hi[d];
    }
    const double mid = lo[dim] + (hi[dim] - lo[dim]) / 2.0;
    chi[0][dim] = mid;
    clo[1][dim] = mid;
    uint64_t i0[3] = {0, 0, 0}, cnt[3] = {1, 1, 1}, nsel[3] = {1, 1, 1};
    double scale = 1.0;
    uint64_t total = 1;
    for (int d = 0; d < g.db; ++d) {
        double a = floor((lo[d] - g.glo[d]) / g.step[d]);
        double b = ceil((hi[d] - g.glo[d]) / g.step[d]);
        if (a < 0) a = 0;
        if (a > double(g.res[d] - 1)) a = double(g.res[d] - 1);
        if (b > double(g.res[d])) b = double(g.res[d]);
        if (b < a + 1) b = a + 1;
        i0[d] = uint64_t(a);
        cnt[d] = uint64_t(b) - i0[d];
        nsel[d] = cnt[d] < uint64_t(BINVAR_MAX_BINS) ? cnt[d] : uint64_t(BINVAR_MAX_BINS);
        scale *= double(cnt[d]) / double(nsel[d]);
        total *= nsel[d];
    }
    const bool kids = v0 != nullptr && v1 != nullptr;
    Vp = 0.0;
    Vc = 0.0;
    for (uint64_t idx = 0; idx < total; ++idx) {
        double a[D], b[D];
        bool empty = false;
        uint64_t rem = idx;
        for (int d = 0; d < D; ++d) {
            a[d] = lo[d];
            b[d] = hi[d];
            if (d < g.db) {
                const uint64_t k = rem % nsel[d];
                rem /= nsel[d];
                const uint64_t bin =
                    i0[d] + (nsel[d] == cnt[d] ? k : uint64_t((double(k) + 0.5) * double(cnt[d]) / double(nsel[d])));
                const double bl = g.glo[d] + double(bin) * g.step[d], bh = g.glo[d] + double(bin + 1) * g.step[d];
                a[d] = vmax(lo[d], bl);
                b[d] = vmin(hi[d], bh);
            }
            if (!(a[d] < b[d])) empty = true;
        }
        if (empty) continue;
        double vol = 1.0;
        for (int d = 0; d < D; ++d) vol *= b[d] - a[d];
        double sp[2] = {0.0, 0.0};
        for (int c = 0; c < 2; ++c) {
            double sa[D], sb[D];
            for (int d = 0; d < D; ++d) {
                sa[d] = a[d];
                sb[d] = b[d];
            }
            sa[dim] = vmax(a[dim], clo[c][dim]);
            sb[dim] = vmin(b[dim], chi[c][dim]);
            if (!(sa[dim] < sb[dim])) continue;
            double sc[2] = {0.0, 0.0};
            piece_sums<D, F64, M>(f, sa, sb, lo, hi, vp, clo[c], chi[c], kids ? (c == 0 ? v0 : v1) : nullptr, sp,
                                  sc);
            if (kids) {
                double cv = 1.0;
                for (int d = 0; d < D; ++d) cv *= sb[d] - sa[d];
                Vc += scale * (sc[1] - sc[0] * sc[0] / cv);
            }
        }
        Vp += scale * (sp[1] - sp[0] * sp[0] / vol);
    }
}

/* training target: log2(e / (s^2 vol)), clamped to [-60, 30] */
VMX_HD inline float error_target(double e, float s, double vol) {
    const double se = double(vmax(s, 0x1p-100f));
    const double q = e / (se * se * vol);
    double t = q > 0x1p-60 ? log2(q) : -60.0;
    if (t > 30.0) t = 30.0;
    return float(t);
}

}  // namespace nn
}  // namespace vmx

#endif

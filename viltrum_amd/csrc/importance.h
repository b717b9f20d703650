/*
 * Importance sampling of a region's polynomial over a box (region_sampling_importance,
 * src/control-variates/region-sampling.h:22-44, with NormDefault): a point is drawn
 * from the density proportional to the Simpson "pdf" of the region's node values,
 * dimension by dimension (the first dim from its marginal, the later dims integrated
 * over the box, last first; then the tensor is conditioned on the drawn coordinate,
 * Region::sample_subrange, src/newton-cotes/region.h:220-265), and weighted by 1/pdf
 * (Region::pdf_subrange, :324-330).  Along one line the pdf is |p| shifted so that the
 * quadratic's interior minimum is 0 (src/newton-cotes/rules.h:104-145); its CDF, a
 * cubic, is inverted by Cardano's formula or Viete's trigonometric one (:166-237).
 *
 * The arithmetic follows the library's mixed float/double expressions as written (the
 * sign of a coefficient is a float, NormDefault::sign at src/control-variates/norm.h:28-29;
 * the cubic's coefficients and the root formulas are double), with cheaper forms of the
 * same values: integer powers (the library's pow(x, 2.) and pow(x, 3.)) as products, its
 * cube root pow(|x|, 1/3.) as cbrt, Viete's three cosines from one sincos by the
 * angle-difference identity, divisions by constants and by one value repeatedly as
 * products with reciprocals, and the line minimum's position in T, each within a few
 * ulps of the library's double result.
 * One deviation: where the pdf integrates to 0 over the box (a region whose nodes are
 * all 0), the reference divides 0 by 0 and its bin becomes NaN; here sample() reports it
 * and the caller samples the box uniformly instead.
 */
#ifndef VMX_IMPORTANCE_H
#define VMX_IMPORTANCE_H

#include "region_math.h"

namespace vmx {
namespace imp {

constexpr int MAX_DIM = 6;  // dims 1..3 keep the 3^D tensors in registers, 4..6 fold depth first

template <typename T>
VMX_HD inline T at3(T t, const T* p) {
    const T c0 = p[0], c1 = -3 * p[0] + 4 * p[1] - p[2], c2 = 2 * p[0] - 4 * p[1] + 2 * p[2];
    return (c2 * t + c1) * t + c0;
}
template <typename T>
VMX_HD inline T subrange3(T a, T b, const T* p) {
    const T c0 = p[0], c1 = -3 * p[0] + 4 * p[1] - p[2], c2 = 2 * p[0] - 4 * p[1] + 2 * p[2];
    return T(((c2 * b * (1.0 / 3.0) + c1 * 0.5) * b + c0) * b - ((c2 * a * (1.0 / 3.0) + c1 * 0.5) * a + c0) * a);
}
template <typename T>
VMX_HD inline T cdf3(T t, const T* p) {
    const T c0 = p[0], c1 = -3 * p[0] + 4 * p[1] - p[2], c2 = 2 * p[0] - 4 * p[1] + 2 * p[2];
    return T(((c2 * t * (1.0 / 3.0) + c1 * 0.5) * t + c0) * t);
}
template <typename T>
VMX_HD inline void pdf_points(const T* p, T* q) {
    q[0] = vabs(p[0]);
    q[1] = vabs(p[1]);
    q[2] = vabs(p[2]);
    const T c1 = -3 * q[0] + 4 * q[1] - q[2], c2 = 2 * q[0] - 4 * q[1] + 2 * q[2], c0 = q[0];
    T ymin = 0;
    if (c2 > 0) {
        const T tmin = -c1 / (T(2) * c2);
        if ((tmin > 0) && (tmin < 1)) {
            const T ytmin = (c2 * tmin + c1) * tmin + c0;
            if (ytmin < ymin) ymin = ytmin;
        }
    }
    for (int i = 0; i < 3; ++i) q[i] -= ymin;
}
template <typename T>
VMX_HD inline T pdf_integral(T a, T b, const T* p) {
    T q[3];
    pdf_points(p, q);
    return subrange3(a, b, q);
}
template <typename T>
VMX_HD inline T pdf_unnormalized(T t, const T* p) {
    T q[3];
    pdf_points(p, q);
    return at3(t, q);
}
VMX_HD inline float sign_of(float v) { return v; }
VMX_HD inline float sign_of(double v) { return float(v); }
VMX_HD inline double cuberoot(double x) { return cbrt(x); }

/* roots of cdf(t) = x in order; returns their count (0..3) */
template <typename T>
VMX_HD inline int inv_cdf(T x, const T* points, T* sol) {
    const T c0 = points[0], c1 = -3 * points[0] + 4 * points[1] - points[2],
            c2 = 2 * points[0] - 4 * points[1] + 2 * points[2];
    double a = sign_of(c2) / 3.;
    double b = sign_of(c1) / 2.;
    float c = sign_of(c0);
    T d = -x;
    const double sum = a + b + c + d, isum = 1.0 / sum;
    a *= isum;
    b *= isum;
    c = float(c * isum);
    d = T(d * isum);
    int n = 0;
    if (fabs(a) < 1.e-3) {
        if (fabs(b) < 1.e-3) {
            if (fabs(c) >= 1.e-3) sol[n++] = T(-d / c);
        } else {
            const double disc = c * c - 4 * b * d;
            if (disc >= 0) {
                sol[n++] = T((-c + sqrt(disc)) / (2 * b));
                sol[n++] = T((-c - sqrt(disc)) / (2 * b));
            }
        }
    } else {
        const double ia = 1.0 / a, ia2 = ia * ia, b2 = b * b, b3 = b2 * b;
        const double pp = c * ia - b2 * ia2 * (1. / 3.);
        const double qq = b3 * ia2 * ia * (2. / 27.) - b * c * ia2 * (1. / 3.) + d * ia;
        const double hq = qq / 2., tp = pp / 3.;
        const double sqr = hq * hq + tp * tp * tp;
        if (sqr >= 0.) {
            const double sq = sqrt(sqr);
            sol[n++] = T(cuberoot(-hq - sq) + cuberoot(-hq + sq) - b * ia * (1. / 3.));
        } else {
            // Viete: 2 sqrt(-p/3) cos(phi - 2 pi i / 3) - b / 3a, phi = acos(...) / 3; the
            // three cosines from one sincos(phi) by the angle-difference identity
            const double ipp = 1.0 / pp;
            const double phi = 1. / 3. * acos(1.5 * qq * ipp * sqrt(-3. * ipp));
            double sphi, cphi;
            sincos(phi, &sphi, &cphi);
            const double r = 2. * sqrt(-pp * (1. / 3.)), sh = b * ia * (1. / 3.);
            const double h3 = 0.86602540378443864676;  // sin(2 pi / 3)
            sol[n++] = T(r * cphi - sh);
            sol[n++] = T(r * (-0.5 * cphi + h3 * sphi) - sh);
            sol[n++] = T(r * (-0.5 * cphi - h3 * sphi) - sh);
        }
    }
    return n;
}

/* The pdf's integral over the normalised box [an, bn] of the 3^D node values v (dim 0
   fastest), without the region volume (Region::pdf_sub, src/newton-cotes/region.h:
   271-283): folds the last dim first, here depth first over the slices, so any D keeps
   3 values per level live; each fold sees the same three values as in the breadth-first
   order.  Call with L = 0 and base = 0. */
template <typename T, int D, int L>
VMX_HD inline T pdf_fold(const T* v, int base, const T* an, const T* bn) {
    constexpr int stride = pow3(L);
    T w[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        if constexpr (L == D - 1)
            w[j] = v[base + j * stride];
        else
            w[j] = pdf_fold<T, D, L + 1>(v, base + j * stride, an, bn);
    }
    return pdf_integral(an[L], bn[L], w);
}

/* one coordinate in [t0, t1] from the line's pdf, s uniform in [0, 1) */
template <typename T>
VMX_HD inline T sample_line(T s, T t0, T t1, const T* p) {
    T q[3];
    pdf_points(p, q);
    const T x = s * cdf3(t1, q) + (T(1) - s) * cdf3(t0, q);
    T sol[3];
    const int n = inv_cdf(x, q, sol);
    for (int i = 0; i < n; ++i) {
        T r = sol[i];
        if ((r < t0) && (vabs(t0 - r) < 1.e-5)) r = t0;
        if ((r > t1) && (vabs(r - t1) < 1.e-5)) r = t1;
        if ((r >= t0) && (r <= t1) && !(r != r)) return r;
    }
    return s * (t1 - t0) + t0;
}

/* A point of the box [slo, shi] inside region [rlo, rhi] (volume rvol, node values v,
   dim 0 fastest) from the uniforms s[D], and its weight (dims 1..3: the tensors in
   registers).
   - strategy_rr false (region_sampling_importance, src/control-variates/
     region-sampling.h:22-44): drawn from the polynomial's pdf, weight 1/pdf (0 where
     pdf < 1e-5).  Returns false where the pdf's integral over the box is not positive
     and finite.
   - strategy_rr true (region_sampling_russian_roulette, region-sampling.h:47-84): with
     probability pdf_integral / (pdf_integral + svol) (svol the box volume, the uniform
     rr in [0, 1) decides) drawn from the pdf, else uniformly (x = slo + s (shi - slo),
     pdf = 1 / svol); weight 1/pdf of the strategy taken, 0 where pdf < 1e-10.  Returns
     true. */
template <typename T, int D>
VMX_HD inline bool sample_regs(const T* v, const T* rlo, const T* rhi, T rvol, const T* slo, const T* shi, const T* s,
                               T* x, T& weight, bool strategy_rr, T rr, T svol) {
    static_assert(D >= 1 && D <= 3, "register tensors: 1..3 dims");
    constexpr int K = pow3(D);
    T an[D], bn[D], un[D], pn[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        an[d] = (rlo[d] >= rhi[d]) ? rlo[d] : (slo[d] - rlo[d]) / (rhi[d] - rlo[d]);
        bn[d] = (rlo[d] >= rhi[d]) ? rlo[d] : (shi[d] - rlo[d]) / (rhi[d] - rlo[d]);
    }
    // normalisation first: the pdf's integral over the box (folds last dim first)
    const T den = rvol * pdf_fold<T, D, 0>(v, 0, an, bn);
    if (strategy_rr && !(rr < den / (den + svol))) {  // the roulette chose uniform sampling
#pragma unroll
        for (int d = 0; d < D; ++d) x[d] = slo[d] + s[d] * (shi[d] - slo[d]);
        const T pdf = T(1) / svol;
        weight = T((pdf < 1.e-10) ? 0.0 : (1.0 / pdf));
        return true;
    }
    if (!(den > T(0)) || !(den < T(INFINITY))) return false;
    // draw: dim k from the marginal, then condition on it
    T t[K];
#pragma unroll
    for (int i = 0; i < K; ++i) t[i] = v[i];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const int tl = pow3(D - k);  // entries of t still live
        T m[K];
#pragma unroll
        for (int i = 0; i < K; ++i)
            if (i < tl) m[i] = t[i];
#pragma unroll
        for (int d = D - 1; d > k; --d) {
            const int ml = pow3(d - k);  // t spans dims k..D-1: folding dim d leaves 3^(d-k) lines
#pragma unroll
            for (int i = 0; i < K / 3; ++i)
                if (i < ml) {
                    const T l[3] = {m[i], m[i + ml], m[i + 2 * ml]};
                    m[i] = pdf_integral(an[d], bn[d], l);
                }
        }
        const T sol = sample_line(s[k], an[k], bn[k], m);
        un[k] = sol;
        if (k + 1 < D) {
#pragma unroll
            for (int j = 0; j < K / 3; ++j)
                if (j < tl / 3) {
                    const T l[3] = {t[3 * j], t[3 * j + 1], t[3 * j + 2]};
                    t[j] = pdf_unnormalized(sol, l);
                }
        }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] = un[d] * (rhi[d] - rlo[d]) + rlo[d];
#pragma unroll
    for (int d = 0; d < D; ++d) pn[d] = (rlo[d] >= rhi[d]) ? rlo[d] : (x[d] - rlo[d]) / (rhi[d] - rlo[d]);
    // pdf at the point: unnormalised folds at its coordinates, last dim first
#pragma unroll
    for (int i = 0; i < K; ++i) t[i] = v[i];
#pragma unroll
    for (int d = D - 1; d > 0; --d) {
        const int len = pow3(d);
#pragma unroll
        for (int i = 0; i < K / 3; ++i)
            if (i < len) {
                const T l[3] = {t[i], t[i + len], t[i + 2 * len]};
                t[i] = pdf_unnormalized(pn[d], l);
            }
    }
    const T pdf = pdf_unnormalized(pn[0], t) / den;
    weight = T((pdf < (strategy_rr ? 1.e-10 : 1.e-5)) ? 0.0 : (1.0 / pdf));
    return true;
}

/* ---- dimensions 4 and 5: the same draw, depth first --------------------------------
   The register-tensor form above keeps 3^D values live; from D = 4 on, the marginal of
   dim k and the pdf at the point are evaluated as nested folds straight from the node
   values (3 live values per fold level), in the reference's nesting: the drawn dims
   0..k-1 conditioned first (dim 0 innermost, pdf_unnormalized at the drawn coordinate),
   then the later dims folded by pdf_integral over the box, the last dim innermost. */

/* conditioned chain over dims L..0 of the block at `base`: pdf_unnormalized(sol[L], .) */
template <typename T, int L>
VMX_HD inline T cond_fold(const T* v, int base, const T* sol) {
    if constexpr (L < 0) {
        return v[base];
    } else {
        T w[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) w[j] = cond_fold<T, L - 1>(v, base + j * pow3(L), sol);
        return pdf_unnormalized(sol[L], w);
    }
}
/* marginal folds over dims L..D-1 (dim K's index fixed in base), leaves the chain over
   dims K-1..0 */
template <typename T, int D, int K, int L>
VMX_HD inline T marg_fold(const T* v, int base, const T* sol, const T* an, const T* bn) {
    if constexpr (L == D) {
        return cond_fold<T, K - 1>(v, base, sol);
    } else {
        T w[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) w[j] = marg_fold<T, D, K, L + 1>(v, base + j * pow3(L), sol, an, bn);
        return pdf_integral(an[L], bn[L], w);
    }
}
/* the unnormalised pdf at normalised coordinates pn: folds last dim first */
template <typename T, int D, int L>
VMX_HD inline T point_fold(const T* v, int base, const T* pn) {
    T w[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        if constexpr (L == D - 1)
            w[j] = v[base + j * pow3(L)];
        else
            w[j] = point_fold<T, D, L + 1>(v, base + j * pow3(L), pn);
    }
    return pdf_unnormalized(pn[L], w);
}
/* draw dims K..D-1 in turn from their marginals given the ones drawn before */
template <typename T, int D, int K>
VMX_HD inline void draw_dims(const T* v, const T* an, const T* bn, const T* s, T* sol) {
    if constexpr (K < D) {
        T m[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) m[i] = marg_fold<T, D, K, K + 1>(v, i * pow3(K), sol, an, bn);
        sol[K] = sample_line(s[K], an[K], bn[K], m);
        draw_dims<T, D, K + 1>(v, an, bn, s, sol);
    }
}

/* sample() for dimensions 4..6: the same arguments and results */
template <typename T, int D>
__attribute__((noinline)) VMX_HD bool sample_deep(const T* v, const T* rlo, const T* rhi, T rvol, const T* slo,
                                                  const T* shi, const T* s, T* x, T& weight, bool strategy_rr, T rr,
                                                  T svol) {
    T an[D], bn[D], un[D], pn[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        an[d] = (rlo[d] >= rhi[d]) ? rlo[d] : (slo[d] - rlo[d]) / (rhi[d] - rlo[d]);
        bn[d] = (rlo[d] >= rhi[d]) ? rlo[d] : (shi[d] - rlo[d]) / (rhi[d] - rlo[d]);
    }
    const T den = rvol * pdf_fold<T, D, 0>(v, 0, an, bn);
    if (strategy_rr && !(rr < den / (den + svol))) {
#pragma unroll
        for (int d = 0; d < D; ++d) x[d] = slo[d] + s[d] * (shi[d] - slo[d]);
        const T pdf = T(1) / svol;
        weight = T((pdf < 1.e-10) ? 0.0 : (1.0 / pdf));
        return true;
    }
    if (!(den > T(0)) || !(den < T(INFINITY))) return false;
    draw_dims<T, D, 0>(v, an, bn, s, un);
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] = un[d] * (rhi[d] - rlo[d]) + rlo[d];
#pragma unroll
    for (int d = 0; d < D; ++d) pn[d] = (rlo[d] >= rhi[d]) ? rlo[d] : (x[d] - rlo[d]) / (rhi[d] - rlo[d]);
    const T pdf = point_fold<T, D, 0>(v, 0, pn) / den;
    weight = T((pdf < (strategy_rr ? 1.e-10 : 1.e-5)) ? 0.0 : (1.0 / pdf));
    return true;
}

/* A point of the box and its weight (sample_regs above for dims 1..3, sample_deep for
   4..6); arguments and results as described at sample_regs. */
template <typename T, int D>
VMX_HD inline bool sample(const T* v, const T* rlo, const T* rhi, T rvol, const T* slo, const T* shi, const T* s, T* x,
                          T& weight, bool strategy_rr = false, T rr = T(0), T svol = T(0)) {
    static_assert(D >= 1 && D <= MAX_DIM, "importance sampling: 1..6 dims");
    if constexpr (D <= 3)
        return sample_regs<T, D>(v, rlo, rhi, rvol, slo, shi, s, x, weight, strategy_rr, rr, svol);
    else
        return sample_deep<T, D>(v, rlo, rhi, rvol, slo, shi, s, x, weight, strategy_rr, rr, svol);
}

}  // namespace imp
}  // namespace vmx

#endif

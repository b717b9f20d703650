/*
 * Integrand families of the MI355X control-variate integrator.
 *
 * The reference integrates any host callable `f(x0, x1, ...)` (src/integrate.h:11-70).
 * A host lambda cannot run on the GPU, so the C-ABI takes an integrand *descriptor*:
 * a family id plus a plain parameter block.  Each family is a small functor usable
 * from host code (g++, oracle and reference driver) and device code (hipcc), so the
 * three sides integrate exactly the same function.
 *
 * Families
 *   VMX_FAMILY_POLY   sum of monomials  sum_m c_m prod_d x_d^e_md  plus an optional
 *                     half-space step  h*[n.x > d].  Only +, * and compares: results
 *                     are bit-identical on CPU and GPU (no FMA contraction).
 *   VMX_FAMILY_BLOBS  2-D or 3-D (x, y[, t]) anisotropic Gaussian blobs that may move
 *                     linearly in t, an optional additive step or multiplicative
 *                     occluder [a x + b y + c t > d], and a separable wave
 *                     A sin(fx x) cos(fy y).  Configs 1-3 are instances.
 *   VMX_FAMILY_LOBES5 5-D: three Gaussian lobes over the integrated dims 2..4, two
 *                     binary half-space steps over all 5 dims and 1 + a cos(2 pi k.x).
 *
 * Parameters in the C structs are double precision; `to_T` converts them once to the
 * evaluation type (float for production, double for the FP64 parity instantiation).
 */
#ifndef VILTRUM_AMD_INTEGRANDS_H
#define VILTRUM_AMD_INTEGRANDS_H

#include <stdint.h>

#if defined(__HIPCC__)
#define VMX_HD __host__ __device__
#else
#define VMX_HD
#endif

#define VMX_MAX_DIM 6
#define VMX_POLY_MAX_TERMS 16
#define VMX_MAX_BLOBS 64

#ifdef __cplusplus
extern "C" {
#endif

enum vmx_family {
    VMX_FAMILY_POLY = 1,
    VMX_FAMILY_BLOBS = 2,
    VMX_FAMILY_LOBES5 = 3,
    VMX_FAMILY_FUBINI = 4,  /* an inner integrand, its trailing dims integrated by Monte Carlo */
    VMX_FAMILY_DYSON = 5,   /* a function of an unbounded sequence: only as the inner integrand of
                               the Fubini form over an infinite range (inner_dim VMX_INFINITE_DIM) */
    VMX_FAMILY_CUSTOM = 6   /* a user-defined device functor (params: vmx_custom_integrand) */
};

/* inner_dim of a Fubini descriptor whose rest is an infinite range (RangeInfinite,
   src/range-infinite.h:16-87): the inner integrand reads an unbounded sequence */
#define VMX_INFINITE_DIM (-1)

typedef struct vmx_poly_params {
    int32_t nterms;                         /* 0..VMX_POLY_MAX_TERMS */
    int32_t has_step;                       /* 1: add step_h*[step_n.x > step_d] */
    double coef[VMX_POLY_MAX_TERMS];
    uint8_t expo[VMX_POLY_MAX_TERMS][8];    /* exponent of x_d (d < VMX_MAX_DIM) */
    double step_n[VMX_MAX_DIM];
    double step_d;
    double step_h;
} vmx_poly_params;

typedef struct vmx_blobs_params {
    int32_t nblobs;                         /* 0..VMX_MAX_BLOBS */
    int32_t step_mode;                      /* 0 none, 1 additive h*[n.x>d], 2 occluder f*[n.x>d] */
    /* per blob: cx, cy, vx, vy, ax, ay, w ;  w*exp(-(ax*dx*dx + ay*dy*dy)),
       dx = x - (cx + vx*t), dy = y - (cy + vy*t), t = x2 (0 in 2-D) */
    double blob[VMX_MAX_BLOBS][7];
    double step[5];                         /* nx, ny, nt, d, h */
    double wave[3];                         /* A, fx, fy */
} vmx_blobs_params;

typedef struct vmx_lobes5_params {
    double centre[3];                       /* lobe centres along dims 2,3,4 */
    double inv2s2[3];                       /* 1/(2 sigma^2) */
    double step_n[2][5];                    /* two visibility half-spaces over all dims */
    double step_d[2];
    double k[5];                            /* integer frequencies of the cosine */
    double amp;                             /* cosine amplitude */
} vmx_lobes5_params;

/* The Fubini form of integrator_crespo2021_infinite<DIM> (src/control-variates/
   integrator-crespo2021.h:24-44): regions and bins over the descriptor's `dim` leading
   dims of an inner integrand of inner_dim dims; its remaining dims are integrated by
   Monte Carlo over [rest_lo, rest_hi] (entries dim .. inner_dim-1), with mc_samples
   uniform samples per region node value and one per residual sample.  With inner_dim
   VMX_INFINITE_DIM the rest is an infinite range (RangeInfinite): the inner integrand
   (VMX_FAMILY_DYSON, or VMX_FAMILY_CUSTOM with a functor of a sequence) reads the region
   point followed by an unbounded sequence of uniform draws, entry k of the sequence in
   [rest_lo[dim + k], rest_hi[dim + k]] for the first rest_count entries and in [0, 1]
   after them (RangeInfinite::min(i) / max(i)); the rest's volume is the product of the
   first rest_count extents.  FP32 only, as the reference. */
typedef struct vmx_fubini_params {
    int32_t inner_family;                   /* VMX_FAMILY_POLY, _BLOBS (3-D), _LOBES5, _CUSTOM or _DYSON */
    int32_t inner_dim;                      /* > the descriptor's dim, <= 5; or VMX_INFINITE_DIM */
    uint32_t mc_samples;                    /* >= 1 */
    uint32_t seed;                          /* stream of the Monte Carlo samples */
    double rest_lo[VMX_MAX_DIM];            /* indexed by inner dim */
    double rest_hi[VMX_MAX_DIM];
    const void* inner;                      /* host pointer to the inner family's parameter block */
    int32_t rest_count;                     /* infinite rest: explicit entries (0 .. VMX_MAX_DIM - dim) */
} vmx_fubini_params;

/* Dyson-series integrand of a sequence s (the reference's infinite-range example,
   main/compilation-tests/crespo21.cc:6-25): sum = 1, range = 1; while the current entry
   is below `decay`, step to the next entry u, x = range * u, range = x,
   sum += x * range_before (x / (1 / range_before)); step again. */
typedef struct vmx_dyson_params {
    double decay;                           /* in (0, 1) */
} vmx_dyson_params;

/* A user-defined integrand (VMX_FAMILY_CUSTOM): a device-evaluable functor for which
   the caller's own HIP translation unit instantiates the pipeline
   (viltrum_amd/device.hpp fills this in).  The library hands each group of integrals
   that share `run` (one functor type and dimension) to run(), which receives the
   library's internal job list (group) and training-pair request (pairs, or null) and
   returns a vmx_status.  `functor` points at the functor's bytes in host memory
   (trivially copyable, at most VMX_CUSTOM_MAX_BYTES: it is staged in LDS).  As the inner
   integrand of the Fubini form (VMX_FAMILY_FUBINI with inner_family VMX_FAMILY_CUSTOM),
   the library calls run_fubini with the form's region dims instead; it is NULL unless
   the functor was instantiated for that form (viltrum_amd::device_integrand_fubini).  A
   functor of a sequence (the form over an infinite range, inner_dim VMX_INFINITE_DIM)
   has only run_fubini (viltrum_amd::device_integrand_infinite); each hook refuses a
   form whose inner_dim is not the one it was built for.
   Not for vmx_eval or vmx_exact_bins. */
#define VMX_CUSTOM_MAX_BYTES 2048
struct vmx_ctx;
typedef struct vmx_custom_integrand {
    int32_t (*run)(struct vmx_ctx* ctx, int32_t scalar, const void* group, void* stream, void* pairs);
    const void* functor;
    int32_t (*run_fubini)(struct vmx_ctx* ctx, int32_t region_dims, const void* group, void* stream);
} vmx_custom_integrand;

/* Integrand descriptor handed to every entry point. */
typedef struct vmx_integrand {
    int32_t family;                         /* enum vmx_family */
    int32_t dim;                            /* D, number of integration dimensions */
    const void* params;                     /* host pointer to the family's parameter block */
} vmx_integrand;

#ifdef __cplusplus
}

namespace vmx {

/* `eval<M, D>` takes a math policy M with static exp, sin and cos overloads for float
   and double; host and device code each supply their own. */

template <typename T>
struct Poly {
    int nterms, has_step;
    T coef[VMX_POLY_MAX_TERMS];
    uint8_t expo[VMX_POLY_MAX_TERMS][8];
    T step_n[VMX_MAX_DIM];
    T step_d, step_h;

    static Poly from(const vmx_poly_params& p) {
        Poly q{};
        q.nterms = p.nterms;
        q.has_step = p.has_step;
        for (int m = 0; m < VMX_POLY_MAX_TERMS; ++m) {
            q.coef[m] = T(p.coef[m]);
            for (int d = 0; d < 8; ++d) q.expo[m][d] = p.expo[m][d];
        }
        for (int d = 0; d < VMX_MAX_DIM; ++d) q.step_n[d] = T(p.step_n[d]);
        q.step_d = T(p.step_d);
        q.step_h = T(p.step_h);
        return q;
    }

    /* Fixed evaluation order: terms in index order, each term the coefficient times
       x_0^e0 * x_1^e1 * ... by repeated multiplication. */
    template <class M, int D>
    VMX_HD T eval(const T* x) const {
        T s = T(0);
        for (int m = 0; m < nterms; ++m) {
            T t = coef[m];
            for (int d = 0; d < D; ++d)
                for (int e = 0; e < expo[m][d]; ++e) t = t * x[d];
            s = s + t;
        }
        if (has_step) {
            T lin = T(0);
            for (int d = 0; d < D; ++d) lin = lin + step_n[d] * x[d];
            if (lin > step_d) s = s + step_h;
        }
        return s;
    }
};

template <typename T>
struct Blobs {
    /* one blob per row, (cx, cy, vx, vy, ax, ay, w, 0): a blob is two (FP32) or four
       (FP64) 16-byte loads */
    alignas(16) T blob[VMX_MAX_BLOBS][8];
    int nblobs, step_mode;
    T snx, sny, snt, sd, sh;
    T wa, wfx, wfy;

    static Blobs from(const vmx_blobs_params& p) {
        Blobs q{};
        q.nblobs = p.nblobs;
        q.step_mode = p.step_mode;
        for (int k = 0; k < VMX_MAX_BLOBS; ++k) {
            for (int i = 0; i < 7; ++i) q.blob[k][i] = T(p.blob[k][i]);
            q.blob[k][7] = T(0);
        }
        q.snx = T(p.step[0]);
        q.sny = T(p.step[1]);
        q.snt = T(p.step[2]);
        q.sd = T(p.step[3]);
        q.sh = T(p.step[4]);
        q.wa = T(p.wave[0]);
        q.wfx = T(p.wave[1]);
        q.wfy = T(p.wave[2]);
        return q;
    }

    /* Sums in blob order.  M::fma(a, b, c) is a*b + c under the policy's rounding; with
       the unfused policies every value is the plain expression
       s + w*exp(-(ax*(dx*dx) + ay*(dy*dy))), dx = x - (cx + vx*t). */
    template <class M, int D>
    VMX_HD T eval(const T* x) const {
        static_assert(D == 2 || D == 3, "blob family is 2-D or 3-D");
        T s = T(0);
        for (int k = 0; k < nblobs; ++k) {
            const T* b = blob[k];
            T dx, dy;
            if constexpr (D == 3) {
                dx = x[0] - M::fma(b[2], x[2], b[0]);
                dy = x[1] - M::fma(b[3], x[2], b[1]);
            } else {  // t = 0: the centre is (cx, cy)
                dx = x[0] - b[0];
                dy = x[1] - b[1];
            }
            s = M::fma(b[6], M::exp(-M::fma(b[5], dy * dy, b[4] * (dx * dx))), s);
        }
        if (wa != T(0)) s = s + wa * (M::sin(wfx * x[0]) * M::cos(wfy * x[1]));
        if (step_mode != 0) {
            T lin = snx * x[0] + sny * x[1];
            if constexpr (D == 3) lin = M::fma(snt, x[2], lin);
            const bool in = lin > sd;
            if (step_mode == 1) {
                if (in) s = s + sh;
            } else {
                if (!in) s = T(0);
            }
        }
        return s;
    }
};

template <typename T>
struct Lobes5 {
    T c[3], a[3];
    T n[2][5];
    T d[2];
    T k[5];
    T amp;

    static Lobes5 from(const vmx_lobes5_params& p) {
        Lobes5 q{};
        for (int j = 0; j < 3; ++j) {
            q.c[j] = T(p.centre[j]);
            q.a[j] = T(p.inv2s2[j]);
        }
        for (int s = 0; s < 2; ++s) {
            for (int j = 0; j < 5; ++j) q.n[s][j] = T(p.step_n[s][j]);
            q.d[s] = T(p.step_d[s]);
        }
        for (int j = 0; j < 5; ++j) q.k[j] = T(p.k[j]);
        q.amp = T(p.amp);
        return q;
    }

    template <class M, int D>
    VMX_HD T eval(const T* x) const {
        static_assert(D == 5, "lobes5 family is 5-D");
        for (int s = 0; s < 2; ++s) {
            T lin = T(0);
            for (int j = 0; j < 5; ++j) lin = M::fma(n[s][j], x[j], lin);
            if (!(lin > d[s])) return T(0);
        }
        T q = T(0);
        for (int j = 0; j < 3; ++j) {
            const T u = x[2 + j] - c[j];
            q = M::fma(a[j], u * u, q);
        }
        T ph = T(0);
        for (int j = 0; j < 5; ++j) ph = M::fma(k[j], x[j], ph);
        return M::exp(-q) * (T(1) + amp * M::cos2pi(ph));
    }
};

/* A function of an unbounded sequence (src/range-infinite.h, the integrand of
   crespo21.cc): `eval_seq(s)` reads entries through a cursor with `cur()` (the current
   entry) and `next()` (step to the following one), as the reference's integrand walks
   a sequence iterator.  The walk stops after 4096 steps, which a decay below 1 reaches
   with probability decay^4096 (never, in practice). */
template <typename T>
struct Dyson {
    T decay;

    static Dyson from(const vmx_dyson_params& p) {
        Dyson q{};
        q.decay = T(p.decay);
        return q;
    }

    /* s: cur() reads the current entry, next() steps to the following one.  At most 4096
       terms (the reference's loop has no bound; reaching it has probability below
       decay^4096). */
    template <class Seq>
    VMX_HD T eval_seq(Seq& s) const {
        T sum = T(1), range = T(1);
        for (int step = 0; step < 4096 && s.cur() < decay; ++step) {
            s.next();
            const T x = range * s.cur();
            const T prob = T(1.0 / double(range));
            range = x;
            sum += x / prob;
            s.next();
        }
        return sum;
    }
};

}  // namespace vmx
#endif /* __cplusplus */

#endif /* VILTRUM_AMD_INTEGRANDS_H */

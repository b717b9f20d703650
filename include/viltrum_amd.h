/*
 * viltrum_amd -- C ABI of the MI355X per-bin control-variate integrator.
 *
 * Plain pointers and sizes only.  One call integrates one problem: adaptive nested
 * Simpson regions, the exact per-bin integral of the piecewise-polynomial control
 * variate, and the residual Monte Carlo accumulated per bin; or a batch of such
 * problems in one launch sequence.
 *
 * Reference interfaces each entry point replaces (adolfomunoz/viltrum):
 *   vmx_integrate / vmx_integrate_device
 *       viltrum::integrate(integrator, bins, resolution, f, range)  src/integrate.h:72-79,
 *       which calls IntegratorRegionBased::integrate
 *       src/newton-cotes/integrator-region-based.h:13-22 with the integrator built by
 *       integrator_crespo2021(iterations, spp, seed)
 *       src/control-variates/integrator-crespo2021.h:7-22  (VMX_FORM_VR_OPTIMIZE), by
 *       regions_integrator_parallel_variance_reduction(rr, alpha, spp, seed)
 *       src/control-variates/regions-integrator-parallel-variance-reduction.h:127-130
 *       (VMX_FORM_VR_FIXED) or by integrator_adaptive_control_variates_parallel(rule,
 *       eh, iterations, spp, seed) src/control-variates/integrator-adaptive-control-variates.h:17-20
 *       (VMX_FORM_CV_ALPHA1).
 *   vmx_integrate_batch
 *       many independent viltrum::integrate calls (no single reference counterpart; the
 *       reference would loop over src/integrate.h:72-79).
 *   vmx_export_regions
 *       the region list the reference hands to Logger::log(seq_regions)
 *       src/newton-cotes/integrator-region-based.h:19, i.e. the output of
 *       RegionsGeneratorAdaptiveHeap::generate src/nested/regions-generator-adaptive-heap.h:18-45.
 *   vmx_integrate_mc
 *       viltrum::integrate(monte_carlo_per_bin(rng, spp), bins, resolution, f, range):
 *       MonteCarloPerBin::integrate src/monte-carlo/monte-carlo-per-bin.h:40-65 (plain
 *       per-bin Monte Carlo; here it also provides the FP64 truth for RMSE).
 *   vmx_eval
 *       the integrand call f(x) of src/integrate.h:11-70 (device evaluation, for tests).
 *
 * Errors: every function returns VMX_OK (0) or a negative vmx_status; the context's
 * message is in vmx_last_error().  The reference has no error reporting (asserts and
 * undefined behaviour, e.g. an empty region list in src/control-variates/region-russian-roulette.h:16).
 */
#ifndef VILTRUM_AMD_H
#define VILTRUM_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "viltrum_amd/integrands.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vmx_ctx vmx_ctx;

enum vmx_status {
    VMX_OK = 0,
    VMX_ERR_INVALID = -1,      /* bad argument */
    VMX_ERR_UNSUPPORTED = -2,  /* family / dimension / form not built */
    VMX_ERR_DEVICE = -3,       /* HIP runtime error */
    VMX_ERR_CAPACITY = -4,     /* internal capacity exceeded (message says which) */
    VMX_ERR_NUMERIC = -5       /* e.g. a bin with no region while spp > 0 */
};

enum vmx_scalar { VMX_F32 = 4, VMX_F64 = 8 };

/* Error heuristic of the region generator. */
enum vmx_heuristic {
    VMX_HEUR_SIZE_RELATIVE = 0,   /* error_heuristic_size(error_metric_relative(), 1e-5) */
    VMX_HEUR_DEFAULT_ABSOLUTE = 1 /* error_heuristic_default(error_metric_absolute()) */
};

/* Per-bin estimator. */
enum vmx_form {
    VMX_FORM_VR_OPTIMIZE = 0, /* bins = residual with cv_optimize_weight; integrator_crespo2021 */
    VMX_FORM_VR_FIXED = 1,    /* bins = residual with cv_fixed_weight(alpha) */
    VMX_FORM_CV_ALPHA1 = 2    /* bins += CV + residual, alpha = 1 */
};

/* Order of the final region list (and so of every per-bin region list). */
enum vmx_order {
    VMX_ORDER_CANONICAL = 0, /* deterministic geometric order, fastest */
    VMX_ORDER_REFERENCE = 1  /* the reference's heap-array order (host replay) */
};

/* Region refinement priority. */
enum vmx_refiner {
    VMX_REFINER_HEURISTIC = 0, /* the error heuristic above */
    VMX_REFINER_LEARNED = 1    /* the trained MLP of the context (vmx_refiner_load) */
};

typedef struct vmx_problem {
    int32_t scalar;       /* VMX_F32 or VMX_F64: evaluation type */
    int32_t bin_dims;     /* DB, 1..3, <= integrand dim */
    int32_t heuristic;    /* enum vmx_heuristic */
    int32_t form;         /* enum vmx_form */
    int32_t order;        /* enum vmx_order */
    int32_t refiner;      /* enum vmx_refiner */
    double alpha;         /* VMX_FORM_VR_FIXED weight */
    double lo[VMX_MAX_DIM], hi[VMX_MAX_DIM]; /* integration range (first integrand.dim used) */
    uint64_t res[3];      /* bin resolution (first bin_dims used) */
    uint64_t regions;     /* region budget N = iterations + 1 (>= 1) */
    uint64_t spp;         /* residual samples per bin */
    uint64_t seed;        /* Philox key */
} vmx_problem;

typedef struct vmx_stats {
    uint64_t regions;          /* final regions */
    uint64_t nodes;            /* tree nodes evaluated (>= 2*regions - 1, speculative rounds) */
    uint64_t rounds;           /* parallel refinement rounds */
    uint64_t pairs;            /* (bin, footprint) pairs of the per-bin lists */
    uint64_t footprints;       /* distinct bin-plane footprints */
    uint64_t evals_regions;    /* integrand evaluations for region nodes */
    uint64_t evals_residual;   /* integrand evaluations for the residual */
    uint64_t ties_unresolved;  /* boundary ties broken by geometric order (canonical only) */
    double ms_generate, ms_select, ms_stratify, ms_bins, ms_total;
} vmx_stats;

int vmx_ctx_create(int device, vmx_ctx** out);
void vmx_ctx_destroy(vmx_ctx* ctx);
const char* vmx_last_error(const vmx_ctx* ctx);

/* Integrate into caller-owned host bins of the evaluation type, C order
   bins[i0][i1][i2] (the vector<vector<T>> accessor of src/integrate.h:139-148).
   VMX_FORM_CV_ALPHA1 reads and accumulates; the other forms assign. */
int vmx_integrate(vmx_ctx* ctx, const vmx_problem* p, const vmx_integrand* f, void* bins, vmx_stats* stats);

/* Same, with bins already in device memory; work is ordered on `stream` (a hipStream_t,
   NULL = the context's stream) and the call returns when it is complete. */
int vmx_integrate_device(vmx_ctx* ctx, const vmx_problem* p, const vmx_integrand* f, void* bins_device,
                         void* stream, vmx_stats* stats);

/* n independent problems in one launch sequence (one cooperative generation kernel per
   (family, dim) group; every problem keeps its own range, resolution, budget, spp,
   seed and form).  All problems share the scalar type, region order and bin_dims.
   bins_device[i]: device bins of problem i; stats: n entries or NULL. */
int vmx_integrate_batch(vmx_ctx* ctx, uint32_t n, const vmx_problem* problems, const vmx_integrand* integrands,
                        void* const* bins_device, void* stream, vmx_stats* stats);

/* Region list of the last vmx_integrate* call on this context, in its final order.
   Pass NULL arrays to query *n.  lo/hi: n*D, nodes: n*3^D, err: n, dim: n. */
int vmx_export_regions(vmx_ctx* ctx, uint64_t* n, double* lo, double* hi, double* nodes, double* err,
                       int32_t* dim);

/* Plain per-bin Monte Carlo with p->spp samples per bin (p->scalar, p->seed, range,
   bin_dims, res; region fields ignored): bins = volume * mean f over the bin's box.
   stratified != 0 jitters samples over a g x g grid of the first two bin dims.
   stderr_out (optional, FP64) receives each bin's standard error.  Host arrays. */
int vmx_integrate_mc(vmx_ctx* ctx, const vmx_problem* p, const vmx_integrand* f, int32_t stratified, void* bins,
                     double* stderr_out);

/* Evaluate the integrand on the device at n points x (n*D, row-major). */
int vmx_eval(vmx_ctx* ctx, const vmx_integrand* f, int32_t scalar, uint64_t n, const double* x, double* out);

#ifdef __cplusplus
}
#endif

#endif /* VILTRUM_AMD_H */

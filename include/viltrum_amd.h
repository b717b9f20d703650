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
    VMX_FORM_CV_ALPHA1 = 2,   /* bins += CV + residual, alpha = 1 */
    /* the two variance-reduction forms with region_sampling_importance in place of
       region_sampling_uniform (src/control-variates/region-sampling.h:22-44): residual
       samples drawn from each region polynomial's |P| pdf over the bin's box, weighted
       by 1/pdf.  FP32, dims 1..6 (the reference compiles it for float regions only). */
    VMX_FORM_VR_OPTIMIZE_IMPORTANCE = 3,
    VMX_FORM_VR_FIXED_IMPORTANCE = 4,
    /* the two variance-reduction forms with rr_integral_region in place of
       rr_uniform_region (src/control-variates/region-russian-roulette.h:30-67): a
       residual sample picks region r of its bin with probability proportional to
       |integral of P_r over the bin's box| (floored at 1% of the mean), weight 1/p.
       A batch holding these forms keeps one footprint per region (no merging). */
    VMX_FORM_VR_OPTIMIZE_RR_INTEGRAL = 5,
    VMX_FORM_VR_FIXED_RR_INTEGRAL = 6,
    /* ... with rr_error_region (region-russian-roulette.h:68-104): probability
       proportional to |Region::error()| times the share of the region inside the bin */
    VMX_FORM_VR_OPTIMIZE_RR_ERROR = 7,
    VMX_FORM_VR_FIXED_RR_ERROR = 8,
    /* ... with rr_pdf_region (region-russian-roulette.h:106-149): probability
       proportional to the integral of the importance pdf over the bin's box
       (Region::pdf_integral_subrange), floored at 1% of the mean */
    VMX_FORM_VR_OPTIMIZE_RR_PDF = 9,
    VMX_FORM_VR_FIXED_RR_PDF = 10,
    /* A roulette and region_sampling_importance together
       (regions_integrator_parallel_variance_reduction(rr, cv, region_sampling_importance(),
       ...), regions-integrator-parallel-variance-reduction.h:94-100, as in
       main/compilation-tests/multiple-parameters-2d.cc:101-103): the region is drawn by
       the roulette, the point in the bin's box from the region's pdf; weight 1/p times
       1/pdf.  FP32 integrands of dimension 1..6, as for importance sampling alone. */
    VMX_FORM_VR_OPTIMIZE_RR_INTEGRAL_IMPORTANCE = 11,
    VMX_FORM_VR_FIXED_RR_INTEGRAL_IMPORTANCE = 12,
    VMX_FORM_VR_OPTIMIZE_RR_ERROR_IMPORTANCE = 13,
    VMX_FORM_VR_FIXED_RR_ERROR_IMPORTANCE = 14,
    VMX_FORM_VR_OPTIMIZE_RR_PDF_IMPORTANCE = 15,
    VMX_FORM_VR_FIXED_RR_PDF_IMPORTANCE = 16,
    /* region_sampling_russian_roulette in place of uniform sampling
       (src/control-variates/region-sampling.h:47-84; uniform region roulette): each
       residual sample chooses importance sampling with probability pdf_integral /
       (pdf_integral + volume) of the bin's box, else uniform sampling, and is weighted
       by 1/pdf of the strategy it took (0 where pdf < 1e-10), as the reference does.
       FP32 integrands of dimension 1..6, as for importance sampling. */
    VMX_FORM_VR_OPTIMIZE_SAMPLING_RR = 17,
    VMX_FORM_VR_FIXED_SAMPLING_RR = 18,
    /* regions_integrator_parallel_variance_reduction_optimized with
       region_stratification_uniform and region_sampling_uniform
       (src/control-variates/regions-integrator-parallel-variance-reduction-optimized.h:54-142,
       region-stratification.h:9-27), cv_optimize_weight / cv_fixed_weight(alpha): a bin's
       spp residual samples are split over its n regions, spp / n each and one more for
       spp % n of them taken cyclically from a random start, each weighted by
       spp / (its region's count).  As
       integrator_adaptive_fubini_variance_reduction_parallel_optimized
       (integrator-adaptive-fubini-variance-reduction-optimized.h:17-23), which the
       reference builds over infinite ranges only: Fubini integrands with inner_dim
       VMX_INFINITE_DIM; the residual sample is f itself, without the rest's volume. */
    VMX_FORM_VR_OPTIMIZE_STRATIFIED = 19,
    VMX_FORM_VR_FIXED_STRATIFIED = 20
};

/* Order of the final region list (and so of every per-bin region list). */
enum vmx_order {
    VMX_ORDER_CANONICAL = 0, /* deterministic geometric order, fastest */
    VMX_ORDER_REFERENCE = 1  /* the reference's heap-array order (host replay) */
};

/* Region refinement priority. */
enum vmx_refiner {
    VMX_REFINER_HEURISTIC = 0, /* the error heuristic above */
    VMX_REFINER_LEARNED = 1    /* the context's priority network (vmx_refiner_train / _set);
                                  FP32, dims 1..5, canonical order */
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

/* vmx_integrate_device that allocates nothing: no hipMalloc, hipFree, hipHostMalloc or
   graph instantiation.  No counterpart in the reference; it is for callers that cannot
   allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
   it).  Every work array of the call is carved from one fixed device array of
   VMX_NOALLOC_SCRATCH_BYTES per GPU (resolved at vmx_ctx_create and held until the
   process ends); a problem needing more returns VMX_ERR_CAPACITY.  Results are those of
   vmx_integrate_device bit for bit.  All calls on one GPU share that array and it keeps
   the last call's contents: calls that overlap on one GPU (other threads, other
   contexts, other streams) race and must be serialised by the caller.  Afterwards
   vmx_export_regions reports no regions. */
#define VMX_NOALLOC_SCRATCH_BYTES (64u << 20)
int vmx_integrate_device_noalloc(vmx_ctx* ctx, const vmx_problem* p, const vmx_integrand* f, void* bins_device,
                                 void* stream, vmx_stats* stats);

/* n independent problems in one launch sequence per (family, dim) group (every problem
   keeps its own range, resolution, budget, spp, seed, form and refiner).  All problems
   share the scalar type, region order and bin_dims.
   bins_device[i]: device bins of problem i; stats: n entries or NULL. */
int vmx_integrate_batch(vmx_ctx* ctx, uint32_t n, const vmx_problem* problems, const vmx_integrand* integrands,
                        void* const* bins_device, void* stream, vmx_stats* stats);

/* The per-bin phase alone over a caller-given region list: control-variate integral and
   residual Monte Carlo per bin, as vmx_integrate with generation skipped.  Replaces
   RegionsIntegratorParallelVarianceReduction::integrate_regions(bins, res, seq_regions,
   f, range, logger) (src/control-variates/regions-integrator-parallel-variance-reduction.h:32-109)
   and its alpha = 1 counterpart (…-parallel-control-variates.h:17-66), which take the
   region sequence a generator produced; here it is the list vmx_export_regions writes
   (lo/hi: n*D, nodes: n*3^D, dim 0 fastest), so a run can be replayed with more samples,
   another seed or another resolution.  The list order is kept (regions are not merged
   into shared footprints).  p->regions and p->order are ignored; p->form, spp, seed, res
   and bin_dims apply.  Regions must be non-empty boxes inside the range; with spp > 0
   every bin must meet one. */
int vmx_integrate_regions(vmx_ctx* ctx, const vmx_problem* p, const vmx_integrand* f, uint64_t n, const double* lo,
                          const double* hi, const double* nodes, void* bins, vmx_stats* stats);

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

/* ---- Learned refinement priority.  No counterpart in the reference: there the error
   heuristic (src/nested/error-heuristic.h:21-47) alone orders splits; it stays the split-
   dimension rule, the fallback for a non-finite network output and the baseline.
   A 4-layer MLP (F -> 64 -> 64 -> 64 -> 1, BF16 weights, FP32 accumulation on the matrix
   cores) scores each new region from its features; it is trained on the device on
   (region, FP64 true error) pairs the method itself generates.  Parameters are FP32,
   laid out W0[KP][64] b0[64] W1[64][64] b1 W2[64][64] b2 w3[64] b3 mu sigma. */

typedef struct vmx_train_opts {
    uint32_t epochs;  /* passes over the training pairs (0: 8) */
    uint32_t batch;   /* pairs per Adam step, a multiple of 64 (0: 1024) */
    float lr;         /* initial step size, cosine-decayed to lr/20 (0: 2e-3) */
    float holdout;    /* fraction of pairs held out for the validation loss (0: 0.05) */
    uint64_t seed;    /* initialisation, split and shuffling */
    int32_t target;   /* FP64 target E: 0 (default) the residual energy, the integral over
                         the region of (f - P)^2; 1 the within-bin residual variance summed
                         over the region's bins; 2 the reduction of 1 by one split along the
                         heuristic's dim */
    uint32_t stride;  /* keep every stride-th generated region as a pair (0: 1) */
} vmx_train_opts;

typedef struct vmx_train_report {
    uint64_t pairs;       /* (region, true error) pairs generated */
    uint64_t steps;       /* Adam steps */
    double loss_first;    /* mean loss 0.5 (y - (t - mu)/sigma)^2 over the first epoch */
    double loss_train;    /* ... over the last epoch */
    double loss_valid;    /* on the held-out pairs after training */
    double target_mean, target_std;  /* mu, sigma */
    double ms_pairs, ms_train;
} vmx_train_report;

/* Sizes for integrands of dimension dim (1..5): parameter count, feature count F and
   the padded feature row width KP of vmx_refiner_pairs / _forward / _grad rows. */
int vmx_refiner_size(int32_t dim, uint64_t* n_params, int32_t* n_features, int32_t* row_width);

/* Training pairs of the regions the method generates for n problems of one family and
   dim (every stride-th node of the refinement rounds), each problem with its own
   refiner: the heuristic, or the context's current network for VMX_REFINER_LEARNED
   (pairs from the trees the network itself builds): features as BF16 bit
   patterns (count x KP) and targets log2(E / (s^2 vol)), E the FP64 target of
   opts->target (bins of each problem's resolution), s = max |node value|.  Only
   opts->target and opts->stride are read (NULL: defaults).  *count is set to the number
   of pairs and min(count, cap) rows are written (NULL features/targets: none).  Rows
   come in the order the device claimed node slots, which varies between calls. */
int vmx_refiner_pairs(vmx_ctx* ctx, uint32_t n, const vmx_problem* problems, const vmx_integrand* integrands,
                      const vmx_train_opts* opts, uint64_t cap, uint16_t* features, float* targets,
                      uint64_t* count);

/* Generate pairs as above and train the network for their dim on the device (Adam,
   minibatches of 64-row tiles, deterministic); the context keeps the parameters. */
int vmx_refiner_train(vmx_ctx* ctx, uint32_t n, const vmx_problem* problems, const vmx_integrand* integrands,
                      const vmx_train_opts* opts, vmx_train_report* report);

/* Copy the context's network for dim out (vmx_refiner_size floats) or in. */
int vmx_refiner_get(vmx_ctx* ctx, int32_t dim, float* params);
int vmx_refiner_set(vmx_ctx* ctx, int32_t dim, const float* params);

/* Raw network outputs y (standardised log2 error) of n feature rows (BF16, n x KP), with
   host params (NULL: the context's network). */
int vmx_refiner_forward(vmx_ctx* ctx, int32_t dim, const float* params, uint64_t n, const uint16_t* features,
                        float* y);

/* Gradient of the training loss (1/n) sum 0.5 (y - (t - mu)/sigma)^2 with respect to the
   trained parameters (all but mu, sigma) at host params, and the loss. */
int vmx_refiner_grad(vmx_ctx* ctx, int32_t dim, const float* params, uint64_t n, const uint16_t* features,
                     const float* targets, float* grad, double* loss);

/* FP64 closed-form bins (bin count * integral over the bin, host array, C order) of a
   2-D static blob integrand with an additive step (step_mode 0 or 1): the truth for
   RMSE where one exists.  The reference has no counterpart. */
int vmx_exact_bins(vmx_ctx* ctx, const vmx_problem* p, const vmx_integrand* f, double* bins);

/* ---- Launch-shape autotuner.  No counterpart in the reference (its parallel loops have
   no launch shapes).  Runs the batch of vmx_integrate_batch once per candidate shape
   (expansion workgroups per CU 2/4/8, refinement rounds per graph launch 4/8, the
   per-bin kernel's workgroup-to-bin mapping contiguous per XCD or round-robin,
   deepened rounds off/on: qualifying children split in the round that creates them,
   which selects the same regions in fewer rounds), keeps
   the fastest for later calls of the same family, dimension, scalar type, refiner,
   batch size and bin count (each within a factor of 2), and leaves that batch's
   results in bins_device.  Launch shapes never change results: the random streams and
   every summation order are independent of them. */
typedef struct vmx_tune_report {
    int32_t gen_blocks_per_cu; /* chosen: expansion workgroups per compute unit */
    int32_t round_chunk;       /* chosen: refinement rounds per graph launch */
    int32_t bins_xcd;          /* chosen: 1 = each XCD takes a contiguous run of bin tiles */
    int32_t deepen;            /* chosen: 1 = children that qualify are split in the same round */
    int32_t trials;            /* shapes timed */
    double ms_best;            /* wall time of one call at the chosen shape */
    double ms_default;         /* ... at the built-in default shape */
} vmx_tune_report;

int vmx_autotune(vmx_ctx* ctx, uint32_t n, const vmx_problem* problems, const vmx_integrand* integrands,
                 void* const* bins_device, void* stream, vmx_tune_report* report);

/* Evaluate the integrand on the device at n points x (n*D, row-major). */
int vmx_eval(vmx_ctx* ctx, const vmx_integrand* f, int32_t scalar, uint64_t n, const double* x, double* out);

#ifdef __cplusplus
}
#endif

#endif /* VILTRUM_AMD_H */

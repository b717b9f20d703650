"""viltrum_amd -- MI355X per-bin control-variate integration.

Python mirror of the reference's call surface over the C ABI (include/viltrum_amd.h):

    integrate(integrator, bins, f, rng)        src/integrate.h:72-79, 139-148
    integrate(integrator, f, rng) -> value     src/integrate.h:105-130
    integrator_crespo2021(iterations, spp, seed)
                                               src/control-variates/integrator-crespo2021.h:7-22
    integrator_adaptive_control_variates_parallel(heuristic, iterations, spp, seed)
                                               src/control-variates/integrator-adaptive-control-variates.h:17-20
    integrator_variance_reduction_fixed(alpha, iterations, spp, seed, heuristic)
                                               regions_integrator_parallel_variance_reduction(rr, alpha, ...)
                                               src/control-variates/regions-integrator-parallel-variance-reduction.h:127-130

`bins` is a numpy array whose shape is the bin resolution (1 to 3 dims, C order like
bins[i0][i1]); float32 bins run the FP32 production path, float64 bins the FP64 path.
Integrands are descriptors from viltrum_amd.integrands (a GPU cannot call a Python
function).  Every call runs on the GPU through libviltrum_amd.so; there is no CPU path.
"""
import ctypes as C
import os

import numpy as np

from . import integrands
from .integrands import Integrand

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libviltrum_amd.so")

F32, F64 = 4, 8
HEUR_SIZE_RELATIVE, HEUR_DEFAULT_ABSOLUTE = 0, 1
FORM_VR_OPTIMIZE, FORM_VR_FIXED, FORM_CV_ALPHA1 = 0, 1, 2
FORM_VR_OPTIMIZE_IMPORTANCE, FORM_VR_FIXED_IMPORTANCE = 3, 4  # with region_sampling_importance
FORM_VR_OPTIMIZE_RR_INTEGRAL, FORM_VR_FIXED_RR_INTEGRAL = 5, 6  # with the rr_integral_region roulette
FORM_VR_OPTIMIZE_RR_ERROR, FORM_VR_FIXED_RR_ERROR = 7, 8  # with the rr_error_region roulette
FORM_VR_OPTIMIZE_RR_PDF, FORM_VR_FIXED_RR_PDF = 9, 10  # with the rr_pdf_region roulette
# a roulette together with region_sampling_importance (optimised / fixed weight)
FORM_VR_OPTIMIZE_RR_INTEGRAL_IMPORTANCE, FORM_VR_FIXED_RR_INTEGRAL_IMPORTANCE = 11, 12
FORM_VR_OPTIMIZE_RR_ERROR_IMPORTANCE, FORM_VR_FIXED_RR_ERROR_IMPORTANCE = 13, 14
FORM_VR_OPTIMIZE_RR_PDF_IMPORTANCE, FORM_VR_FIXED_RR_PDF_IMPORTANCE = 15, 16
# with region_sampling_russian_roulette (importance or uniform sampling by a roulette)
FORM_VR_OPTIMIZE_SAMPLING_RR, FORM_VR_FIXED_SAMPLING_RR = 17, 18
# region_stratification_uniform (the optimised integrator; the Fubini form over an infinite range)
FORM_VR_OPTIMIZE_STRATIFIED, FORM_VR_FIXED_STRATIFIED = 19, 20
ORDER_CANONICAL, ORDER_REFERENCE = 0, 1
REFINER_HEURISTIC, REFINER_LEARNED = 0, 1

SYMBOLS = ["vmx_ctx_create", "vmx_ctx_destroy", "vmx_last_error", "vmx_integrate", "vmx_integrate_device",
           "vmx_integrate_batch", "vmx_export_regions", "vmx_integrate_mc", "vmx_eval", "vmx_refiner_size",
           "vmx_refiner_pairs", "vmx_refiner_train", "vmx_refiner_get", "vmx_refiner_set", "vmx_refiner_forward",
           "vmx_refiner_grad", "vmx_exact_bins", "vmx_autotune", "vmx_integrate_regions",
           "vmx_integrate_device_noalloc"]


class VmxError(RuntimeError):
    pass


class Problem(C.Structure):
    _fields_ = [
        ("scalar", C.c_int32), ("bin_dims", C.c_int32), ("heuristic", C.c_int32), ("form", C.c_int32),
        ("order", C.c_int32), ("refiner", C.c_int32), ("alpha", C.c_double),
        ("lo", C.c_double * 6), ("hi", C.c_double * 6), ("res", C.c_uint64 * 3),
        ("regions", C.c_uint64), ("spp", C.c_uint64), ("seed", C.c_uint64),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("regions", C.c_uint64), ("nodes", C.c_uint64), ("rounds", C.c_uint64), ("pairs", C.c_uint64),
        ("footprints", C.c_uint64), ("evals_regions", C.c_uint64), ("evals_residual", C.c_uint64),
        ("ties_unresolved", C.c_uint64), ("ms_generate", C.c_double), ("ms_select", C.c_double),
        ("ms_stratify", C.c_double), ("ms_bins", C.c_double), ("ms_total", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


TARGET_ENERGY, TARGET_BINVAR, TARGET_GAIN = 0, 1, 2


class TrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float), ("holdout", C.c_float),
                ("seed", C.c_uint64), ("target", C.c_int32), ("stride", C.c_uint32)]


class TrainReport(C.Structure):
    _fields_ = [("pairs", C.c_uint64), ("steps", C.c_uint64), ("loss_first", C.c_double),
                ("loss_train", C.c_double), ("loss_valid", C.c_double), ("target_mean", C.c_double),
                ("target_std", C.c_double), ("ms_pairs", C.c_double), ("ms_train", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneReport(C.Structure):
    _fields_ = [("gen_blocks_per_cu", C.c_int32), ("round_chunk", C.c_int32), ("bins_xcd", C.c_int32),
                ("deepen", C.c_int32), ("trials", C.c_int32), ("ms_best", C.c_double), ("ms_default", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """The HIP library; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VmxError("libviltrum_amd.so not built: run `python -m viltrum_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.vmx_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        L.vmx_ctx_destroy.argtypes = [C.c_void_p]
        L.vmx_last_error.argtypes = [C.c_void_p]
        L.vmx_last_error.restype = C.c_char_p
        L.vmx_integrate.argtypes = [C.c_void_p, C.POINTER(Problem), C.c_void_p, C.c_void_p, C.POINTER(Stats)]
        L.vmx_integrate_device.argtypes = [C.c_void_p, C.POINTER(Problem), C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.POINTER(Stats)]
        L.vmx_integrate_device_noalloc.argtypes = [C.c_void_p, C.POINTER(Problem), C.c_void_p, C.c_void_p,
                                                   C.c_void_p, C.POINTER(Stats)]
        L.vmx_integrate_batch.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p]
        L.vmx_autotune.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.POINTER(TuneReport)]
        L.vmx_export_regions.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]
        L.vmx_integrate_regions.argtypes = [C.c_void_p, C.POINTER(Problem), C.c_void_p, C.c_uint64, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Stats)]
        L.vmx_eval.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint64, C.c_void_p, C.c_void_p]
        L.vmx_integrate_mc.argtypes = [C.c_void_p, C.POINTER(Problem), C.c_void_p, C.c_int32, C.c_void_p,
                                       C.c_void_p]
        L.vmx_refiner_size.argtypes = [C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32)]
        L.vmx_refiner_pairs.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.POINTER(TrainOpts),
                                        C.c_uint64, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]
        L.vmx_refiner_train.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.POINTER(TrainOpts),
                                        C.POINTER(TrainReport)]
        L.vmx_refiner_get.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.vmx_refiner_set.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.vmx_refiner_forward.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]
        L.vmx_refiner_grad.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.POINTER(C.c_double)]
        L.vmx_exact_bins.argtypes = [C.c_void_p, C.POINTER(Problem), C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def refiner_size(dim):
    """(parameter count, feature count F, padded row width KP) of the priority network."""
    n, f, kp = C.c_uint64(), C.c_int32(), C.c_int32()
    rc = lib().vmx_refiner_size(dim, C.byref(n), C.byref(f), C.byref(kp))
    if rc != 0:
        raise VmxError("no priority network for dimension %d (status %d)" % (dim, rc))
    return n.value, f.value, kp.value


class Context:
    """Owns device buffers and a stream on one GPU (vmx_ctx)."""

    def __init__(self, device=0):
        self.ptr = C.c_void_p()
        rc = lib().vmx_ctx_create(device, C.byref(self.ptr))
        if rc != 0:
            raise VmxError("vmx_ctx_create failed (%d): no usable GPU %d" % (rc, device))

    def close(self):
        if self.ptr:
            lib().vmx_ctx_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise VmxError("%s (status %d)" % (lib().vmx_last_error(self.ptr).decode(), rc))

    def integrate_host(self, problem, f, bins):
        """bins: contiguous numpy array of the evaluation type, filled in place."""
        st = Stats()
        self._check(lib().vmx_integrate(self.ptr, C.byref(problem), C.cast(f.ptr(), C.c_void_p),
                                        bins.ctypes.data_as(C.c_void_p), C.byref(st)))
        return st

    def integrate_device(self, problem, f, bins_ptr, stream=None):
        st = Stats()
        self._check(lib().vmx_integrate_device(self.ptr, C.byref(problem), C.cast(f.ptr(), C.c_void_p),
                                               C.c_void_p(bins_ptr), C.c_void_p(stream or 0), C.byref(st)))
        return st

    def integrate_device_noalloc(self, problem, f, bins_ptr, stream=None):
        """integrate_device without any allocation: work memory comes from the library's
        fixed per-GPU scratch array (VMX_NOALLOC_SCRATCH_BYTES); calls on one GPU must
        not overlap."""
        st = Stats()
        self._check(lib().vmx_integrate_device_noalloc(self.ptr, C.byref(problem), C.cast(f.ptr(), C.c_void_p),
                                                       C.c_void_p(bins_ptr), C.c_void_p(stream or 0), C.byref(st)))
        return st

    def integrate_batch(self, problems, fs, bins_ptrs, stream=None):
        """n problems in one launch sequence; bins_ptrs are device addresses."""
        n = len(problems)
        P = (Problem * n)(*problems)
        F = (integrands.IntegrandDesc * n)(*[f.desc for f in fs])
        B = (C.c_void_p * n)(*bins_ptrs)
        S = (Stats * n)()
        self._check(lib().vmx_integrate_batch(self.ptr, n, C.cast(P, C.c_void_p), C.cast(F, C.c_void_p),
                                              C.cast(B, C.c_void_p), C.c_void_p(stream or 0), C.cast(S, C.c_void_p)))
        return list(S)

    def autotune(self, problems, fs, bins_ptrs, stream=None):
        """Time the batch at each candidate launch shape and keep the fastest for this
        workload class (vmx_autotune); the batch's results stay in bins_ptrs."""
        n = len(problems)
        P = (Problem * n)(*problems)
        F = (integrands.IntegrandDesc * n)(*[f.desc for f in fs])
        B = (C.c_void_p * n)(*bins_ptrs)
        rep = TuneReport()
        self._check(lib().vmx_autotune(self.ptr, n, C.cast(P, C.c_void_p), C.cast(F, C.c_void_p),
                                       C.cast(B, C.c_void_p), C.c_void_p(stream or 0), C.byref(rep)))
        return rep.as_dict()

    def integrate_regions_host(self, problem, f, regions, bins):
        """The per-bin phase over a given region list (dict lo, hi, nodes as export_regions
        returns); bins: contiguous numpy array of the evaluation type, filled in place."""
        lo = np.ascontiguousarray(regions["lo"], dtype=np.float64)
        hi = np.ascontiguousarray(regions["hi"], dtype=np.float64)
        nodes = np.ascontiguousarray(regions["nodes"], dtype=np.float64)
        st = Stats()
        self._check(lib().vmx_integrate_regions(self.ptr, C.byref(problem), C.cast(f.ptr(), C.c_void_p),
                                                lo.shape[0], lo.ctypes.data_as(C.c_void_p),
                                                hi.ctypes.data_as(C.c_void_p), nodes.ctypes.data_as(C.c_void_p),
                                                bins.ctypes.data_as(C.c_void_p), C.byref(st)))
        return st

    def export_regions(self, D):
        n = C.c_uint64(0)
        L = lib()
        self._check(L.vmx_export_regions(self.ptr, C.byref(n), None, None, None, None, None))
        N = n.value
        lo, hi = np.zeros((N, D)), np.zeros((N, D))
        nodes, err = np.zeros((N, 3 ** D)), np.zeros(N)
        dim = np.zeros(N, dtype=np.int32)
        self._check(L.vmx_export_regions(self.ptr, C.byref(n), lo.ctypes.data_as(C.c_void_p),
                                         hi.ctypes.data_as(C.c_void_p), nodes.ctypes.data_as(C.c_void_p),
                                         err.ctypes.data_as(C.c_void_p), dim.ctypes.data_as(C.c_void_p)))
        return dict(lo=lo, hi=hi, nodes=nodes, err=err, dim=dim)

    def integrate_mc(self, f, res, spp, seed, rng=None, dtype=np.float64, stratified=True):
        """Plain per-bin Monte Carlo (monte_carlo_per_bin); returns (bins, stderr)."""
        rng = rng or range_primary(f.dim)
        p = Integrator(0, spp, seed).problem(f, res, rng, F32 if dtype == np.float32 else F64)
        bins = np.zeros(res, dtype=dtype)
        se = np.zeros(res, dtype=np.float64)
        self._check(lib().vmx_integrate_mc(self.ptr, C.byref(p), C.cast(f.ptr(), C.c_void_p), int(stratified),
                                           bins.ctypes.data_as(C.c_void_p), se.ctypes.data_as(C.c_void_p)))
        return bins, se

    # ---- learned refinement priority
    @staticmethod
    def _problems(problems, fs):
        n = len(problems)
        return n, (Problem * n)(*problems), (integrands.IntegrandDesc * n)(*[f.desc for f in fs])

    def refiner_pairs(self, problems, fs, target=TARGET_ENERGY, stride=1):
        """(features uint16 [n, KP] BF16 bits, targets float32 [n]) of the regions the
        heuristic method generates for the problems (one family and dim)."""
        n, P, F = self._problems(problems, fs)
        o = TrainOpts(0, 0, 0.0, 0.0, 0, target, stride)
        cnt = C.c_uint64()
        L = lib()
        self._check(L.vmx_refiner_pairs(self.ptr, n, C.cast(P, C.c_void_p), C.cast(F, C.c_void_p), C.byref(o), 0,
                                        None, None, C.byref(cnt)))
        kp = refiner_size(fs[0].dim)[2]
        cap = cnt.value + 4096
        X = np.zeros((cap, kp), dtype=np.uint16)
        t = np.zeros(cap, dtype=np.float32)
        self._check(L.vmx_refiner_pairs(self.ptr, n, C.cast(P, C.c_void_p), C.cast(F, C.c_void_p), C.byref(o),
                                        cap, X.ctypes.data_as(C.c_void_p), t.ctypes.data_as(C.c_void_p),
                                        C.byref(cnt)))
        m = min(cnt.value, cap)
        return X[:m], t[:m]

    def refiner_train(self, problems, fs, epochs=8, batch=1024, lr=2e-3, holdout=0.05, seed=0,
                      target=TARGET_ENERGY, stride=1):
        n, P, F = self._problems(problems, fs)
        o = TrainOpts(epochs, batch, lr, holdout, seed, target, stride)
        r = TrainReport()
        self._check(lib().vmx_refiner_train(self.ptr, n, C.cast(P, C.c_void_p), C.cast(F, C.c_void_p), C.byref(o),
                                            C.byref(r)))
        return r.as_dict()

    def refiner_get(self, dim):
        p = np.zeros(refiner_size(dim)[0], dtype=np.float32)
        self._check(lib().vmx_refiner_get(self.ptr, dim, p.ctypes.data_as(C.c_void_p)))
        return p

    def refiner_set(self, dim, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != refiner_size(dim)[0]:
            raise VmxError("expected %d parameters" % refiner_size(dim)[0])
        self._check(lib().vmx_refiner_set(self.ptr, dim, p.ctypes.data_as(C.c_void_p)))

    def refiner_forward(self, dim, X, params=None):
        X = np.ascontiguousarray(X, dtype=np.uint16)
        y = np.zeros(X.shape[0], dtype=np.float32)
        pp = None
        if params is not None:
            params = np.ascontiguousarray(params, dtype=np.float32)
            pp = params.ctypes.data_as(C.c_void_p)
        self._check(lib().vmx_refiner_forward(self.ptr, dim, pp, X.shape[0], X.ctypes.data_as(C.c_void_p),
                                              y.ctypes.data_as(C.c_void_p)))
        return y

    def refiner_grad(self, dim, params, X, t):
        params = np.ascontiguousarray(params, dtype=np.float32)
        X = np.ascontiguousarray(X, dtype=np.uint16)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(params.size - 2, dtype=np.float32)
        loss = C.c_double()
        self._check(lib().vmx_refiner_grad(self.ptr, dim, params.ctypes.data_as(C.c_void_p), X.shape[0],
                                           X.ctypes.data_as(C.c_void_p), t.ctypes.data_as(C.c_void_p),
                                           g.ctypes.data_as(C.c_void_p), C.byref(loss)))
        return g, loss.value

    def exact_bins(self, f, res, rng=None):
        """FP64 closed-form bins of a 2-D static blob integrand (the RMSE truth)."""
        rng = rng or range_primary(f.dim)
        p = Integrator(0, 0, 0).problem(f, res, rng, F64)
        out = np.zeros(res, dtype=np.float64)
        self._check(lib().vmx_exact_bins(self.ptr, C.byref(p), C.cast(f.ptr(), C.c_void_p),
                                         out.ctypes.data_as(C.c_void_p)))
        return out

    def eval(self, f, x, scalar=F64):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, f.dim)
        out = np.zeros(x.shape[0])
        self._check(lib().vmx_eval(self.ptr, C.cast(f.ptr(), C.c_void_p), scalar, x.shape[0],
                                   x.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p)))
        return out


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Range:
    """Box [lo, hi] (src/range.h:17-139)."""

    def __init__(self, lo, hi):
        self.lo = [float(v) for v in lo]
        self.hi = [float(v) for v in hi]
        if len(self.lo) != len(self.hi):
            raise ValueError("lo/hi dims differ")

    @property
    def dims(self):
        return len(self.lo)


def range_primary(D):
    """[0,1]^D (src/range.h:373-376)."""
    return Range([0.0] * D, [1.0] * D)


class RangeInfinite:
    """A range of unbounded dimension (src/range-infinite.h:16-87): explicit entries
    [lo[i], hi[i]], then [0, 1] for every entry after them."""
    dims = integrands.INFINITE_DIM

    def __init__(self, lo=(), hi=()):
        self.lo = [float(v) for v in lo]
        self.hi = [float(v) for v in hi]
        if len(self.lo) != len(self.hi):
            raise ValueError("lo/hi entries differ")

    def min(self, i):
        return self.lo[i] if i < len(self.lo) else 0.0

    def max(self, i):
        return self.hi[i] if i < len(self.hi) else 1.0


def range_primary_infinite():
    """[0, 1] in every entry (src/range-infinite.h:129-132)."""
    return RangeInfinite()


def range_infinite(lo, hi):
    """Explicit entries [lo[i], hi[i]], [0, 1] after them (src/range-infinite.h:106-126)."""
    return RangeInfinite(lo, hi)


class Integrator:
    """Parameters of a region-based CV integrator (IntegratorRegionBased,
    src/newton-cotes/integrator-region-based.h:5-29)."""

    def __init__(self, iterations, spp, seed, heuristic=HEUR_SIZE_RELATIVE, form=FORM_VR_OPTIMIZE, alpha=1.0,
                 order=ORDER_CANONICAL, refiner=REFINER_HEURISTIC):
        self.iterations = int(iterations)
        self.spp = int(spp)
        self.seed = int(seed)
        self.heuristic = heuristic
        self.form = form
        self.alpha = float(alpha)
        self.order = order
        self.refiner = refiner
        self.last_stats = None

    def prepare(self, f, rng):
        """(integrand descriptor, range) the problem is posed on."""
        return f, rng

    def problem(self, f, res, rng, scalar):
        if rng.dims != f.dim:
            raise ValueError("range has %d dims, integrand %d" % (rng.dims, f.dim))
        if not 1 <= len(res) <= min(3, f.dim):
            raise ValueError("bins must have 1..min(3, D) dims")
        p = Problem()
        p.scalar = scalar
        p.bin_dims = len(res)
        p.heuristic = self.heuristic
        p.form = self.form
        p.order = self.order
        p.refiner = self.refiner
        p.alpha = self.alpha
        for d in range(f.dim):
            p.lo[d] = rng.lo[d]
            p.hi[d] = rng.hi[d]
        for i, r in enumerate(res):
            p.res[i] = int(r)
        p.regions = self.iterations + 1
        p.spp = self.spp
        p.seed = self.seed
        return p


class FubiniIntegrator(Integrator):
    """integrator_crespo2021_infinite<rd>(iterations, mc_samples, spp, seed)
    (src/control-variates/integrator-crespo2021.h:24-44): crespo2021 over the first rd
    dims of the range, the remaining dims integrated by Monte Carlo (mc_samples per region
    node value, one per residual sample).  FP32 bins only, as the reference."""

    def __init__(self, rd, iterations, mc_samples, spp, seed=0, **kw):
        super().__init__(iterations, spp, seed, HEUR_SIZE_RELATIVE, FORM_VR_OPTIMIZE, **kw)
        self.rd = int(rd)
        self.mc_samples = int(mc_samples)

    # seed of the node values' Monte Carlo: monte_carlo(mc_samples, 2 * seed + 1)
    def mc_seed(self):
        return 2 * self.seed + 1

    def prepare(self, f, rng):
        if isinstance(rng, RangeInfinite) or f.dim == integrands.INFINITE_DIM:
            if not (isinstance(rng, RangeInfinite) and f.dim == integrands.INFINITE_DIM):
                raise ValueError("an integrand of a sequence goes with an infinite range")
            rd = self.rd
            lo = [rng.min(i) for i in range(rd)]
            hi = [rng.max(i) for i in range(rd)]
            g = integrands.fubini_infinite(f, rd, self.mc_samples, self.mc_seed(), lo, hi, rng.lo[rd:], rng.hi[rd:])
            return g, Range(lo, hi)
        if rng.dims != f.dim:
            raise ValueError("range has %d dims, integrand %d" % (rng.dims, f.dim))
        g = integrands.fubini(f, self.rd, self.mc_samples, self.mc_seed(), rng.lo, rng.hi)
        return g, Range(rng.lo[:self.rd], rng.hi[:self.rd])


def integrator_crespo2021_infinite(rd, iterations, mc_samples, spp, seed=0, **kw):
    return FubiniIntegrator(rd, iterations, mc_samples, spp, seed, **kw)


class FubiniOptimizedIntegrator(FubiniIntegrator):
    """integrator_adaptive_fubini_variance_reduction_parallel_optimized<rd>(nested(simpson,
    trapezoidal), error_heuristic_size(error_metric_relative(), 1e-5), iterations,
    mc_samples, region_stratification_uniform(), cv_optimize_weight() or
    cv_fixed_weight(alpha), region_sampling_uniform(), spp, seed)
    (src/control-variates/integrator-adaptive-fubini-variance-reduction-optimized.h:17-23):
    node values Monte Carlo from seed, residual samples split evenly over each bin's
    regions from seed + 1.  The reference builds it over infinite ranges only."""

    def __init__(self, rd, iterations, mc_samples, spp, seed=0, alpha=None, **kw):
        super().__init__(rd, iterations, mc_samples, spp, seed, **kw)
        self.form = FORM_VR_OPTIMIZE_STRATIFIED if alpha is None else FORM_VR_FIXED_STRATIFIED
        self.alpha = 1.0 if alpha is None else float(alpha)

    def mc_seed(self):
        return self.seed

    def prepare(self, f, rng):
        if not isinstance(rng, RangeInfinite):
            raise ValueError("the optimised Fubini integrator needs an infinite range (as the reference)")
        return super().prepare(f, rng)

    def problem(self, f, res, rng, scalar):
        p = super().problem(f, res, rng, scalar)
        p.seed = self.seed + 1
        return p


def integrator_adaptive_fubini_variance_reduction_parallel_optimized(rd, iterations, mc_samples, spp, seed=0,
                                                                     alpha=None, **kw):
    return FubiniOptimizedIntegrator(rd, iterations, mc_samples, spp, seed, alpha=alpha, **kw)


def integrator_crespo2021(iterations, spp, seed=0, **kw):
    """Nested Simpson/trapezoid regions, error_heuristic_size(relative, 1e-5), per-bin
    residual with cv_optimize_weight (src/control-variates/integrator-crespo2021.h:7-22)."""
    return Integrator(iterations, spp, seed, HEUR_SIZE_RELATIVE, FORM_VR_OPTIMIZE, **kw)


def integrator_adaptive_control_variates_parallel(iterations, spp, seed=0, heuristic=HEUR_DEFAULT_ABSOLUTE, **kw):
    """alpha = 1 control variates, bins accumulated
    (src/control-variates/integrator-adaptive-control-variates.h:17-20)."""
    return Integrator(iterations, spp, seed, heuristic, FORM_CV_ALPHA1, **kw)


def integrator_variance_reduction_importance(iterations, spp, seed=0, alpha=None, heuristic=HEUR_SIZE_RELATIVE, **kw):
    """regions_integrator_parallel_variance_reduction(rr_uniform_region(), cv_optimize_weight()
    or cv_fixed_weight(alpha), region_sampling_importance(), ...)
    (src/control-variates/region-sampling.h:22-44): residual samples from each region
    polynomial's |P| pdf over the bin's box, weighted by 1/pdf.  FP32, dims 1..6."""
    if alpha is None:
        return Integrator(iterations, spp, seed, heuristic, FORM_VR_OPTIMIZE_IMPORTANCE, **kw)
    return Integrator(iterations, spp, seed, heuristic, FORM_VR_FIXED_IMPORTANCE, alpha=alpha, **kw)


def integrator_variance_reduction_sampling_russian_roulette(iterations, spp, seed=0, alpha=None,
                                                            heuristic=HEUR_SIZE_RELATIVE, **kw):
    """regions_integrator_parallel_variance_reduction(rr_uniform_region(), cv_optimize_weight()
    or cv_fixed_weight(alpha), region_sampling_russian_roulette(), ...)
    (src/control-variates/region-sampling.h:47-84): each residual sample is drawn from the
    region polynomial's pdf with probability pdf_integral / (pdf_integral + volume) of the
    bin's box, else uniformly, and weighted by 1/pdf of the strategy taken.  FP32, dims 1..6."""
    if alpha is None:
        return Integrator(iterations, spp, seed, heuristic, FORM_VR_OPTIMIZE_SAMPLING_RR, **kw)
    return Integrator(iterations, spp, seed, heuristic, FORM_VR_FIXED_SAMPLING_RR, alpha=alpha, **kw)


def integrator_variance_reduction_rr_integral(iterations, spp, seed=0, alpha=None, heuristic=HEUR_SIZE_RELATIVE,
                                              importance=False, **kw):
    """regions_integrator_parallel_variance_reduction(rr_integral_region(), cv_optimize_weight()
    or cv_fixed_weight(alpha), region_sampling_uniform(), ...)
    (src/control-variates/region-russian-roulette.h:30-67): a residual sample picks the
    bin's region r with probability ~ |integral of P_r over the bin's box|, weight 1/p."""
    if importance:  # with region_sampling_importance as the point sampler (float, dims 1..6)
        if alpha is None:
            return Integrator(iterations, spp, seed, heuristic, FORM_VR_OPTIMIZE_RR_INTEGRAL_IMPORTANCE, **kw)
        return Integrator(iterations, spp, seed, heuristic, FORM_VR_FIXED_RR_INTEGRAL_IMPORTANCE, alpha=alpha, **kw)
    if alpha is None:
        return Integrator(iterations, spp, seed, heuristic, FORM_VR_OPTIMIZE_RR_INTEGRAL, **kw)
    return Integrator(iterations, spp, seed, heuristic, FORM_VR_FIXED_RR_INTEGRAL, alpha=alpha, **kw)


def integrator_variance_reduction_rr_error(iterations, spp, seed=0, alpha=None, heuristic=HEUR_SIZE_RELATIVE,
                                           importance=False, **kw):
    """... with rr_error_region() (src/control-variates/region-russian-roulette.h:68-104):
    probability ~ |Region::error()| times the share of the region inside the bin."""
    if importance:  # with region_sampling_importance as the point sampler (float, dims 1..6)
        if alpha is None:
            return Integrator(iterations, spp, seed, heuristic, FORM_VR_OPTIMIZE_RR_ERROR_IMPORTANCE, **kw)
        return Integrator(iterations, spp, seed, heuristic, FORM_VR_FIXED_RR_ERROR_IMPORTANCE, alpha=alpha, **kw)
    if alpha is None:
        return Integrator(iterations, spp, seed, heuristic, FORM_VR_OPTIMIZE_RR_ERROR, **kw)
    return Integrator(iterations, spp, seed, heuristic, FORM_VR_FIXED_RR_ERROR, alpha=alpha, **kw)


def integrator_variance_reduction_rr_pdf(iterations, spp, seed=0, alpha=None, heuristic=HEUR_SIZE_RELATIVE,
                                         importance=False, **kw):
    """... with rr_pdf_region() (src/control-variates/region-russian-roulette.h:106-149):
    probability ~ the integral of the region's importance pdf over the bin's box."""
    if importance:  # with region_sampling_importance as the point sampler (float, dims 1..6)
        if alpha is None:
            return Integrator(iterations, spp, seed, heuristic, FORM_VR_OPTIMIZE_RR_PDF_IMPORTANCE, **kw)
        return Integrator(iterations, spp, seed, heuristic, FORM_VR_FIXED_RR_PDF_IMPORTANCE, alpha=alpha, **kw)
    if alpha is None:
        return Integrator(iterations, spp, seed, heuristic, FORM_VR_OPTIMIZE_RR_PDF, **kw)
    return Integrator(iterations, spp, seed, heuristic, FORM_VR_FIXED_RR_PDF, alpha=alpha, **kw)


def integrator_variance_reduction_fixed(alpha, iterations, spp, seed=0, heuristic=HEUR_SIZE_RELATIVE, **kw):
    """cv_fixed_weight(alpha) residual
    (src/control-variates/regions-integrator-parallel-variance-reduction.h:127-130)."""
    return Integrator(iterations, spp, seed, heuristic, FORM_VR_FIXED, alpha=alpha, **kw)


def integrate_regions(integrator, bins, f, rng, regions, ctx=None):
    """RegionsIntegrator::integrate_regions: the integrator's per-bin phase (form, spp,
    seed) over a given region list (dict lo, hi, nodes, e.g. from Context.export_regions)
    instead of generating one; fills bins."""
    ctx = ctx or default_context()
    if bins.dtype not in (np.float32, np.float64) or not bins.flags.c_contiguous:
        raise TypeError("bins must be a C-contiguous float32 or float64 array")
    scalar = F32 if bins.dtype == np.float32 else F64
    g, grng = integrator.prepare(f, rng)
    p = integrator.problem(g, bins.shape, grng, scalar)
    integrator.last_stats = ctx.integrate_regions_host(p, g, regions, bins)
    return bins


def integrate(integrator, *args, ctx=None):
    """integrate(integrator, bins, f, range)  -> fills bins (numpy, float32/float64)
    integrate(integrator, f, range[, dtype])  -> scalar (one bin)."""
    ctx = ctx or default_context()
    if isinstance(args[0], np.ndarray):
        bins, f, rng = args[:3]
        if bins.dtype not in (np.float32, np.float64) or not bins.flags.c_contiguous:
            raise TypeError("bins must be a C-contiguous float32 or float64 array")
        scalar = F32 if bins.dtype == np.float32 else F64
        g, grng = integrator.prepare(f, rng)
        p = integrator.problem(g, bins.shape, grng, scalar)
        integrator.last_stats = ctx.integrate_host(p, g, bins)
        return bins
    f, rng = args[:2]
    dtype = args[2] if len(args) > 2 else np.float64
    one = np.zeros(1, dtype=dtype)
    integrate(integrator, one, f, rng, ctx=ctx)
    return float(one[0])

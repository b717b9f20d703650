"""ctypes access to the checkers: the oracle (oracle/liboracle.so, a CPU restatement of
the reference algorithm) and, when it has been built here, the reference itself
(oracle/_ref/libviltrum_ref.so, the reference's own headers instantiated by
oracle/ref_driver.cc).  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libviltrum_ref.so")

INT_VR_OPTIMIZE, INT_VR_FIXED, INT_CV_ALPHA1 = 0, 1, 2
INT_VR_OPTIMIZE_IMPORTANCE, INT_VR_FIXED_IMPORTANCE = 3, 4  # with region_sampling_importance
INT_VR_OPTIMIZE_RR_INTEGRAL, INT_VR_FIXED_RR_INTEGRAL = 5, 6  # with the rr_integral_region roulette
INT_VR_OPTIMIZE_RR_ERROR, INT_VR_FIXED_RR_ERROR = 7, 8  # with the rr_error_region roulette
INT_VR_OPTIMIZE_RR_PDF, INT_VR_FIXED_RR_PDF = 9, 10  # with the rr_pdf_region roulette
# a roulette together with region_sampling_importance (optimised / fixed weight)
INT_VR_OPTIMIZE_RR_INTEGRAL_IMPORTANCE, INT_VR_FIXED_RR_INTEGRAL_IMPORTANCE = 11, 12
INT_VR_OPTIMIZE_RR_ERROR_IMPORTANCE, INT_VR_FIXED_RR_ERROR_IMPORTANCE = 13, 14
INT_VR_OPTIMIZE_RR_PDF_IMPORTANCE, INT_VR_FIXED_RR_PDF_IMPORTANCE = 15, 16
# region_sampling_russian_roulette as the point sampler (uniform region roulette)
INT_VR_OPTIMIZE_SAMPLING_RR, INT_VR_FIXED_SAMPLING_RR = 17, 18
HEUR_SIZE_RELATIVE, HEUR_DEFAULT_ABSOLUTE = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint64)


class OracleStats(C.Structure):
    _fields_ = [("regions", C.c_uint64), ("pairs", C.c_uint64), ("evals_regions", C.c_uint64),
                ("evals_residual", C.c_uint64), ("t_generate", C.c_double), ("t_stratify", C.c_double),
                ("t_bins", C.c_double)]


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


def _load(path):
    if not os.path.exists(path):
        return None
    return C.CDLL(path)


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        lib = _load(ORACLE_SO)
        if lib is None:
            raise RuntimeError("oracle not built: run `make -C oracle`")
        lib.vmo_last_error.restype = C.c_char_p
        _oracle = lib
    return _oracle


def ref():
    """The reference build, or None where it was not built (no /root/reference)."""
    global _ref
    if _ref is None:
        lib = _load(REF_SO)
        if lib is None:
            return None
        lib.vmr_last_error.restype = C.c_char_p
        _ref = lib
    return _ref


def n_evals(D, iterations):
    return 3 ** D + iterations * 2 * 3 ** (D - 1)


class Regions:
    def __init__(self, lo, hi, err, dim, nodes=None, evals=None):
        self.lo, self.hi, self.err, self.dim, self.nodes, self.evals = lo, hi, err, dim, nodes, evals

    def __len__(self):
        return len(self.err)


def _box(D, lo, hi):
    lo = np.zeros(D) if lo is None else np.asarray(lo, dtype=np.float64)
    hi = np.ones(D) if hi is None else np.asarray(hi, dtype=np.float64)
    return np.ascontiguousarray(lo), np.ascontiguousarray(hi)


def generate(f, iterations, scalar=8, heuristic=HEUR_SIZE_RELATIVE, lo=None, hi=None, use_ref=False, evals=False):
    """Region list in the reference's heap-array order."""
    D = f.dim
    lo, hi = _box(D, lo, hi)
    N = iterations + 1
    olo, ohi = np.zeros(N * D), np.zeros(N * D)
    err, dim = np.zeros(N), np.zeros(N, dtype=np.int32)
    ev = np.zeros(n_evals(D, iterations) * (D + 1)) if evals else None
    nev = C.c_uint64(0)
    if use_ref:
        lib = ref()
        nodes = None
        rc = lib.vmr_generate(C.c_int(scalar), f.ptr(), _ptr(lo), _ptr(hi), C.c_int(heuristic),
                              C.c_uint64(iterations), _ptr(olo), _ptr(ohi), _ptr(err), _ptr(dim, _ip),
                              _ptr(ev), C.byref(nev))
        if rc < 0:
            raise RuntimeError(lib.vmr_last_error().decode())
    else:
        lib = oracle()
        nodes = np.zeros(N * 3 ** D)
        rc = lib.vmo_generate(C.c_int(scalar), f.ptr(), _ptr(lo), _ptr(hi), C.c_int(heuristic),
                              C.c_uint64(iterations), _ptr(olo), _ptr(ohi), _ptr(nodes), _ptr(err),
                              _ptr(dim, _ip), _ptr(ev), C.byref(nev))
        if rc < 0:
            raise RuntimeError(lib.vmo_last_error().decode())
        nodes = nodes.reshape(N, 3 ** D)
    if ev is not None:
        ev = ev[: nev.value * (D + 1)].reshape(-1, D + 1)
    return Regions(olo.reshape(N, D), ohi.reshape(N, D), err, dim, nodes, ev)


def integrate(f, res, iterations, spp, seed, scalar=8, integrator=INT_VR_OPTIMIZE, heuristic=HEUR_SIZE_RELATIVE,
              alpha=1.0, lo=None, hi=None, nthreads=1, use_ref=False, bins_in=None):
    """Bins in C order (bins[i0][i1]); returns (bins, stats or None)."""
    D = f.dim
    lo, hi = _box(D, lo, hi)
    res = tuple(int(r) for r in res)
    r = np.ascontiguousarray(np.array(res, dtype=np.uint64))
    bins = np.zeros(int(np.prod(res))) if bins_in is None else np.array(bins_in, dtype=np.float64).ravel().copy()
    if use_ref:
        lib = ref()
        rc = lib.vmr_integrate(C.c_int(scalar), C.c_int(integrator), C.c_int(heuristic), C.c_double(alpha), f.ptr(),
                               _ptr(lo), _ptr(hi), C.c_int(len(res)), _ptr(r, _up), C.c_uint64(iterations),
                               C.c_uint64(spp), C.c_uint64(seed), _ptr(bins))
        if rc < 0:
            raise RuntimeError(lib.vmr_last_error().decode())
        return bins.reshape(res), None
    lib = oracle()
    st = OracleStats()
    rc = lib.vmo_integrate(C.c_int(scalar), C.c_int(integrator), C.c_int(heuristic), C.c_double(alpha), f.ptr(),
                           _ptr(lo), _ptr(hi), C.c_int(len(res)), _ptr(r, _up), C.c_uint64(iterations),
                           C.c_uint64(spp), C.c_uint64(seed), C.c_int(nthreads), _ptr(bins), C.byref(st))
    if rc < 0:
        raise RuntimeError(lib.vmo_last_error().decode())
    return bins.reshape(res), st


FUBINI_GEN_COPIES, FUBINI_BIN_COPIES = 3, 1  # copy-constructions of the Monte Carlo objects (see oracle.hpp)


def integrate_fubini(f, rd, res, iterations, mc, spp, seed, lo=None, hi=None, nthreads=1, use_ref=False,
                     gen_copies=FUBINI_GEN_COPIES, bin_copies=FUBINI_BIN_COPIES):
    """integrator_crespo2021_infinite<rd>(iterations, mc, spp, seed) over f's float range:
    regions and bins over the first rd dims.  Bins in C order; returns (bins, stats)."""
    D = f.dim
    lo, hi = _box(D, lo, hi)
    res = tuple(int(r) for r in res)
    r = np.ascontiguousarray(np.array(res, dtype=np.uint64))
    bins = np.zeros(int(np.prod(res)))
    if use_ref:
        lib = ref()
        rc = lib.vmr_integrate_fubini(C.c_int(rd), f.ptr(), _ptr(lo), _ptr(hi), C.c_int(len(res)), _ptr(r, _up),
                                      C.c_uint64(iterations), C.c_uint64(mc), C.c_uint64(spp), C.c_uint64(seed),
                                      _ptr(bins))
        if rc < 0:
            raise RuntimeError(lib.vmr_last_error().decode())
        return bins.reshape(res), None
    lib = oracle()
    st = OracleStats()
    rc = lib.vmo_integrate_fubini(C.c_int(rd), f.ptr(), _ptr(lo), _ptr(hi), C.c_int(len(res)), _ptr(r, _up),
                                  C.c_uint64(iterations), C.c_uint64(mc), C.c_uint64(spp), C.c_uint64(seed),
                                  C.c_int(gen_copies), C.c_int(bin_copies), C.c_int(nthreads), _ptr(bins),
                                  C.byref(st))
    if rc < 0:
        raise RuntimeError(lib.vmo_last_error().decode())
    return bins.reshape(res), st


class DysonParams(C.Structure):
    _fields_ = [("decay", C.c_double)]


# forms of integrate_fubini_inf
INF_CRESPO2021, INF_OPTIMIZE_STRATIFIED, INF_FIXED_STRATIFIED = 0, 1, 2
# copy-constructions of the Monte Carlo objects over an infinite range, per form (see oracle.hpp)
FUBINI_INF_GEN_COPIES = {INF_CRESPO2021: 3, INF_OPTIMIZE_STRATIFIED: 3, INF_FIXED_STRATIFIED: 3}
FUBINI_INF_BIN_COPIES = 1


def integrate_fubini_inf(decay, rd, res, iterations, mc, spp, seed, form=INF_CRESPO2021, alpha=1.0, lo=None,
                         hi=None, rest_lo=(), rest_hi=(), nthreads=1, use_ref=False, gen_copies=None,
                         bin_copies=FUBINI_INF_BIN_COPIES):
    """The Dyson series (decay) over an infinite range: region dims [lo, hi] (rd entries),
    then the rest's explicit entries [rest_lo, rest_hi] and [0, 1] after them.  form 0:
    integrator_crespo2021_infinite<rd>; 1 / 2: the optimised Fubini integrator with
    region_stratification_uniform and cv_optimize_weight / cv_fixed_weight(alpha).
    Bins in C order over the first len(res) region dims; returns (bins, stats)."""
    lo, hi = _box(rd, lo, hi)
    rl = np.ascontiguousarray(np.array(rest_lo, dtype=np.float64).reshape(-1))
    rh = np.ascontiguousarray(np.array(rest_hi, dtype=np.float64).reshape(-1))
    assert rl.size == rh.size
    dp = DysonParams(float(decay))
    res = tuple(int(r) for r in res)
    r = np.ascontiguousarray(np.array(res, dtype=np.uint64))
    bins = np.zeros(int(np.prod(res)))
    rlp = _ptr(rl) if rl.size else None
    rhp = _ptr(rh) if rh.size else None
    if use_ref:
        lib = ref()
        rc = lib.vmr_integrate_fubini_inf(C.c_int(rd), C.c_int(form), C.byref(dp), _ptr(lo), _ptr(hi),
                                          C.c_int(rl.size), rlp, rhp, C.c_int(len(res)), _ptr(r, _up),
                                          C.c_uint64(iterations), C.c_uint64(mc), C.c_uint64(spp), C.c_uint64(seed),
                                          C.c_double(alpha), _ptr(bins))
        if rc < 0:
            raise RuntimeError(lib.vmr_last_error().decode())
        return bins.reshape(res), None
    if gen_copies is None:
        gen_copies = FUBINI_INF_GEN_COPIES[form]
    lib = oracle()
    st = OracleStats()
    rc = lib.vmo_integrate_fubini_inf(C.c_int(rd), C.c_int(form), C.byref(dp), _ptr(lo), _ptr(hi), C.c_int(rl.size),
                                      rlp, rhp, C.c_int(len(res)), _ptr(r, _up), C.c_uint64(iterations),
                                      C.c_uint64(mc), C.c_uint64(spp), C.c_uint64(seed), C.c_double(alpha),
                                      C.c_int(gen_copies), C.c_int(bin_copies), C.c_int(nthreads), _ptr(bins),
                                      C.byref(st))
    if rc < 0:
        raise RuntimeError(lib.vmo_last_error().decode())
    return bins.reshape(res), st


def integrate_regions(f, regs_lo, regs_hi, regs_nodes, res, spp, seed, scalar=8, integrator=INT_VR_OPTIMIZE,
                      alpha=1.0, lo=None, hi=None, nthreads=1, bins_in=None):
    """Per-bin phase of the oracle on a caller-given region list."""
    D = f.dim
    lo, hi = _box(D, lo, hi)
    res = tuple(int(x) for x in res)
    r = np.ascontiguousarray(np.array(res, dtype=np.uint64))
    bins = np.zeros(int(np.prod(res))) if bins_in is None else np.array(bins_in, dtype=np.float64).ravel().copy()
    rl = np.ascontiguousarray(regs_lo, dtype=np.float64)
    rh = np.ascontiguousarray(regs_hi, dtype=np.float64)
    rn = np.ascontiguousarray(regs_nodes, dtype=np.float64)
    st = OracleStats()
    lib = oracle()
    rc = lib.vmo_integrate_regions(C.c_int(scalar), C.c_int(integrator), C.c_double(alpha), f.ptr(), _ptr(lo),
                                   _ptr(hi), C.c_uint64(rl.shape[0]), _ptr(rl), _ptr(rh), _ptr(rn),
                                   C.c_int(len(res)), _ptr(r, _up), C.c_uint64(spp), C.c_uint64(seed),
                                   C.c_int(nthreads), _ptr(bins), C.byref(st))
    if rc < 0:
        raise RuntimeError(lib.vmo_last_error().decode())
    return bins.reshape(res), st


def evaluate(f, x, scalar=8):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, f.dim)
    out = np.zeros(x.shape[0])
    lib = oracle()
    rc = lib.vmo_eval(C.c_int(scalar), f.ptr(), C.c_uint64(x.shape[0]), _ptr(x), _ptr(out))
    if rc < 0:
        raise RuntimeError(lib.vmo_last_error().decode())
    return out


def heuristic(D, lo, hi, nodes, scalar=8, heuristic=HEUR_SIZE_RELATIVE):
    lo = np.ascontiguousarray(lo, dtype=np.float64).reshape(-1, D)
    hi = np.ascontiguousarray(hi, dtype=np.float64).reshape(-1, D)
    nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(lo.shape[0], -1)
    err = np.zeros(lo.shape[0])
    dim = np.zeros(lo.shape[0], dtype=np.int32)
    lib = oracle()
    rc = lib.vmo_heuristic(C.c_int(scalar), C.c_int(D), C.c_int(heuristic), C.c_uint64(lo.shape[0]), _ptr(lo),
                           _ptr(hi), _ptr(nodes), _ptr(err), _ptr(dim, _ip))
    if rc < 0:
        raise RuntimeError(lib.vmo_last_error().decode())
    return err, dim


def blobs2d_exact_bins(f, res0, res1, lo=(0.0, 0.0), hi=(1.0, 1.0)):
    lib = oracle()
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    out = np.zeros(res0 * res1)
    rc = lib.vmo_blobs2d_exact_bins(C.byref(f.params), _ptr(lo), _ptr(hi), C.c_uint64(res0), C.c_uint64(res1),
                                    _ptr(out))
    if rc < 0:
        raise RuntimeError(lib.vmo_last_error().decode())
    return out.reshape(res0, res1)


# ---------------------------------------------------------------- learned priority

def nn_dims(D):
    F, KP = C.c_int(), C.c_int()
    if oracle().vmo_nn_dims(C.c_int(D), C.byref(F), C.byref(KP)) != 0:
        raise ValueError("bad D")
    return F.value, KP.value


def f2bf(x):
    """float32 array -> BF16 bit patterns (round to nearest even), via the oracle."""
    lib = oracle()
    lib.vmo_f2bf.restype = C.c_uint16
    lib.vmo_f2bf.argtypes = [C.c_float]
    x = np.asarray(x, dtype=np.float32)
    return np.array([lib.vmo_f2bf(float(v)) for v in x.ravel()], dtype=np.uint16).reshape(x.shape)


def nn_forward(kp, params, X, emulate=True):
    params = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.uint16)
    y = np.zeros(X.shape[0])
    rc = oracle().vmo_nn_forward(C.c_int(kp), params.ctypes.data_as(C.c_void_p), C.c_uint64(X.shape[0]),
                                 X.ctypes.data_as(C.c_void_p), C.c_int(int(emulate)), _ptr(y))
    if rc != 0:
        raise RuntimeError("vmo_nn_forward failed")
    return y


def nn_grad(kp, params, X, t, emulate=True):
    params = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.uint16)
    t = np.ascontiguousarray(t, dtype=np.float32)
    g = np.zeros(params.size - 2)
    loss = C.c_double()
    rc = oracle().vmo_nn_grad(C.c_int(kp), params.ctypes.data_as(C.c_void_p), C.c_uint64(X.shape[0]),
                              X.ctypes.data_as(C.c_void_p), t.ctypes.data_as(C.c_void_p), C.c_int(int(emulate)),
                              _ptr(g), C.byref(loss))
    if rc != 0:
        raise RuntimeError("vmo_nn_grad failed")
    return g, loss.value


def lres_of(res, D):
    """log2 of the bin counts per dim as the library computes it (float of the double
    log2), 0 for integrated dims."""
    out = np.zeros(D, dtype=np.float32)
    for d, r in enumerate(res):
        out[d] = np.float32(np.log2(float(r)))
    return out


def region_features(D, lo, hi, rlo, rhi, v, code, lres=None):
    F, KP = nn_dims(D)
    x = np.zeros(KP, dtype=np.float32)
    s = C.c_float()
    depth = C.c_int()
    a = [np.ascontiguousarray(z, dtype=np.float64) for z in (lo, hi, rlo, rhi, v)]
    code = np.ascontiguousarray(code, dtype=np.uint32)
    lres = np.ascontiguousarray(np.zeros(D) if lres is None else lres, dtype=np.float32)
    rc = oracle().vmo_region_features(C.c_int(D), *[_ptr(z) for z in a], code.ctypes.data_as(C.c_void_p),
                                      lres.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p), C.byref(s),
                                      C.byref(depth))
    if rc != 0:
        raise RuntimeError("vmo_region_features failed")
    return x, s.value, depth.value


def true_error(f, lo, hi, v, Q):
    lib = oracle()
    lib.vmo_true_error.restype = C.c_double
    a = [np.ascontiguousarray(z, dtype=np.float64) for z in (lo, hi, v)]
    return lib.vmo_true_error(f.ptr(), *[_ptr(z) for z in a], C.c_int(Q))


def eval_f64(f, x):
    lib = oracle()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = C.c_double()
    if lib.vmo_eval_f64(f.ptr(), _ptr(x), C.byref(y)) != 0:
        raise RuntimeError("vmo_eval_f64 failed")
    return y.value


def bin_variances(f, res, lo, hi, vp, dim, v0=None, v1=None, glo=None, ghi=None, maxb=4):
    """(V of the region, V of its children or None) on the bin grid res over [glo, ghi]."""
    D = f.dim
    db = len(res)
    glo = np.zeros(db) if glo is None else np.asarray(glo, dtype=np.float64)
    ghi = np.ones(db) if ghi is None else np.asarray(ghi, dtype=np.float64)
    step = np.ascontiguousarray((ghi - glo) / np.asarray(res, dtype=np.float64))
    glo = np.ascontiguousarray(glo)
    r = np.ascontiguousarray(np.asarray(res, dtype=np.uint64))
    a = [np.ascontiguousarray(z, dtype=np.float64) for z in (lo, hi, vp)]
    kids = v0 is not None
    k0 = np.ascontiguousarray(v0, dtype=np.float64) if kids else None
    k1 = np.ascontiguousarray(v1, dtype=np.float64) if kids else None
    Vp, Vc = C.c_double(), C.c_double()
    rc = oracle().vmo_bin_variances(f.ptr(), C.c_int(db), _ptr(glo), _ptr(step), _ptr(r, _up), *[_ptr(z) for z in a],
                                    C.c_int(dim), _ptr(k0), _ptr(k1), C.c_int(maxb), C.byref(Vp), C.byref(Vc))
    if rc != 0:
        raise RuntimeError("vmo_bin_variances failed")
    return Vp.value, (Vc.value if kids else None)

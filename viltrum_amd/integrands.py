"""Integrand descriptors: ctypes mirrors of include/viltrum_amd/integrands.h and the
seeded workload generators for the benchmark configurations.

The reference integrates arbitrary host callables (src/integrate.h:11-70); a GPU
kernel needs a closed description instead, so an integrand is a family id plus a
plain parameter block.  Parameters of the seeded workloads come from SplitMix64
(Steele, Lea, Flood 2014) with the draw order documented in each generator, so the
same seed yields the same integrand in every process.
"""
import ctypes as C
import math

MAX_DIM = 6
POLY_MAX_TERMS = 16
MAX_BLOBS = 64

FAMILY_POLY = 1
FAMILY_BLOBS = 2
FAMILY_LOBES5 = 3
FAMILY_FUBINI = 4
FAMILY_DYSON = 5
INFINITE_DIM = -1  # the dimension of an integrand of an unbounded sequence


class PolyParams(C.Structure):
    _fields_ = [
        ("nterms", C.c_int32),
        ("has_step", C.c_int32),
        ("coef", C.c_double * POLY_MAX_TERMS),
        ("expo", (C.c_uint8 * 8) * POLY_MAX_TERMS),
        ("step_n", C.c_double * MAX_DIM),
        ("step_d", C.c_double),
        ("step_h", C.c_double),
    ]


class BlobsParams(C.Structure):
    _fields_ = [
        ("nblobs", C.c_int32),
        ("step_mode", C.c_int32),
        ("blob", (C.c_double * 7) * MAX_BLOBS),
        ("step", C.c_double * 5),
        ("wave", C.c_double * 3),
    ]


class Lobes5Params(C.Structure):
    _fields_ = [
        ("centre", C.c_double * 3),
        ("inv2s2", C.c_double * 3),
        ("step_n", (C.c_double * 5) * 2),
        ("step_d", C.c_double * 2),
        ("k", C.c_double * 5),
        ("amp", C.c_double),
    ]


class FubiniParams(C.Structure):
    _fields_ = [
        ("inner_family", C.c_int32),
        ("inner_dim", C.c_int32),
        ("mc_samples", C.c_uint32),
        ("seed", C.c_uint32),
        ("rest_lo", C.c_double * MAX_DIM),
        ("rest_hi", C.c_double * MAX_DIM),
        ("inner", C.c_void_p),
        ("rest_count", C.c_int32),
    ]


class DysonParams(C.Structure):
    _fields_ = [("decay", C.c_double)]


class IntegrandDesc(C.Structure):
    _fields_ = [("family", C.c_int32), ("dim", C.c_int32), ("params", C.c_void_p)]


class Integrand:
    """A family id, its dimension and a parameter block kept alive with the descriptor."""

    def __init__(self, family, dim, params, name=""):
        self.family = family
        self.dim = dim
        self.params = params
        self.name = name
        self.desc = IntegrandDesc(family, dim, C.cast(C.pointer(params), C.c_void_p))

    def ptr(self):
        return C.byref(self.desc)

    def __repr__(self):
        return "Integrand(%s, D=%d)" % (self.name or self.family, self.dim)


def poly(terms, dim, step=None, name="poly"):
    """terms: list of (coef, (e0, e1, ...)); step: (normal[dim], d, h) adds h*[n.x > d]."""
    if len(terms) > POLY_MAX_TERMS:
        raise ValueError("at most %d terms" % POLY_MAX_TERMS)
    p = PolyParams()
    p.nterms = len(terms)
    for m, (c, e) in enumerate(terms):
        p.coef[m] = c
        for d, ed in enumerate(e):
            p.expo[m][d] = ed
    if step is not None:
        n, d, h = step
        p.has_step = 1
        for i, v in enumerate(n):
            p.step_n[i] = v
        p.step_d = d
        p.step_h = h
    return Integrand(FAMILY_POLY, dim, p, name)


def blobs(blob_list, dim=2, step_mode=0, step=(0, 0, 0, 0, 0), wave=(0, 0, 0), name="blobs"):
    """blob_list: (cx, cy, vx, vy, ax, ay, w) per blob; step = (nx, ny, nt, d, h)."""
    if len(blob_list) > MAX_BLOBS:
        raise ValueError("at most %d blobs" % MAX_BLOBS)
    if dim not in (2, 3):
        raise ValueError("blob family is 2-D or 3-D")
    p = BlobsParams()
    p.nblobs = len(blob_list)
    p.step_mode = step_mode
    for k, b in enumerate(blob_list):
        for j in range(7):
            p.blob[k][j] = b[j]
    for j in range(5):
        p.step[j] = step[j]
    for j in range(3):
        p.wave[j] = wave[j]
    return Integrand(FAMILY_BLOBS, dim, p, name)


def lobes5(centre, inv2s2, step_n, step_d, k, amp, name="lobes5"):
    p = Lobes5Params()
    for j in range(3):
        p.centre[j] = centre[j]
        p.inv2s2[j] = inv2s2[j]
    for s in range(2):
        for j in range(5):
            p.step_n[s][j] = step_n[s][j]
        p.step_d[s] = step_d[s]
    for j in range(5):
        p.k[j] = k[j]
    p.amp = amp
    return Integrand(FAMILY_LOBES5, 5, p, name)


def fubini(inner, rd, mc_samples, seed, rest_lo, rest_hi, name=None):
    """The Fubini form of integrator_crespo2021_infinite<rd>: regions and bins over the
    first rd dims of `inner`; its other dims (rest_lo/hi, indexed by inner dim) are
    integrated by Monte Carlo, mc_samples per node value and one per residual sample."""
    if not 1 <= rd < inner.dim:
        raise ValueError("region dims must be 1..%d" % (inner.dim - 1))
    p = FubiniParams()
    p.inner_family = inner.family
    p.inner_dim = inner.dim
    p.mc_samples = int(mc_samples)
    p.seed = int(seed) & 0xFFFFFFFF
    for d in range(inner.dim):
        p.rest_lo[d] = rest_lo[d]
        p.rest_hi[d] = rest_hi[d]
    p.inner = C.cast(C.pointer(inner.params), C.c_void_p)
    f = Integrand(FAMILY_FUBINI, rd, p, name or "fubini(%s)" % inner.name)
    f.inner = inner  # keeps the inner parameter block alive
    return f


def dyson(decay=0.75, name="dyson"):
    """The Dyson series of a sequence (the reference's infinite-range example integrand,
    main/compilation-tests/crespo21.cc:6-25): sum = 1, range = 1; while the current entry
    is below `decay`: step, x = range * entry, sum += x * range, range = x; step.  Its
    integral over the primary infinite range is 1 + 3 decay / (2 (3 - decay))."""
    if not 0.0 < decay < 1.0:
        raise ValueError("decay must be in (0, 1)")
    return Integrand(FAMILY_DYSON, INFINITE_DIM, DysonParams(float(decay)), name)


def fubini_infinite(inner, rd, mc_samples, seed, lo, hi, rest_lo=(), rest_hi=(), name=None):
    """The Fubini form over an infinite range: regions and bins over [lo, hi] (rd dims, the
    range's first entries), the sequence after them Monte Carlo, entry j in
    [rest_lo[j], rest_hi[j]] for the explicit entries and in [0, 1] after them."""
    if inner.family != FAMILY_DYSON:
        raise ValueError("an integrand of a sequence (the Dyson series) is needed")
    if not 1 <= rd <= 4:
        raise ValueError("region dims must be 1..4")
    if len(rest_lo) != len(rest_hi) or rd + len(rest_lo) > MAX_DIM:
        raise ValueError("at most %d explicit rest entries" % (MAX_DIM - rd))
    p = FubiniParams()
    p.inner_family = inner.family
    p.inner_dim = INFINITE_DIM
    p.mc_samples = int(mc_samples)
    p.seed = int(seed) & 0xFFFFFFFF
    for d in range(rd):
        p.rest_lo[d] = lo[d]
        p.rest_hi[d] = hi[d]
    for j in range(len(rest_lo)):
        p.rest_lo[rd + j] = rest_lo[j]
        p.rest_hi[rd + j] = rest_hi[j]
    p.rest_count = len(rest_lo)
    p.inner = C.cast(C.pointer(inner.params), C.c_void_p)
    f = Integrand(FAMILY_FUBINI, rd, p, name or "fubini_infinite(%s)" % inner.name)
    f.inner = inner
    return f


# ---------------------------------------------------------------- seeded generators

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & _M64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def log_uniform(self, a, b):
        return math.exp(self.uniform(math.log(a), math.log(b)))

    def integer(self, a, b):
        """Uniform integer in [a, b]."""
        return a + int(self.uniform() * (b - a + 1)) if b > a else a


def config1():
    """exp(-((x-.3)^2+(y-.6)^2)/.02) + 0.5*[x+y>1] + 0.1*sin(20x)*cos(13y) on [0,1]^2."""
    return blobs([(0.3, 0.6, 0.0, 0.0, 50.0, 50.0, 1.0)], dim=2, step_mode=1,
                 step=(1.0, 1.0, 0.0, 1.0, 0.5), wave=(0.1, 20.0, 13.0), name="config1")


def _halfspace(rng, dim):
    """Random normal in [-1,1]^dim through a random point of [.25,.75]^dim.
    Draw order: normal components, then the point components."""
    n = [rng.uniform(-1.0, 1.0) for _ in range(dim)]
    q = [rng.uniform(0.25, 0.75) for _ in range(dim)]
    return n, sum(a * b for a, b in zip(n, q))


def moving_blobs(seed, nblobs, dim=3, name=None):
    """Config-2 family: nblobs anisotropic Gaussians, centres U[0,1]^2, sigmas
    logU[.01,.1], weights U[.2,1], velocities U[-.2,.2]^2 (t = x2), times the occluder
    [a x + b y + c t > d].  Draw order per blob: cx, cy, sx, sy, w, vx, vy; then the
    occluder (normal a, b, c then its anchor point)."""
    rng = SplitMix64(seed)
    bl = []
    for _ in range(nblobs):
        cx, cy = rng.uniform(), rng.uniform()
        sx, sy = rng.log_uniform(0.01, 0.1), rng.log_uniform(0.01, 0.1)
        w = rng.uniform(0.2, 1.0)
        vx, vy = rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2)
        if dim == 2:
            vx = vy = 0.0
        bl.append((cx, cy, vx, vy, 1.0 / (2 * sx * sx), 1.0 / (2 * sy * sy), w))
    n, d = _halfspace(rng, 3)
    if dim == 2:
        n[2] = 0.0
    return blobs(bl, dim=dim, step_mode=2, step=(n[0], n[1], n[2], d, 0.0),
                 name=name or ("blobs%dd_s%d_k%d" % (dim, seed, nblobs)))


def config2(seed=0):
    """3-D, 32 moving anisotropic blobs times a seeded step occluder."""
    return moving_blobs(seed, 32, dim=3, name="config2")


def c1_family(seed, nblobs):
    """Config-1 style 2-D integrand with nblobs static isotropic/anisotropic blobs, an
    additive step and a wave.  Draw order per blob: cx, cy, sx, sy, w; then the
    step (normal, anchor, height) and the wave (A, fx, fy)."""
    rng = SplitMix64(seed)
    bl = []
    for _ in range(nblobs):
        cx, cy = rng.uniform(), rng.uniform()
        sx, sy = rng.log_uniform(0.03, 0.2), rng.log_uniform(0.03, 0.2)
        w = rng.uniform(0.2, 1.0)
        bl.append((cx, cy, 0.0, 0.0, 1.0 / (2 * sx * sx), 1.0 / (2 * sy * sy), w))
    n, d = _halfspace(rng, 2)
    h = rng.uniform(0.1, 0.6)
    wave = (rng.uniform(0.0, 0.2), rng.uniform(5.0, 25.0), rng.uniform(5.0, 25.0))
    return blobs(bl, dim=2, step_mode=1, step=(n[0], n[1], 0.0, d, h), wave=wave,
                 name="c1fam_s%d_k%d" % (seed, nblobs))


def config4_integrand(seed):
    """Config-1 style integrand of the learned-refinement workload: 1..8 blobs (count drawn
    first from SplitMix64(seed)), then c1_family with the next draw as its seed."""
    rng = SplitMix64(seed)
    nb = rng.integer(1, 8)
    return c1_family(rng.next() & 0xFFFFFFFF, nb)


def config5(seed):
    """5-D: 3 Gaussian lobes over dims 2..4 (sigma U[.05,.3], centre U[.2,.8]), two
    half-space steps over all dims, 1 + 0.3 cos(2 pi k.x), k integer in [1,8]^5.
    Draw order: centres, sigmas, step 0, step 1, k."""
    rng = SplitMix64(seed)
    c = [rng.uniform(0.2, 0.8) for _ in range(3)]
    s = [rng.uniform(0.05, 0.3) for _ in range(3)]
    steps = [_halfspace(rng, 5) for _ in range(2)]
    k = [float(rng.integer(1, 8)) for _ in range(5)]
    return lobes5(c, [1.0 / (2 * v * v) for v in s], [st[0] for st in steps], [st[1] for st in steps], k, 0.3,
                  name="config5_s%d" % seed)


def config3_batch(seed=0, count=256):
    """256 integrals: res side U{16,32,64,128,256}, budget logU[256,32768], spp
    U{4,16,64}, D U{2,3}; 2-D ones are config-1 style, 3-D ones moving blobs, 4..64
    blobs each.  Returns list of dicts (integrand, res, regions, spp)."""
    rng = SplitMix64(seed)
    out = []
    for i in range(count):
        side = [16, 32, 64, 128, 256][rng.integer(0, 4)]
        regions = int(round(rng.log_uniform(256, 32768)))
        spp = [4, 16, 64][rng.integer(0, 2)]
        dim = 2 + rng.integer(0, 1)
        nb = rng.integer(4, 64)
        sub = rng.next() & 0xFFFFFFFF
        f = c1_family(sub, nb) if dim == 2 else moving_blobs(sub, nb, dim=3)
        out.append(dict(integrand=f, res=(side, side), regions=regions, spp=spp))
    return out


def flops_per_eval(f):
    """Static FP32 operation count of one integrand evaluation, read off the source in
    include/viltrum_amd/integrands.h (add/sub/mul/neg/compare = 1, a*b + c = 2 whether
    or not the FP32 path fuses it).  Returns (flops, special) where `special`
    counts exp/sin/cos calls, never converted to flops."""
    if f.family == FAMILY_FUBINI:  # a residual sample evaluates the inner integrand once
        return flops_per_eval(f.inner)
    if f.family == FAMILY_DYSON:  # per term: compare, multiply, divide, multiply-add
        return 5, 0
    if f.family == FAMILY_BLOBS:
        p = f.params
        flops = 14 * p.nblobs + 1  # per blob: 6 (centre, dx, dy) + 8 (quadratic form, exp arg, w*, s+)
        special = p.nblobs
        if p.wave[0] != 0:
            flops += 5
            special += 2
        if p.step_mode != 0:
            flops += 6 + (1 if p.step_mode == 1 else 0)
        return flops, special
    if f.family == FAMILY_LOBES5:
        return 49, 2
    if f.family == FAMILY_POLY:
        p = f.params
        flops = sum(1 + sum(p.expo[m][d] for d in range(f.dim)) for m in range(p.nterms))
        if p.has_step:
            flops += 2 * f.dim + 2
        return flops, 0
    raise ValueError("unknown family")

"""The oracle (CPU restatement of the reference algorithm) against the reference's own
outputs: golden fixtures written by the reference build (tests/golden/make_golden.py)
and, where it is built here, the reference build itself.  Plus closed forms that pin
what no reference output covers."""
import os

import numpy as np
import pytest

import oracle_lib as O
from viltrum_amd import integrands as I

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "reference_fixtures.npz"))


def cases():
    import importlib.util
    spec = importlib.util.spec_from_file_location("mg", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg


MG = cases()
FS = MG.integrand_cases()


@pytest.mark.parametrize("case", MG.REGION_CASES, ids=lambda c: "%s-s%d-h%d-it%d" % c)
def test_generate_matches_reference_fixture(case):
    name, scalar, heur, it = case
    g = O.generate(FS[name], it, scalar=scalar, heuristic=heur, evals=True)
    key = "regions/%s/s%d/h%d/it%d" % (name, scalar, heur, it)
    # heap-array order, boxes, (err, dim) and every integrand evaluation in call order
    assert np.array_equal(g.lo, FIX[key + "/lo"]) and np.array_equal(g.hi, FIX[key + "/hi"])
    assert np.array_equal(g.err, FIX[key + "/err"]) and np.array_equal(g.dim, FIX[key + "/dim"])
    assert np.array_equal(g.evals, FIX[key + "/evals"])


@pytest.mark.parametrize("case", MG.BIN_CASES, ids=lambda c: "%s-s%d-f%d-r%s-spp%d" % (c[0], c[1], c[2], c[4], c[6]))
def test_integrate_matches_reference_fixture(case):
    name, scalar, form, heur, res, it, spp, seed, alpha = case
    b, _ = O.integrate(FS[name], res, it, spp, seed, scalar=scalar, integrator=form, heuristic=heur, alpha=alpha,
                       nthreads=3)
    key = "bins/%s/s%d/f%d/h%d/r%s/it%d/spp%d/seed%d/a%g" % (name, scalar, form, heur, "x".join(map(str, res)), it,
                                                            spp, seed, alpha)
    assert np.array_equal(b, FIX[key], equal_nan=True)


@pytest.mark.parametrize("case", MG.FUBINI_CASES, ids=lambda c: "%s-rd%d-mc%d-spp%d" % (c[0], c[1], c[4], c[5]))
def test_fubini_matches_reference_fixture(case):
    """integrator_crespo2021_infinite: node values Monte Carlo over the dims after the
    regions', every engine seeded as the reference's Monte Carlo objects are copied."""
    name, rd, res, it, mc, spp, seed = case
    b, st = O.integrate_fubini(FS[name], rd, res, it, mc, spp, seed, nthreads=2)
    assert st.regions == it + 1
    assert np.array_equal(b, FIX[MG.fubini_key(name, rd, res, it, mc, spp, seed)])


@pytest.mark.skipif(not os.path.exists(O.REF_SO), reason="reference build absent (no /root/reference where this runs)")
@pytest.mark.parametrize("seed", range(3))
def test_fubini_oracle_matches_live_reference(seed):
    rng = np.random.default_rng(40 + seed)
    f = I.moving_blobs(50 + seed, int(rng.integers(1, 9)))
    res = (int(rng.integers(3, 17)), int(rng.integers(3, 17)))
    it, mc, spp = int(rng.integers(0, 250)), int(rng.integers(1, 5)), int(rng.integers(0, 9))
    a, _ = O.integrate_fubini(f, 2, res, it, mc, spp, seed)
    b, _ = O.integrate_fubini(f, 2, res, it, mc, spp, seed, use_ref=True)
    assert np.array_equal(a, b)


def test_fubini_with_a_tensor_quadratic_in_the_region_dims_is_exact():
    """f(x, y, t) = x^2 y + y + 3 t with regions over (x, y): each node value is
    x^2 y + y plus a Monte Carlo estimate of the integral of 3 t (1.5), so with spp = 0
    the mean bin is 1/6 + 1/2 + 1.5 up to the node values' Monte Carlo error."""
    f = I.poly([(1.0, (2, 1, 0)), (1.0, (0, 1, 0)), (3.0, (0, 0, 1))], 3)
    b, _ = O.integrate_fubini(f, 2, (4, 4), 30, 4000, 0, 3)
    # sum over bins / B = integral over [0,1]^2 of x^2 y + y + E[3 t] = 1/6 + 1/2 + 1.5
    assert abs(b.mean() - (1.0 / 6 + 0.5 + 1.5)) < 5e-3


@pytest.mark.parametrize("case", MG.FUBINI_INF_CASES,
                         ids=lambda c: "f%d-rd%d-r%s-spp%d" % (c[0], c[2], "x".join(map(str, c[3])), c[6]))
def test_fubini_infinite_matches_reference_fixture(case):
    """The Dyson series over an infinite range: integrator_crespo2021_infinite (form 0) and
    the optimised Fubini integrator with region_stratification_uniform (forms 1, 2)."""
    b, st = MG.run_fubini_inf(case, nthreads=2)
    assert st.regions == case[4] + 1
    assert np.array_equal(b, FIX[MG.fubini_inf_key(*case)])


@pytest.mark.skipif(not os.path.exists(O.REF_SO), reason="reference build absent (no /root/reference where this runs)")
@pytest.mark.parametrize("seed", range(8))
def test_fubini_infinite_oracle_matches_live_reference(seed):
    rng = np.random.default_rng(70 + seed)
    form = seed % 3
    rd, res = [(1, (7,)), (2, (5, 3)), (4, (4, 2)), (3, (3, 2, 2))][seed % 4]
    it, mc, spp = int(rng.integers(5, 300)), int(rng.integers(1, 6)), int(rng.integers(1, 30))
    nr = int(rng.integers(0, 3))
    rl = rng.uniform(0, 0.4, nr)
    rh = rl + rng.uniform(0.2, 0.6, nr)
    lo = rng.uniform(-1, 0, rd)
    hi = lo + rng.uniform(0.5, 2, rd)
    kw = dict(form=form, alpha=0.6, lo=lo, hi=hi, rest_lo=tuple(rl), rest_hi=tuple(rh))
    decay = float(rng.uniform(0.3, 0.9))
    a, _ = O.integrate_fubini_inf(decay, rd, res, it, mc, spp, 100 + seed, nthreads=3, **kw)
    b, _ = O.integrate_fubini_inf(decay, rd, res, it, mc, spp, 100 + seed, use_ref=True, **kw)
    assert np.array_equal(a, b)


def dyson_bins_closed_form(decay, rd, res):
    """Per-bin value (bin count times the integral over the bin) of the Dyson series over
    the primary infinite range with regions over the first rd entries, bins over the
    first len(res) of them.  The series reads entry 0 (continue while below d), then
    entry 1 (its first term), then alternately; the region coordinates are its first
    entries and the rest are uniform on [0, 1].  Term k after the first multiplies the
    previous range squared by a fresh uniform and survives with probability d^k, so with
    G = (1/2) / (1 - d/3):
      rd 1: E[f | x0] = 1 + [x0 < d] G
      rd 2: E[f | x0, x1] = 1 + [x0 < d] (x1 + d G x1^2)."""
    d = decay
    G = 0.5 / (1.0 - d / 3.0)
    if rd == 1:
        (n,) = res
        e = np.linspace(0, 1, n + 1)
        below = np.clip(np.minimum(e[1:], d) - e[:-1], 0, None) * n  # share of the bin below d
        return 1.0 + below * G
    n0, n1 = res
    e0, e1 = np.linspace(0, 1, n0 + 1), np.linspace(0, 1, n1 + 1)
    below = np.clip(np.minimum(e0[1:], d) - e0[:-1], 0, None) * n0
    m1 = (e1[1:] ** 2 - e1[:-1] ** 2) / 2 * n1  # mean of x1 over the bin
    m2 = (e1[1:] ** 3 - e1[:-1] ** 3) / 3 * n1  # mean of x1^2
    return 1.0 + below[:, None] * (m1 + d * G * m2)[None, :]


@pytest.mark.parametrize("form,rd,res", [(0, 1, (8,)), (1, 1, (8,)), (0, 2, (4, 4)), (2, 2, (4, 4))])
def test_fubini_infinite_oracle_is_unbiased_against_the_closed_form(form, rd, res):
    """Mean over seeds of the oracle's bins against the closed form of the Dyson series:
    pins the sequence order (region coordinates first), the node values' Monte Carlo and
    both residual estimators (a dropped term, a swapped entry or a wrong weight moves
    the mean by far more than the tolerance)."""
    d = 0.75
    truth = dyson_bins_closed_form(d, rd, res)
    runs = np.array([O.integrate_fubini_inf(d, rd, res, 60, 2, 16, s, form=form, alpha=0.5, nthreads=4)[0]
                     for s in range(64)])
    se = runs.std(0, ddof=1) / np.sqrt(runs.shape[0])
    flat = se == 0  # bins past the decay: the series is 1 there, every estimate exact
    np.testing.assert_allclose(runs.mean(0)[flat], truth[flat], rtol=1e-6)
    z = (runs.mean(0) - truth)[~flat] / se[~flat]
    assert np.max(np.abs(z)) < 4.5 and abs(z.mean()) < 4.0 / np.sqrt(z.size) + 0.1, z
    # the total: 1 + 3 d / (2 (3 - d)) = 1.5 at d = 0.75
    assert abs(runs.mean() - 1.5) < 5 * runs.mean(axis=tuple(range(1, runs.ndim))).std() / np.sqrt(64)


def test_stratified_estimator_is_biased_with_fewer_samples_than_regions():
    """region_stratification_uniform with fewer samples than a bin's regions: a region
    that draws no sample adds nothing to the residual (the weights spp / count cover the
    sampled regions only), so the estimate is biased; with enough samples every region
    is sampled and the bins are unbiased again.  This is the reference's estimator (the
    fixtures above are its output); the GPU keeps it."""
    d, res = 0.75, (2, 2)
    truth = dyson_bins_closed_form(d, 2, res)
    few = np.array([O.integrate_fubini_inf(d, 2, res, 300, 2, 16, 700 + s, form=1, nthreads=4)[0] for s in range(64)])
    many = np.array([O.integrate_fubini_inf(d, 2, res, 300, 2, 400, 700 + s, form=1, nthreads=4)[0] for s in range(64)])
    z_few = (few.mean(0) - truth) / (few.std(0, ddof=1) / 8)
    z_many = (many.mean(0) - truth) / (many.std(0, ddof=1) / 8)
    assert z_few.min() < -4.0, z_few
    assert np.max(np.abs(z_many)) < 4.5, z_many


@pytest.mark.skipif(not os.path.exists(O.REF_SO), reason="reference build absent (no /root/reference where this runs)")
@pytest.mark.parametrize("seed", range(4))
def test_oracle_matches_live_reference(seed):
    rng = np.random.default_rng(seed)
    f = I.c1_family(100 + seed, int(rng.integers(1, 8)))
    res = (int(rng.integers(3, 20)), int(rng.integers(3, 20)))
    it, spp = int(rng.integers(0, 300)), int(rng.integers(0, 9))
    for scalar in (4, 8):
        for form in range(19):
            heur = 1 if form == 2 else 0
            if (form in (3, 4) or form >= 11) and scalar == 8:
                continue  # the importance and roulette samplers: the reference compiles them for float only
            a, _ = O.integrate(f, res, it, spp, seed, scalar=scalar, integrator=form, heuristic=heur, alpha=0.3)
            b, _ = O.integrate(f, res, it, spp, seed, scalar=scalar, integrator=form, heuristic=heur, alpha=0.3,
                               use_ref=True)
            assert np.array_equal(a, b)


def test_known_answer_x2_plus_y2():
    """main/doc/integrators.cc:96-97: integrator_crespo2021(16, 64) on x^2+y^2 = 2/3."""
    f = I.poly([(1.0, (2, 0)), (1.0, (0, 2))], 2)
    for scalar in (8, 4):
        b, _ = O.integrate(f, (1,), 16, 64, 5, scalar=scalar)
        assert abs(b[0] - 2.0 / 3.0) < (1e-15 if scalar == 8 else 1e-7)


def test_simpson_exact_on_cubics_single_region():
    """One region over [0,1]: the bin is the Simpson integral, exact for cubics."""
    f = I.poly([(1.0, (3,)), (-2.0, (2,)), (0.5, (1,)), (3.0, (0,))], 1)
    b, _ = O.integrate(f, (1,), 0, 0, 0, scalar=8)
    assert abs(b[0] - (0.25 - 2.0 / 3.0 + 0.25 + 3.0)) < 1e-15


def test_tensor_quadratic_bins_exact_for_any_tree():
    """A tensor quadratic is reproduced exactly by every region's interpolant, so the
    spp = 0 bins equal the exact per-bin integrals (no sliver drops: region and bin
    edges are dyadic and coincide)."""
    f = I.poly([(1.0, (2, 2)), (0.5, (1, 1)), (-1.0, (2, 0))], 2)
    b, _ = O.integrate(f, (8, 8), 12, 0, 0, scalar=8)
    e = np.linspace(0, 1, 9)
    F2 = lambda a, c: (c ** 3 - a ** 3) / 3.0
    F1 = lambda a, c: (c ** 2 - a ** 2) / 2.0
    exact = np.array([[F2(e[i], e[i + 1]) * F2(e[j], e[j + 1]) + 0.5 * F1(e[i], e[i + 1]) * F1(e[j], e[j + 1])
                       - F2(e[i], e[i + 1]) * (e[j + 1] - e[j]) for j in range(8)] for i in range(8)]) * 64
    np.testing.assert_allclose(b, exact, rtol=1e-12, atol=1e-14)


def test_cv_bins_partition_sum_equals_region_integrals():
    """Sum over bins of the spp=0 bins = B * sum of region integrals when no region
    edge splits a bin (coarse tree on a power-of-two grid)."""
    f = I.config1()
    res = (16, 16)
    it = 15
    b, _ = O.integrate(f, res, it, 0, 0, scalar=8)
    g = O.generate(f, it, scalar=8)
    # region integral = vol * Simpson fold of the nodes = full subrange
    one, _ = O.integrate_regions(f, g.lo, g.hi, g.nodes, (1,), 0, 0, scalar=8)
    assert abs(b.sum() / 256 - one[0]) < 1e-12


def test_blobs2d_closed_form_against_quadrature():
    """The FP64 config-1 oracle (erf / polygon / separable trig) against brute-force
    midpoint quadrature on rectangles."""
    f = I.config1()
    rects = [(0.0, 0.125, 0.5, 0.625), (0.4, 0.7, 0.2, 0.65), (0.9, 1.0, 0.0, 0.3), (0.0, 1.0, 0.0, 1.0)]
    x = np.linspace(0, 1, 8 * 8 + 1)
    for x0, x1, y0, y1 in rects:
        n = 1500
        xs = x0 + (np.arange(n) + 0.5) * (x1 - x0) / n
        ys = y0 + (np.arange(n) + 0.5) * (y1 - y0) / n
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        F = np.exp(-((X - 0.3) ** 2 + (Y - 0.6) ** 2) * 50.0) + 0.5 * (X + Y > 1.0) + 0.1 * np.sin(20 * X) * np.cos(13 * Y)
        quad = F.mean() * (x1 - x0) * (y1 - y0)
        exact = O.oracle().vmo_blobs2d_exact_rect
        import ctypes as C
        exact.restype = C.c_double
        v = exact(C.byref(f.params), C.c_double(x0), C.c_double(x1), C.c_double(y0), C.c_double(y1))
        assert abs(v - quad) < 2e-5 * (x1 - x0) * (y1 - y0) * 10 + 1e-6


def test_config1_cv_converges_to_closed_form():
    """FP64 control variate with a large budget against the closed-form bins: the
    residual of the method (and its dropped slivers) is small."""
    f = I.config1()
    b, _ = O.integrate(f, (32, 32), 8191, 0, 0, scalar=8)
    ex = O.blobs2d_exact_bins(f, 32, 32)
    assert np.sqrt(np.mean((b - ex) ** 2)) < 5e-3
    assert abs(b.mean() - ex.mean()) < 2e-3


def test_mt19937_10000th_output():
    """The baseline RNG algorithm: the C++ standard fixes the 10000th output of a
    default-seeded (5489) mt19937 at 4123659995 (numpy's MT19937 is the same
    generator).  libstdc++'s engine and distributions, as the reference uses them, are
    pinned by the spp > 0 fixtures above."""
    rs = np.random.RandomState(5489)
    v = rs.randint(0, 2 ** 32, size=10000, dtype=np.uint64)
    assert int(v[-1]) == 4123659995

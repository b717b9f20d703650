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

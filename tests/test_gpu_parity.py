"""GPU parity against the oracle (CPU restatement of the reference, itself pinned
bit-exactly against the reference build).  Runs through the C ABI (libviltrum_amd.so)."""
import numpy as np
import pytest

import oracle_lib as O
import viltrum_amd as V
from viltrum_amd import integrands as I

pytestmark = pytest.mark.gpu


def poly_cases():
    return [
        I.poly([(1.0, (2, 0)), (1.0, (0, 2))], 2, name="x2y2"),
        I.poly([(1.0, (3, 1)), (-2.0, (1, 4)), (0.5, (0, 0))], 2, step=([1.0, 1.0], 1.0, 0.5), name="poly2step"),
        I.poly([(1.0, (2, 1, 3)), (0.25, (0, 5, 1))], 3, step=([1.0, -0.5, 0.7], 0.3, 1.0), name="poly3step"),
        I.poly([(1.0, (1, 2, 0, 3, 1)), (2.0, (4, 0, 1, 0, 0))], 5, step=([1, 1, 1, 1, 1], 2.5, 0.25), name="poly5"),
    ]


def _x(D, n, seed):
    return np.random.default_rng(seed).random((n, D))


@pytest.mark.parametrize("f", poly_cases() + [I.config1(), I.config2(), I.config5(3)], ids=lambda f: f.name)
def test_eval_matches_host(ctx, f):
    x = _x(f.dim, 2048, 1)
    g64 = ctx.eval(f, x, V.F64)
    o64 = O.evaluate(f, x, scalar=8)
    if f.family == I.FAMILY_POLY:
        assert np.array_equal(g64, o64)
    else:
        # ocml vs glibc: an ulp or so on each term of an O(1) sum
        np.testing.assert_allclose(g64, o64, rtol=4e-15, atol=1e-15)
    g32 = ctx.eval(f, x, V.F32)
    o32 = O.evaluate(f, x, scalar=4)
    if f.family == I.FAMILY_POLY:
        assert np.array_equal(g32, o32)
    else:
        np.testing.assert_allclose(g32, o32, rtol=2e-5, atol=2e-6)


def _gpu_regions(ctx, f, iterations, scalar, order=V.ORDER_REFERENCE, heuristic=V.HEUR_SIZE_RELATIVE):
    integ = V.Integrator(iterations, 0, 0, heuristic=heuristic, order=order)
    bins = np.zeros((4, 4) if f.dim >= 2 else (4,), dtype=np.float64 if scalar == 8 else np.float32)
    V.integrate(integ, bins, f, V.range_primary(f.dim), ctx=ctx)
    return ctx.export_regions(f.dim), integ.last_stats


@pytest.mark.parametrize("f", poly_cases(), ids=lambda f: f.name)
@pytest.mark.parametrize("heuristic", [V.HEUR_SIZE_RELATIVE, V.HEUR_DEFAULT_ABSOLUTE])
@pytest.mark.parametrize("iterations", [0, 1, 5, 63, 500])
def test_regions_fp64_reference_order_bitwise(ctx, f, heuristic, iterations):
    g, st = _gpu_regions(ctx, f, iterations, 8, heuristic=heuristic)
    o = O.generate(f, iterations, scalar=8, heuristic=heuristic)
    assert len(g["err"]) == iterations + 1
    assert np.array_equal(g["lo"], o.lo) and np.array_equal(g["hi"], o.hi)
    assert np.array_equal(g["err"], o.err) and np.array_equal(g["dim"], o.dim)
    assert np.array_equal(g["nodes"], o.nodes)


@pytest.mark.parametrize("f", [I.config1(), I.moving_blobs(5, 8), I.c1_family(7, 6)], ids=lambda f: f.name)
def test_regions_fp64_transcendental(ctx, f):
    """Transcendental integrands: GPU libm and glibc may differ by an ulp, so compare
    the region *set* (boxes) and errors to 1e-13."""
    g, st = _gpu_regions(ctx, f, 300, 8)
    o = O.generate(f, 300, scalar=8)
    ko = sorted(map(tuple, np.hstack([o.lo, o.hi])))
    kg = sorted(map(tuple, np.hstack([g["lo"], g["hi"]])))
    assert ko == kg
    np.testing.assert_allclose(np.sort(g["err"]), np.sort(o.err), rtol=1e-12)


@pytest.mark.parametrize("f,res", [(poly_cases()[1], (16, 16)), (poly_cases()[2], (8, 8)), (poly_cases()[3], (6, 5)),
                                   (I.config1(), (32, 32)), (I.moving_blobs(5, 8), (16, 16))],
                         ids=lambda v: getattr(v, "name", str(v)))
@pytest.mark.parametrize("form", [V.FORM_VR_OPTIMIZE, V.FORM_CV_ALPHA1])
def test_cv_bins_fp64_spp0(ctx, f, res, form):
    """spp = 0: bins are the pure control-variate integral; reference order makes the
    FP64 per-bin sums follow the reference's order."""
    it = 255
    heur = V.HEUR_SIZE_RELATIVE if form == V.FORM_VR_OPTIMIZE else V.HEUR_DEFAULT_ABSOLUTE
    integ = V.Integrator(it, 0, 0, heuristic=heur, form=form, order=V.ORDER_REFERENCE)
    bins = np.zeros(res, dtype=np.float64)
    V.integrate(integ, bins, f, V.range_primary(f.dim), ctx=ctx)
    ob, _ = O.integrate(f, res, it, 0, 0, scalar=8, integrator=form, heuristic=heur)
    if f.family == I.FAMILY_POLY:
        np.testing.assert_array_equal(bins, ob)
    else:
        np.testing.assert_allclose(bins, ob, rtol=1e-12, atol=1e-12)


def test_known_answer_x2_plus_y2(ctx):
    """main/doc/integrators.cc:96-97: integrator_crespo2021(16, 64) on x^2 + y^2 over
    [0,1]^2 is 2/3 exactly (quadratic => zero residual)."""
    f = poly_cases()[0]
    for dt in (np.float64, np.float32):
        v = V.integrate(V.integrator_crespo2021(16, 64, seed=3), f, V.range_primary(2), dt, ctx=ctx)
        assert abs(v - 2.0 / 3.0) < (1e-15 if dt == np.float64 else 1e-6)

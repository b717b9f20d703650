"""rr_integral_region, rr_error_region and rr_pdf_region
(src/control-variates/region-russian-roulette.h:30-149): residual samples pick a bin's
region with probability proportional to |integral of its polynomial over the bin's box|,
to |Region::error()| times the region's share inside the bin, or to the integral of the
region's importance pdf over the bin's box (floored at 1% of the mean), weight 1/p; also
each of them with region_sampling_importance drawing the point.  The GPU's
Philox stream and the oracle's mt19937 + std::discrete_distribution (the oracle is
pinned bit for bit to the reference build) run the same estimator: per-bin means over
independent seeds agree within their standard errors, with the same spread."""
import numpy as np
import pytest

import oracle_lib as O
import viltrum_amd as V
from viltrum_amd import integrands as I

from test_gpu_stats import SEEDS, _check
from test_gpu_parity import poly_cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("f,res,it,spp", [
    (I.config1(), (16, 16), 255, 8),
    (I.moving_blobs(4, 6), (8, 8), 200, 8),
    (poly_cases()[2], (6, 6), 150, 8),
    (poly_cases()[3], (4, 4), 400, 8),
], ids=lambda v: getattr(v, "name", str(v)))
@pytest.mark.parametrize("form", [V.FORM_VR_FIXED_RR_INTEGRAL, V.FORM_VR_OPTIMIZE_RR_INTEGRAL, V.FORM_VR_FIXED_RR_ERROR,
                                  V.FORM_VR_OPTIMIZE_RR_ERROR, V.FORM_VR_FIXED_RR_PDF, V.FORM_VR_OPTIMIZE_RR_PDF],
                         ids=["int_fixed1", "int_optimize", "err_fixed1", "err_optimize", "pdf_fixed1", "pdf_optimize"])
@pytest.mark.parametrize("scalar", [4, 8])
def test_rr_integral_matches_oracle(ctx, f, res, it, spp, form, scalar):
    dt = np.float32 if scalar == 4 else np.float64
    g, o = [], []
    for s in range(SEEDS):
        b = np.zeros(res, dtype=dt)
        V.integrate(V.Integrator(it, spp, 4000 + s, form=form, alpha=1.0), b, f, V.range_primary(f.dim), ctx=ctx)
        g.append(b.astype(np.float64))
        ob, _ = O.integrate(f, res, it, spp, 4000 + s, scalar=scalar, integrator=form, alpha=1.0)
        o.append(ob)
    # the roulette's 1/p reaches 100 n for a region at the 1% weight floor: per-bin
    # variances over 48 seeds are heavy-tailed, so the spread check is wider here
    _check(np.array(g), np.array(o), rel_floor=1e-6, spread=(0.6, 1.6))


def test_rr_integral_is_unbiased(ctx):
    f = I.config1()
    res = (16, 16)
    truth = O.blobs2d_exact_bins(f, *res)
    runs = []
    for s in range(SEEDS):
        b = np.zeros(res, dtype=np.float32)
        V.integrate(V.integrator_variance_reduction_rr_integral(255, 4, 5000 + s, alpha=1.0), b, f, V.range_primary(2),
                    ctx=ctx)
        runs.append(b.astype(np.float64))
    runs = np.array(runs)
    se = runs.std(0, ddof=1) / np.sqrt(SEEDS)
    ok = se > 0
    z = (runs.mean(0) - truth)[ok] / se[ok]
    assert abs(z.mean()) < 5 / np.sqrt(z.size) + 0.05 and np.max(np.abs(z)) < 5.5


def test_roulettes_importance_and_uniform_share_a_batch(ctx):
    """One batch holding a roulette, importance sampling, both, and neither: each job's
    bins are exactly what it gives alone (the roulette jobs keep one footprint per
    region, the others merge footprints as they do alone)."""
    from viltrum_amd import hip
    f = I.config1()
    res = (8, 8)
    forms = [V.FORM_VR_OPTIMIZE_RR_INTEGRAL, V.FORM_VR_OPTIMIZE_IMPORTANCE, V.FORM_VR_OPTIMIZE_RR_PDF_IMPORTANCE,
             V.FORM_VR_OPTIMIZE]
    integs = [V.Integrator(63, 4, 10 + k, form=fm) for k, fm in enumerate(forms)]
    ps = [ig.problem(f, res, V.range_primary(2), V.F32) for ig in integs]
    bufs = [hip.DeviceBuffer(64 * 4) for _ in ps]
    ctx.integrate_batch(ps, [f] * len(ps), [b.address for b in bufs])
    got = [b.to_host(np.zeros(res, dtype=np.float32)) for b in bufs]
    for b in bufs:
        b.free()
    for ig, g in zip(integs, got):
        alone = np.zeros(res, dtype=np.float32)
        V.integrate(ig, alone, f, V.range_primary(2), ctx=ctx)
        assert np.array_equal(g, alone), ig.form


def test_rr_pdf_is_unbiased(ctx):
    f = I.config1()
    res = (16, 16)
    truth = O.blobs2d_exact_bins(f, *res)
    runs = []
    for s in range(SEEDS):
        b = np.zeros(res, dtype=np.float32)
        V.integrate(V.integrator_variance_reduction_rr_pdf(255, 4, 6000 + s, alpha=1.0), b, f, V.range_primary(2),
                    ctx=ctx)
        runs.append(b.astype(np.float64))
    runs = np.array(runs)
    se = runs.std(0, ddof=1) / np.sqrt(SEEDS)
    ok = se > 0
    z = (runs.mean(0) - truth)[ok] / se[ok]
    assert abs(z.mean()) < 5 / np.sqrt(z.size) + 0.05 and np.max(np.abs(z)) < 5.5


def test_rr_pdf_in_five_dimensions_matches_oracle(ctx):
    """The depth-first pdf fold at D = 5 (243 node values per region)."""
    f = I.config5(3)
    res = (6, 6)
    g, o = [], []
    for s in range(SEEDS):
        b = np.zeros(res, dtype=np.float64)
        V.integrate(V.Integrator(60, 8, 7000 + s, form=V.FORM_VR_FIXED_RR_PDF, alpha=1.0), b, f,
                    V.range_primary(f.dim), ctx=ctx)
        g.append(b)
        ob, _ = O.integrate(f, res, 60, 8, 7000 + s, scalar=8, integrator=V.FORM_VR_FIXED_RR_PDF, alpha=1.0)
        o.append(ob)
    _check(np.array(g), np.array(o), rel_floor=1e-6, spread=(0.6, 1.6))


@pytest.mark.parametrize("f,res,it,spp", [
    (I.config1(), (8, 8), 127, 8),
    (poly_cases()[1], (8, 8), 100, 8),
    (poly_cases()[2], (5, 5), 150, 8),
    (I.config1(), (20,), 60, 8),
], ids=lambda v: getattr(v, "name", str(v)))
@pytest.mark.parametrize("form", [V.FORM_VR_FIXED_RR_INTEGRAL_IMPORTANCE, V.FORM_VR_OPTIMIZE_RR_INTEGRAL_IMPORTANCE,
                                  V.FORM_VR_FIXED_RR_ERROR_IMPORTANCE, V.FORM_VR_OPTIMIZE_RR_ERROR_IMPORTANCE,
                                  V.FORM_VR_FIXED_RR_PDF_IMPORTANCE, V.FORM_VR_OPTIMIZE_RR_PDF_IMPORTANCE],
                         ids=["int_fixed1", "int_optimize", "err_fixed1", "err_optimize", "pdf_fixed1", "pdf_optimize"])
def test_roulette_with_importance_matches_oracle(ctx, f, res, it, spp, form):
    """A roulette picks the region, region_sampling_importance the point (FP32)."""
    g, o = [], []
    for s in range(SEEDS):
        b = np.zeros(res, dtype=np.float32)
        V.integrate(V.Integrator(it, spp, 8000 + s, form=form, alpha=1.0), b, f, V.range_primary(f.dim), ctx=ctx)
        g.append(b.astype(np.float64))
        ob, _ = O.integrate(f, res, it, spp, 8000 + s, scalar=4, integrator=form, alpha=1.0)
        o.append(ob)
    g, o = np.array(g), np.array(o)
    assert np.all(np.isfinite(g)) and np.all(np.isfinite(o))
    _check(g, o, rel_floor=1e-6, spread=(0.6, 1.6))


def test_full_importance_sampling_is_unbiased(ctx):
    """rr_pdf_region with region_sampling_importance, weight 1, against the closed form."""
    f = I.config1()
    res = (16, 16)
    truth = O.blobs2d_exact_bins(f, *res)
    runs = []
    for s in range(SEEDS):
        b = np.zeros(res, dtype=np.float32)
        V.integrate(V.integrator_variance_reduction_rr_pdf(255, 4, 9000 + s, alpha=1.0, importance=True), b, f,
                    V.range_primary(2), ctx=ctx)
        runs.append(b.astype(np.float64))
    runs = np.array(runs)
    se = runs.std(0, ddof=1) / np.sqrt(SEEDS)
    ok = se > 0
    z = (runs.mean(0) - truth)[ok] / se[ok]
    assert abs(z.mean()) < 5 / np.sqrt(z.size) + 0.05 and np.max(np.abs(z)) < 5.5


def test_roulette_with_importance_rejects_fp64(ctx):
    f = I.config1()
    with pytest.raises(V.VmxError):
        V.integrate(V.integrator_variance_reduction_rr_pdf(63, 4, 0, importance=True), np.zeros((4, 4), np.float64), f,
                    V.range_primary(2), ctx=ctx)


def test_unknown_forms_are_rejected(ctx):
    f = I.config1()
    for form in (-1, 21):
        with pytest.raises((ValueError, V.VmxError)):
            V.integrate(V.Integrator(63, 4, 0, form=form), np.zeros((4, 4), np.float32), f, V.range_primary(2),
                        ctx=ctx)

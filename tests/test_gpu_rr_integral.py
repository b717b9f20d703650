"""rr_integral_region, rr_error_region and rr_pdf_region
(src/control-variates/region-russian-roulette.h:30-149): residual samples pick a bin's
region with probability proportional to |integral of its polynomial over the bin's box|,
to |Region::error()| times the region's share inside the bin, or to the integral of the
region's importance pdf over the bin's box (floored at 1% of the mean), weight 1/p.  The GPU's
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


def test_rr_integral_and_importance_do_not_share_a_batch(ctx):
    from viltrum_amd import hip
    f = I.config1()
    ps = [V.Integrator(63, 4, 0, form=V.FORM_VR_OPTIMIZE_RR_INTEGRAL).problem(f, (8, 8), V.range_primary(2), V.F32),
          V.Integrator(63, 4, 0, form=V.FORM_VR_OPTIMIZE_IMPORTANCE).problem(f, (8, 8), V.range_primary(2), V.F32)]
    bufs = [hip.DeviceBuffer(64 * 4) for _ in ps]
    with pytest.raises(V.VmxError):
        ctx.integrate_batch(ps, [f, f], [b.address for b in bufs])
    for b in bufs:
        b.free()


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

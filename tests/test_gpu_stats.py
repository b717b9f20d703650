"""Residual Monte Carlo parity: the GPU (Philox) and the oracle (the reference's own
mt19937 stream) run the same estimator, so per-bin means over independent seeds must
agree within their standard errors.  Aggregate checks (a per-bin 4-sigma rule over
B bins has false positives): mean z ~ 0, share of |z| > 4 below a binomial bound,
and equal spread."""
import numpy as np
import pytest

import oracle_lib as O
import viltrum_amd as V
from viltrum_amd import integrands as I

pytestmark = pytest.mark.gpu

SEEDS = 48


def _runs(ctx, f, res, it, spp, form, heur, alpha, scalar):
    g, o = [], []
    dt = np.float64 if scalar == 8 else np.float32
    for s in range(SEEDS):
        integ = V.Integrator(it, spp, 1000 + s, heuristic=heur, form=form, alpha=alpha)
        b = np.zeros(res, dtype=dt)
        V.integrate(integ, b, f, V.range_primary(f.dim), ctx=ctx)
        g.append(b.astype(np.float64))
        ob, _ = O.integrate(f, res, it, spp, 1000 + s, scalar=8, integrator=form, heuristic=heur, alpha=alpha)
        o.append(ob)
    return np.array(g), np.array(o)


def _check(g, o, rel_floor=0.0, spread=(0.8, 1.25), flat_rel=1e-12):
    mg, mo = g.mean(0), o.mean(0)
    vg, vo = g.var(0, ddof=1), o.var(0, ddof=1)
    se = np.sqrt((vg + vo) / SEEDS)
    # bins whose spread is at the rounding level are compared by value (last assertion)
    ok = se > flat_rel * (1 + np.abs(mo))
    z = (mg - mo)[ok] / se[ok]
    assert ok.sum() > 0.5 * ok.size
    B = z.size
    # mean of z: standard error 1/sqrt(B) if bins were independent
    assert abs(z.mean()) < 5.0 / np.sqrt(B) + 0.05, z.mean()
    # share beyond 4 sigma: expected 6.3e-5 under the null, allow a generous bound
    assert (np.abs(z) > 4).mean() <= 0.005 + 5 * np.sqrt(6.3e-5 / B), (np.abs(z) > 4).mean()
    # spread of the estimator: same per-bin variance (median ratio is robust to the
    # heavy tails of samples landing in tiny regions; oracle against itself on these
    # cases gives 0.93..1.06)
    pos = ok & (vg > 0) & (vo > 0)
    r = np.median(vg[pos] / vo[pos])
    assert spread[0] < r < spread[1], r
    # and bins where the variance vanishes agree exactly-ish
    np.testing.assert_allclose(mg[~ok], mo[~ok], rtol=1e-9, atol=1e-9 + rel_floor)


@pytest.mark.parametrize("f,res,it,spp", [
    (I.config1(), (16, 16), 127, 16),
    (I.c1_family(11, 5), (12, 10), 200, 8),
    (I.moving_blobs(4, 6), (8, 8), 150, 16),
    # 5-D: every bin column meets hundreds of regions, so bin lists are far longer than
    # the per-bin kernel's LDS cache and the region lookup searches the stored prefixes
    (I.poly([(1.0, (1, 2, 0, 3, 1)), (2.0, (4, 0, 1, 0, 0))], 5, step=([1, 1, 1, 1, 1], 2.5, 0.25), name="poly5"),
     (4, 4), 400, 8),
], ids=["config1", "c1fam", "blobs3d", "poly5_long_lists"])
@pytest.mark.parametrize("form,alpha", [(V.FORM_VR_OPTIMIZE, 1.0), (V.FORM_VR_FIXED, 0.5), (V.FORM_CV_ALPHA1, 1.0)],
                         ids=["vr_opt", "vr_fixed", "cv_alpha1"])
def test_residual_statistics_fp64(ctx, f, res, it, spp, form, alpha):
    heur = V.HEUR_DEFAULT_ABSOLUTE if form == V.FORM_CV_ALPHA1 else V.HEUR_SIZE_RELATIVE
    g, o = _runs(ctx, f, res, it, spp, form, heur, alpha, 8)
    _check(g, o)


def test_residual_statistics_fp32_production(ctx):
    """FP32 evaluation with FP64 assembly against the FP64 oracle."""
    f = I.config1()
    g, o = _runs(ctx, f, (16, 16), 127, 16, V.FORM_VR_OPTIMIZE, V.HEUR_SIZE_RELATIVE, 1.0, 4)
    _check(g, o, rel_floor=1e-5)


def test_seed_determinism_and_independence(ctx):
    f = I.config1()
    a, b, c = (np.zeros((32, 32), dtype=np.float32) for _ in range(3))
    V.integrate(V.integrator_crespo2021(255, 16, seed=5), a, f, V.range_primary(2), ctx=ctx)
    V.integrate(V.integrator_crespo2021(255, 16, seed=5), b, f, V.range_primary(2), ctx=ctx)
    V.integrate(V.integrator_crespo2021(255, 16, seed=6), c, f, V.range_primary(2), ctx=ctx)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)


def test_single_region_and_invalid_arguments(ctx):
    """One region (iterations = 0) covers every bin; bad shapes are rejected."""
    f = I.config1()
    b = np.zeros((3, 5))
    V.integrate(V.integrator_crespo2021(0, 4, seed=1), b, f, V.range_primary(2), ctx=ctx)
    assert np.all(np.isfinite(b))
    with pytest.raises((ValueError, V.VmxError)):
        V.integrate(V.integrator_crespo2021(10, 4), np.zeros((4, 4, 4, 4)), f, V.range_primary(2), ctx=ctx)

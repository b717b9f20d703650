"""The Fubini form over an infinite range (RangeInfinite): the Dyson series integrated by
integrator_crespo2021_infinite<rd> and by the optimised Fubini integrator with
region_stratification_uniform.  The GPU draws the sequences from Philox streams keyed by
the point, the oracle (bit for bit the reference build, see test_oracle.py) from the
reference's engines, so parity is statistical: per-bin means within their standard
errors over independent seeds, equal spread; plus the closed form of the series."""
import numpy as np
import pytest

import oracle_lib as O
import viltrum_amd as V
from viltrum_amd import hip
from viltrum_amd import integrands as I

from test_gpu_stats import SEEDS, _check
from test_oracle import dyson_bins_closed_form

pytestmark = pytest.mark.gpu

CRESPO, OPT, FIXED = O.INF_CRESPO2021, O.INF_OPTIMIZE_STRATIFIED, O.INF_FIXED_STRATIFIED

# (form, decay, region dims, res, iterations, mc, spp, alpha, region box or None, rest entries)
CASES = [
    (CRESPO, 0.6, 1, (32,), 100, 3, 8, 1.0, None, ((0.1,), (0.8,))),
    (CRESPO, 0.75, 2, (6, 5), 200, 2, 16, 1.0, ((-0.5, 0.0), (1.0, 2.0)), ((), ())),
    (CRESPO, 0.75, 4, (16,), 300, 4, 16, 1.0, None, ((), ())),
    (OPT, 0.75, 2, (4, 4), 150, 4, 32, 1.0, None, ((), ())),
    (OPT, 0.6, 4, (6, 4), 120, 2, 5, 1.0, None, ((0.2, 0.0), (0.7, 0.5))),  # spp below the bins' region counts
    (FIXED, 0.8, 1, (24,), 60, 5, 12, 0.4, None, ((), ())),
    (FIXED, 0.75, 2, (16,), 300, 3, 7, 0.7, ((0.0, 0.25), (1.0, 0.75)), ((0.3,), (0.9,))),
    (OPT, 0.7, 3, (4, 3, 2), 200, 2, 8, 1.0, None, ((), ())),
]


def gpu_integrator(form, rd, it, mc, spp, seed, alpha, order=V.ORDER_CANONICAL):
    if form == CRESPO:
        return V.integrator_crespo2021_infinite(rd, it, mc, spp, seed=seed, order=order)
    return V.integrator_adaptive_fubini_variance_reduction_parallel_optimized(
        rd, it, mc, spp, seed=seed, alpha=None if form == OPT else alpha, order=order)


def gpu_range(rd, box, rest):
    lo, hi = ([0.0] * rd, [1.0] * rd) if box is None else (list(box[0]), list(box[1]))
    return V.range_infinite(lo + list(rest[0]), hi + list(rest[1]))


def gpu_bins(ctx, case, seed, order=V.ORDER_CANONICAL):
    form, decay, rd, res, it, mc, spp, alpha, box, rest = case
    integ = gpu_integrator(form, rd, it, mc, spp, seed, alpha, order)
    b = np.zeros(res, dtype=np.float32)
    V.integrate(integ, b, I.dyson(decay), gpu_range(rd, box, rest), ctx=ctx)
    assert integ.last_stats.regions == it + 1
    return b.astype(np.float64)


@pytest.mark.parametrize("case", CASES,
                         ids=lambda c: "f%d-rd%d-r%s-spp%d" % (c[0], c[2], "x".join(map(str, c[3])), c[6]))
def test_infinite_statistics_match_oracle(ctx, case):
    """The stratified forms give a bin's samples to consecutive regions of its list from
    a random start, so with fewer samples than regions the spread depends on the list's
    order: they run in the reference's region order here, which the oracle's lists
    follow (the reference's own lists are filled by a parallel loop in no fixed order)."""
    form, decay, rd, res, it, mc, spp, alpha, box, rest = case
    lo, hi = (None, None) if box is None else box
    order = V.ORDER_CANONICAL if form == CRESPO else V.ORDER_REFERENCE
    g, o = [], []
    for s in range(SEEDS):
        g.append(gpu_bins(ctx, case, 500 + s, order))
        ob, _ = O.integrate_fubini_inf(decay, rd, res, it, mc, spp, 500 + s, form=form, alpha=alpha, lo=lo, hi=hi,
                                       rest_lo=rest[0], rest_hi=rest[1], nthreads=4)
        o.append(ob)
    g, o = np.array(g), np.array(o)
    assert np.all(np.isfinite(g))
    # past the decay the series is the constant 1 (times the rest's volume): such bins
    # vary only by FP32 rounding of the node sums, which differs between the two
    _check(g, o, rel_floor=1e-5, spread=(0.75, 1.33), flat_rel=1e-6)


@pytest.mark.parametrize("form,rd,res,it,spp", [(CRESPO, 1, (8,), 60, 16), (OPT, 1, (8,), 60, 16),
                                                (CRESPO, 2, (4, 4), 60, 16), (FIXED, 2, (4, 4), 60, 16),
                                                (OPT, 2, (2, 2), 300, 400)],
                         ids=["crespo_rd1", "opt_rd1", "crespo_rd2", "fixed_rd2", "opt_many_regions"])
def test_infinite_is_unbiased_against_the_closed_form(ctx, form, rd, res, it, spp):
    """Mean over seeds against the Dyson series' closed form per bin (every region of a
    bin gets samples: the stratified forms are unbiased only then, see below)."""
    d = 0.75
    truth = dyson_bins_closed_form(d, rd, res)
    runs = np.array([gpu_bins(ctx, (form, d, rd, res, it, 2, spp, 0.5, None, ((), ())), 700 + s)
                     for s in range(64)])
    se = runs.std(0, ddof=1) / np.sqrt(runs.shape[0])
    flat = se < 1e-6  # bins past the decay: the series is 1 there, estimates exact up to rounding
    np.testing.assert_allclose(runs.mean(0)[flat], truth[flat], rtol=1e-5)
    z = (runs.mean(0) - truth)[~flat] / se[~flat]
    assert np.max(np.abs(z)) < 4.5 and abs(z.mean()) < 4.0 / np.sqrt(z.size) + 0.1, z


def test_stratified_fewer_samples_than_regions_keeps_the_reference_bias(ctx):
    """With fewer samples than regions the reference's stratified estimator leaves the
    regions that draw no sample out of the residual (weights spp / count, counted over
    the sampled regions only), so its bins are biased (test_oracle.py shows the oracle's
    bias against the closed form).  The GPU reproduces that bias: its means agree with
    the oracle's, not with the closed form."""
    d, res = 0.75, (2, 2)
    truth = dyson_bins_closed_form(d, 2, res)
    case = (OPT, d, 2, res, 300, 2, 16, 1.0, None, ((), ()))
    g = np.array([gpu_bins(ctx, case, 700 + s, V.ORDER_REFERENCE) for s in range(64)])
    o = np.array([O.integrate_fubini_inf(d, 2, res, 300, 2, 16, 700 + s, form=OPT, nthreads=4)[0]
                  for s in range(64)])
    se = np.sqrt((g.var(0, ddof=1) + o.var(0, ddof=1)) / 64)
    assert np.max(np.abs(g.mean(0) - o.mean(0)) / se) < 4.0
    zg = (g.mean(0) - truth) / (g.std(0, ddof=1) / 8)
    assert zg.min() < -4.0, zg  # the bias is there


def test_stratified_canonical_order_is_unbiased(ctx):
    """In the canonical region order (the default) the stratified form with fewer samples
    than regions keeps the oracle's means; only its spread may differ."""
    case = (OPT, 0.6, 4, (6, 4), 120, 2, 5, 1.0, None, ((0.2, 0.0), (0.7, 0.5)))
    form, decay, rd, res, it, mc, spp, alpha, box, rest = case
    g = np.array([gpu_bins(ctx, case, 900 + s) for s in range(SEEDS)])
    o = np.array([O.integrate_fubini_inf(decay, rd, res, it, mc, spp, 900 + s, form=form, alpha=alpha,
                                         rest_lo=rest[0], rest_hi=rest[1], nthreads=4)[0] for s in range(SEEDS)])
    _check(g, o, rel_floor=1e-5, spread=(0.5, 2.0), flat_rel=1e-6)


def test_infinite_batch_equals_single_calls(ctx):
    """Jobs of both integrators and several region dims in one batch give exactly what
    they give alone."""
    cases = [(CRESPO, 0.75, 2, (8, 8), 150, 2, 8, 1.0, None, ((), ())),
             (OPT, 0.6, 2, (5, 7), 90, 3, 16, 1.0, None, ((0.2,), (0.9,))),
             (FIXED, 0.75, 2, (6, 4), 200, 1, 4, 0.3, None, ((), ())),
             (OPT, 0.75, 4, (3, 5), 40, 4, 8, 1.0, None, ((), ()))]  # one bin_dims per batch
    singles = [gpu_bins(ctx, c, 40 + k).astype(np.float32) for k, c in enumerate(cases)]
    probs, descs, bufs = [], [], []
    dy = [I.dyson(c[1]) for c in cases]
    for k, (form, decay, rd, res, it, mc, spp, alpha, box, rest) in enumerate(cases):
        integ = gpu_integrator(form, rd, it, mc, spp, 40 + k, alpha)
        g, grng = integ.prepare(dy[k], gpu_range(rd, box, rest))
        probs.append(integ.problem(g, res, grng, V.F32))
        descs.append(g)
        bufs.append(hip.DeviceBuffer(int(np.prod(res)) * 4))
    ctx.integrate_batch(probs, descs, [b.address for b in bufs])
    for k, c in enumerate(cases):
        assert np.array_equal(bufs[k].to_host(np.zeros(c[3], dtype=np.float32)), singles[k]), k
    for b in bufs:
        b.free()


def test_infinite_limits(ctx):
    dy = I.dyson(0.75)
    # the stratified forms exist over infinite ranges only, as in the reference
    for form in (V.FORM_VR_OPTIMIZE_STRATIFIED, V.FORM_VR_FIXED_STRATIFIED):
        with pytest.raises(V.VmxError):
            V.integrate(V.Integrator(40, 4, 1, form=form), np.zeros((4, 4), np.float32), I.config1(),
                        V.range_primary(2), ctx=ctx)
    # FP32 bins only
    with pytest.raises(V.VmxError):
        V.integrate(V.integrator_crespo2021_infinite(2, 10, 1, 1), np.zeros((4, 4)), dy, V.range_primary_infinite(),
                    ctx=ctx)
    # an integrand of a sequence needs an infinite range and the reverse
    with pytest.raises(ValueError):
        V.integrate(V.integrator_crespo2021_infinite(2, 10, 1, 1), np.zeros((4, 4), np.float32), dy,
                    V.range_primary(3), ctx=ctx)
    with pytest.raises(ValueError):
        V.integrate(V.integrator_adaptive_fubini_variance_reduction_parallel_optimized(2, 10, 1, 1),
                    np.zeros((4, 4), np.float32), I.moving_blobs(1, 2), V.range_primary(3), ctx=ctx)
    # the C entry points check the descriptor themselves
    integ = V.integrator_crespo2021_infinite(2, 10, 1, 1)
    g, grng = integ.prepare(dy, V.range_primary_infinite())
    g.inner.params.decay = 1.5  # outside (0, 1)
    with pytest.raises(V.VmxError):
        ctx.integrate_host(integ.problem(g, (4, 4), grng, V.F32), g, np.zeros((4, 4), np.float32))
    g.inner.params.decay = 0.75
    g.params.rest_count = 5  # 2 region dims + 5 explicit entries exceed 6
    with pytest.raises(V.VmxError):
        ctx.integrate_host(integ.problem(g, (4, 4), grng, V.F32), g, np.zeros((4, 4), np.float32))
    g.params.rest_count = 0
    integ = V.integrator_crespo2021_infinite(2, 100, 8, 16)
    g, grng = integ.prepare(dy, V.range_primary_infinite())
    b = np.zeros((4, 4), np.float32)
    ctx.integrate_host(integ.problem(g, (4, 4), grng, V.F32), g, b)
    assert np.all(np.isfinite(b)) and abs(b.mean() - 1.5) < 0.1
    # region roulettes and importance sampling stay off the Fubini form
    for form in (V.FORM_VR_OPTIMIZE_RR_INTEGRAL, V.FORM_VR_OPTIMIZE_IMPORTANCE, V.FORM_VR_OPTIMIZE_SAMPLING_RR):
        p = V.Integrator(10, 4, 1, form=form).problem(g, (4, 4), grng, V.F32)
        with pytest.raises(V.VmxError):
            ctx.integrate_host(p, g, np.zeros((4, 4), np.float32))

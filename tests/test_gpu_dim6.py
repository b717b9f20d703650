"""Regions of six dimensions, the reference's cap (VILTRUM_MAX_DIMENSIONS_REGION,
src/newton-cotes/region.h:16-18,432): the GPU pipeline against the reference's own region
lists (golden fixtures written by the reference build) and against the oracle."""
import os

import numpy as np
import pytest

import oracle_lib as O
import test_gpu_stats as S
import viltrum_amd as V
from viltrum_amd import integrands as I

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "reference_fixtures.npz"))
F6 = I.poly([(1.0, (1, 2, 0, 1, 0, 2)), (0.5, (0, 0, 3, 0, 1, 0)), (0.25, (2, 0, 0, 0, 0, 1))], 6,
            step=([1, 1, 1, 1, 1, 1], 3.0, 0.5), name="poly6")


def _regions(ctx, it, heuristic, scalar=8):
    integ = V.Integrator(it, 0, 0, heuristic=heuristic, order=V.ORDER_REFERENCE)
    V.integrate(integ, np.zeros((3, 2), dtype=np.float64 if scalar == 8 else np.float32), F6, V.range_primary(6),
                ctx=ctx)
    return ctx.export_regions(6)


@pytest.mark.parametrize("heuristic,it", [(V.HEUR_SIZE_RELATIVE, 60), (V.HEUR_DEFAULT_ABSOLUTE, 30)])
def test_regions_match_reference_fixture(ctx, heuristic, it):
    """FP64, reference order: boxes, errors and split dims bit for bit as the reference
    wrote them (a polynomial: the GPU and the host evaluate it identically)."""
    g = _regions(ctx, it, heuristic)
    key = "regions/poly6/s8/h%d/it%d" % (heuristic, it)
    assert np.array_equal(g["lo"], FIX[key + "/lo"]) and np.array_equal(g["hi"], FIX[key + "/hi"])
    assert np.array_equal(g["err"], FIX[key + "/err"]) and np.array_equal(g["dim"], FIX[key + "/dim"])


def test_regions_match_oracle_at_300(ctx):
    g = _regions(ctx, 300, V.HEUR_SIZE_RELATIVE)
    o = O.generate(F6, 300, scalar=8)
    assert np.array_equal(g["lo"], o.lo) and np.array_equal(g["hi"], o.hi)
    assert np.array_equal(g["err"], o.err) and np.array_equal(g["nodes"], o.nodes)


@pytest.mark.parametrize("res", [(4, 3), (3, 2, 2)])
def test_cv_bins_fp64_spp0_match_oracle(ctx, res):
    """spp = 0, FP64, reference order: the per-bin control-variate integrals bit for bit."""
    b = np.zeros(res)
    V.integrate(V.Integrator(200, 0, 0, order=V.ORDER_REFERENCE), b, F6, V.range_primary(6), ctx=ctx)
    ob, _ = O.integrate(F6, res, 200, 0, 0, scalar=8)
    np.testing.assert_array_equal(b, ob)


def test_residual_statistics_fp32(ctx):
    """FP32 production path at D = 6 against the oracle over 48 seeds."""
    g, o = S._runs(ctx, F6, (6, 6), 150, 8, V.FORM_VR_OPTIMIZE, V.HEUR_SIZE_RELATIVE, 1.0, 4)
    S._check(g, o, rel_floor=1e-5)


def test_batch_equals_single_and_fixed_weight_is_unbiased(ctx):
    """Canonical order, a batch of two D = 6 problems gives what each gives alone; with
    weight 1 the mean over seeds matches the closed-form integral of the polynomial."""
    from viltrum_amd import hip
    integs = [V.Integrator(120, 4, 7), V.Integrator(80, 6, 8)]
    ps = [ig.problem(F6, (3, 3), V.range_primary(6), V.F32) for ig in integs]
    bufs = [hip.DeviceBuffer(9 * 4) for _ in ps]
    ctx.integrate_batch(ps, [F6, F6], [b.address for b in bufs])
    got = [b.to_host(np.zeros((3, 3), np.float32)) for b in bufs]
    for b in bufs:
        b.free()
    for ig, g in zip(integs, got):
        alone = np.zeros((3, 3), np.float32)
        V.integrate(ig, alone, F6, V.range_primary(6), ctx=ctx)
        assert np.array_equal(g, alone)
    # x0 x1^2 x3 x5^2 / ... : the monomial integrals over [0,1]^6 plus the step's volume
    runs = []
    for s in range(24):
        b = np.zeros(1, np.float64)
        V.integrate(V.integrator_variance_reduction_fixed(1.0, 100, 64, seed=100 + s), b, F6, V.range_primary(6),
                    ctx=ctx)
        runs.append(b[0])
    mono = 1.0 / (2 * 3 * 2 * 3) + 0.5 / (4 * 2) + 0.25 / (3 * 2)
    # the step 0.5 * [sum x > 3] covers half of [0,1]^6 by symmetry
    exact = mono + 0.5 * 0.5
    se = np.std(runs, ddof=1) / np.sqrt(len(runs))
    assert abs(np.mean(runs) - exact) < 5 * se + 1e-12, (np.mean(runs), exact, se)

"""Replay of a region list: the per-bin phase alone (vmx_integrate_regions) over the
regions a run exported, or any other list of boxes with their node values
(RegionsIntegrator::integrate_regions,
src/control-variates/regions-integrator-parallel-variance-reduction.h:32-109)."""
import numpy as np
import pytest

import oracle_lib as O
import viltrum_amd as V
from viltrum_amd import integrands as I

from test_gpu_parity import poly_cases

pytestmark = pytest.mark.gpu


def _run_and_export(ctx, integ, f, res, dtype):
    bins = np.zeros(res, dtype=dtype)
    V.integrate(integ, bins, f, V.range_primary(f.dim), ctx=ctx)
    return bins, ctx.export_regions(f.dim)


@pytest.mark.parametrize("f,res", [(I.config1(), (64, 64)), (I.moving_blobs(5, 8), (24, 24)),
                                   (poly_cases()[3], (6, 5))], ids=lambda v: getattr(v, "name", str(v)))
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_replay_reproduces_the_run(ctx, f, res, dtype):
    """In the reference order no footprints are merged, so replaying the exported list
    with the same integrator gives the run's bins bit for bit."""
    integ = V.Integrator(1023, 8, 3, order=V.ORDER_REFERENCE)
    bins, regs = _run_and_export(ctx, integ, f, res, dtype)
    again = np.zeros(res, dtype=dtype)
    V.integrate_regions(integ, again, f, V.range_primary(f.dim), regs, ctx=ctx)
    assert integ.last_stats.regions == regs["lo"].shape[0] and integ.last_stats.rounds == 0
    np.testing.assert_array_equal(again, bins)
    back = ctx.export_regions(f.dim)
    for key in ("lo", "hi", "nodes"):
        np.testing.assert_array_equal(back[key], regs[key])


@pytest.mark.parametrize("f,res", [(poly_cases()[1], (16, 16)), (poly_cases()[2], (8, 8)), (I.config1(), (32, 32))],
                         ids=lambda v: getattr(v, "name", str(v)))
@pytest.mark.parametrize("form", [V.FORM_VR_OPTIMIZE, V.FORM_CV_ALPHA1])
def test_replay_fp64_matches_oracle(ctx, f, res, form):
    """spp = 0, FP64: the per-bin control-variate integral over the given list equals the
    oracle's integrate_regions on the same list (bit for bit for polynomials)."""
    heur = V.HEUR_SIZE_RELATIVE if form == V.FORM_VR_OPTIMIZE else V.HEUR_DEFAULT_ABSOLUTE
    gen = V.Integrator(300, 0, 0, heuristic=heur, form=form, order=V.ORDER_REFERENCE)
    _, regs = _run_and_export(ctx, gen, f, res, np.float64)
    # replay at another resolution than the one the list was made at
    res2 = tuple(2 * r for r in res)
    got = np.zeros(res2, dtype=np.float64)
    V.integrate_regions(V.Integrator(0, 0, 0, heuristic=heur, form=form), got, f, V.range_primary(f.dim), regs,
                        ctx=ctx)
    want, _ = O.integrate_regions(f, regs["lo"], regs["hi"], regs["nodes"], res2, 0, 0, scalar=8, integrator=form)
    if f.family == I.FAMILY_POLY:
        np.testing.assert_array_equal(got, want)
    else:
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_progressive_replay_lowers_the_error(ctx):
    """More samples over the same regions: unbiased (weight 1) and the error falls as
    1/sqrt(spp) against the closed-form bins."""
    f = I.config1()
    res = (64, 64)
    truth = O.blobs2d_exact_bins(f, *res)
    _, regs = _run_and_export(ctx, V.integrator_crespo2021(1023, 0, 0), f, res, np.float32)
    rmse = []
    for spp in (4, 64):
        bins = np.zeros(res, dtype=np.float32)
        V.integrate_regions(V.integrator_variance_reduction_fixed(1.0, 0, spp, 7), bins, f, V.range_primary(2), regs,
                            ctx=ctx)
        err = bins.astype(np.float64) - truth
        assert abs(err.mean()) < 5 * err.std() / np.sqrt(err.size) + 1e-7
        rmse.append(float(np.sqrt(np.mean(err ** 2))))
    assert rmse[1] < 0.4 * rmse[0], rmse


def test_replay_rejects_bad_lists(ctx):
    f = I.config1()
    _, regs = _run_and_export(ctx, V.integrator_crespo2021(63, 0, 0), f, (8, 8), np.float32)
    integ = V.integrator_crespo2021(0, 4, 0)
    bad = {k: v.copy() for k, v in regs.items()}
    bad["hi"][3, 0] = 1.5  # outside the range
    with pytest.raises(V.VmxError):
        V.integrate_regions(integ, np.zeros((8, 8), np.float32), f, V.range_primary(2), bad, ctx=ctx)
    bad = {k: v.copy() for k, v in regs.items()}
    bad["hi"][5, 1] = bad["lo"][5, 1]  # empty
    with pytest.raises(V.VmxError):
        V.integrate_regions(integ, np.zeros((8, 8), np.float32), f, V.range_primary(2), bad, ctx=ctx)
    bad = {k: v.copy() for k, v in regs.items()}
    bad["nodes"][2, 4] = np.nan
    with pytest.raises(V.VmxError):
        V.integrate_regions(integ, np.zeros((8, 8), np.float32), f, V.range_primary(2), bad, ctx=ctx)
    # a list that leaves bins uncovered, with samples to draw in them
    part = {k: v[: v.shape[0] // 2] for k, v in regs.items()}
    with pytest.raises(V.VmxError):
        V.integrate_regions(integ, np.zeros((8, 8), np.float32), f, V.range_primary(2), part, ctx=ctx)

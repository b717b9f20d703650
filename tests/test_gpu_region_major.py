"""The region-major per-bin kernel (FP32 regions of 4..6 dims, uniform region choice and
uniform points): per-bin statistics against the oracle for one, two and three bin dims,
sample counts that are not a multiple of its pass, both weights and the alpha = 1 form;
and batches that mix its jobs with jobs of the lane-per-bin kernel (roulettes,
importance sampling) give every job exactly what it gives alone."""
import numpy as np
import pytest

import oracle_lib as O
import viltrum_amd as V
from viltrum_amd import hip
from viltrum_amd import integrands as I

from test_gpu_stats import SEEDS, _check

pytestmark = pytest.mark.gpu

POLY4 = I.poly([(1.0, (1, 2, 0, 3)), (0.5, (2, 0, 1, 0)), (0.25, (0, 0, 0, 0))], 4, step=([1, 1, 1, 1], 2.0, 0.5),
               name="poly4step")
POLY5 = I.poly([(1.0, (1, 2, 0, 3, 1)), (2.0, (4, 0, 1, 0, 0))], 5, step=([1, 1, 1, 1, 1], 2.5, 0.25), name="poly5")


def _runs(ctx, f, res, it, spp, form, alpha, heur=V.HEUR_SIZE_RELATIVE):
    g, o = [], []
    for s in range(SEEDS):
        b = np.zeros(res, dtype=np.float32)
        V.integrate(V.Integrator(it, spp, 9100 + s, heuristic=heur, form=form, alpha=alpha), b, f,
                    V.range_primary(f.dim), ctx=ctx)
        g.append(b.astype(np.float64))
        ob, _ = O.integrate(f, res, it, spp, 9100 + s, scalar=4, integrator=form, heuristic=heur, alpha=alpha)
        o.append(ob)
    return np.array(g), np.array(o)


@pytest.mark.parametrize("f,res,it,spp", [
    (POLY4, (20,), 150, 21),          # one bin dim (runs of 256 bins), 21 = 16 + 5 samples
    (POLY4, (9, 7), 150, 5),          # 16 x 16 tiles cut at the grid's edge, fewer samples than a pass
    (POLY5, (4, 3, 3), 200, 16),      # three bin dims
    (I.config5(4), (10, 10), 300, 33),
], ids=["poly4_db1_spp21", "poly4_db2_spp5", "poly5_db3", "config5_spp33"])
@pytest.mark.parametrize("form,alpha", [(V.FORM_VR_OPTIMIZE, 1.0), (V.FORM_VR_FIXED, 0.5)], ids=["optimize", "fixed"])
def test_region_major_statistics_match_oracle(ctx, f, res, it, spp, form, alpha):
    g, o = _runs(ctx, f, res, it, spp, form, alpha)
    assert np.all(np.isfinite(g))
    _check(g, o, rel_floor=1e-6, spread=(0.6, 1.67))


def test_region_major_alpha1_matches_oracle(ctx):
    g, o = _runs(ctx, POLY5, (5, 5), 200, 12, V.FORM_CV_ALPHA1, 1.0, heur=V.HEUR_DEFAULT_ABSOLUTE)
    _check(g, o, rel_floor=1e-6, spread=(0.6, 1.67))


@pytest.mark.parametrize("f,res", [(POLY4, (12, 10)), (I.config5(1), (16, 16))], ids=["poly4", "config5"])
def test_region_major_and_lane_per_bin_jobs_share_a_batch(ctx, f, res):
    """One batch: two uniform jobs (region-major kernel), a roulette and importance
    sampling (lane-per-bin kernel); each job's bins equal its single call's bit for bit."""
    forms = [(V.FORM_VR_OPTIMIZE, 1.0), (V.FORM_VR_FIXED_RR_PDF, 1.0), (V.FORM_VR_FIXED, 0.3),
             (V.FORM_VR_OPTIMIZE_IMPORTANCE, 1.0)]
    integs = [V.Integrator(180 + 40 * k, 9 + 4 * k, 70 + k, form=fm, alpha=al) for k, (fm, al) in enumerate(forms)]
    ps = [ig.problem(f, res, V.range_primary(f.dim), V.F32) for ig in integs]
    nb = int(np.prod(res))
    bufs = [hip.DeviceBuffer(nb * 4) for _ in ps]
    ctx.integrate_batch(ps, [f] * len(ps), [b.address for b in bufs])
    got = [b.to_host(np.zeros(res, dtype=np.float32)) for b in bufs]
    for b in bufs:
        b.free()
    for ig, g in zip(integs, got):
        alone = np.zeros(res, dtype=np.float32)
        V.integrate(ig, alone, f, V.range_primary(f.dim), ctx=ctx)
        assert np.all(np.isfinite(alone))
        assert np.array_equal(g, alone), ig.form


def test_region_major_is_deterministic(ctx):
    """Samples are grouped by region with LDS atomics; the bins do not depend on it."""
    f = I.config5(2)
    outs = []
    for _ in range(3):
        b = np.zeros((64, 64), dtype=np.float32)
        V.integrate(V.integrator_crespo2021(4095, 24, seed=5), b, f, V.range_primary(5), ctx=ctx)
        outs.append(b)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("res", [(48, 40), (10, 9, 7)], ids=["9_tiles", "3_tiles_3_bin_dims"])
def test_region_major_tile_map_does_not_change_results(ctx, monkeypatch, res):
    """The XCD tile map (each XCD takes one contiguous run of tiles) only reorders which
    workgroup takes which tile.  9 tiles leave a remainder over the 8 XCDs; 3 tiles are
    fewer than the XCDs."""
    f = I.config5(4)
    outs = []
    for x in ("0", "1"):
        monkeypatch.setenv("VMX_BINS_XCD", x)
        b = np.zeros(res, dtype=np.float32)
        V.integrate(V.integrator_crespo2021(4095, 24, seed=7), b, f, V.range_primary(5), ctx=ctx)
        assert np.all(np.isfinite(b))
        outs.append(b)
    assert np.array_equal(outs[0], outs[1])

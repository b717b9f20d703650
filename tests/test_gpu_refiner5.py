"""GPU checks of the learned priority at dimension 5 (the 288-wide first layer): the
training pairs of a 5-D lobes x steps x cosine integrand against the oracle's features
and FP64 true error, training on the device, and learned refinement giving a valid,
unbiased partition."""
import numpy as np
import pytest

import oracle_lib as O
import viltrum_amd as V
from viltrum_amd import integrands as I
from viltrum_amd import refiner as R

from test_gpu_refiner import _codes, deepen_check, fused_priority_check

pytestmark = pytest.mark.gpu


def test_pairs_match_oracle_d5(ctx):
    f = I.config5(3)
    res = (32, 32)
    p = V.integrator_crespo2021(255, 0, 0).problem(f, res, V.range_primary(5), V.F32)
    X, t = ctx.refiner_pairs([p], [f], target=V.TARGET_ENERGY)
    F, kp = O.nn_dims(5)
    assert kp == 288 and X.shape[1] == kp and X.shape[0] >= 2 * 256 - 1
    assert np.all(np.isfinite(t))
    assert np.all(X[:, F:] == 0)  # zero padding up to the MFMA K step
    bins = np.zeros(res, dtype=np.float32)
    ctx.integrate_host(p, f, bins)
    regs = ctx.export_regions(5)
    assert regs["lo"].shape[0] == 256
    codes = _codes(regs["lo"], regs["hi"], [0] * 5, [1] * 5)
    keys = {tuple(X[i]): i for i in range(X.shape[0])}
    lres = O.lres_of(res, 5)
    checked = 0
    for r in range(0, regs["lo"].shape[0], 5):
        lo, hi, v = regs["lo"][r], regs["hi"][r], regs["nodes"][r]
        x_o, s_o, d_o = O.region_features(5, lo, hi, [0] * 5, [1] * 5, v, codes[r], lres=lres)
        xb = O.f2bf(x_o)
        assert tuple(xb) in keys, r  # the device features of this region, bit for bit
        i = keys[tuple(xb)]
        vol = float(np.prod(hi - lo))
        se = max(float(s_o), 2.0 ** -100)
        E = O.true_error(f, lo, hi, v, 1)  # one 3-point Gauss cell per dim at D = 5
        E_dev = 2.0 ** float(t[i]) * se * se * vol
        floor = 2.0 ** -60 * se * se * vol
        if E <= floor:
            assert t[i] == -60.0 or E_dev <= 2 * floor
        else:
            assert abs(E_dev - E) <= 1e-6 * E + floor, (r, E_dev, E)
        checked += 1
    assert checked >= 40


@pytest.fixture(scope="module")
def trained5(ctx):
    return R.train(ctx, count=16, regions=2048, res=(64, 64), epochs=4, batch=1024, lr=2e-3, seed=3, dim=5)


def test_training_learns_d5(trained5):
    rep = trained5
    assert rep["pairs"] >= 16 * 1000
    # standardised targets: predicting the mean gives a loss of 0.5
    assert rep["loss_valid"] < 0.4, rep
    assert rep["loss_train"] < rep["loss_first"], rep


@pytest.mark.parametrize("seed", [0, 2])
def test_learned_refinement_d5_is_unbiased_partition(ctx, trained5, seed):
    """Weight alpha = 1 (variance_reduction_fixed): the per-bin estimate is then exactly
    unbiased for any partition.  crespo2021's optimised weight is estimated from the
    same samples (a ratio estimator, cv_optimize_weight), whose bias falls as 1/spp."""
    f = R.held_out(3, dim=5)[seed]
    res = (64, 64)
    out = {}
    for learned in (False, True):
        bins = np.zeros(res, dtype=np.float32)
        it = V.integrator_variance_reduction_fixed(1.0, 2047, 4, 5,
                                                   refiner=V.REFINER_LEARNED if learned else V.REFINER_HEURISTIC)
        V.integrate(it, bins, f, V.range_primary(5), ctx=ctx)
        assert it.last_stats.regions == 2048 and np.all(np.isfinite(bins))
        regs = ctx.export_regions(5)
        vol = np.prod(regs["hi"] - regs["lo"], axis=1)
        assert abs(vol.sum() - 1.0) < 1e-5
        if learned:
            assert np.all(regs["err"] > 0)
        out[learned] = (bins.astype(np.float64), regs)
    # the two partitions differ (the network reorders splits) ...
    assert not np.array_equal(out[False][1]["lo"], out[True][1]["lo"])
    # ... and both estimate the same integral: with the same sample positions the
    # difference per bin is the Monte Carlo error of (P_heuristic - P_learned), mean 0
    d = out[True][0] - out[False][0]
    se = np.std(d) / np.sqrt(d.size)
    assert abs(d.mean()) < 5 * se + 1e-7, (d.mean(), se)


def test_fused_priority_matches_forward_d5(ctx, trained5):
    share = fused_priority_check(ctx, 5, R.held_out(1, dim=5)[0], (64, 64), 2048)
    print("uncapped share", share)
    assert share > 0.5


def test_learned_deepened_rounds_d5(ctx, trained5, monkeypatch):
    deepen_check(ctx, monkeypatch, 5, R.held_out(2, dim=5)[1], (32, 32), 2048)

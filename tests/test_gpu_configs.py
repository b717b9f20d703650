"""The five workloads on their own integrands and sizes, on the GPU, against the oracle
(the CPU restatement of the reference, itself pinned bit for bit against the reference
build) and against the reference's own outputs (golden fixtures).

Where the oracle finishes in seconds the comparison is at full size; where it does not
(config 5: 216 s per run on 8 cores), at a cropped bin grid, plus size-independent
properties at full size (unbiasedness against an FP64 truth, error below the
heuristic's, no non-finite bins)."""
import os

import numpy as np
import pytest

import oracle_lib as O
import test_gpu_stats as S
import viltrum_amd as V
from viltrum_amd import hip as H
from viltrum_amd import integrands as I
from viltrum_amd import refiner as R

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "reference_fixtures.npz"))


# ---------------------------------------------------------------- config 1
def test_config1_regions_match_reference_at_1024(ctx):
    """Config 1's budget (1023 iterations, 1024 regions), FP64, reference order: the
    region list the reference itself wrote (golden fixture).  The integrand is
    transcendental and the GPU's libm may differ from glibc by an ulp, so boxes and split
    dims must be identical and errors and node values agree to 1e-12."""
    integ = V.Integrator(1023, 0, 0, order=V.ORDER_REFERENCE)
    V.integrate(integ, np.zeros((4, 4)), I.config1(), V.range_primary(2), ctx=ctx)
    g = ctx.export_regions(2)
    key = "regions/config1/s8/h0/it1023"
    assert len(g["err"]) == 1024
    assert np.array_equal(g["lo"], FIX[key + "/lo"]) and np.array_equal(g["hi"], FIX[key + "/hi"])
    assert np.array_equal(g["dim"], FIX[key + "/dim"])
    np.testing.assert_allclose(g["err"], FIX[key + "/err"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(g["nodes"], O.generate(I.config1(), 1023, scalar=8).nodes, rtol=1e-13, atol=1e-15)


def test_config1_cv_bins_match_reference_at_64x64(ctx):
    """Config 1's grid and budget with spp = 0 (pure control variate, deterministic),
    FP64, reference order: the reference's own bins (golden fixture)."""
    b = np.zeros((64, 64))
    V.integrate(V.Integrator(1023, 0, 0, order=V.ORDER_REFERENCE), b, I.config1(), V.range_primary(2), ctx=ctx)
    np.testing.assert_allclose(b, FIX["bins/config1/s8/f0/h0/r64x64/it1023/spp0/seed0/a1"], rtol=1e-12, atol=1e-13)


def test_config1_workload_statistics(ctx):
    """The config-1 workload itself (64x64 FP32 bins, 1024 regions, 16 spp) against the
    oracle over 48 seeds: per-bin means within their standard errors, equal spread."""
    g, o = S._runs(ctx, I.config1(), (64, 64), 1023, 16, V.FORM_VR_OPTIMIZE, V.HEUR_SIZE_RELATIVE, 1.0, 4)
    S._check(g, o, rel_floor=1e-5)


# ---------------------------------------------------------------- config 2
def _pair_variance(a, b):
    """Per-bin variance of an estimator from two independent runs, averaged over bins."""
    return float(np.mean((a - b) ** 2) / 2)


def test_config2_full_size_matches_oracle(ctx):
    """Config 2 at full size (3-D, 32 moving blobs x occluder, 512x512 FP32 bins, 16384
    regions, 64 spp), GPU against the oracle, two seeds each.
    - Error level: the per-bin variance of the two estimators (from each side's seed
      pair) agrees within 10% (the oracle against itself: within 2%).
    - No bias between them: the bin-averaged difference of the seed means is within 5
      standard errors.
    - Both are unbiased for the integral: against FP64 stratified Monte Carlo (4096 spp)
      the bin-averaged difference is within 5 standard errors."""
    f = I.config2()
    res = (512, 512)
    g = []
    for s in (11, 12):
        b = np.zeros(res, dtype=np.float32)
        V.integrate(V.integrator_crespo2021(16383, 64, seed=s), b, f, V.range_primary(3), ctx=ctx)
        assert np.all(np.isfinite(b))
        g.append(b.astype(np.float64))
    o = [O.integrate(f, res, 16383, 64, s, scalar=4, nthreads=8)[0].astype(np.float64) for s in (11, 12)]
    vg, vo = _pair_variance(*g), _pair_variance(*o)
    assert 0.9 < vg / vo < 1.1, (vg, vo)
    B = res[0] * res[1]
    d = (g[0] + g[1]) / 2 - (o[0] + o[1]) / 2
    se = np.sqrt((vg + vo) / 2 / B)
    assert abs(d.mean()) < 5 * se, (d.mean(), se)
    truth, tse = ctx.integrate_mc(f, res, 4096, seed=777, stratified=True)
    dt = (g[0] + g[1]) / 2 - truth
    se_t = np.sqrt((vg / 2 + np.mean(tse ** 2)) / B)
    assert abs(dt.mean()) < 5 * se_t, (dt.mean(), se_t)
    # the estimator's RMSE (its variance plus the squared bias against the truth, truth
    # noise removed) is the oracle's: 7.8e-4 at these settings
    rmse = np.sqrt(vg)
    assert 6e-4 < rmse < 1e-3, rmse


def test_config2_statistics_vs_oracle_48_seeds(ctx):
    """The config-2 integrand at 16x16 bins, 2048 regions, 64 spp: per-bin means and
    spreads against the oracle over 48 seeds (FP32 GPU, FP64 oracle)."""
    g, o = S._runs(ctx, I.config2(), (16, 16), 2047, 64, V.FORM_VR_OPTIMIZE, V.HEUR_SIZE_RELATIVE, 1.0, 4)
    S._check(g, o, rel_floor=1e-5)


# ---------------------------------------------------------------- config 3
def _config3(ctx):
    items = I.config3_batch(seed=0, count=256)
    probs, fs, bufs = [], [], []
    for k, it in enumerate(items):
        f = it["integrand"]
        probs.append(V.integrator_crespo2021(it["regions"] - 1, it["spp"], seed=k).problem(
            f, it["res"], V.range_primary(f.dim), V.F32))
        fs.append(f)
        bufs.append(H.DeviceBuffer(int(np.prod(it["res"])) * 4))
    return items, probs, fs, bufs


def test_config3_batch_equals_single_calls(ctx):
    """All 256 integrals of config 3 (uneven resolution, budget, spp and dimension) in one
    batched call give bit for bit what each gives alone."""
    items, probs, fs, bufs = _config3(ctx)
    try:
        ctx.integrate_batch(probs, fs, [b.address for b in bufs])
        batched = [b.to_host(np.zeros(it["res"], dtype=np.float32)) for b, it in zip(bufs, items)]
        for k, (p, f, b, it) in enumerate(zip(probs, fs, bufs, items)):
            b.zero()
            ctx.integrate_device(p, f, b.address)
            single = b.to_host(np.zeros(it["res"], dtype=np.float32))
            assert np.all(np.isfinite(single)), k
            assert np.array_equal(batched[k], single), k
    finally:
        for b in bufs:
            b.free()


def test_config3_sample_statistics_vs_oracle(ctx):
    """Integrals of the config-3 batch, as the batch poses them (their own resolution,
    budget and spp), against the oracle over 48 seeds: the three smallest 2-D and the two
    smallest 3-D ones."""
    items = I.config3_batch(seed=0, count=256)
    by_cost = sorted(range(len(items)), key=lambda k: (np.prod(items[k]["res"]) * items[k]["spp"], k))
    two = [k for k in by_cost if items[k]["integrand"].dim == 2][:3]
    three = [k for k in by_cost if items[k]["integrand"].dim == 3][:2]
    for k in two + three:
        it = items[k]
        g, o = S._runs(ctx, it["integrand"], it["res"], it["regions"] - 1, it["spp"], V.FORM_VR_OPTIMIZE,
                       V.HEUR_SIZE_RELATIVE, 1.0, 4)
        S._check(g, o, rel_floor=1e-5)


# ---------------------------------------------------------------- config 5
def test_config5_statistics_vs_oracle(ctx):
    """crespo2021 (uniform region roulette, optimised weight) on config-5 integrands at
    8x8 bins, 2048 regions, 32 spp: per-bin means and spreads against the oracle over 48
    seeds (FP32 GPU, FP64 oracle)."""
    for s in (0, 3):
        g, o = S._runs(ctx, I.config5(s), (8, 8), 2047, 32, V.FORM_VR_OPTIMIZE, V.HEUR_SIZE_RELATIVE, 1.0, 4)
        S._check(g, o, rel_floor=1e-6, spread=(0.75, 1.33))


def _blocks(res, blk=8, n=4):
    """n x n blocks of blk x blk bins spread over the grid (their top-left corners)."""
    step = res[0] // n
    return [(i * step + step // 2 - blk // 2, j * step + step // 2 - blk // 2) for i in range(n) for j in range(n)]


def test_config5_full_size_learned_batched(ctx):
    """Config 5's full size (5-D, 1024x1024 FP32 bins, 65536 regions, 32 spp) on two
    integrals in one batched call with the learned priority (network trained on the GPU
    on seeds disjoint from the evaluated ones), and the same batch with the heuristic.
    Against FP64 stratified Monte Carlo at 4096 spp on 16 blocks of 8x8 bins per
    integral: both paths unbiased (bin-averaged difference within 5 standard errors)
    and the learned path's error below the heuristic's (truth noise removed)."""
    R.train(ctx, count=16, regions=32768, res=(1024, 1024), dim=5, epochs=4)
    fs = [I.config5(s) for s in (0, 1)]
    res = (1024, 1024)
    bufs = [H.DeviceBuffer(4 * res[0] * res[1]) for _ in fs]
    try:
        truths = []
        for f in fs:
            blocks = []
            for (i0, j0) in _blocks(res):
                lo = [i0 / res[0], j0 / res[1], 0.0, 0.0, 0.0]
                hi = [(i0 + 8) / res[0], (j0 + 8) / res[1], 1.0, 1.0, 1.0]
                t, se = ctx.integrate_mc(f, (8, 8), 4096, seed=99, rng=V.Range(lo, hi), stratified=True)
                # the crop's bins are (crop volume) x (bin mean), the full grid's (bin
                # count) x (bin integral) = bin mean: factor (1024 / 8)^2
                scale = (res[0] / 8) * (res[1] / 8)
                blocks.append((i0, j0, t * scale, se * scale))
            truths.append(blocks)
        out = {}
        for learned in (False, True):
            it = R.integrator(65535, 32, seed=5, learned=learned)
            ps = [it.problem(f, res, V.range_primary(5), V.F32) for f in fs]
            ctx.integrate_batch(ps, fs, [b.address for b in bufs])
            d, se2, sq = [], [], []
            for b, blocks in zip(bufs, truths):
                full = b.to_host(np.zeros(res, dtype=np.float32)).astype(np.float64)
                assert np.all(np.isfinite(full))
                for (i0, j0, t, se) in blocks:
                    e = full[i0:i0 + 8, j0:j0 + 8] - t
                    d.append(e.ravel())
                    se2.append((se ** 2).ravel())
            d, se2 = np.concatenate(d), np.concatenate(se2)
            n = d.size
            # unbiased: mean difference against the truth's noise plus the estimator's
            # spread (estimated from the differences themselves)
            assert abs(d.mean()) < 5 * np.sqrt(np.var(d) / n), (learned, d.mean())
            out[learned] = np.sqrt(max(0.0, np.mean(d ** 2) - np.mean(se2)))
        assert out[True] < out[False], out
    finally:
        for b in bufs:
            b.free()

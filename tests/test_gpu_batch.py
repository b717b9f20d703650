"""Batching: many integrals of uneven resolution, budget, spp and form in one launch
sequence give, integral by integral, exactly what each gives alone."""
import numpy as np
import pytest

import viltrum_amd as V
from viltrum_amd import hip
from viltrum_amd import integrands as I

pytestmark = pytest.mark.gpu


def _problems():
    specs = [  # (integrand, res, iterations, spp, seed, form)
        (I.config1(), (16, 16), 255, 16, 1, V.FORM_VR_OPTIMIZE),
        (I.c1_family(3, 9), (32, 20), 600, 4, 2, V.FORM_VR_OPTIMIZE),
        (I.moving_blobs(7, 12), (24, 24), 900, 8, 3, V.FORM_VR_OPTIMIZE),
        (I.c1_family(4, 2), (8, 64), 50, 64, 4, V.FORM_VR_FIXED),
        (I.moving_blobs(8, 4), (16, 16), 0, 3, 5, V.FORM_CV_ALPHA1),
        (I.config1(), (64, 64), 2000, 0, 6, V.FORM_VR_OPTIMIZE),
        (I.c1_family(5, 30), (48, 48), 3000, 16, 7, V.FORM_VR_OPTIMIZE),
    ]
    out = []
    for f, res, it, spp, seed, form in specs:
        heur = V.HEUR_DEFAULT_ABSOLUTE if form == V.FORM_CV_ALPHA1 else V.HEUR_SIZE_RELATIVE
        integ = V.Integrator(it, spp, seed, heuristic=heur, form=form, alpha=0.6)
        out.append((f, res, integ))
    return out


@pytest.mark.parametrize("scalar", [V.F32, V.F64])
def test_batch_equals_individual_runs(ctx, scalar):
    dt = np.float32 if scalar == V.F32 else np.float64
    items = _problems()
    singles = []
    for f, res, integ in items:
        b = np.zeros(res, dtype=dt)
        V.integrate(integ, b, f, V.range_primary(f.dim), ctx=ctx)
        singles.append(b)
    bufs = [hip.DeviceBuffer(int(np.prod(res)) * np.dtype(dt).itemsize) for _, res, _ in items]
    for buf in bufs:
        buf.zero()
    probs = [integ.problem(f, res, V.range_primary(f.dim), scalar) for f, res, integ in items]
    stats = ctx.integrate_batch(probs, [f for f, _, _ in items], [b.address for b in bufs])
    for k, (f, res, integ) in enumerate(items):
        got = bufs[k].to_host(np.zeros(res, dtype=dt))
        assert np.array_equal(got, singles[k]), k
        assert stats[k].regions == integ.iterations + 1
    for b in bufs:
        b.free()


def test_batch_of_one_hundred_small_integrals(ctx):
    rng = np.random.default_rng(0)
    items = []
    for k in range(100):
        f = I.c1_family(1000 + k, int(rng.integers(1, 10)))
        side = int(rng.choice([4, 8, 16]))
        items.append((f, (side, side), V.integrator_crespo2021(int(rng.integers(0, 400)), int(rng.choice([2, 4, 8])),
                                                                seed=k)))
    bufs = [hip.DeviceBuffer(int(np.prod(res)) * 4) for _, res, _ in items]
    probs = [integ.problem(f, res, V.range_primary(2), V.F32) for f, res, integ in items]
    ctx.integrate_batch(probs, [f for f, _, _ in items], [b.address for b in bufs])
    for k in (0, 37, 99):
        f, res, integ = items[k]
        b = np.zeros(res, dtype=np.float32)
        V.integrate(integ, b, f, V.range_primary(2), ctx=ctx)
        assert np.array_equal(bufs[k].to_host(np.zeros(res, dtype=np.float32)), b)
    for b in bufs:
        b.free()


def test_repeated_calls_replay_the_same_rounds(ctx):
    """The refinement rounds run as a captured graph that later calls with the same
    buffers replay: repeated calls, with other work in between, give identical bins."""
    f = I.moving_blobs(21, 10)
    integ = V.integrator_crespo2021(1500, 8, seed=4)
    first = np.zeros((40, 40), dtype=np.float32)
    V.integrate(integ, first, f, V.range_primary(3), ctx=ctx)
    for k in range(4):
        if k == 2:  # a different problem in between (other buffers, a new capture)
            other = np.zeros((8, 8), dtype=np.float32)
            V.integrate(V.integrator_crespo2021(300, 4, seed=1), other, I.c1_family(5, 3), V.range_primary(2),
                        ctx=ctx)
        again = np.zeros((40, 40), dtype=np.float32)
        V.integrate(integ, again, f, V.range_primary(3), ctx=ctx)
        assert np.array_equal(again, first), k


def _run_batch(ctx, items, scalar=V.F32):
    dt = np.float32 if scalar == V.F32 else np.float64
    bufs = [hip.DeviceBuffer(int(np.prod(res)) * np.dtype(dt).itemsize) for _, res, _ in items]
    probs = [integ.problem(f, res, V.range_primary(f.dim), scalar) for f, res, integ in items]
    ctx.integrate_batch(probs, [f for f, _, _ in items], [b.address for b in bufs])
    out = [b.to_host(np.zeros(res, dtype=dt)) for b, (_, res, _) in zip(bufs, items)]
    for b in bufs:
        b.free()
    return out


@pytest.mark.parametrize("env", [{"VMX_BINS_XCD": "1"}, {"VMX_GEN_PER_CU": "2", "VMX_ROUND_CHUNK": "4"},
                                 {"VMX_GEN_PER_CU": "8", "VMX_BINS_XCD": "1"}, {"VMX_DEEPEN": "1"},
                                 {"VMX_DEEPEN": "1", "VMX_GEN_PER_CU": "2", "VMX_ROUND_CHUNK": "4"}],
                         ids=["xcd", "g2r4", "g8xcd", "deepen", "deepen_g2r4"])
def test_launch_shapes_do_not_change_results(ctx, monkeypatch, env):
    items = [it for it in _problems() if it[0].dim == 2]
    base = _run_batch(ctx, items)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    other = _run_batch(ctx, items)
    for k in range(len(items)):
        assert np.array_equal(base[k], other[k]), k


@pytest.mark.parametrize("scalar", [V.F32, V.F64])
def test_deepened_rounds_give_the_same_regions_and_bins(ctx, monkeypatch, scalar):
    """Splitting qualifying children in the round that creates them adds only splits a
    later round would consider: the selected regions (every box, error and node value)
    and the bins are unchanged, in fewer rounds.  All problems, 2-D and 3-D, FP32 and
    FP64."""
    items = _problems()
    runs = []
    for deepen in ("0", "1"):
        monkeypatch.setenv("VMX_DEEPEN", deepen)
        regs, bins, rounds = [], [], []
        for f, res, integ in items:
            dt = np.float32 if scalar == V.F32 else np.float64
            b = np.zeros(res, dtype=dt)
            V.integrate(integ, b, f, V.range_primary(f.dim), ctx=ctx)
            bins.append(b)
            rounds.append(integ.last_stats.rounds)
            regs.append(ctx.export_regions(f.dim))
        runs.append((regs, bins, rounds))
    for k in range(len(items)):
        for key in ("lo", "hi", "nodes", "err"):
            assert np.array_equal(runs[0][0][k][key], runs[1][0][k][key]), (k, key)
        assert np.array_equal(runs[0][1][k], runs[1][1][k]), k
    assert sum(runs[1][2]) < sum(runs[0][2]), (runs[0][2], runs[1][2])


def test_autotune_picks_a_shape_and_keeps_results(ctx):
    items = [it for it in _problems() if it[0].dim == 2]
    base = _run_batch(ctx, items)
    bufs = [hip.DeviceBuffer(int(np.prod(res)) * 4) for _, res, _ in items]
    probs = [integ.problem(f, res, V.range_primary(f.dim), V.F32) for f, res, integ in items]
    rep = ctx.autotune(probs, [f for f, _, _ in items], [b.address for b in bufs])
    assert rep["trials"] == 24 and rep["gen_blocks_per_cu"] in (2, 4, 8) and rep["round_chunk"] in (4, 8)
    assert rep["bins_xcd"] in (0, 1) and rep["deepen"] in (0, 1) and 0 < rep["ms_best"] <= rep["ms_default"]
    for k, (_, res, _) in enumerate(items):
        assert np.array_equal(bufs[k].to_host(np.zeros(res, dtype=np.float32)), base[k]), k
    for b in bufs:
        b.free()
    # later calls of the same class run at the chosen shape, with the same results
    again = _run_batch(ctx, items)
    for k in range(len(items)):
        assert np.array_equal(again[k], base[k]), k

"""Deferred rounds (a round expands a job's candidates with b >= max(theta, tau) and keeps
those in [theta, tau) for later rounds) select exactly the regions of the undeferred
rounds: region lists and bins bit for bit, for any tau (VMX_DEFER = k sets tau near the
(N/k)-th largest candidate; 0 turns deferral off), in FP64 and FP32, for single calls
and batches, with the heuristic and the learned priority, with the candidates sorted by
each job's theta block (jobs of up to 32768 regions) and by the grid-wide filter."""
import os

import numpy as np
import pytest

import viltrum_amd as V
from viltrum_amd import hip
from viltrum_amd import integrands as I
from viltrum_amd import refiner as R

pytestmark = pytest.mark.gpu


@pytest.fixture
def defer():
    old = os.environ.get("VMX_DEFER")

    def set_k(k):
        os.environ["VMX_DEFER"] = str(k)

    yield set_k
    if old is None:
        os.environ.pop("VMX_DEFER", None)
    else:
        os.environ["VMX_DEFER"] = old


def _single(ctx, f, res, it, spp, scalar, dt):
    b = np.zeros(res, dtype=dt)
    V.integrate(V.Integrator(it, spp, 5, order=V.ORDER_CANONICAL), b, f, V.range_primary(f.dim), ctx=ctx)
    return b, ctx.export_regions(f.dim)


@pytest.mark.parametrize("f,res,it,scalar", [
    (I.config1(), (32, 32), 4095, V.F64),
    (I.poly([(1.0, (1, 2, 0)), (0.5, (2, 0, 1))], 3, step=([1, 1, 1], 1.5, 0.25)), (12, 10), 3000, V.F64),
    (I.config5(3), (16, 16), 4095, V.F32),
    (I.config1(), (64, 64), 39999, V.F64),  # above 32768 regions: the grid-wide filter kernel
], ids=["config1_fp64", "poly3_fp64", "config5_fp32", "config1_fp64_40k"])
def test_single_call_regions_match(ctx, defer, f, res, it, scalar):
    dt = np.float64 if scalar == V.F64 else np.float32
    defer(0)
    b0, r0 = _single(ctx, f, res, it, 0 if scalar == V.F64 else 4, scalar, dt)
    for k in (2, 16, 256):
        defer(k)
        b1, r1 = _single(ctx, f, res, it, 0 if scalar == V.F64 else 4, scalar, dt)
        for key in ("lo", "hi", "nodes"):
            assert np.array_equal(r0[key], r1[key]), (k, key)
        assert np.array_equal(b0, b1), k


def _batch(ctx, fs, its, res):
    ps = [V.Integrator(n, 6, 30 + i).problem(f, res, V.range_primary(f.dim), V.F32) for i, (f, n) in
          enumerate(zip(fs, its))]
    nb = int(np.prod(res))
    bufs = [hip.DeviceBuffer(nb * 4) for _ in ps]
    st = ctx.integrate_batch(ps, fs, [b.address for b in bufs])
    out = [b.to_host(np.zeros(res, dtype=np.float32)) for b in bufs]
    for b in bufs:
        b.free()
    return out, st


def test_batch_bins_match_and_fewer_nodes(ctx, defer):
    fs = R.held_out(6)
    its = [8191, 4095, 12000, 2047, 8191, 6000]
    defer(0)
    b0, s0 = _batch(ctx, fs, its, (64, 64))
    defer(32)
    b1, s1 = _batch(ctx, fs, its, (64, 64))
    for x, y in zip(b0, b1):
        assert np.array_equal(x, y)
    assert sum(s.nodes for s in s1) < sum(s.nodes for s in s0)


def test_learned_priority_regions_match(defer):
    """Any network gives a priority to refine by: a fixed random one here, in a context of
    its own (the shared one keeps its networks)."""
    ctx = V.Context(0)
    n = V.refiner_size(2)[0]
    rng = np.random.default_rng(7)
    p = (rng.standard_normal(n) * 0.1).astype(np.float32)
    p[-2], p[-1] = 0.0, 1.0  # mu, sigma
    ctx.refiner_set(2, p)
    f = R.held_out(1)[0]
    res = (64, 64)
    outs = []
    for k in (0, 8, 64):
        defer(k)
        b = np.zeros(res, dtype=np.float32)
        V.integrate(V.Integrator(8191, 4, 9, refiner=V.REFINER_LEARNED), b, f, V.range_primary(2), ctx=ctx)
        outs.append((b, ctx.export_regions(2)))
    ctx.close()
    for b, r in outs[1:]:
        assert np.array_equal(b, outs[0][0])
        assert np.array_equal(r["lo"], outs[0][1]["lo"]) and np.array_equal(r["hi"], outs[0][1]["hi"])

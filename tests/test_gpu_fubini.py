"""The Fubini control variate (integrator_crespo2021_infinite<rd>): regions over the first
rd dims, the rest integrated by Monte Carlo.  The GPU draws its rest-dim samples from
Philox streams keyed by the point, the oracle (bit for bit the reference build, see
test_oracle.py) from the reference's engines, so parity is statistical: per-bin means
within their standard errors over independent seeds, equal spread."""
import numpy as np
import pytest

import oracle_lib as O
import viltrum_amd as V
from viltrum_amd import hip
from viltrum_amd import integrands as I

from test_gpu_stats import SEEDS, _check

pytestmark = pytest.mark.gpu

CASES = [  # (integrand, region dims, res, iterations, mc, spp)
    (I.moving_blobs(4, 6), 2, (8, 8), 150, 2, 8),
    (I.poly([(1.0, (2, 1, 3)), (0.25, (0, 2, 1))], 3, step=([1.0, -0.5, 0.7], 0.3, 1.0), name="poly3step"), 2,
     (6, 6), 80, 3, 4),
    (I.config5(1), 2, (4, 4), 60, 2, 4),
    (I.poly([(1.0, (2, 3)), (0.5, (1, 0))], 2, step=([1.0, 1.0], 1.0, 0.5), name="poly2step"), 1, (12,), 60, 4, 6),
]


@pytest.mark.parametrize("f,rd,res,it,mc,spp", CASES, ids=["blobs3d_rd2", "poly3_rd2", "lobes5_rd2", "poly2_rd1"])
def test_fubini_statistics_match_oracle(ctx, f, rd, res, it, mc, spp):
    g, o = [], []
    for s in range(SEEDS):
        integ = V.integrator_crespo2021_infinite(rd, it, mc, spp, seed=300 + s)
        b = np.zeros(res, dtype=np.float32)
        V.integrate(integ, b, f, V.range_primary(f.dim), ctx=ctx)
        assert integ.last_stats.regions == it + 1
        g.append(b.astype(np.float64))
        ob, _ = O.integrate_fubini(f, rd, res, it, mc, spp, 300 + s)
        o.append(ob)
    # small grids with few samples: the median variance ratio of the 3-D polynomial case
    # (36 bins, 4 spp) ranged over 0.88..1.26 across four seed bases
    _check(np.array(g), np.array(o), rel_floor=1e-5, spread=(0.75, 1.33))


def test_fubini_is_unbiased_against_full_monte_carlo(ctx):
    """Bins of the Fubini form against FP64 stratified Monte Carlo of the whole 3-D
    integrand (the same bin boxes), averaged over seeds."""
    f = I.moving_blobs(9, 8)
    res = (16, 16)
    truth, tse = ctx.integrate_mc(f, res, 4096, seed=5, rng=V.range_primary(3), stratified=True)
    acc = np.zeros(res)
    n = 8
    for s in range(n):
        b = np.zeros(res, dtype=np.float32)
        V.integrate(V.integrator_crespo2021_infinite(2, 400, 4, 16, seed=s), b, f, V.range_primary(3), ctx=ctx)
        acc += b
    err = acc / n - truth
    assert abs(err.mean()) < 5 * err.std() / np.sqrt(err.size) + 1e-4, (err.mean(), err.std())


def test_fubini_batch_equals_single_calls(ctx):
    items = [(I.moving_blobs(20 + k, 3 + k), (8 + 4 * k, 8), V.integrator_crespo2021_infinite(2, 100 + 50 * k, 1 + k,
                                                                                                4 * k, seed=k))
             for k in range(3)]
    singles = []
    for f, res, integ in items:
        b = np.zeros(res, dtype=np.float32)
        V.integrate(integ, b, f, V.range_primary(3), ctx=ctx)
        singles.append(b)
    probs, descs, bufs = [], [], []
    for f, res, integ in items:
        g, grng = integ.prepare(f, V.range_primary(3))
        probs.append(integ.problem(g, res, grng, V.F32))
        descs.append(g)
        bufs.append(hip.DeviceBuffer(int(np.prod(res)) * 4))
    ctx.integrate_batch(probs, descs, [b.address for b in bufs])
    for k, (_, res, _) in enumerate(items):
        assert np.array_equal(bufs[k].to_host(np.zeros(res, dtype=np.float32)), singles[k]), k
    for b in bufs:
        b.free()


def test_fubini_rejects_fp64_and_bad_dims(ctx):
    f = I.moving_blobs(1, 2)
    with pytest.raises(V.VmxError):
        V.integrate(V.integrator_crespo2021_infinite(2, 10, 1, 1), np.zeros((4, 4)), f, V.range_primary(3), ctx=ctx)
    with pytest.raises(ValueError):
        V.integrator_crespo2021_infinite(3, 10, 1, 1).prepare(f, V.range_primary(3))

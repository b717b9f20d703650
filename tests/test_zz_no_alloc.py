"""vmx_integrate_device_noalloc (the entry point that allocates nothing: its work memory
is the library's fixed per-GPU scratch array) against vmx_integrate_device, through the
C ABI: bit for bit, on two calls in a row, the second with a smaller problem."""
import ctypes as C

import numpy as np
import pytest

import viltrum_amd as V
from viltrum_amd import hip as H
from viltrum_amd import integrands as I

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20


def run(ctx, f, integ, res, noalloc):
    rng = V.range_primary(f.dim)
    scalar = V.F32
    p = integ.problem(f, res, rng, scalar)
    n = int(np.prod(res))
    buf = H.DeviceBuffer(4 * n)
    try:
        buf.zero()
        if noalloc:
            st = ctx.integrate_device_noalloc(p, f, buf.address)
        else:
            st = ctx.integrate_device(p, f, buf.address)
        out = buf.to_host(np.zeros(res, dtype=np.float32))
    finally:
        buf.free()
    return out, st


def test_noalloc_matches_device_entry_point_on_two_calls():
    ctx = V.Context(0)
    try:
        # first call larger, second smaller: the scratch keeps the first call's contents
        cases = [(I.config1(), V.integrator_crespo2021(16383, 16, seed=3), (512, 512)),
                 (I.c1_family(7, 5), V.integrator_crespo2021(300, 8, seed=4), (40, 24))]
        for f, integ, res in cases:
            a, sa = run(ctx, f, integ, res, noalloc=True)
            b, sb = run(ctx, f, integ, res, noalloc=False)
            print("noalloc case", f.dim, res, integ.iterations, flush=True)
            assert np.all(np.isfinite(a))
            assert np.array_equal(a, b)
            assert sa.regions == sb.regions and sa.pairs == sb.pairs
    finally:
        ctx.close()


def test_noalloc_other_forms_and_dims():
    """α = 1 form (bins accumulated), a 3-D integrand and a 5-D one, in a fresh context
    whose own arenas were never grown."""
    ctx = V.Context(0)
    try:
        cases = [(I.moving_blobs(5, 6), V.integrator_crespo2021(4095, 8, seed=1), (64, 64)),
                 (I.config5(2), V.integrator_crespo2021(1023, 4, seed=2), (32, 32)),
                 (I.config1(), V.integrator_variance_reduction_fixed(0.5, 500, 4, seed=9), (33, 17))]
        for f, integ, res in cases:
            a, _ = run(ctx, f, integ, res, noalloc=True)
            b, _ = run(ctx, f, integ, res, noalloc=False)
            assert np.array_equal(a, b)
    finally:
        ctx.close()


def test_noalloc_refuses_problems_beyond_the_scratch(ctx):
    f = I.config1()
    p = V.integrator_crespo2021(2 ** 20, 4).problem(f, (1024, 1024), V.range_primary(2), V.F32)
    buf = H.DeviceBuffer(4 * 1024 * 1024)
    try:
        with pytest.raises(V.VmxError, match="no-alloc scratch"):
            ctx.integrate_device_noalloc(p, f, buf.address)
    finally:
        buf.free()
    # the context still works, with and without allocation
    small = np.zeros((8, 8), dtype=np.float32)
    V.integrate(V.integrator_crespo2021(63, 4), small, f, V.range_primary(2), ctx=ctx)
    assert np.all(np.isfinite(small))

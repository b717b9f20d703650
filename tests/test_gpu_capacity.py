"""Capacity limits of the device pipeline that must end in an error, not in
out-of-bounds writes."""
import numpy as np
import pytest

import viltrum_amd as V
from viltrum_amd import hip as H
from viltrum_amd import integrands as I

pytestmark = pytest.mark.gpu


def test_pair_total_beyond_32_bits_is_refused(ctx):
    """A 3-D integrand varying only along t over 8192 x 4096 bins: about 1100 regions
    each covering millions of bins.  In the reference's region order every region is its
    own footprint (no merging of regions with equal bin-dim boxes), which makes 6.5e9
    (bin, footprint) pairs (counted on the host in tests/test_host_arith.py).  The
    engine sums the pair counts in 64 bits and refuses the call before it sizes or fills
    the bin lists."""
    f = I.poly([(1.0, (0, 0, 4))], 3)
    integ = V.integrator_crespo2021(1099, 4, seed=0, order=V.ORDER_REFERENCE)
    res = (8192, 4096)
    p = integ.problem(f, res, V.range_primary(3), V.F32)
    buf = H.DeviceBuffer(4 * res[0] * res[1])
    try:
        with pytest.raises(V.VmxError, match="pairs"):
            ctx.integrate_device(p, f, buf.address)
        # the context stays usable
        b = np.zeros((16, 16), dtype=np.float32)
        V.integrate(integ, b, f, V.range_primary(3), ctx=ctx)
        assert abs(float(b.mean()) - 0.2) < 0.01
    finally:
        buf.free()

"""The C++ drop-in header used the way the reference's own examples are written
(tests/cpp/dropin.cpp) gives what the Python mirror gives."""
import os
import subprocess

import numpy as np
import pytest

import viltrum_amd as V
from viltrum_amd import integrands as I

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dropin")


def test_cpp_dropin_matches_python(ctx):
    assert os.path.exists(EXE), "build tests/cpp/dropin (python tests/cpp/build.py)"
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120, check=True).stdout.split("\n")
    scalar = float(out[0].split()[1])
    assert abs(scalar - 2.0 / 3.0) < 1e-15
    b = np.zeros((32, 32), dtype=np.float32)
    V.integrate(V.integrator_crespo2021(255, 16, seed=1), b, I.config1(), V.range_primary(2), ctx=ctx)
    vals = [float(x) for x in out[1].split()[1:]]
    # 9 significant digits identify a float32 exactly
    assert np.float32(vals[0]) == b[3, 7] and np.float32(vals[1]) == b[20, 11]
    assert abs(vals[2] - float(b.astype(np.float64).sum())) < 1e-6 * abs(vals[2])
    # the Fubini form through the drop-in header and through the Python mirror
    f3 = I.blobs([(0.4, 0.5, 0.2, -0.1, 40.0, 60.0, 1.0)], dim=3)
    fb = np.zeros((16, 16), dtype=np.float32)
    V.integrate(V.integrator_crespo2021_infinite(2, 200, 2, 8, seed=4), fb, f3, V.range_primary(3), ctx=ctx)
    vals = [float(x) for x in out[2].split()[1:]]
    assert np.float32(vals[0]) == fb[5, 9] and np.float32(vals[1]) == fb[12, 3]
    assert abs(vals[2] - float(fb.astype(np.float64).sum())) < 1e-6 * abs(vals[2])
    # three bin dims (vector<vector<vector<float>>>)
    b3 = np.zeros((8, 8, 4), dtype=np.float32)
    V.integrate(V.integrator_crespo2021(300, 4, seed=2), b3, f3, V.range_primary(3), ctx=ctx)
    vals = [float(x) for x in out[3].split()[1:]]
    assert np.float32(vals[0]) == b3[1, 2, 3] and np.float32(vals[1]) == b3[6, 5, 0]
    assert abs(vals[2] - float(b3.astype(np.float64).sum())) < 1e-6 * abs(vals[2])
    # the logged regions replayed with more samples, through the header and through Python
    b1 = np.zeros((32, 32), dtype=np.float32)
    V.integrate(V.integrator_crespo2021(255, 0, seed=1), b1, I.config1(), V.range_primary(2), ctx=ctx)
    regs = ctx.export_regions(2)
    b2 = np.zeros((32, 32), dtype=np.float32)
    V.integrate_regions(V.integrator_crespo2021(0, 64, seed=9), b2, I.config1(), V.range_primary(2), regs, ctx=ctx)
    parts = out[4].split()
    assert int(parts[1]) == regs["lo"].shape[0] == 256
    vals = [float(x) for x in parts[2:]]
    assert np.float32(vals[0]) == b2[3, 7] and np.float32(vals[1]) == b2[20, 11]
    assert abs(vals[2] - float(b2.astype(np.float64).sum())) < 1e-6 * abs(vals[2])
    # the Dyson series over infinite ranges, both Fubini integrators
    dy = I.dyson(0.75)
    bi = np.zeros(10, dtype=np.float32)
    V.integrate(V.integrator_crespo2021_infinite(4, 1280, 8, 256), bi, dy, V.range_primary_infinite(), ctx=ctx)
    vals = [float(x) for x in out[5].split()[1:]]
    assert np.float32(vals[0]) == bi[0] and np.float32(vals[1]) == bi[9]
    assert abs(vals[2] - float(bi.astype(np.float64).sum())) < 1e-6 * abs(vals[2])
    bo = np.zeros(10, dtype=np.float32)
    V.integrate(V.integrator_adaptive_fubini_variance_reduction_parallel_optimized(2, 300, 4, 32, seed=7), bo, dy,
                V.range_infinite([0.0, 0.0, 0.1], [1.0, 1.0, 0.9]), ctx=ctx)
    vals = [float(x) for x in out[6].split()[1:]]
    assert np.float32(vals[0]) == bo[0] and np.float32(vals[1]) == bo[9]
    assert abs(vals[2] - float(bo.astype(np.float64).sum())) < 1e-6 * abs(vals[2])

"""The GPU's importance sampler (viltrum_amd/csrc/importance.h, compiled for the host by
tests/native/build.py) against the oracle's restatement of the reference's
(Rules::sample_line, Region::sample_subrange and pdf_subrange, src/newton-cotes/rules.h:
216-237 and region.h:220-330; the oracle is pinned bit for bit to the reference build by
test_oracle.py).  The product evaluates the same root formulas in cheaper forms
(products for pow, cbrt, one sincos), so points and weights agree to a tolerance, not
bitwise: drawn coordinates within 6e-5, weights within 5e-4 relative."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle_lib as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "native"))
import build as native_build  # noqa: E402

H = C.CDLL(native_build.build())
FP = C.POINTER(C.c_float)
H.vmh_imp_sample_line.restype = C.c_float
H.vmh_imp_sample_line.argtypes = [C.c_float, C.c_float, C.c_float, FP]
H.vmh_imp_sample.argtypes = [C.c_int, FP, FP, FP, C.c_float, FP, FP, FP, FP, FP]


def _oracle():
    L = O.oracle()
    L.vmo_imp_sample_line.restype = C.c_float
    L.vmo_imp_sample_line.argtypes = [C.c_float, C.c_float, C.c_float, FP]
    L.vmo_imp_sample.argtypes = [C.c_int, FP, FP, FP, FP, FP, FP, FP, FP]
    return L


def _p(a):
    return a.ctypes.data_as(FP)


def _line(rng, i):
    """node values of a line: generic, tiny, an interior zero of the pdf, a spike"""
    kind = i % 4
    if kind == 0:
        p = rng.normal(size=3)
    elif kind == 1:
        p = rng.random(3) * 10.0 ** rng.uniform(-6, 2)
    elif kind == 2:
        p = np.array([rng.normal(), rng.normal() * 1e-3, rng.normal()])
    else:
        p = np.array([0.0, rng.normal(), 0.0]) + rng.normal(size=3) * 1e-4
    return np.ascontiguousarray(p, dtype=np.float32)


def test_sample_line_matches_oracle():
    L = _oracle()
    rng = np.random.default_rng(0)
    for i in range(40000):
        p = _line(rng, i)
        t0, t1 = (np.float32(0), np.float32(1)) if i % 3 == 0 else np.sort(rng.random(2)).astype(np.float32)
        s = np.float32(rng.random())
        a = H.vmh_imp_sample_line(s, t0, t1, _p(p))
        b = L.vmo_imp_sample_line(s, t0, t1, _p(p))
        assert abs(a - b) <= 6e-5, (a, b, s, t0, t1, p.tolist())


@pytest.mark.parametrize("D,n", [(1, 6000), (2, 6000), (3, 3000), (4, 1500), (5, 600), (6, 200)])
def test_region_sample_matches_oracle(D, n):
    L = _oracle()
    rng = np.random.default_rng(D)
    K = 3 ** D
    checked = 0
    for i in range(n):
        v = rng.normal(size=K)
        if i % 3 == 1:
            v = v * 10.0 ** rng.uniform(-4, 1)
        if i % 3 == 2:
            v[rng.random(K) < 0.5] = 0
        v = np.ascontiguousarray(v, dtype=np.float32)
        rlo = (rng.random(D) * 0.5).astype(np.float32)
        rhi = (rlo + 0.05 + rng.random(D) * 0.45).astype(np.float32)
        a, b = rng.random(D), rng.random(D)
        slo = (rlo + np.minimum(a, b) * (rhi - rlo)).astype(np.float32)
        shi = (rlo + np.maximum(a, b) * (rhi - rlo)).astype(np.float32)
        if i % 4 == 0:
            slo, shi = rlo.copy(), rhi.copy()
        s = rng.random(D).astype(np.float32)
        rvol = np.float32(np.prod(rhi - rlo, dtype=np.float32))
        xg, wg = np.zeros(D, np.float32), np.zeros(1, np.float32)
        xo, wo = np.zeros(D, np.float32), np.zeros(1, np.float32)
        ok = H.vmh_imp_sample(D, _p(v), _p(rlo), _p(rhi), rvol, _p(slo), _p(shi), _p(s), _p(xg), _p(wg))
        L.vmo_imp_sample(D, _p(v), _p(rlo), _p(rhi), _p(slo), _p(shi), _p(s), _p(xo), _p(wo))
        if ok != 1:  # the pdf integrates to 0 over the box: the kernel samples it uniformly
            continue
        checked += 1
        assert np.max(np.abs(xg - xo)) <= 6e-5, (xg, xo)
        assert abs(float(wg[0]) - float(wo[0])) <= 5e-4 * abs(float(wo[0])), (wg, wo)
    assert checked > 0.8 * n

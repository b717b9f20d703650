"""The product's device arithmetic (viltrum_amd/csrc/region_math.h, philox.h, features.h) compiled
for the host, against the oracle and against published Philox known answers."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from viltrum_amd import integrands as I

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "native"))
import build as native_build  # noqa: E402

H = C.CDLL(native_build.build())
H.vmh_subrange.restype = C.c_double
H.vmh_monomial.restype = C.c_double
dp = C.POINTER(C.c_double)


def P(a, t=dp):
    return np.ascontiguousarray(a).ctypes.data_as(t)


# Random123 known-answer vectors (Salmon et al., SC'11): counter, key, output
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    H.vmh_philox(P(c, C.POINTER(C.c_uint32)), P(k, C.POINTER(C.c_uint32)), P(out, C.POINTER(C.c_uint32)))
    assert tuple(int(v) for v in out) == want


@pytest.mark.parametrize("D", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("scalar", [4, 8])
@pytest.mark.parametrize("heur", [0, 1])
def test_heuristic_matches_oracle(D, scalar, heur):
    rng = np.random.default_rng(D * 10 + scalar + heur)
    n = 300
    lo = rng.random((n, D)) * 0.5
    hi = lo + rng.random((n, D)) * 0.5 + 1e-3
    nodes = rng.normal(size=(n, 3 ** D))
    nodes[: n // 4] = np.abs(nodes[: n // 4])  # same-sign lines
    nodes[n // 4: n // 3] = 0.0                 # all-zero regions (relative metric branch)
    if scalar == 4:
        lo, hi, nodes = (a.astype(np.float32).astype(np.float64) for a in (lo, hi, nodes))
    e_o, d_o = O.heuristic(D, lo, hi, nodes, scalar=scalar, heuristic=heur)
    e_h, d_h = np.zeros(n), np.zeros(n, dtype=np.int32)
    H.vmh_heuristic(scalar, D, heur, n, P(lo), P(hi), P(nodes), P(e_h), P(d_h, C.POINTER(C.c_int32)))
    assert np.array_equal(e_h, e_o) and np.array_equal(d_h, d_o)


@pytest.mark.parametrize("f", [I.poly([(1.0, (2, 1, 1)), (2.0, (0, 3, 1))], 3, name="p3"), I.config1(),
                               I.config5(1)], ids=lambda f: f.name)
@pytest.mark.parametrize("scalar", [4, 8])
def test_split_coordinates_match_reference_evaluation_order(f, scalar):
    """The first split's new nodes: the harness's coordinates are exactly the points
    the oracle (pinned to the reference's evaluation log) evaluates, in order."""
    D = f.dim
    K, KM = 3 ** D, 3 ** (D - 1)
    g0 = O.generate(f, 0, scalar=scalar)
    g1 = O.generate(f, 1, scalar=scalar, evals=True)
    dim = int(g0.dim[0])
    boxes, xs = np.zeros(4 * D), np.zeros(2 * KM * D)
    idxs, cv = np.zeros(2 * KM, dtype=np.int32), np.zeros(2 * K)
    H.vmh_split(scalar, D, dim, P(g0.lo[0]), P(g0.hi[0]), P(g0.nodes[0]), P(boxes), P(xs),
                P(idxs, C.POINTER(C.c_int32)), P(cv))
    assert np.array_equal(xs.reshape(2 * KM, D), g1.evals[K:, :D])
    # inherited planes and the new values rebuild the oracle's two children
    for c in range(2):
        child = cv[c * K:(c + 1) * K].copy()
        child[idxs[c * KM:(c + 1) * KM]] = g1.evals[K + c * KM: K + (c + 1) * KM, D]
        match = [r for r in range(2) if np.array_equal(g1.lo[r], boxes[2 * c * D: 2 * c * D + D])]
        assert match and np.array_equal(child, g1.nodes[match[0]])


@pytest.mark.parametrize("D,db", [(2, 2), (3, 2), (3, 1), (5, 2), (3, 3)])
def test_bin_subrange_matches_oracle_fp64(D, db):
    """vol * marginal_subrange == the oracle's integral_subrange (bit for bit in FP64):
    one region, bins of an uneven grid, spp = 0."""
    rng = np.random.default_rng(D + 7 * db)
    lo, hi = np.zeros(D), np.ones(D)
    rlo = rng.random(D) * 0.3
    rhi = rlo + 0.3 + rng.random(D) * 0.4
    nodes = rng.normal(size=3 ** D)
    res = tuple([5, 7, 3][:db])
    f = I.poly([(0.0, (0,) * D)], D)
    bins, _ = O.integrate_regions(f, rlo[None], rhi[None], nodes[None], res, 0, 0, scalar=8)
    nb = int(np.prod(res))
    box = []
    for i in range(db):  # pixels_in_region through the product's pixel_range
        se = np.zeros(2, dtype=np.uint64)
        H.vmh_pixel_range(8, C.c_double(rlo[i]), C.c_double(rhi[i]), C.c_double(0.0), C.c_double(1.0),
                          C.c_uint64(res[i]), P(se, C.POINTER(C.c_uint64)))
        box.append((int(se[0]), int(se[1])))
    for idx in np.ndindex(*res):
        if not all(box[i][0] <= idx[i] < box[i][1] for i in range(db)):
            assert bins[idx] == 0.0
            continue
        slo, shi = lo.copy(), hi.copy()
        for i in range(db):
            slo[i] = idx[i] * (1.0 / res[i])
            shi[i] = (idx[i] + 1) * (1.0 / res[i])
        a = np.maximum(slo, rlo)
        b = np.maximum(a, np.minimum(shi, rhi))
        if np.any(a >= b):
            continue
        v = nb * H.vmh_subrange(8, D, db, P(rlo), P(rhi), P(nodes), P(a), P(b))
        assert v == bins[idx]


@pytest.mark.parametrize("D", [1, 2, 3, 5])
def test_monomial_form_interpolates_nodes(D):
    rng = np.random.default_rng(D)
    rlo = rng.random(D) * 0.3
    rhi = rlo + 0.5
    nodes = rng.normal(size=3 ** D)
    for k in range(3 ** D):
        x = np.array([rlo[d] + ((k // 3 ** d) % 3) / 2.0 * (rhi[d] - rlo[d]) for d in range(D)])
        v = H.vmh_monomial(8, D, P(rlo), P(rhi), P(nodes), P(x))
        assert abs(v - nodes[k]) < 1e-12 * (1 + np.abs(nodes).max())


def test_pixel_range_clamps_and_matches_rule():
    se = np.zeros(2, dtype=np.uint64)
    up = C.POINTER(C.c_uint64)
    H.vmh_pixel_range.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64, up]
    H.vmh_pixel_range(8, 0.25, 0.5, 0.0, 1.0, 8, P(se, up))
    assert tuple(se) == (2, 4)
    H.vmh_pixel_range(8, 0.25, 0.5 + 0.004 / 8, 0.0, 1.0, 8, P(se, up))   # < 1% into bin 4: dropped
    assert tuple(se) == (2, 4)
    H.vmh_pixel_range(8, 0.25, 0.5 + 0.02 / 8, 0.0, 1.0, 8, P(se, up))    # > 1%: kept
    assert tuple(se) == (2, 5)
    H.vmh_pixel_range(4, 0.99999994, 1.0, 0.0, 1.0, 1 << 24, P(se, up))   # rounding lands on res: clamped
    assert se[0] < (1 << 24) and se[1] <= (1 << 24)


@pytest.mark.parametrize("D", [1, 2, 3, 4])
def test_region_features_match_oracle(D):
    """The kernels' feature function (features.h, host build) against the oracle's
    restatement, bit for bit, on random FP32 regions."""
    rng = np.random.default_rng(100 + D)
    F, kp = O.nn_dims(D)
    H.vmh_region_features.restype = C.c_int
    for trial in range(40):
        depth = rng.integers(0, 8, D)
        code = np.array([(1 << int(n)) | int(rng.integers(0, 1 << int(n))) if n else 1 for n in depth],
                        dtype=np.uint32)
        rlo, rhi = np.full(D, -0.5), np.full(D, 1.5)
        ext = (rhi - rlo) / 2.0 ** depth
        lo = rlo + (code - (1 << depth)) * ext
        hi = lo + ext
        v = rng.normal(0, 1, 3 ** D).astype(np.float32).astype(np.float64) * (10.0 ** rng.integers(-20, 5))
        if trial == 0:
            v[:] = 0.0  # scale floor
        res = [int(r) for r in rng.integers(1, 700, min(D, 3))]
        lres = np.zeros(D, dtype=np.float32)
        lres[:len(res)] = O.lres_of(res, len(res))
        x_o, s_o, d_o = O.region_features(D, lo, hi, rlo, rhi, v, code, lres=lres)
        x_h = np.zeros(kp, dtype=np.float32)
        s_h, d_h = C.c_float(), C.c_int()
        assert H.vmh_region_features(D, P(lo), P(hi), P(rlo), P(rhi), P(v), P(code, C.POINTER(C.c_uint32)),
                                     P(lres, C.POINTER(C.c_float)), P(x_h, C.POINTER(C.c_float)), C.byref(s_h),
                                     C.byref(d_h)) == 0
        assert d_h.value == d_o == int(depth.sum())
        assert s_h.value == s_o
        np.testing.assert_array_equal(x_h, x_o)


def test_bf16_rounding_matches_oracle():
    H.vmh_f2bf.restype = C.c_uint16
    H.vmh_f2bf.argtypes = [C.c_float]
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.normal(0, 1, 500).astype(np.float32),
                         np.array([0.0, -0.0, 1.00390625, 1.005859375, 3.0e38, -3.4e38, np.inf, -np.inf],
                                  dtype=np.float32)])
    want = O.f2bf(xs)
    got = np.array([H.vmh_f2bf(float(v)) for v in xs], dtype=np.uint16)
    np.testing.assert_array_equal(got, want)
    # round to nearest even at the halfway points
    assert H.vmh_f2bf(1.00390625) == 0x3F80 and H.vmh_f2bf(1.01171875) == 0x3F82


# a 3-D integrand that varies only along the integrated dim t: its regions stay wide in
# the bin dims, so at a fine resolution each region covers millions of bins
PAIRS_CASE = dict(terms=[(1.0, (0, 0, 4))], iterations=1099, res=(8192, 4096))


def pair_total(g, res):
    se = np.zeros(2, dtype=np.uint64)
    up = C.POINTER(C.c_uint64)
    H.vmh_pixel_range.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64, up]
    total = 0
    for lo, hi in zip(g.lo, g.hi):
        n = 1
        for d in range(2):
            H.vmh_pixel_range(4, lo[d], hi[d], 0.0, 1.0, res[d], P(se, up))
            n *= int(se[1] - se[0])
        total += n
    return total


def test_pair_total_guard_covers_the_t_only_case():
    """The (bin, region) pair total of the t-only case, one footprint per region as in
    the reference's region order, passes what 32-bit list offsets hold (so the engine
    must refuse it; tests/test_gpu_capacity.py checks that it does), while the same
    regions at 1024^2 bins stay far below."""
    H.vmh_max_pairs.restype = C.c_uint64
    limit = H.vmh_max_pairs()
    assert limit == 2 ** 32 - 2
    f = I.poly(PAIRS_CASE["terms"], 3)
    g = O.generate(f, PAIRS_CASE["iterations"], scalar=4)
    big = pair_total(g, PAIRS_CASE["res"])
    assert big > limit
    assert big % 2 ** 32 < big  # a 32-bit count would wrap
    assert pair_total(g, (1024, 1024)) < limit // 16

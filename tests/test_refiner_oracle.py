"""CPU checks of the oracle's learned-priority pieces (the reference has no learned
component, so these are pinned by mathematics): back-propagation against central
finite differences, the forward pass against an independent numpy evaluation,
features against a hand-computed region, and the true-error quadrature against closed
forms."""
import numpy as np
import pytest

import oracle_lib as O
from viltrum_amd import integrands as I

W = 64


def layout(kp):
    o = {}
    o["w0"] = 0
    o["b0"] = kp * W
    o["w1"] = o["b0"] + W
    o["b1"] = o["w1"] + W * W
    o["w2"] = o["b1"] + W
    o["b2"] = o["w2"] + W * W
    o["w3"] = o["b2"] + W
    o["b3"] = o["w3"] + W
    o["trained"] = o["b3"] + 1
    return o


def random_net(kp, F, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    L = layout(kp)
    p = np.zeros(L["trained"] + 2, dtype=np.float32)
    w0 = rng.uniform(-1, 1, (kp, W)) * np.sqrt(6.0 / F) * scale
    w0[F:] = 0
    p[L["w0"]:L["b0"]] = w0.ravel()
    p[L["b0"]:L["w1"]] = rng.uniform(-0.1, 0.1, W)
    p[L["w1"]:L["b1"]] = (rng.uniform(-1, 1, W * W) * np.sqrt(6.0 / W)).ravel()
    p[L["b1"]:L["w2"]] = rng.uniform(-0.1, 0.1, W)
    p[L["w2"]:L["b2"]] = (rng.uniform(-1, 1, W * W) * np.sqrt(6.0 / W)).ravel()
    p[L["b2"]:L["w3"]] = rng.uniform(-0.1, 0.1, W)
    p[L["w3"]:L["b3"]] = rng.uniform(-1, 1, W) * np.sqrt(3.0 / W)
    p[L["b3"]] = 0.3
    p[L["trained"]] = -7.0  # mu
    p[L["trained"] + 1] = 3.0  # sigma
    return p


def random_rows(n, kp, F, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, kp), dtype=np.float32)
    x[:, :F] = rng.normal(0, 0.6, (n, F))
    return O.f2bf(x)


def numpy_forward(kp, p, X):
    L = layout(kp)
    x = (X.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    relu = lambda z: np.maximum(z, 0.0)
    h = relu(x @ p[L["w0"]:L["b0"]].astype(np.float64).reshape(kp, W) + p[L["b0"]:L["w1"]])
    h = relu(h @ p[L["w1"]:L["b1"]].astype(np.float64).reshape(W, W) + p[L["b1"]:L["w2"]])
    h = relu(h @ p[L["w2"]:L["b2"]].astype(np.float64).reshape(W, W) + p[L["b2"]:L["w3"]])
    return h @ p[L["w3"]:L["b3"]].astype(np.float64) + p[L["b3"]]


def test_forward_exact_matches_numpy():
    F, kp = O.nn_dims(2)
    p = random_net(kp, F, 1)
    X = random_rows(50, kp, F, 2)
    y = O.nn_forward(kp, p, X, emulate=False)
    np.testing.assert_allclose(y, numpy_forward(kp, p, X), rtol=1e-12, atol=1e-12)


def test_forward_emulated_close_to_exact():
    F, kp = O.nn_dims(3)
    p = random_net(kp, F, 3)
    X = random_rows(200, kp, F, 4)
    ye = O.nn_forward(kp, p, X, emulate=True)
    yx = O.nn_forward(kp, p, X, emulate=False)
    assert np.max(np.abs(ye - yx)) < 0.05 * (1 + np.max(np.abs(yx)))
    assert np.any(ye != yx)  # the rounding is really applied


@pytest.mark.parametrize("D", [2, 3, 5])
def test_grad_exact_matches_finite_differences(D):
    F, kp = O.nn_dims(D)
    p = random_net(kp, F, 10 + D)
    X = random_rows(7, kp, F, 20 + D)
    t = np.random.default_rng(30 + D).normal(-7, 3, 7).astype(np.float32)
    g, loss = O.nn_grad(kp, p, X, t, emulate=False)
    # loss agrees with the forward pass
    y = O.nn_forward(kp, p, X, emulate=False)
    tn = (t.astype(np.float64) - p[-2]) / p[-1]
    assert abs(loss - np.mean(0.5 * (y - tn) ** 2)) < 1e-12 * max(1.0, loss)
    L = layout(kp)
    rng = np.random.default_rng(40 + D)
    picks = []
    for a, b in [("w0", "b0"), ("b0", "w1"), ("w1", "b1"), ("b1", "w2"), ("w2", "b2"), ("b2", "w3"), ("w3", "b3")]:
        lo, hi = L[a], L[b]
        if a == "w0":  # live input rows only
            hi = lo + F * W
        picks += list(rng.choice(np.arange(lo, hi), 12, replace=False))
    picks.append(L["b3"])
    ok = 0
    for i in picks:
        h = 1e-4 * max(1.0, abs(float(p[i])))
        pp, pm = p.copy(), p.copy()
        pp[i] = np.float32(p[i] + h)
        pm[i] = np.float32(p[i] - h)
        lp = O.nn_grad(kp, pp, X, t, emulate=False)[1]
        lm = O.nn_grad(kp, pm, X, t, emulate=False)[1]
        fd = (lp - lm) / (float(pp[i]) - float(pm[i]))
        if abs(fd - g[i]) <= 1e-5 + 1e-4 * abs(fd):
            ok += 1
    # a ReLU kink inside +-h can spoil an occasional difference quotient
    assert ok >= len(picks) - 2, (ok, len(picks))


def test_grad_emulated_close_to_exact():
    """BF16 rounding moves a single row's gradient by about its precision (2^-8); a
    pre-activation near 0 can flip a ReLU mask, which changes a row's input-layer
    gradient a lot, so the batch sum is held to a looser bound than the typical row."""
    F, kp = O.nn_dims(2)
    p = random_net(kp, F, 5)
    X = random_rows(150, kp, F, 6)
    t = np.random.default_rng(7).normal(-7, 3, 150).astype(np.float32)
    ge, le = O.nn_grad(kp, p, X, t, emulate=True)
    gx, lx = O.nn_grad(kp, p, X, t, emulate=False)
    assert abs(le - lx) < 0.02 * lx
    L = layout(kp)
    for a, b in [("w0", "b0"), ("w1", "b1"), ("w2", "b2"), ("w3", "b3")]:
        d = np.linalg.norm(ge[L[a]:L[b]] - gx[L[a]:L[b]]) / np.linalg.norm(gx[L[a]:L[b]])
        assert d < 0.15, (a, d)
    rows = []
    for r in range(30):
        g1 = O.nn_grad(kp, p, X[r:r + 1], t[r:r + 1], emulate=True)[0]
        g2 = O.nn_grad(kp, p, X[r:r + 1], t[r:r + 1], emulate=False)[0]
        rows.append(np.linalg.norm(g1 - g2) / np.linalg.norm(g2))
    assert np.median(rows) < 0.02, np.median(rows)


def test_region_features_hand_computed():
    # f = x^2 on the box [0.25, 0.5] x [0.5, 0.75] of [0,1]^2 (depth 2 in both dims),
    # bins 8 x 16 (log2: 3, 4)
    lo, hi = np.array([0.25, 0.5]), np.array([0.5, 0.75])
    v = np.zeros(9)
    for k in range(9):
        x = lo[0] + (k % 3) * 0.5 * (hi[0] - lo[0])
        v[k] = np.float32(x * x)
    code = np.array([0b101, 0b110], dtype=np.uint32)
    x, s, depth = O.region_features(2, lo, hi, [0, 0], [1, 1], v, code, lres=[3.0, 4.0])
    F, kp = O.nn_dims(2)
    assert F == 18 and kp == 32
    assert depth == 4 and s == np.float32(0.25)
    np.testing.assert_allclose(x[:6], [0.375, 0.625, -2 / 16, -2 / 16, 1 / 16, 2 / 16], rtol=1e-7)
    np.testing.assert_allclose(x[6:15], v / 0.25, rtol=1e-7)
    # |S - T| of a line p0, p1, p2 is |2 p1 - p0 - p2| / 3; Simpson weights sum to 1
    st = abs(2 * 0.375 ** 2 - 0.25 ** 2 - 0.5 ** 2) / 3
    np.testing.assert_allclose(x[15], st / 0.25, rtol=1e-6)
    assert x[16] == 0.0
    np.testing.assert_allclose(x[17], 4 / 32, rtol=1e-7)
    assert np.all(x[F:] == 0)


def _nodes(f, lo, hi):
    pts = np.array([[lo[d] + ((k // 3 ** d) % 3) * 0.5 * (hi[d] - lo[d]) for d in range(len(lo))]
                    for k in range(3 ** len(lo))])
    return O.evaluate(f, pts)


def _quad_at(v, lo, hi, x):
    """tensor quadratic through nodes v on [lo, hi] at points x (n, 2)"""
    t = (x - lo) / (hi - lo)
    L = [lambda u: 2 * (u - 0.5) * (u - 1), lambda u: -4 * u * (u - 1), lambda u: 2 * u * (u - 0.5)]
    out = np.zeros(x.shape[0])
    for k in range(9):
        out += v[k] * L[k % 3](t[:, 0]) * L[k // 3](t[:, 1])
    return out


def _brute_binvar(f, v, lo, hi, bins, step, m=96):
    """sum over the given bins of within-bin variance of f - P, midpoint rule, m x m per bin"""
    V = 0.0
    for (i, j) in bins:
        a = np.maximum(lo, [i * step[0], j * step[1]])
        b = np.minimum(hi, [(i + 1) * step[0], (j + 1) * step[1]])
        if np.any(a >= b):
            continue
        u = (np.arange(m) + 0.5) / m
        X, Y = np.meshgrid(a[0] + u * (b[0] - a[0]), a[1] + u * (b[1] - a[1]), indexing="ij")
        pts = np.stack([X.ravel(), Y.ravel()], 1)
        g = O.evaluate(f, pts) - _quad_at(v, lo, hi, pts)
        w = np.prod(b - a) / (m * m)
        V += np.sum(w * g * g) - (np.sum(w * g)) ** 2 / np.prod(b - a)
    return V


def test_bin_variances_against_brute_force():
    """Within-bin residual variance of a region and of its two children against
    midpoint sums on fine grids (a smooth integrand, so both rules converge)."""
    f = I.blobs([(0.45, 0.3, 0, 0, 18.0, 30.0, 1.0), (0.7, 0.1, 0, 0, 9.0, 4.0, 0.6)], dim=2, step_mode=0,
                wave=(0.2, 7.0, 5.0))
    lo, hi = np.array([0.25, 0.0]), np.array([0.75, 0.5])
    res = (8, 8)
    step = np.array([1 / 8, 1 / 8])
    v = _nodes(f, lo, hi)
    bins = [(i, j) for i in range(2, 6) for j in range(0, 4)]
    Vp, _ = O.bin_variances(f, res, lo, hi, v, 0, maxb=64)
    want = _brute_binvar(f, v, lo, hi, bins, step)
    assert abs(Vp - want) < 2e-3 * want, (Vp, want)
    # children split along dim 1
    mid = 0.25
    c0lo, c0hi = lo.copy(), np.array([0.75, mid])
    c1lo, c1hi = np.array([0.25, mid]), hi.copy()
    v0, v1 = _nodes(f, c0lo, c0hi), _nodes(f, c1lo, c1hi)
    Vp2, Vc = O.bin_variances(f, res, lo, hi, v, 1, v0, v1, maxb=64)
    assert Vp2 == Vp
    want_c = _brute_binvar(f, v0, c0lo, c0hi, bins, step) + _brute_binvar(f, v1, c1lo, c1hi, bins, step)
    assert abs(Vc - want_c) < 2e-3 * want_c, (Vc, want_c)
    assert Vc < Vp
    # sampled bins: 4 of the 16 bins per dim at i0 + floor((k + 1/2) * 16 / 4), scaled by 4 per dim
    res2 = (32, 32)
    Vs, _ = O.bin_variances(f, res2, lo, hi, v, 0, maxb=4)
    sel_i = [8 + int((k + 0.5) * 16 / 4) for k in range(4)]
    sel_j = [0 + int((k + 0.5) * 16 / 4) for k in range(4)]
    want_s = 16 * _brute_binvar(f, v, lo, hi, [(i, j) for i in sel_i for j in sel_j], np.array([1 / 32, 1 / 32]))
    assert abs(Vs - want_s) < 2e-3 * want_s, (Vs, want_s)


def test_true_error_closed_forms():
    # f = x^3 on [0,1]^2: f - P = x (x - 1/2) (x - 1), so e = int_0^1 (x(x-1/2)(x-1))^2 dx
    f = I.poly([(1.0, (3, 0))], 2)
    lo, hi = np.zeros(2), np.ones(2)
    v = np.array([((k % 3) * 0.5) ** 3 for k in range(9)])
    g = np.polynomial.Polynomial([0, 0.5, -1.5, 1.0])
    exact = (g * g).integ()(1.0)
    assert abs(O.true_error(f, lo, hi, v, 40) - exact) < 1e-9 * exact
    assert abs(O.true_error(f, lo, hi, v, 3) - exact) < 1e-3 * exact
    # a tensor quadratic is its own control variate: e = 0
    f2 = I.poly([(2.0, (2, 2)), (-1.0, (1, 0)), (0.5, (0, 0))], 2)
    v2 = np.array([2 * ((k % 3) * 0.5) ** 2 * ((k // 3) * 0.5) ** 2 - (k % 3) * 0.5 + 0.5 for k in range(9)])
    assert O.true_error(f2, lo, hi, v2, 3) < 1e-28
    # scaling: the box [0, 1/2]^2 shrinks e of x^3 by vol * (1/2)^6
    v3 = np.array([((k % 3) * 0.25) ** 3 for k in range(9)])
    e3 = O.true_error(f, lo, hi * 0.5, v3, 40)
    assert abs(e3 - exact * 0.25 * 0.5 ** 6) < 1e-9 * e3

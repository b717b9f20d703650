#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference build
(oracle/_ref/libviltrum_ref.so: adolfomunoz/viltrum's own headers, compiled in this
container by oracle/Makefile).  Only numbers are stored (npz, float64 holds every FP32
and FP64 value exactly); nothing of the reference travels.

    make -C oracle && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle_lib as O  # noqa: E402
from viltrum_amd import integrands as I  # noqa: E402


def integrand_cases():
    return {
        "config1": I.config1(),
        "blobs3d": I.moving_blobs(3, 8),
        "c1fam": I.c1_family(5, 6),
        "poly2": I.poly([(1.0, (3, 1)), (-2.0, (1, 4)), (0.5, (0, 0))], 2, step=([1.0, 1.0], 1.0, 0.5)),
        "poly3": I.poly([(1.0, (2, 1, 3)), (0.25, (0, 5, 1))], 3, step=([1.0, -0.5, 0.7], 0.3, 1.0)),
        "lobes5": I.config5(2),
        "poly6": I.poly([(1.0, (1, 2, 0, 1, 0, 2)), (0.5, (0, 0, 3, 0, 1, 0)), (0.25, (2, 0, 0, 0, 0, 1))], 6,
                        step=([1, 1, 1, 1, 1, 1], 3.0, 0.5)),
    }


REGION_CASES = [  # (integrand, scalar, heuristic, iterations)
    ("config1", 8, 0, 0), ("config1", 8, 0, 1), ("config1", 8, 0, 2), ("config1", 8, 0, 7), ("config1", 8, 0, 63),
    ("config1", 8, 0, 1023), ("config1", 4, 0, 1023), ("config1", 8, 1, 63), ("config1", 4, 1, 255),
    ("blobs3d", 8, 0, 300), ("blobs3d", 4, 0, 300), ("c1fam", 8, 0, 400), ("poly2", 8, 0, 200),
    ("poly2", 8, 1, 200), ("poly3", 8, 0, 150), ("lobes5", 8, 0, 40), ("lobes5", 4, 0, 40),
    ("poly6", 8, 0, 60), ("poly6", 4, 0, 40), ("poly6", 8, 1, 30),
]

BIN_CASES = [  # (integrand, scalar, form, heuristic, res, iterations, spp, seed, alpha)
    ("config1", 8, 0, 0, (16, 16), 1023, 0, 0, 1.0),
    ("config1", 8, 0, 0, (64, 64), 1023, 0, 0, 1.0),
    ("config1", 4, 0, 0, (64, 64), 1023, 0, 0, 1.0),
    ("config1", 8, 0, 0, (32, 32), 255, 16, 7, 1.0),
    ("config1", 4, 0, 0, (32, 32), 255, 16, 7, 1.0),
    ("config1", 8, 1, 0, (24, 20), 200, 8, 3, 0.5),
    ("config1", 8, 2, 1, (16, 16), 127, 4, 11, 1.0),
    ("config1", 8, 0, 0, (40,), 100, 5, 2, 1.0),
    ("blobs3d", 8, 0, 0, (8, 8), 300, 4, 5, 1.0),
    ("blobs3d", 8, 0, 0, (4, 4, 4), 100, 3, 5, 1.0),
    ("poly3", 8, 0, 0, (8, 8), 150, 0, 0, 1.0),
    ("lobes5", 8, 0, 0, (6, 6), 40, 2, 9, 1.0),
    # region_sampling_importance (forms 3 and 4: optimised and fixed weight; float only)
    ("config1", 4, 3, 0, (16, 16), 255, 8, 5, 1.0),
    ("config1", 4, 4, 0, (8, 8), 63, 16, 3, 0.7),
    ("config1", 4, 3, 0, (40,), 100, 5, 2, 1.0),
    ("poly2", 4, 3, 0, (8, 8), 100, 8, 5, 1.0),
    ("poly3", 4, 4, 0, (4, 4), 200, 4, 2, 1.0),
    ("blobs3d", 4, 3, 0, (8, 8), 300, 8, 5, 1.0),  # regions of zeros: the reference's 0/0 pdf gives NaN bins
    # rr_integral_region (forms 5 and 6: optimised and fixed weight)
    ("config1", 8, 5, 0, (16, 16), 255, 8, 5, 1.0),
    ("config1", 4, 6, 0, (8, 8), 63, 16, 3, 0.7),
    ("blobs3d", 8, 5, 0, (8, 8), 300, 4, 5, 1.0),
    ("lobes5", 4, 5, 0, (3, 3), 60, 4, 2, 1.0),
    # rr_error_region (forms 7 and 8)
    ("config1", 4, 7, 0, (16, 16), 255, 8, 5, 1.0),
    ("poly3", 8, 8, 0, (4, 4), 200, 4, 2, 0.7),
    ("lobes5", 8, 7, 0, (3, 3), 60, 4, 2, 1.0),
    # rr_pdf_region (forms 9 and 10)
    ("config1", 4, 9, 0, (16, 16), 255, 8, 5, 1.0),
    ("poly3", 8, 10, 0, (4, 4), 200, 4, 2, 0.7),
    ("blobs3d", 4, 9, 0, (8, 8), 300, 4, 5, 1.0),
    ("lobes5", 8, 9, 0, (3, 3), 60, 4, 2, 1.0),
    # a roulette with region_sampling_importance (forms 11..16; float only)
    ("config1", 4, 11, 0, (16, 16), 255, 8, 5, 1.0),
    ("poly3", 4, 14, 0, (4, 4), 200, 4, 2, 0.7),
    ("config1", 4, 15, 0, (16, 16), 255, 8, 5, 1.0),
    ("blobs3d", 4, 16, 0, (8, 8), 300, 4, 5, 0.5),
    ("config1", 4, 15, 0, (40,), 100, 5, 2, 1.0),
    # region_sampling_russian_roulette (forms 17 and 18; float only)
    ("config1", 4, 17, 0, (16, 16), 255, 8, 5, 1.0),
    ("config1", 4, 18, 0, (8, 8), 63, 16, 3, 0.7),
    ("poly3", 4, 17, 0, (4, 4), 200, 4, 2, 1.0),
    ("blobs3d", 4, 18, 0, (8, 8), 300, 4, 5, 0.5),
    ("config1", 4, 17, 0, (40,), 100, 5, 2, 1.0),
    # importance sampling and the sampling roulette at D = 5
    ("lobes5", 4, 3, 0, (3, 3), 60, 4, 2, 1.0),
    ("lobes5", 4, 18, 0, (3, 3), 60, 4, 2, 0.7),
    ("lobes5", 4, 15, 0, (3, 3), 60, 4, 2, 1.0),
    # importance sampling and the sampling roulette at D = 6 (the reference's region cap)
    ("poly6", 4, 3, 0, (2, 2), 30, 3, 2, 1.0),
    ("poly6", 4, 18, 0, (2, 2), 30, 3, 2, 0.7),
]

FUBINI_CASES = [  # (integrand, region dims, res, iterations, mc samples, spp, seed): integrator_crespo2021_infinite
    ("blobs3d", 2, (8, 8), 150, 3, 4, 11),
    ("blobs3d", 2, (12, 10), 60, 1, 0, 4),
    ("lobes5", 2, (5, 6), 60, 2, 3, 5),
    ("poly2", 1, (12,), 80, 5, 6, 2),
]


def fubini_key(name, rd, res, it, mc, spp, seed):
    return "fubini/%s/rd%d/r%s/it%d/mc%d/spp%d/seed%d" % (name, rd, "x".join(map(str, res)), it, mc, spp, seed)


# the Dyson series over an infinite range: (form (0 integrator_crespo2021_infinite, 1 / 2
# the optimised Fubini integrator with cv_optimize_weight / cv_fixed_weight(alpha)),
# decay, region dims, res, iterations, mc samples, spp, seed, alpha, region range
# (lo, hi) or None for [0, 1], explicit rest entries (lo, hi))
FUBINI_INF_CASES = [
    (0, 0.75, 4, (10,), 1280, 8, 256, 3, 1.0, None, ((), ())),  # main/compilation-tests/crespo21.cc
    (0, 0.6, 1, (12,), 100, 3, 8, 5, 1.0, None, ((0.1,), (0.8,))),
    (0, 0.75, 2, (6, 5), 200, 2, 16, 7, 1.0, ((-0.5, 0.0), (1.0, 2.0)), ((), ())),
    (1, 0.75, 2, (4, 4), 150, 4, 32, 9, 1.0, None, ((), ())),
    (1, 0.5, 4, (3, 2), 120, 2, 5, 2, 1.0, None, ((0.2, 0.0), (0.7, 0.5))),  # spp below most bins' region counts
    (2, 0.8, 1, (9,), 60, 5, 12, 4, 0.4, None, ((), ())),
    (2, 0.75, 2, (8,), 300, 3, 7, 11, 0.7, ((0.0, 0.25), (1.0, 0.75)), ((0.3,), (0.9,))),
]


def fubini_inf_key(form, decay, rd, res, it, mc, spp, seed, alpha, box, rest):
    return "fubini_inf/f%d/d%g/rd%d/r%s/it%d/mc%d/spp%d/seed%d/a%g/box%s/rest%s" % (
        form, decay, rd, "x".join(map(str, res)), it, mc, spp, seed, alpha, "" if box is None else repr(box),
        repr(rest))


def run_fubini_inf(case, **kw):
    form, decay, rd, res, it, mc, spp, seed, alpha, box, rest = case
    lo, hi = (None, None) if box is None else box
    return O.integrate_fubini_inf(decay, rd, res, it, mc, spp, seed, form=form, alpha=alpha, lo=lo, hi=hi,
                                  rest_lo=rest[0], rest_hi=rest[1], **kw)


def main():
    if O.ref() is None:
        sys.exit("reference build missing: make -C oracle (needs /root/reference)")
    fs = integrand_cases()
    out = {}
    for name, scalar, heur, it in REGION_CASES:
        f = fs[name]
        g = O.generate(f, it, scalar=scalar, heuristic=heur, use_ref=True, evals=True)
        key = "regions/%s/s%d/h%d/it%d" % (name, scalar, heur, it)
        out[key + "/lo"] = g.lo
        out[key + "/hi"] = g.hi
        out[key + "/err"] = g.err
        out[key + "/dim"] = g.dim
        out[key + "/evals"] = g.evals
    for name, scalar, form, heur, res, it, spp, seed, alpha in BIN_CASES:
        f = fs[name]
        b, _ = O.integrate(f, res, it, spp, seed, scalar=scalar, integrator=form, heuristic=heur, alpha=alpha,
                           use_ref=True)
        key = "bins/%s/s%d/f%d/h%d/r%s/it%d/spp%d/seed%d/a%g" % (name, scalar, form, heur, "x".join(map(str, res)),
                                                                it, spp, seed, alpha)
        out[key] = b
    for name, rd, res, it, mc, spp, seed in FUBINI_CASES:
        b, _ = O.integrate_fubini(fs[name], rd, res, it, mc, spp, seed, use_ref=True)
        out[fubini_key(name, rd, res, it, mc, spp, seed)] = b
    for case in FUBINI_INF_CASES:
        b, _ = run_fubini_inf(case, use_ref=True)
        out[fubini_inf_key(*case)] = b
    path = os.path.join(HERE, "reference_fixtures.npz")
    np.savez_compressed(path, **out)
    print("wrote %s (%d arrays, %.0f KB)" % (path, len(out), os.path.getsize(path) / 1024))


if __name__ == "__main__":
    main()

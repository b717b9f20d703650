"""Learned refinement priority: training sets of the learned-refinement workload and
helpers over the C ABI (vmx_refiner_*).

The network is trained on the GPU by the library's own kernels on (region, FP64 true
error) pairs that the heuristic method generates for seeded config-1 style integrands;
held-out integrands use a disjoint seed range.  There is no reference counterpart (the
reference orders splits by its error heuristic only, src/nested/error-heuristic.h:21-47).
"""
import numpy as np

from . import F32, REFINER_HEURISTIC, REFINER_LEARNED, integrator_crespo2021, range_primary
from . import integrands as I

TRAIN_SEED0 = 1_000_000   # training integrands: seeds TRAIN_SEED0 + i
EVAL_SEED0 = 2_000_000    # held-out integrands: seeds EVAL_SEED0 + i


def training_set(count, regions, res=(512, 512), seed0=TRAIN_SEED0):
    """Problems (heuristic method) and integrands whose generated regions form the
    training pairs.  Budgets cycle through regions/4, regions/2, regions, 2*regions so
    the pairs cover region sizes around the evaluation budgets; `res` is the bin grid
    the targets are measured on."""
    fs = [I.config4_integrand(seed0 + i) for i in range(count)]
    ps = []
    for i, f in enumerate(fs):
        n = max(2, int(regions * (0.25, 0.5, 1.0, 2.0)[i % 4]))
        ps.append(integrator_crespo2021(n - 1, 0, 0).problem(f, res, range_primary(2), F32))
    return ps, fs


def train(ctx, count=512, regions=4096, res=(512, 512), epochs=8, batch=1024, lr=2e-3, seed=0, target=0, stride=1):
    """Trains the D=2 network of `ctx` on the pairs of `count` integrands; returns the
    training report."""
    ps, fs = training_set(count, regions, res)
    return ctx.refiner_train(ps, fs, epochs=epochs, batch=batch, lr=lr, seed=seed, target=target, stride=stride)


def held_out(count, seed0=EVAL_SEED0):
    return [I.config4_integrand(seed0 + i) for i in range(count)]


def integrator(iterations, spp, seed=0, learned=True):
    """integrator_crespo2021 with the learned priority (or the heuristic)."""
    return integrator_crespo2021(iterations, spp, seed, refiner=REFINER_LEARNED if learned else REFINER_HEURISTIC)


def save(path, params):
    np.save(path, np.asarray(params, dtype=np.float32), allow_pickle=False)


def load(path):
    return np.load(path, allow_pickle=False).astype(np.float32)

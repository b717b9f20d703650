"""Learned refinement priority: training sets of the learned-refinement workload and
helpers over the C ABI (vmx_refiner_*).

The network is trained on the GPU by the library's own kernels on (region, FP64 true
error) pairs that the heuristic method generates for seeded integrands of the workload's
family: 2-D config-1 style blobs (dim 2) or 5-D lobes x steps x cosine (dim 5); held-out
integrands use a disjoint seed range.  There is no reference counterpart (the
reference orders splits by its error heuristic only, src/nested/error-heuristic.h:21-47).
"""
import numpy as np

from . import F32, REFINER_HEURISTIC, REFINER_LEARNED, integrator_crespo2021, range_primary
from . import integrands as I

TRAIN_SEED0 = 1_000_000   # training integrands: seeds TRAIN_SEED0 + i
EVAL_SEED0 = 2_000_000    # held-out integrands: seeds EVAL_SEED0 + i


def integrator(iterations, spp, seed=0, learned=True):
    """integrator_crespo2021 with the learned priority (or the heuristic)."""
    return integrator_crespo2021(iterations, spp, seed, refiner=REFINER_LEARNED if learned else REFINER_HEURISTIC)


def family(dim):
    """The seeded integrand family the dim-`dim` network is trained and evaluated on."""
    if dim == 2:
        return I.config4_integrand
    if dim == 5:
        return I.config5
    raise ValueError("training sets exist for dim 2 (blobs, step, wave) and dim 5 (lobes x steps x cosine)")


def training_set(count, regions, res=(512, 512), seed0=TRAIN_SEED0, dim=2, learned_odd=False):
    """Problems and integrands whose generated regions form the training pairs.  Budgets
    cycle through regions/4, regions/2, regions, 2*regions so the pairs cover region
    sizes around the evaluation budgets; `res` is the bin grid the targets are measured
    on.  Every problem uses the heuristic priority, or with learned_odd every other one
    uses the context's current network (its own regions become pairs)."""
    make = family(dim)
    fs = [make(seed0 + i) for i in range(count)]
    ps = []
    for i, f in enumerate(fs):
        n = max(2, int(regions * (0.25, 0.5, 1.0, 2.0)[i % 4]))
        it = integrator(n - 1, 0, 0, learned=learned_odd and i % 2 == 1)
        ps.append(it.problem(f, res, range_primary(dim), F32))
    return ps, fs


def train(ctx, count=512, regions=4096, res=(512, 512), epochs=8, batch=1024, lr=2e-3, seed=0, target=0, stride=1,
          dim=2, on_policy=0):
    """Trains the dim-`dim` network of `ctx` on the pairs of `count` integrands; returns
    the training report (times summed over rounds).  Each of the `on_policy` further
    rounds retrains from scratch on pairs of which half come from trees the previous
    round's network builds: the network then also sees the regions its own priority
    keeps, which the heuristic's trees rarely hold."""
    total = {"ms_pairs": 0.0, "ms_train": 0.0}
    for r in range(on_policy + 1):
        ps, fs = training_set(count, regions, res, dim=dim, learned_odd=r > 0)
        rep = ctx.refiner_train(ps, fs, epochs=epochs, batch=batch, lr=lr, seed=seed + r, target=target,
                                stride=stride)
        total["ms_pairs"] += rep["ms_pairs"]
        total["ms_train"] += rep["ms_train"]
    rep.update(total)
    rep["rounds"] = on_policy + 1
    return rep


def held_out(count, seed0=EVAL_SEED0, dim=2):
    make = family(dim)
    return [make(seed0 + i) for i in range(count)]


def save(path, params):
    np.save(path, np.asarray(params, dtype=np.float32), allow_pickle=False)


def load(path):
    return np.load(path, allow_pickle=False).astype(np.float32)

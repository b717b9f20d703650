#!/usr/bin/env python3
"""Benchmark of the per-bin control-variate path on one MI355X.  Prints ONE JSON line.

A step is one full integration pass (adaptive regions, control-variate integral and
residual Monte Carlo per bin) over the workload, integrand parameters resident and bins
written to device memory.

Workloads:
  config4 (default)  64 held-out 2-D config-1 style integrands at 512x512 FP32 bins:
                     the learned-priority batched path against the plain GPU path at
                     equal RMSE vs FP64 closed-form bins (value = learned+batched bins/s)
  config2            3-D moving-blob integrand (32 blobs x step occluder), 512x512 bins,
                     16384 regions, 64 spp, plain GPU path
  config3            256 uneven integrals in one batched call vs one call each
  config5            16 5-D integrals at 1024x1024 bins, 65536 regions, 32 spp, one
                     batched call with the learned priority (heuristic beside it)
  fubini             integrator_crespo2021_infinite<2> on a 3-D integrand
  importance         importance-sampled residual vs uniform, 64 integrals at 256x256 bins
  config1            2-D config-1 integrand, 64x64 bins, 1024 regions, 16 spp
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "bins/sec at fixed RMSE vs FP64 oracle: learned+batched vs plain GPU path"
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 peak
# HBM bytes per launch of k_bins1 from two rocprofv3 PMC passes (FETCH_SIZE x2.0,
# WRITE_SIZE x1.0: the factors a 256 MiB device copy needs to read its known size), at
# the profiled point of each workload; the commands and raw counts are in the profiles
# named in TRAFFIC_SOURCE
TRAFFIC_BYTES = {"config4": (131929 * 2.0 + 69097) * 1024, "config2": (8274 * 2.0 + 3072) * 1024}
TRAFFIC_SOURCE = {
    "config4": "PMC FETCH_SIZE/WRITE_SIZE of k_bins1 at this line's point (4096 regions, 16 spp) with the network "
               "this bench trains (profiles/k_bins1_config4_learned_batched.md)",
    "config2": "PMC FETCH_SIZE/WRITE_SIZE of k_bins1 at 16384 regions, 64 spp (profiles/k_bins1_config2_plain.md)",
}
PEAK_HBM_GBS = 8000.0


def workloads():
    from viltrum_amd import integrands as I
    return {
        "config2": dict(desc="3-D moving-blob integrand (32 blobs x step occluder), 512x512 fp32 bins, "
                             "16384 regions, 64 residual spp, seed 0, plain GPU path",
                        f=I.config2(), res=(512, 512), regions=16384, spp=64, seed=0),
        "config1": dict(desc="2-D config-1 integrand, 64x64 fp32 bins, 1024 regions, 16 spp, seed 0",
                        f=I.config1(), res=(64, 64), regions=1024, spp=16, seed=0),
    }


TRUTH_REGIONS, TRUTH_SPP = 1 << 18, 1024


def fp64_truth(ctx, V, f, res):
    """FP64 truth of the per-bin integrals of f on the grid res.
    - 2-D static blob integrands (blobs, additive step, separable wave): closed form.
    - Otherwise: the mean of two FP64 control-variate estimates with weight 1 (exactly
      unbiased: the control variate is the exact integral of each region's polynomial),
      2^18 regions and 1024 spp, seeds 7001 and 7002.  Its per-bin variance, estimated
      from the seed pair as mean((t1 - t2)^2) / 4, is the truth's own noise; the RMSE
      figures subtract it.
    Returns (truth, mean per-bin variance of the truth, description)."""
    from viltrum_amd import integrands as I
    p = f.params
    if f.family == I.FAMILY_BLOBS and f.dim == 2 and all(p.blob[k][2] == 0 and p.blob[k][3] == 0
                                                         for k in range(p.nblobs)) and p.step_mode != 2:
        return ctx.exact_bins(f, res), 0.0, {"kind": "FP64 closed-form per-bin integrals (erf, polygon area, "
                                                     "separable wave)"}
    runs = []
    for seed in (7001, 7002):
        b = np.zeros(res, dtype=np.float64)
        V.integrate(V.integrator_variance_reduction_fixed(1.0, TRUTH_REGIONS - 1, TRUTH_SPP, seed=seed), b, f,
                    V.range_primary(f.dim), ctx=ctx)
        runs.append(b)
    var = float(np.mean((runs[0] - runs[1]) ** 2) / 4)
    return (runs[0] + runs[1]) / 2, var, {
        "kind": "FP64 control-variate estimate with weight 1 (unbiased), %d regions, %d spp, mean of 2 seeds; "
                "its own per-bin variance (from the seed pair) is subtracted in rmse_less_truth_noise"
                % (TRUTH_REGIONS, TRUTH_SPP), "stderr_rms": float(np.sqrt(var))}


def rmse_pair(est, truth, truth_var):
    """(RMSE against the truth, the same with the truth's own noise removed)."""
    mse = float(np.mean((np.asarray(est, dtype=np.float64) - truth) ** 2))
    return float(np.sqrt(mse)), float(np.sqrt(max(0.0, mse - truth_var)))


def run_config3(args, V, hip):
    """256 integrals of uneven resolution, budget and spp (2-D config-1 family and 3-D
    moving blobs, 4..64 blobs) in one batched call, against the same integrals one call
    each.  value = total bins / batched time."""
    from viltrum_amd import integrands as I
    items = I.config3_batch(seed=0, count=256)
    ctx = V.Context(0)
    probs, fs, bufs = [], [], []
    for it in items:
        f = it["integrand"]
        integ = V.integrator_crespo2021(it["regions"] - 1, it["spp"], seed=len(probs))
        probs.append(integ.problem(f, it["res"], V.range_primary(f.dim), V.F32))
        fs.append(f)
        bufs.append(hip.DeviceBuffer(int(np.prod(it["res"])) * 4))
    stream = hip.Stream()
    total_bins = sum(int(np.prod(it["res"])) for it in items)
    ptrs = [b.address for b in bufs]
    tune = ctx.autotune(probs, fs, ptrs, stream.address)
    for _ in range(args.warmup):
        ctx.integrate_batch(probs, fs, ptrs, stream.address)
    hip.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = ctx.integrate_batch(probs, fs, ptrs, stream.address)
    hip.device_synchronize()
    ms_batch = (time.perf_counter() - t0) * 1e3 / args.steps
    t0 = time.perf_counter()
    for p, f, ptr in zip(probs, fs, ptrs):
        ctx.integrate_device(p, f, ptr, stream.address)
    hip.device_synchronize()
    ms_seq = (time.perf_counter() - t0) * 1e3
    evals = sum(x.evals_regions + x.evals_residual for x in st)
    # RMSE per integral against its FP64 truth (closed form for the 2-D ones), and pooled
    # over all bins of the batch
    ctx.integrate_batch(probs, fs, ptrs, stream.address)
    hip.device_synchronize()
    per, se_all, var_all, n_all = [], 0.0, 0.0, 0
    for it, f, b in zip(items, fs, bufs):
        out = b.to_host(np.zeros(it["res"], dtype=np.float32))
        truth, tvar, _ = fp64_truth(ctx, V, f, it["res"])
        r, rl = rmse_pair(out, truth, tvar)
        nb = out.size
        se_all += r * r * nb
        var_all += tvar * nb
        n_all += nb
        per.append({"dim": f.dim, "res": it["res"][0], "regions": it["regions"], "spp": it["spp"],
                    "rmse": r, "rmse_less_truth_noise": rl})
    pooled = float(np.sqrt(se_all / n_all))
    pooled_less = float(np.sqrt(max(0.0, (se_all - var_all) / n_all)))
    res = {
        "metric": METRIC, "value": total_bins / (ms_batch * 1e-3), "unit": "bins/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_batch, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "256 seeded integrals (2-D config-1 family / 3-D moving blobs, 4..64 blobs), "
                               "res 16^2..256^2, budgets logU[256,32768], spp {4,16,64}, one batched call",
                   "integrals": len(items), "bins": total_bins, "path": "batched"},
        "evals_per_s": evals / (ms_batch * 1e-3),
        "sequential": {"ms": ms_seq, "bins_per_s": total_bins / (ms_seq * 1e-3)},
        "batch_speedup": ms_seq / ms_batch,
        "rmse_vs_fp64": {"pooled": pooled, "pooled_less_truth_noise": pooled_less,
                         "per_integral_less_truth_noise": {
                             "median": float(np.median([q["rmse_less_truth_noise"] for q in per])),
                             "max": float(max(q["rmse_less_truth_noise"] for q in per))},
                         "per_integral": per},
        "fp64_truth": {"kind": "2-D integrals: FP64 closed-form bins; 3-D integrals: FP64 control-variate "
                               "estimate with weight 1, %d regions, %d spp, mean of 2 seeds, its own noise "
                               "removed in the *_less_truth_noise figures" % (TRUTH_REGIONS, TRUTH_SPP)},
        "autotune": tune,
        "phase_ms": {k: float(getattr(st[0], k)) for k in ("ms_generate", "ms_select", "ms_stratify", "ms_bins")},
    }
    print(json.dumps(res), flush=True)


def trained(ctx, R, args, dim, train):
    """The context's dim-`dim` priority network: trained on the GPU (`train()`), or loaded
    from --load-net (a file an earlier --save-net run wrote: training is deterministic, so
    it is the network the same command trains).  Returns the training report."""
    if args.load_net:
        ctx.refiner_set(dim, R.load(args.load_net))
        return {"loaded": args.load_net}
    rep = train()
    if args.save_net:
        R.save(args.save_net, ctx.refiner_get(dim))
    return rep


def run_config4(args, V, hip):
    """Learned refinement + batching against the plain GPU path at equal RMSE.

    Setup (untimed): train the priority network on the GPU (pairs the heuristic method
    generates on 512 seeded config-1 style integrands), then take 64 held-out integrands
    at 512x512 FP32 bins with their FP64 closed-form bins.  Target RMSE: the plain path
    (heuristic priority, one integral per call) at 16384 regions and 16 spp.  Each path
    then picks, over region budgets {2048, 4096, 8192, 16384}, the spp that reaches the target
    (RMSE ~ 1/sqrt(spp) from a 16-spp probe, then checked and stepped) and the faster
    of the budgets.  A step is one learned+batched pass over the 64 integrals at its
    chosen (budget, spp); value = their bins / step time.  The plain path at its chosen
    point is timed the same way beside it, and so are the two changes apart: batching
    with the heuristic priority at plain's point, and the learned priority one call per
    integral at its own chosen point."""
    from viltrum_amd import integrands as I
    from viltrum_amd import refiner as R
    ctx = V.Context(0)
    rep = trained(ctx, R, args, 2, lambda: R.train(ctx, count=args.train_integrands, regions=args.train_regions,
                                                   epochs=args.train_epochs, seed=0, stride=args.train_stride))
    fs = R.held_out(args.integrals)
    res = (512, 512)
    nb = res[0] * res[1]
    truths = [ctx.exact_bins(f, res) for f in fs]
    bufs = [hip.DeviceBuffer(nb * 4) for _ in fs]
    ptrs = [b.address for b in bufs]
    stream = hip.Stream()

    def problems(regions, spp, learned):
        it = R.integrator(regions - 1, spp, seed=0, learned=learned)
        return [it.problem(f, res, V.range_primary(2), V.F32) for f in fs]

    def run_plain(ps):
        return [ctx.integrate_device(p, f, ptr, stream.address) for p, f, ptr in zip(ps, fs, ptrs)]

    def run_batch(ps):
        return ctx.integrate_batch(ps, fs, ptrs, stream.address)

    def rmse():
        hip.device_synchronize()
        se = 0.0
        out = np.zeros(res, dtype=np.float32)
        for b, t in zip(bufs, truths):
            b.to_host(out)
            se += float(np.sum((out.astype(np.float64) - t) ** 2))
        return float(np.sqrt(se / (nb * len(fs))))

    def timed(fn, ps, steps):
        for _ in range(args.warmup):
            fn(ps)
        hip.device_synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st = fn(ps)
        hip.device_synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, st

    if args.only_timed:
        # the learned+batched step alone at a given (regions, spp), launch shapes from the
        # VMX_* environment (no tuning runs): for kernel profiles of the timed step
        regions, spp = (int(x) for x in args.only_timed.split(","))
        ms, st = timed(run_batch, problems(regions, spp, True), args.steps)
        flops = sum(st[i].evals_residual * I.flops_per_eval(f)[0] for i, f in enumerate(fs))
        ms_k = float(st[0].ms_bins)
        print(json.dumps({
            "metric": METRIC, "value": nb * len(fs) / (ms * 1e-3), "unit": "bins/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "learned+batched step only: %d held-out integrands, 512x512 fp32 bins, "
                                   "%d regions, %d spp" % (len(fs), regions, spp), "path": "learned+batched"},
            "rmse_vs_fp64": rmse(),
            "phase_ms": {k: float(getattr(st[0], k)) for k in ("ms_generate", "ms_select", "ms_stratify",
                                                                "ms_bins", "ms_total")},
            "roofline": {"bound": "valu", "kernel": "k_bins1 (per-bin CV + residual Monte Carlo, batched)",
                         "achieved": flops / (ms_k * 1e-3) / 1e12, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "frac": flops / (ms_k * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
                         "traffic": TRAFFIC_BYTES["config4"],
                         "kernel_ms": ms_k}}), flush=True)
        return
    budgets = (2048, 4096, 8192, 16384)
    # launch shapes of each path's workload class, picked by the library's autotuner
    # (setup, untimed, like the training)
    tune = {"learned_batched": ctx.autotune(problems(4096, 16, True), fs, ptrs, stream.address),
            "heuristic_batched": ctx.autotune(problems(4096, 16, False), fs, ptrs, stream.address),
            "plain": ctx.autotune(problems(4096, 16, False)[:1], fs[:1], ptrs[:1], stream.address),
            "learned_unbatched": ctx.autotune(problems(4096, 16, True)[:1], fs[:1], ptrs[:1], stream.address)}
    run_plain(problems(16384, 16, False))
    target = rmse()

    def choose(fn, learned):
        """(regions, spp, rmse, ms) per budget: the smallest spp whose RMSE <= target."""
        pts = {}
        for n in budgets:
            seen = {}
            fn(problems(n, 16, learned))
            seen[16] = rmse()
            spp = max(1, int(np.ceil(16 * (seen[16] / target) ** 2)))
            for _ in range(12):
                if spp not in seen:
                    fn(problems(n, spp, learned))
                    seen[spp] = rmse()
                if seen[spp] <= target:
                    if spp > 1 and (spp - 1) not in seen:
                        fn(problems(n, spp - 1, learned))
                        seen[spp - 1] = rmse()
                    if spp > 1 and seen[spp - 1] <= target:
                        spp -= 1
                        continue
                    break
                spp = max(spp + 1, int(np.ceil(spp * 1.1)))
            ms, _ = timed(fn, problems(n, spp, learned), 6)  # enough steps to rank budgets a few % apart
            pts[n] = dict(regions=n, spp=spp, rmse=seen[spp], ms=ms)
        return pts

    plain_pts = choose(run_plain, False)
    learn_pts = choose(run_batch, True)
    # the learned priority alone (one call per integral) at its own equal-RMSE point
    lone_pts = choose(run_plain, True)
    bp = min(plain_pts.values(), key=lambda x: x["ms"])
    bl = min(learn_pts.values(), key=lambda x: x["ms"])
    bo = min(lone_pts.values(), key=lambda x: x["ms"])
    p_learn = problems(bl["regions"], bl["spp"], True)
    ms_learn, st = timed(run_batch, p_learn, args.steps)
    rmse_learn = rmse()
    p_plain = problems(bp["regions"], bp["spp"], False)
    ms_plain, st_plain = timed(run_plain, p_plain, max(1, args.steps // 2))
    rmse_plain = rmse()
    # the two changes apart: batching alone (heuristic priority, plain's point) and the
    # learned priority alone (one call per integral, at its own point and at the batched
    # path's point)
    ms_bheur, _ = timed(run_batch, p_plain, max(1, args.steps // 2))
    ms_lone, _ = timed(run_plain, problems(bo["regions"], bo["spp"], True), max(1, args.steps // 2))
    rmse_lone = rmse()
    ms_lone_b, _ = timed(run_plain, p_learn, max(1, args.steps // 2))
    total_bins = nb * len(fs)
    value = total_bins / (ms_learn * 1e-3)
    plain_bps = total_bins / (ms_plain * 1e-3)
    flops = sum(st[i].evals_residual * I.flops_per_eval(f)[0] for i, f in enumerate(fs))
    ms_k = float(st[0].ms_bins)
    achieved = flops / (ms_k * 1e-3) / 1e12
    evals = sum(x.evals_regions + x.evals_residual for x in st)
    out = {
        "metric": METRIC, "value": value, "unit": "bins/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_learn, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "learned refinement + batching vs plain GPU path at equal RMSE: %d held-out "
                               "config-1 style integrands (1..8 blobs, step, wave), 512x512 fp32 bins"
                               % len(fs),
                   "integrals": len(fs), "bins": total_bins, "target": "plain path at 16384 regions, 16 spp",
                   "learned_batched": {"regions": bl["regions"], "spp": bl["spp"]},
                   "plain": {"regions": bp["regions"], "spp": bp["spp"]}, "path": "learned+batched"},
        "plain_bins_per_s": plain_bps,
        "speedup_vs_plain": value / plain_bps,
        "batched_heuristic_bins_per_s": total_bins / (ms_bheur * 1e-3),
        "learned_unbatched_bins_per_s": total_bins / (ms_lone * 1e-3),
        "learned_unbatched_point": {"regions": bo["regions"], "spp": bo["spp"]},
        "learned_unbatched_at_batched_point_bins_per_s": total_bins / (ms_lone_b * 1e-3),
        "rmse_vs_fp64": {"target": target, "learned_batched": rmse_learn, "plain": rmse_plain,
                         "learned_unbatched": rmse_lone},
        "sweep": {"plain": list(plain_pts.values()), "learned_batched": list(learn_pts.values()),
                  "learned_unbatched": list(lone_pts.values())},
        "fp64_truth": {"kind": "FP64 closed-form per-bin integrals (erf, polygon area, separable wave)"},
        "evals_per_s": evals / (ms_learn * 1e-3),
        "phase_ms": {k: float(getattr(st[0], k)) for k in ("ms_generate", "ms_select", "ms_stratify", "ms_bins",
                                                            "ms_total")},
        "plain_phase_ms_per_integral": {k: float(np.mean([getattr(x, k) for x in st_plain])) for k in
                                        ("ms_generate", "ms_select", "ms_stratify", "ms_bins", "ms_total")},
        "roofline": {"bound": "valu", "kernel": "k_bins1 (per-bin CV + residual Monte Carlo, batched)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": TRAFFIC_BYTES["config4"],
                     "traffic_source": TRAFFIC_SOURCE["config4"], "kernel_ms": ms_k},
        "training": {k: rep[k] for k in ("pairs", "steps", "loss_first", "loss_train", "loss_valid", "ms_pairs",
                                         "ms_train", "loaded") if k in rep},
        "autotune": tune,
    }
    if not args.no_cpu_baseline:
        ws = [dict(f=f, res=res, regions=bp["regions"], spp=bp["spp"], seed=0) for f in fs]
        cb = cpu_baseline(ws, min(16, os.cpu_count() or 1))
        cb["sample"] = "the held-out integrals in turn, plain method at the plain path's point: " + cb["sample"]
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


def run_fubini(args, V, hip):
    """Fubini control variate (integrator_crespo2021_infinite<2>) on the config-2 integrand:
    16384 regions over (x, y), the time dim integrated by Monte Carlo (4 samples per
    region node value, one per residual sample), 64 spp, 512x512 FP32 bins.  RMSE against
    the same FP64 truth as config 2.  value = bins / step time."""
    from viltrum_amd import integrands as I
    ctx = V.Context(0)
    f = I.config2()
    res = (512, 512)
    nb = res[0] * res[1]
    integ = V.integrator_crespo2021_infinite(2, 16383, 4, 64, seed=0)
    g, grng = integ.prepare(f, V.range_primary(3))
    p = integ.problem(g, res, grng, V.F32)
    buf = hip.DeviceBuffer(nb * 4)
    stream = hip.Stream()
    tune = ctx.autotune([p], [g], [buf.address], stream.address)
    for _ in range(args.warmup):
        ctx.integrate_device(p, g, buf.address, stream.address)
    hip.device_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = ctx.integrate_device(p, g, buf.address, stream.address)
    hip.device_synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    out = np.zeros(res, dtype=np.float32)
    buf.to_host(out)
    truth, tvar, tdesc = fp64_truth(ctx, V, f, res)
    rmse, rmse_less = rmse_pair(out, truth, tvar)
    flops = st.evals_residual * I.flops_per_eval(g)[0]
    ms_k = float(st.ms_bins)
    achieved = flops / (ms_k * 1e-3) / 1e12
    print(json.dumps({
        "metric": METRIC, "value": nb / (ms * 1e-3), "unit": "bins/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "Fubini CV: config-2 integrand, regions over (x, y), t by Monte Carlo (4 per node, "
                               "1 per residual sample), 512x512 fp32 bins, 16384 regions, 64 spp",
                   "bins": nb, "regions": 16384, "mc_samples": 4, "spp": 64, "path": "fubini"},
        "rmse_vs_fp64": rmse,
        "rmse_vs_fp64_less_truth_noise": rmse_less,
        "fp64_truth": tdesc,
        "evals_per_s": (st.evals_regions + st.evals_residual) / (ms * 1e-3),
        "phase_ms": {k: float(getattr(st, k)) for k in ("ms_generate", "ms_select", "ms_stratify", "ms_bins",
                                                         "ms_total")},
        "roofline": {"bound": "valu", "kernel": "k_bins1 (per-bin CV + residual Monte Carlo, Fubini form)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None, "kernel_ms": ms_k},
        "autotune": tune,
    }), flush=True)


def run_importance(args, V, hip):
    """Importance-sampled residual (region_sampling_importance) against uniform sampling,
    the region roulettes (rr_integral_region, rr_error_region, rr_pdf_region, and
    rr_pdf_region with importance sampling) and region_sampling_russian_roulette: 64 held-out config-1 style integrands at
    256x256 FP32 bins (bin boxes above the sampler's 1e-5 volume threshold), 4096 regions,
    16 spp, one batched call per variant.  RMSE per bin against the FP64 closed-form
    bins.  value = bins/s with importance sampling."""
    from viltrum_amd import integrands as I
    from viltrum_amd import refiner as R
    ctx = V.Context(0)
    fs = R.held_out(args.integrals)
    res = (256, 256)
    nb = res[0] * res[1]
    truths = [ctx.exact_bins(f, res) for f in fs]
    bufs = [hip.DeviceBuffer(nb * 4) for _ in fs]
    ptrs = [b.address for b in bufs]
    stream = hip.Stream()
    out = {}
    variants = (("uniform", V.FORM_VR_OPTIMIZE), ("importance", V.FORM_VR_OPTIMIZE_IMPORTANCE),
                ("rr_integral", V.FORM_VR_OPTIMIZE_RR_INTEGRAL), ("rr_error", V.FORM_VR_OPTIMIZE_RR_ERROR),
                ("rr_pdf", V.FORM_VR_OPTIMIZE_RR_PDF), ("rr_pdf_importance", V.FORM_VR_OPTIMIZE_RR_PDF_IMPORTANCE),
                ("sampling_russian_roulette", V.FORM_VR_OPTIMIZE_SAMPLING_RR))
    sts = {}
    for name, form in variants:
        ps = [V.Integrator(4095, 16, 0, form=form).problem(f, res, V.range_primary(2), V.F32) for f in fs]
        tune = ctx.autotune(ps, fs, ptrs, stream.address)
        for _ in range(args.warmup):
            ctx.integrate_batch(ps, fs, ptrs, stream.address)
        hip.device_synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = ctx.integrate_batch(ps, fs, ptrs, stream.address)
        hip.device_synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps

        def rmse():
            se2 = 0.0
            h = np.zeros(res, dtype=np.float32)
            for b, t in zip(bufs, truths):
                b.to_host(h)
                se2 += float(np.sum((h.astype(np.float64) - t) ** 2))
            return float(np.sqrt(se2 / (nb * len(fs))))

        # the 1/pdf weights make the importance samplers' error heavy-tailed (one bin can
        # carry most of it), so the RMSE is taken over seeds 0..3 and the median reported
        per_seed = [rmse()]
        for seed in (1, 2, 3):
            ctx.integrate_batch([V.Integrator(4095, 16, seed, form=form).problem(f, res, V.range_primary(2), V.F32)
                                 for f in fs], fs, ptrs, stream.address)
            hip.device_synchronize()
            per_seed.append(rmse())
        out[name] = {"ms": ms, "bins_per_s": nb * len(fs) / (ms * 1e-3), "rmse": float(np.median(per_seed)),
                     "rmse_per_seed": per_seed, "kernel_ms": float(st[0].ms_bins), "autotune": tune}
        sts[name] = st
        print("variant %s: %.3f ms, rmse %.3e" % (name, ms, out[name]["rmse"]), file=sys.stderr, flush=True)
    imp = out["importance"]
    st = sts["importance"]
    flops = sum(x.evals_residual for x in st) * I.flops_per_eval(fs[0])[0]
    achieved = flops / (imp["kernel_ms"] * 1e-3) / 1e12
    res_out = {
        "metric": METRIC, "value": imp["bins_per_s"], "unit": "bins/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": imp["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "importance-sampled residual: %d held-out config-1 style integrands, 256x256 fp32 "
                               "bins, 4096 regions, 16 spp, one batched call" % len(fs),
                   "integrals": len(fs), "bins": nb * len(fs), "regions": 4096, "spp": 16,
                   "path": "batched, region_sampling_importance"},
        "rmse_vs_fp64": {k: v["rmse"] for k, v in out.items()},
        "fp64_truth": {"kind": "FP64 closed-form per-bin integrals (erf, polygon area, separable wave)"},
        "uniform": out["uniform"],
        "variants": {k: {"bins_per_s": v["bins_per_s"], "ms": v["ms"], "kernel_ms": v["kernel_ms"], "rmse": v["rmse"],
                         "rmse_per_seed": v["rmse_per_seed"]} for k, v in out.items()},
        "rmse_note": "median over seeds 0..3 of the RMSE against the closed-form bins",
        "roofline": {"bound": "valu", "kernel": "k_bins1 (importance-sampling instantiation, batched)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None, "kernel_ms": imp["kernel_ms"]},
    }
    print(json.dumps(res_out), flush=True)


def c5_truth(ctx, V, fs, res, spp=4096, side=8, grid=4):
    """FP64 stratified per-bin Monte Carlo (spp samples per bin) on grid x grid blocks of
    side x side bins spread evenly over the bin grid of each integral: the same bin boxes
    as the full grid.  integrate_mc's bins are (block range volume) x (bin mean), the full
    grid's (bin count) x (bin integral): factor (res / side)^2.  Returns per integral a
    list of (i0, j0, truth block), the summed squared standard error and the bin count."""
    starts = [[int((g + 0.5) * r / grid) - side // 2 for g in range(grid)] for r in res]
    scale = (res[0] / side) * (res[1] / side)
    out, tse, n = [], 0.0, 0
    for f in fs:
        blocks = []
        for i0 in starts[0]:
            for j0 in starts[1]:
                rng = V.Range([i0 / res[0], j0 / res[1], 0.0, 0.0, 0.0],
                              [(i0 + side) / res[0], (j0 + side) / res[1], 1.0, 1.0, 1.0])
                t, se = ctx.integrate_mc(f, (side, side), spp, seed=777, rng=rng, stratified=True)
                blocks.append((i0, j0, t * scale))
                tse += float(np.sum((se * scale) ** 2))
                n += side * side
        out.append(blocks)
    return out, tse, n


def run_config5(args, V, hip):
    """5-D stress: 16 integrals (3 Gaussian lobes over the integrated dims x 2 half-space
    steps x a cosine, seeds 0..15) at 1024x1024 FP32 bins, 65536 regions, 32 spp, in one
    batched call with the learned priority.  Setup (untimed): the D=5 network is trained
    on the GPU on the pairs the heuristic method generates for seeded integrands of the
    same family (seeds disjoint from the evaluated ones).  The same batch with the
    heuristic priority is timed beside it.  RMSE per bin against FP64 stratified Monte
    Carlo at 4096 spp on 16 blocks of 8x8 bins spread over the grid of every integral.  value = total bins /
    learned+batched time."""
    from viltrum_amd import integrands as I
    from viltrum_amd import refiner as R
    ctx = V.Context(0)
    rep = trained(ctx, R, args, 5, lambda: R.train(ctx, count=args.c5_train_integrands,
                                                   regions=args.c5_train_regions, res=(1024, 1024),
                                                   epochs=args.train_epochs, seed=0, stride=1, dim=5))
    n = args.integrals if args.integrals != 64 else 16
    fs = [I.config5(s) for s in range(n)]
    res = (1024, 1024)
    nb = res[0] * res[1]
    bufs = [hip.DeviceBuffer(nb * 4) for _ in fs]
    ptrs = [b.address for b in bufs]
    stream = hip.Stream()

    def timed(learned, regions=65536, spp=32, tune=True):
        it = R.integrator(regions - 1, spp, seed=0, learned=learned)
        ps = [it.problem(f, res, V.range_primary(5), V.F32) for f in fs]
        # the sweep points share the tuned shape of their workload class
        tune = ctx.autotune(ps, fs, ptrs, stream.address) if tune else None
        for _ in range(args.warmup):
            ctx.integrate_batch(ps, fs, ptrs, stream.address)
        hip.device_synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = ctx.integrate_batch(ps, fs, ptrs, stream.address)
        hip.device_synchronize()
        return (time.perf_counter() - t0) * 1e3 / args.steps, st, tune

    if args.only_timed:
        # the learned+batched step alone at a given (regions, spp), launch shapes from the
        # VMX_* environment (no tuning runs): for kernel profiles of the timed step
        regions, spp = (int(x) for x in args.only_timed.split(","))
        ms, st, _ = timed(True, regions, spp, tune=False)
        print(json.dumps({"workload": "config5 learned+batched", "regions": regions, "spp": spp, "ms_per_step": ms,
                          "kernel_ms": float(st[0].ms_bins),
                          "phase_ms": {k: float(getattr(st[0], k)) for k in ("ms_generate", "ms_select",
                                                                              "ms_stratify", "ms_bins")},
                          "rounds": int(st[0].rounds), "nodes": int(sum(x.nodes for x in st))}), flush=True)
        return
    truths, tse, n_sq = c5_truth(ctx, V, fs, res)

    def rmse():
        out = np.zeros(res, dtype=np.float32)
        se2 = 0.0
        for b, blocks in zip(bufs, truths):
            b.to_host(out)
            for (i0, j0, t) in blocks:
                se2 += float(np.sum((out[i0:i0 + t.shape[0], j0:j0 + t.shape[1]].astype(np.float64) - t) ** 2))
        return float(np.sqrt(se2 / n_sq))

    noise2 = tse / n_sq

    def less_noise(r):
        return float(np.sqrt(max(0.0, r * r - noise2)))

    ms_h, st_h, tune_h = timed(False)
    rmse_h = rmse()
    ms, st, tune = timed(True)
    rmse_l = rmse()
    total_bins = nb * len(fs)
    # equal RMSE: the fastest smaller learned+batched point whose RMSE (truth noise
    # removed) does not exceed the heuristic's at the full budget
    sweep = []
    for regions, spp in ((32768, 16), (16384, 16), (16384, 8), (8192, 8)):
        ms_p, _, _ = timed(True, regions, spp, tune=False)
        sweep.append({"regions": regions, "spp": spp, "ms": ms_p, "rmse": rmse(),
                      "rmse_less_truth_noise": None})
        sweep[-1]["rmse_less_truth_noise"] = less_noise(sweep[-1]["rmse"])
    ok = [p for p in sweep if p["rmse_less_truth_noise"] <= less_noise(rmse_h)]
    best = min(ok, key=lambda p: p["ms"]) if ok else None
    equal = {"target_rmse_less_truth_noise": less_noise(rmse_h), "heuristic_batched_ms": ms_h, "sweep": sweep,
             "learned_batched": best,
             "speedup_vs_heuristic_batched": ms_h / best["ms"] if best else None}
    flops = sum(x.evals_residual for x in st) * I.flops_per_eval(fs[0])[0]
    ms_k = float(st[0].ms_bins)
    achieved = flops / (ms_k * 1e-3) / 1e12
    phases = ("ms_generate", "ms_select", "ms_stratify", "ms_bins", "ms_total")
    res_out = {
        "metric": METRIC, "value": total_bins / (ms * 1e-3), "unit": "bins/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "5-D lobes x steps x cosine, %d integrals, 1024x1024 fp32 bins, 65536 regions, "
                               "32 spp, one batched call, learned priority" % len(fs),
                   "integrals": len(fs), "bins": total_bins, "regions": 65536, "spp": 32, "path": "learned+batched"},
        "rmse_vs_fp64": {"learned_batched": rmse_l, "heuristic_batched": rmse_h},
        # the truth's own Monte Carlo error adds in quadrature: removed here
        "rmse_vs_fp64_less_truth_noise": {"learned_batched": less_noise(rmse_l),
                                          "heuristic_batched": less_noise(rmse_h)},
        "equal_rmse": equal,
        "fp64_truth": {"kind": "GPU FP64 stratified per-bin MC on 16 blocks of 8x8 bins spread over the grid "
                               "of every integral (1024 bins each)", "spp": 4096,
                       "stderr_rms": float(np.sqrt(tse / n_sq))},
        "heuristic_batched_bins_per_s": total_bins / (ms_h * 1e-3),
        "evals_per_s": sum(x.evals_regions + x.evals_residual for x in st) / (ms * 1e-3),
        "phase_ms": {k: float(getattr(st[0], k)) for k in phases},
        "heuristic_phase_ms": {k: float(getattr(st_h[0], k)) for k in phases},
        "roofline": {"bound": "valu", "kernel": "k_bins_rm (region-major per-bin CV + residual Monte Carlo, batched)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None, "kernel_ms": ms_k},
        "stats": {"nodes": int(sum(x.nodes for x in st)), "rounds": int(st[0].rounds),
                  "rounds_heuristic": int(st_h[0].rounds)},
        "training": rep,
        "autotune": {"learned_batched": tune, "heuristic_batched": tune_h},
    }
    print(json.dumps(res_out), flush=True)


def cpu_baseline(ws, nthreads, budget_s=10.0, min_runs=3):
    """The oracle (C++ restatement of integrator_crespo2021, FP32, per-bin mt19937,
    threads over dim-0 rows) on the host cores, over a bounded sample: one warm-up run,
    then the workloads of `ws` in turn (cycling) until about budget_s seconds of CPU work
    and at least min_runs runs; value = bins integrated / time."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    if isinstance(ws, dict):
        ws = [ws]

    def run(w):
        O.integrate(w["f"], w["res"], w["regions"] - 1, w["spp"], w["seed"], scalar=4,
                    integrator=O.INT_VR_OPTIMIZE, heuristic=O.HEUR_SIZE_RELATIVE, nthreads=nthreads)

    run(ws[0])
    total_t, total_bins, n = 0.0, 0, 0
    while n < min_runs or total_t < budget_s:
        w = ws[n % len(ws)]
        t0 = time.perf_counter()
        run(w)
        total_t += time.perf_counter() - t0
        total_bins += int(np.prod(w["res"]))
        n += 1
    w0 = ws[0]
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": total_bins / total_t, "unit": "bins/s", "cores": nthreads, "kind": "port",
            "cpu_model": model, "host_cpus": os.cpu_count(),
            "sample": "%d runs (%d distinct integrals) of %s bins, %d regions, %d spp after one warm-up: "
                      "%.1f s of CPU work" % (n, min(n, len(ws)), "x".join(map(str, w0["res"])), w0["regions"],
                                              w0["spp"], total_t)}


def run_mlp(args, V, hip):
    """The refinement network's kernels alone on synthetic BF16 feature rows: k_mlp_forward
    (inference) and k_mlp_grad (one backward pass, 64-row tiles), at dims 2 and 5.  Prints
    their MFMA work per row (the v_mfma_f32_16x16x32_bf16 tiles each issues) for a
    roofline against a kernel trace; wall times include the host round trip of the C
    entry points."""
    ctx = V.Context(0)
    n = args.mlp_rows
    out = {"workload": "mlp", "rows": n, "steps": args.steps, "dims": {}}
    W = 64
    for D in (2, 5):
        P, F, kp = V.refiner_size(D)
        rng = np.random.default_rng(100 + D)
        p = rng.normal(0.0, 0.05, P).astype(np.float32)
        p[-2], p[-1] = -7.0, 3.0  # target offset and scale
        x = np.zeros((n, kp), dtype=np.float32)
        x[:, :F] = rng.normal(0.0, 0.6, (n, F))
        X = (x.view(np.uint32) >> 16).astype(np.uint16)  # BF16 by truncation: values do not matter here
        t = rng.normal(-7.0, 3.0, n).astype(np.float32)
        ctx.refiner_forward(D, X, params=p)
        ctx.refiner_grad(D, p, X, t)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.refiner_forward(D, X, params=p)
        tf = (time.perf_counter() - t0) / args.steps
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.refiner_grad(D, p, X, t)
        tg = (time.perf_counter() - t0) / args.steps
        fwd = 2 * W * (kp + 2 * W)  # three MFMA layers: kp -> 64 -> 64 -> 64
        grad = fwd + 2 * 2 * 2 * W * W + 2 * kp * W  # forward, dW and g_in of layers 2 and 1, dW of layer 0
        out["dims"][str(D)] = {"kp": kp, "fwd_mfma_flop_per_row": fwd, "grad_mfma_flop_per_row": grad,
                               "fwd_mfma_flop_per_launch": fwd * n, "grad_mfma_flop_per_launch": grad * n,
                               "wall_ms_forward": tf * 1e3, "wall_ms_grad": tg * 1e3}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="config4")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--integrals", type=int, default=64, help="config4: held-out integrals")
    ap.add_argument("--train-integrands", type=int, default=512, help="config4: training integrands")
    ap.add_argument("--train-regions", type=int, default=4096,
                    help="config4: training budgets cycle through 1/4, 1/2, 1, 2 times this")
    ap.add_argument("--train-epochs", type=int, default=8)
    ap.add_argument("--train-stride", type=int, default=1, help="config4: keep every n-th generated region")
    ap.add_argument("--c5-train-integrands", type=int, default=32, help="config5: training integrands")
    ap.add_argument("--c5-train-regions", type=int, default=32768,
                    help="config5: training budgets cycle through 1/4, 1/2, 1, 2 times this")
    ap.add_argument("--only-timed", default="", metavar="REGIONS,SPP",
                    help="config4, config5: run only the learned+batched step at this point (for kernel profiles)")
    ap.add_argument("--save-net", default="", metavar="PATH", help="config4, config5: save the trained network")
    ap.add_argument("--load-net", default="", metavar="PATH",
                    help="config4, config5: load the network instead of training it (e.g. for counter passes)")
    ap.add_argument("--mlp-rows", type=int, default=1 << 20, help="mlp: feature rows per launch (<= 2^20)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (a known-size kernel for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import viltrum_amd as V
    from viltrum_amd import hip
    from viltrum_amd.integrands import flops_per_eval

    if args.calibrate_copy > 0:
        nbytes = args.calibrate_copy << 20
        src, dst = hip.DeviceBuffer(nbytes), hip.DeviceBuffer(nbytes)
        src.zero()
        hip.copy_device(dst, src, nbytes)
        hip.device_synchronize()
        src.free()
        dst.free()

    if args.workload == "config3":
        return run_config3(args, V, hip)
    if args.workload == "config4":
        return run_config4(args, V, hip)
    if args.workload == "config5":
        return run_config5(args, V, hip)
    if args.workload == "fubini":
        return run_fubini(args, V, hip)
    if args.workload == "importance":
        return run_importance(args, V, hip)
    if args.workload == "mlp":
        return run_mlp(args, V, hip)
    w = workloads()[args.workload]
    f = w["f"]
    ctx = V.Context(0)
    integ = V.integrator_crespo2021(w["regions"] - 1, w["spp"], seed=w["seed"])
    prob = integ.problem(f, w["res"], V.range_primary(f.dim), V.F32)
    nb = int(np.prod(w["res"]))
    bins = hip.DeviceBuffer(nb * 4)
    stream = hip.Stream()

    for _ in range(args.warmup):
        ctx.integrate_device(prob, f, bins.address, stream.address)
    stream.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    stats = []
    hip.device_synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        stats.append(ctx.integrate_device(prob, f, bins.address, stream.address).as_dict())
    e1.record(stream)
    e1.synchronize()
    hip.device_synchronize()
    wall = time.perf_counter() - t0
    ms_step = max(wall * 1e3, e0.elapsed_ms(e1)) / args.steps

    out = np.zeros(w["res"], dtype=np.float32)
    bins.to_host(out)
    truth, tvar, tdesc = fp64_truth(ctx, V, f, w["res"])
    rmse, rmse_less = rmse_pair(out, truth, tvar)

    st = stats[-1]
    flops, special = flops_per_eval(f)
    ms_k = float(np.mean([s["ms_bins"] for s in stats]))
    achieved = st["evals_residual"] * flops / (ms_k * 1e-3) / 1e12
    evals = st["evals_regions"] + st["evals_residual"]
    res = {
        "metric": METRIC,
        "value": nb / (ms_step * 1e-3),
        "unit": "bins/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": w["desc"], "bins": list(w["res"]), "regions": w["regions"], "spp": w["spp"],
                   "path": "plain"},
        "evals_per_s": evals / (ms_step * 1e-3),
        "rmse_vs_fp64": rmse,
        "rmse_vs_fp64_less_truth_noise": rmse_less,
        "fp64_truth": tdesc,
        "phase_ms": {k: float(np.mean([s[k] for s in stats])) for k in
                     ("ms_generate", "ms_select", "ms_stratify", "ms_bins", "ms_total")},
        "roofline": {"bound": "valu", "kernel": "k_bins1 (per-bin CV + residual Monte Carlo)", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": TRAFFIC_BYTES.get(args.workload),
                     "traffic_source": TRAFFIC_SOURCE.get(args.workload), "flops_per_eval": flops,
                     "special_per_eval": special,
                     "kernel_ms": ms_k},
        "stats": {k: st[k] for k in ("regions", "nodes", "rounds", "pairs", "evals_regions", "evals_residual",
                                     "ties_unresolved")},
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(w, min(16, os.cpu_count() or 1))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark of the FP32-scaled pruning likelihood on one MI355X.

Workload: GTR+Gamma4, 128 taxa (seeded Yule tree, mean branch 0.1),
100,000 unique simulated patterns with weights U{1..5}.  One step is one
full likelihood evaluation (every edge dirty): update list + edge times
uploaded, P(t) built on the GPU, post-order pruning over all 127 internal
nodes for 4 categories, FP64 root / category mix / weighted sum, the total
copied back.  Also reported: the same evaluation on the FP64 plain GPU
kernel (the baseline of the speed-up target), partial evaluations (one edge
to the root), the error against the FP64 CPU oracle, and the oracle's own
throughput on the host cores (cpu_baseline).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --workload c1                        (16-taxon HKY chain, 10,000 steps)
       python bench.py --workload batch [--instances 256]   (batched HKY+Gamma4)
       python bench.py --workload codon                     (GY94 61-state, MFMA)
       python bench.py --workload mcmc                      (ESS/s, delayed acceptance)
       python bench.py --workload hetero                    (HeteroLikelihood, relaxed clocks)
       python bench.py --workload storage                   (partial vs propagation storage)
       python bench.py --workload ingest                    (FASTA parse + pattern dedup)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "logL evals/s vs plain FP64 at <=1e-6 rel err; HBM GB/s; ESS/s vs plain MCMC"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def read_traffic(workload, kernel="prune_f32"):
    """HBM bytes per launch of the pruning kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy
    kernel calibration), if one exists for this workload."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def time_evals(calc, lists, times, rates, steps, warmup, sync_each=True, prewarm_s=0.0):
    """Time ``steps`` evaluations after ``prewarm_s`` seconds of back-to-back
    evaluations (so the device is at steady clocks whatever ``warmup`` is)
    and ``warmup`` more.  Returns (wall seconds, last total, mean kernel ms,
    per-launch kernel ms)."""
    internals, children = lists
    # every evaluation is followed by accept(), as in an MCMC step: b3's FP64
    # scale counts are not idempotent across proposes without one (cpu4.rs:78-89)
    t_end = time.perf_counter() + prewarm_s
    while time.perf_counter() < t_end:
        for _ in range(10):
            calc.propose_times(internals, children, times, rates)
            calc.total()
            calc.accept()
    for _ in range(warmup):
        calc.propose_times(internals, children, times, rates)
        calc.total()
        calc.accept()
    calc.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        calc.propose_times(internals, children, times, rates)
        if sync_each:
            calc.total()
        calc.accept()
    total = calc.total()
    dt = time.perf_counter() - t0
    each = calc.profile_launch_times()
    calc.profile_enable(False)
    return dt, total, float(each.mean()) if each.size else 0.0, each


def cpu_threads(K, P):
    """Threads of the CPU baseline: b3 runs K category calculators at once
    (gamma.rs:69,143-152), each with min(cores, ceil(P / 1000)) threads
    (cpu4.rs:127-132).  On a host with fewer cores than that product the
    threads per category are cut so the total does not exceed the cores:
    oversubscribing the cores does not add throughput."""
    ncores = len(os.sched_getaffinity(0))
    tpc = min(ncores, (P + 999) // 1000)
    tpc = max(1, min(tpc, ncores // K))
    return ncores, tpc


def bench_c1(args):
    """The reference's CPU-runnable case: 16 taxa (seeded Yule tree), HKY with
    kappa 2 and equal frequencies, 1,000 simulated sites compressed to unique
    patterns, 10,000 plain MCMC steps with every node recomputed at each
    likelihood call -- timed end to end with the native chain driver
    (NodeSlide, TreeScale, NarrowExchange at b3's weights; the substitution
    model fixed) and its GPU likelihood.  cpu_baseline: the same chain on the
    CPU in FP64 (oracle/chain.py: b3's step and operators over the oracle's
    tree and likelihood, one thread as cpu4.rs picks below 1,000 patterns),
    timed end to end over a bounded number of steps.  At this size a step is
    launch- and synchronisation-bound on the GPU."""
    from aspartik_amd import workloads
    from aspartik_amd.b3 import Chain

    wl = workloads.hky_cpu_mcmc(seed=4)
    L, P = wl["codes"].shape
    steps = args.c1_steps
    log(f"[bench] C1: {L} taxa, {P} patterns (1,000 sites), {steps} steps")
    c = Chain(wl["codes"], wl["parents"], wl["children"], wl["heights"], wl["gtr_rates"],
              wl["freqs"], 1.0, weights=wl["weights"], categories=1, w_rates=0.0, w_alpha=0.0,
              full_evaluations=1, seed=4)
    c.run(200)
    t0 = time.perf_counter()
    _, st, _ = c.run(steps)
    dt = time.perf_counter() - t0
    loglik = c.state()[0]
    par, ch, h, _, _ = c.params()
    c.close()
    log(f"[bench] C1 chain: {dt:.2f} s for {steps} steps; accepted {st['accepted']}")
    cpu, err = None, None
    if not args.no_cpu:
        from oracle import chain as OC
        from oracle import oracle as O
        ot = O.Tree(L, par, ch.astype(np.uint32), h)
        om = O.Model.hky(wl["kappa"], wl["freqs"])
        lik = O.GammaLikelihood(ot, om, wl["codes"], P, wl["weights"], num_categories=1,
                                threads_per_category=1)
        ref = lik.likelihood()
        err = {"final_state_rel": abs(loglik - ref) / abs(ref), "ref": ref, "gpu": loglik}
        # the same chain on the CPU in FP64: b3's step and the three operators
        # over the oracle's tree and likelihood, every node recomputed per step,
        # from the same starting state (one thread, as cpu4.rs picks below
        # 1,000 patterns)
        oc = OC.Chain(L, wl["parents"], wl["children"], wl["heights"], om, wl["codes"],
                      wl["weights"], birth_rate=1.0, seed=4)
        oc.run(200)
        budget = int(min(steps, max(1000, steps * args.cpu_seconds / 30.0)))
        cst = oc.run(budget)
        dtc = cst["seconds"]
        cpu = {"value": dtc / budget * 10_000, "unit": "s per 10,000 steps", "cores": 1,
               "kind": "port", "accepted_fraction": cst["accepted"] / cst["steps"],
               "sample": f"{budget} steps of the same chain (NodeSlide, TreeScale, "
                         f"NarrowExchange at 4:1:2; Yule prior; FP64 full evaluations) on the "
                         f"CPU ({dtc:.2f} s), scaled to 10,000; GPU chain accepted "
                         f"{st['accepted'] / steps:.3f} of its steps, CPU chain "
                         f"{cst['accepted'] / cst['steps']:.3f}"}
        log(f"[bench] C1 CPU chain: {cpu['value']:.2f} s per 10,000 steps; final-state "
            f"ln L rel err {err['final_state_rel']:.1e}")
    print(json.dumps({
        "metric": "seconds per 10,000 plain MCMC steps (16 taxa, HKY, 1,000 sites)",
        "value": dt / steps * 10_000, "unit": "s", "n_gpus": 1, "steps": steps, "warmup": 200,
        "ms_per_step": dt / steps * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Yule tree, sequences simulated along it)",
        "config": {"workload": wl["name"], "taxa": L, "patterns": P, "sites": 1000,
                   "model": "HKY kappa 2, equal frequencies", "categories": 1,
                   "evaluation": "full (every node) at each step",
                   "operators": "NodeSlide 4, TreeScale 1, NarrowExchange 2"},
        "roofline": None, "cpu_baseline": cpu, "error_vs_fp64_oracle": err,
        "chain_stats": {k: st[k] for k in ("steps", "accepted", "exact_evals")},
    }), flush=True)


def bench_batch(args):
    """256 independent HKY+Gamma4 instances (50-500 taxa, 2k-50k patterns),
    every instance fully evaluated in one persistent launch per step,
    replayed from a captured HIP graph."""
    from aspartik_amd import workloads
    from aspartik_amd.b3 import Batch, Calculator, DatedTree, autotune_batch

    log("[bench] building batched workload")
    wl = workloads.hky_gamma4_batch(count=args.instances)
    insts = wl["instances"]
    K, rates = wl["K"], wl["rates"]
    ups, B = [], 0
    for x in insts:
        t = DatedTree(x["L"], x["parents"], x["children"], x["heights"])
        internals, children = workloads.full_lists(t)
        ups.append((internals, children, workloads.edge_times(t, children), rates))
        B += workloads.store_bytes(internals, children, x["L"], K, x["P"], tree_children=t.children)
    data = [(x["codes"], None) for x in insts]
    models = [(x["model"].q_matrix(), x["model"].frequencies(), True) for x in insts]
    if args.batch_tile:
        parts = [int(v) for v in args.batch_tile.split("x")]
        tile, lanes, states = (parts + [4])[:3]
        trials = []
    else:
        log("[bench] batched: tuning tiles")
        (tile, lanes, states), trials = autotune_batch(data, models, ups, K, evals=6)
    for tr in trials:
        log(f"[bench]   tile {tr.tile_patterns} x {tr.lane_patterns} patterns/lane x "
            f"{tr.lane_states} states/lane: {tr.ms:.2f} ms per batch, kernel {tr.kernel_ms:.2f} ms")
    b = Batch(data, num_categories=K, single_slot=True, tile_patterns=tile, lane_patterns=lanes,
              lane_states=states)
    for i, (q, f, special) in enumerate(models):
        b.set_model(i, q, f, special)
    log("[bench] batched: warmup")
    for _ in range(args.warmup):
        b.propose_times(ups, graph=True)
        b.totals()
        b.accept()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.propose_times(ups, graph=True)
        totals = b.totals()
        b.accept()
    dt = time.perf_counter() - t0
    b.profile_enable(True)
    for _ in range(3):
        b.propose_times(ups)
        b.totals()
        b.accept()
    kms, kn = b.profile_read()
    b.profile_enable(False)
    device_bytes = b.device_bytes()
    kernel_ms = kms / max(1, kn)
    log(f"[bench] batched: {args.steps / dt:.2f} batch evals/s, kernel {kernel_ms:.2f} ms")
    b.close()
    # the plain FP64 GPU path over the same instances, one at a time (both
    # of its slots for every instance at once would exceed the HBM): each
    # instance is created, evaluated once untimed and then timed over
    # `f64_evals` full evaluations; the batch's FP64 time is the sum
    f64_evals, f64_s, f64_err = 2, 0.0, 0.0
    if not args.only_f32:
        log("[bench] batched: plain FP64 GPU path, instance by instance")
    for i, x in enumerate([] if args.only_f32 else insts):
        c64 = Calculator.HIPPlainF64().make(x["codes"], x["P"], K)
        c64.set_weights(np.ones(x["P"], np.uint32))
        c64.set_model(*models[i])
        c64.propose_times(*ups[i])
        ref64 = c64.total()
        c64.accept()
        f64_err = max(f64_err, abs(totals[i] - ref64) / abs(ref64))
        t1 = time.perf_counter()
        for _ in range(f64_evals):
            c64.propose_times(*ups[i])
            c64.total()
            c64.accept()
        f64_s += (time.perf_counter() - t1) / f64_evals
        c64.close()
    if f64_s > 0:
        log(f"[bench] batched: FP64 plain {f64_s * 1e3:.1f} ms per batch, FP32 batched "
            f"speed-up {f64_s / (dt / args.steps):.2f}x, max total rel err vs FP64 "
            f"{f64_err:.2e}")
    cpu, err = None, None
    if not args.no_cpu:
        from oracle import oracle as O
        order = np.random.default_rng(1).permutation(len(insts))
        ncores = len(os.sched_getaffinity(0))
        rel, nev, t_cpu, work, sample = [], 0, 0.0, 0, []
        threads_used = 1
        for i in order:
            if t_cpu > args.cpu_seconds or len(sample) >= 24:
                break
            i = int(i)
            sample.append(i)
            x = insts[i]
            ot = O.Tree(x["L"], x["parents"], x["children"], x["heights"])
            om = O.Model.hky(x["kappa"], np.full(4, 0.25))
            _, tpc = cpu_threads(K, x["P"])
            threads_used = max(threads_used, tpc * K)
            lik = O.GammaLikelihood(ot, om, x["codes"], x["P"], np.ones(x["P"], np.uint32),
                                    num_categories=K, alpha=wl["alpha"],
                                    threads_per_category=tpc, parallel_categories=True)
            ref = lik.likelihood()
            rel.append(abs(totals[i] - ref) / abs(ref))
            t1 = time.perf_counter()
            for _ in range(3):
                ot.mark_all_edges_updated()
                lik.likelihood()
                lik.accept()
                ot.accept()
            t_cpu += time.perf_counter() - t1
            nev += 3
            work += 3 * (x["L"] - 1) * x["P"]
        all_work = sum((x["L"] - 1) * x["P"] for x in insts)
        # instance evaluations per second at the average instance size
        cpu = {"value": (work / t_cpu) / (all_work / len(insts)), "unit": "instance logL evals/s",
               "cores": threads_used, "host_cores": ncores, "kind": "port",
               "sample": f"{nev} full evaluations of {len(sample)} randomly chosen instances "
                         f"({t_cpu:.1f} s), scaled to the batch's mean instance size by "
                         f"node x pattern count"}
        err = {"max_total_rel": float(max(rel)), "instances_checked": sample}
    print(json.dumps({
        "metric": METRIC, "value": args.instances * args.steps / dt,
        "unit": "instance logL evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Yule trees, sequences simulated along them)",
        "config": {"workload": f"hky_gamma4_batch{args.instances}", "instances": args.instances,
                   "taxa": "U{50..500}", "patterns": "U{2000..50000}", "categories": K,
                   "evaluation": "full, all instances per step, graph replay, single slot",
                   "tile_patterns": tile, "lane_patterns": lanes, "lane_states": states},
        "roofline": {"bound": "hbm", "achieved": B / (kernel_ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": B / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": read_traffic(f"hky_gamma4_batch{args.instances}",
                                             "prune_f32_batch")},
        "cpu_baseline": cpu, "kernel_ms": kernel_ms, "algorithmic_bytes_per_eval": B,
        "batch_evals_per_s": args.steps / dt, "error_vs_fp64_oracle": err,
        "fp64_plain_gpu": None if f64_s == 0 else {"ms_per_batch": f64_s * 1e3,
                           "instance_evals_per_s": len(insts) / f64_s,
                           "speedup_of_batched_fp32": f64_s / (dt / args.steps),
                           "max_total_rel_err_fp32_vs_fp64": f64_err,
                           "method": f"each instance alone on the plain FP64 kernel, "
                                     f"{f64_evals} timed full evaluations after one untimed"},
        "device_bytes": device_bytes,
        "tile_trials": [{"tile_patterns": tr.tile_patterns, "lane_patterns": tr.lane_patterns,
                         "lane_states": tr.lane_states, "ms_per_batch": tr.ms,
                         "kernel_ms": tr.kernel_ms} for tr in trials],
    }), flush=True)


def bench_codon(args):
    """GY94 codon likelihood (61 states, F3x4, Gamma4) on the MFMA path:
    64 taxa, 20,000 unique simulated codon patterns, one step = one full
    evaluation with the transition matrices built on the GPU."""
    from aspartik_amd import workloads
    from aspartik_amd.b3 import Calculator, DatedTree

    log("[bench] building codon workload")
    wl = workloads.gy94_codon(args.codon_taxa, args.codon_patterns)
    L, P, K = args.codon_taxa, args.codon_patterns, wl["K"]
    tree = DatedTree(L, wl["parents"], wl["children"], wl["heights"])
    full = workloads.full_lists(tree)
    times_full = workloads.edge_times(tree, full[1])
    rates, model = wl["rates"], wl["model"]
    pi = model.frequencies()
    calc = Calculator.HIPCodon(61).make(wl["codes"], P, K)
    calc.set_model(model.q_matrix(), pi)
    dt, total, kernel_mean_ms, each = time_evals(calc, full, times_full, rates, args.steps,
                                                 args.warmup, prewarm_s=args.prewarm)
    kernel_ms = float(np.median(each))
    log(f"[bench] codon full: {args.steps / dt:.1f} evals/s, kernel median {kernel_ms * 1e3:.1f} us")
    site_gpu = np.array([calc.likelihoods(k) for k in range(K)])
    mixed_gpu = calc.site_mixed()
    # one-edge partial evaluations: one random non-root edge and its path to
    # the root (propose, read the total, accept), then the full state back
    rng = np.random.default_rng(99)
    edges = [int(e) for e in rng.choice(2 * L - 2, size=32) if int(e) != tree.root]
    plists = [workloads.path_lists(tree, e) for e in edges]
    ptimes = [workloads.edge_times(tree, pl[1]) for pl in plists]
    for i in range(50):
        calc.propose_times(*plists[i % len(plists)], ptimes[i % len(plists)], rates)
        calc.total()
        calc.accept()
    nparts = max(500, args.steps)
    t0 = time.perf_counter()
    for i in range(nparts):
        j = i % len(plists)
        calc.propose_times(*plists[j], ptimes[j], rates)
        calc.total()
        calc.accept()
    dt_part = time.perf_counter() - t0
    mean_path = float(np.mean([pl[0].size for pl in plists]))
    log(f"[bench] codon partial: {nparts / dt_part:.1f} evals/s (mean path {mean_path:.1f} nodes)")
    calc.propose_times(*full, times_full, rates)
    calc.total()
    calc.accept()
    dev_bytes = calc.device_bytes()
    calc.close()
    root = tree.root
    B = workloads.store_bytes(full[0], full[1], L, K, P, unit=130, reread_in_pass=True,
                              epilogue=False)
    B_ref = workloads.algorithmic_bytes(full[0], full[1], L, root, K, P, states=64, elem=2)
    F = workloads.algorithmic_flops(full[0], full[1], L, K, P, states=61)
    achieved = B / (kernel_ms * 1e-3) / 1e9
    cpu, err = None, None
    if not args.no_cpu:
        from oracle import oracle as O
        log("[bench] FP64 CPU oracle (cpu_baseline)")
        otree = O.Tree(L, wl["parents"], wl["children"], wl["heights"])
        om = O.Model.reversible(model.exchangeability(), pi)
        flat = np.asarray(full[1]).reshape(-1)
        mats = np.zeros((K, flat.size, 61, 61))
        for c in range(K):
            for m in range(flat.size):
                mats[c, m] = om.transition(otree.edge_length(int(flat[m])) * rates[c])
        ncores = len(os.sched_getaffinity(0))
        calcs = [O.CalcN(wl["codes"], P, 61, 30, ncores) for _ in range(K)]
        nev, t0 = 0, time.perf_counter()
        while True:
            ref = []
            for c in range(K):
                calcs[c].propose(full[0], full[1], mats[c], pi)
                ref.append(calcs[c].likelihoods())
                calcs[c].accept()
            nev += 1
            if time.perf_counter() - t0 > args.cpu_seconds or nev >= 50:
                break
        dtc = time.perf_counter() - t0
        ref = np.array(ref)
        mx = ref.max(0)
        ref_mixed = mx + np.log(np.exp(ref - mx).sum(0)) + np.log(1.0 / K)
        cpu = {"value": nev / dtc, "unit": "logL evals/s", "cores": ncores, "kind": "port",
               "sample": f"{nev} full evaluations of the same workload by the FP64 oracle "
                         f"(S-state cpu4.rs recursion, {K} categories in turn, {ncores} "
                         f"threads each, host-built matrices excluded; {dtc:.1f} s)"}
        err = {"total_rel": abs(total - ref_mixed.sum()) / abs(ref_mixed.sum()),
               "total_abs": abs(total - ref_mixed.sum()),
               "max_site_mixed_abs": float(np.abs(mixed_gpu - ref_mixed).max()),
               "mean_site_mixed_abs": float(np.abs(mixed_gpu - ref_mixed).mean()),
               "max_site_category_abs": float(np.abs(site_gpu - ref).max())}
        log(f"[bench] codon oracle {cpu['value']:.3f} evals/s; total rel err {err['total_rel']:.2e}")
    tflops = F / (kernel_ms * 1e-3) / 1e12
    print(json.dumps({
        "metric": METRIC, "value": args.steps / dt, "unit": "logL evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f16",
        "data": "synthetic (seeded Yule tree, codon sequences simulated along it)",
        "config": {"workload": wl["name"], "taxa": L, "patterns": P, "categories": K,
                   "model": "GY94 omega 0.3 kappa 2 F3x4 + Gamma4 alpha 0.5",
                   "evaluation": "full (all edges dirty), P(t) built on the GPU",
                   "precision": "FP16 MFMA inputs (P split hi+lo), FP32 accumulation, "
                                "FP64 root/mix/sum", "seed": wl["seed"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": read_traffic(wl["name"], "prune_codon")},
        "mfma": {"useful_tflops": tflops, "issued_tflops": tflops * 2 * (64 / 61) ** 2,
                 "peak_fp16_tflops": 2500.0},
        "cpu_baseline": cpu, "kernel_us": kernel_ms * 1e3, "kernel_us_mean": kernel_mean_ms * 1e3,
        "algorithmic_bytes_per_eval": B, "reference_algorithm_bytes_per_eval": B_ref,
        "algorithmic_flops_per_eval": F, "error_vs_fp64_oracle": err,
        "device_bytes": dev_bytes,
        "partial_evals_per_s": nparts / dt_part, "partial_mean_path_nodes": mean_path,
    }), flush=True)


def ess_batch_means(x, batches=30):
    """Batch-means effective sample size, a cross-check of Geyer's estimate:
    n s^2 / (b s_b^2) with b = n // batches samples per batch and s_b^2 the
    variance of the batch means.  Batches must be longer than the trace's
    autocorrelation time or the estimate is too high: at 1,000,000 steps
    the log posterior's is thousands of trace rows, far above the sqrt(n)
    rows of the usual choice, so the trace is cut into a few long batches."""
    x = np.asarray(x, np.float64)
    n = x.size
    b = max(1, n // batches)
    m = n // b
    if m < 2:
        return float("nan")
    means = x[: m * b].reshape(m, b).mean(1)
    var_bm = b * means.var(ddof=1)
    return float(n * x.var(ddof=1) / var_bm) if var_bm > 0 else float("nan")


def _run_chunked(chain, steps, mode, name, chunk=50_000, max_rows=100_000):
    """Run a chain in chunks of ``chunk`` steps on the same handle (the chain
    state carries over, so the steps are those of one run), logging each
    chunk; returns the concatenated trace, recorded every
    ceil(steps / max_rows) steps, and the summed statistics."""
    every = -(-steps // max_rows)
    chunk = max(every, chunk // every * every)
    traces, total = [], None
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        tr, st, _ = chain.run(n, mode=mode, record_every=every)
        traces.append(tr)
        if total is None:
            total = st
        else:
            for k, v in st.items():
                if k == "tuning":
                    total[k] = v
                elif isinstance(v, list):
                    total[k] = [a + b for a, b in zip(total[k], v)]
                else:
                    total[k] += v
        done += n
        log(f"[bench] {name}: {done} of {steps} steps, {total['seconds']:.1f} s")
    return np.concatenate(traces), total


def bench_mcmc(args):
    """ESS per second of delayed-acceptance MH against plain MH: GTR+Gamma4,
    200 taxa, 200,000 sites.  Burn-in with plain MH collects (features,
    exact ln L) pairs; the MLP surrogate (404 -> 256 -> 256 -> 1) is trained
    on them; then plain and delayed-acceptance chains run the same number of
    steps from the burn-in's final state (topology moves go straight to the
    exact test; with --screen-topology 1 NarrowExchange passes the
    surrogate's first stage too; one delayed chain per first-stage
    temperature in --betas, the first being the headline).  NarrowExchange
    takes its reversible form by default (--narrow-reversible), so both
    chains keep the same posterior and their summaries can be compared.
    ESS is Geyer's estimator on the log posterior, tree length and alpha
    traces (recorded every ceil(steps / 100,000) steps), cross-checked by
    batch means; ESS/s uses each chain's wall clock (likelihood, surrogate
    and host work)."""
    from aspartik_amd import workloads
    from aspartik_amd.b3 import Chain, Surrogate, ess

    log("[bench] building MCMC workload")
    wl = workloads.gtr_gamma4_mcmc(args.mcmc_taxa, args.mcmc_sites)
    L, P = args.mcmc_taxa, args.mcmc_sites
    kw = dict(birth_rate=float((L - 1) / (wl["heights"][wl["parents"][:-1].astype(np.int64)]
                                          - wl["heights"][:-1]).sum()),
              narrow_reversible=args.narrow_reversible)
    # operator tuning during burn-in (b3's LinearOptimizer schedule), frozen after
    burn = Chain(wl["codes"], wl["parents"], wl["children"], wl["heights"], wl["gtr_rates"],
                 wl["freqs"], wl["alpha"], seed=101, tune_every=max(50, args.burnin // 40),
                 tune_cutoff=args.burnin, **kw)
    t0 = time.perf_counter()
    thin = max(1, args.collect_every)
    collect = args.collect if 0 < args.collect < args.burnin else args.burnin
    block = 100
    if thin > 1:
        collect = min(collect, args.burnin // thin) // block * block
        if collect == 0:
            raise SystemExit("--collect-every: the burn-in is too short for one block of 100 "
                             "collected steps")
    span = collect * thin  # the burn-in steps the training pairs come from
    if span < args.burnin:   # training pairs from the last `span` steps only
        burn.run(args.burnin - span)
    if thin == 1:
        _, bstats, (X, y) = burn.run(collect, collect=collect)
    else:
        # a chain at 200,000 sites drifts for tens of thousands of steps
        # between independent states: pairs spread over a longer stretch
        # cover more of what the measured chains visit
        xs, ys = [], []
        for _ in range(collect // block):
            _, bstats, (xb, yb) = burn.run(block, collect=block)
            xs.append(xb)
            ys.append(yb)
            _, bstats, _ = burn.run(block * (thin - 1))
        X, y = np.concatenate(xs), np.concatenate(ys)
    log(f"[bench] burn-in {args.burnin} steps: {time.perf_counter() - t0:.1f} s, "
        f"{len(y)} exact evaluations collected from the last {span}")
    state = burn.params()
    nf = burn.num_features
    tuned = dict(scale_factor=bstats["tuning"][1], delta_tuning=bstats["tuning"][2],
                 alpha_factor=bstats["tuning"][3])
    kw.update(tuned)
    burn.close()
    # the surrogate must resolve differences near the posterior's bulk; far
    # below it, any value that rejects will do: clip the targets there
    y = np.maximum(y, y.max() - args.clip)
    # surrogate: train on 90 %, report held-out error on 10 %
    order = np.random.default_rng(7).permutation(len(y))
    ntr = int(0.9 * len(y))
    tr, te = order[:ntr], order[ntr:]
    sur = Surrogate(nf, 256, seed=3)
    sur.fit_scaling(X[tr], y[tr])
    t0 = time.perf_counter()
    # Adam with a step-decayed learning rate: at the same 100 epochs, held-out
    # RMSE 1.87 against 2.28 at a constant 1e-3 on this workload
    for part in args.train_schedule.split(","):
        ep, lr = part.split("@")
        tr_loss = sur.train(X[tr], y[tr], epochs=int(ep), batch=256, lr=float(lr))
    t_train = time.perf_counter() - t0
    pred = sur.forward(X[te])
    top = te[y[te] > y.max() - args.clip / 4]
    held = {"rmse_loglik": float(np.sqrt(np.mean((pred - y[te]) ** 2))),
            "rmse_top_quarter": float(np.sqrt(np.mean((sur.forward(X[top]) - y[top]) ** 2)))
            if top.size else None,
            "loglik_std": float(np.std(y[te])), "train_loss_standardised": tr_loss,
            "train_pairs": int(ntr), "train_seconds": t_train, "clip": args.clip,
            "train_schedule": args.train_schedule,
            "tuned_operators": tuned}
    log(f"[bench] surrogate trained in {t_train:.1f} s; held-out RMSE {held['rmse_loglik']:.2f} "
        f"(ln L std {held['loglik_std']:.2f})")
    par, ch, h, r, a = state
    res = {}
    betas = [float(b) for b in args.betas.split(",")]
    chains = [("plain", "plain", args.chain_seed, 1.0)] + [
        ("delayed" if i == 0 else f"delayed_beta{b:g}", "delayed", args.chain_seed + 1 + i, b)
        for i, b in enumerate(betas)]
    for name, mode, seed, beta in chains:
        c = Chain(wl["codes"], par, ch, h, r, wl["freqs"], a, seed=seed,
                  screen_topology=args.screen_topology, surrogate_beta=beta, **kw)
        if mode == "delayed":
            c.set_surrogate(sur)
        trace, st = _run_chunked(c, args.mcmc_steps, mode, name)
        c.close()
        cols = (("posterior", 1), ("tree_length", 5), ("alpha", 6))
        e = {k: ess(trace[:, col]) for k, col in cols}
        ebm = {k: ess_batch_means(trace[:, col]) for k, col in cols}
        res[name] = dict(stats=st, ess=e, ess_batch_means=ebm, min_ess=min(e.values()),
                         ess_per_s=min(e.values()) / st["seconds"],
                         ess_per_s_by_statistic={k: v / st["seconds"] for k, v in e.items()},
                         means={k: float(trace[:, col].mean()) for k, col in cols},
                         sds={k: float(trace[:, col].std()) for k, col in cols})
        log(f"[bench] {name}: {st['seconds']:.1f} s (likelihood {st['likelihood_seconds']:.1f} s, "
            f"surrogate {st['surrogate_seconds']:.1f} s), accepted {st['op_accepted']} of "
            f"{st['op_proposed']}, exact {st['exact_evals']}, ESS {e}")
    ratio = res["delayed"]["ess_per_s"] / res["plain"]["ess_per_s"]
    ratio_by_stat = {k: res["delayed"]["ess_per_s_by_statistic"][k] /
                     res["plain"]["ess_per_s_by_statistic"][k]
                     for k in res["plain"]["ess_per_s_by_statistic"]}
    ratio_bm = {k: (res["delayed"]["ess_batch_means"][k] / res["delayed"]["stats"]["seconds"]) /
                (res["plain"]["ess_batch_means"][k] / res["plain"]["stats"]["seconds"])
                for k in res["plain"]["ess_batch_means"]}
    log(f"[bench] ESS/s ratios (Geyer) {ratio_by_stat}; batch means {ratio_bm}")
    # posterior summaries: each statistic's mean, delayed minus plain, in
    # units of the two chains' combined Monte Carlo standard error (ESS-based)
    agree = {}
    pr, dr = res["plain"], res["delayed"]
    for k in pr["means"]:
        se = math.sqrt(pr["sds"][k] ** 2 / max(pr["ess"][k], 1.0) +
                       dr["sds"][k] ** 2 / max(dr["ess"][k], 1.0))
        agree[k] = (dr["means"][k] - pr["means"][k]) / se if se > 0 else 0.0
    log(f"[bench] posterior means, delayed - plain in combined MCSE units: {agree}")
    ratio_incl_training = (res["delayed"]["min_ess"] /
                           (res["delayed"]["stats"]["seconds"] + t_train)) / res["plain"]["ess_per_s"]
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        ncores = len(os.sched_getaffinity(0))
        otree = O.Tree(L, wl["parents"], wl["children"], wl["heights"])
        om = O.Model.gtr(wl["gtr_rates"], wl["freqs"])
        _, tpc = cpu_threads(4, P)
        lik = O.GammaLikelihood(otree, om, wl["codes"], P, wl["weights"], num_categories=4,
                                alpha=wl["alpha"], threads_per_category=tpc,
                                parallel_categories=True)
        nev, t0 = 0, time.perf_counter()
        while True:
            otree.mark_all_edges_updated()
            lik.likelihood()
            lik.accept()
            otree.accept()
            nev += 1
            if time.perf_counter() - t0 > args.cpu_seconds or nev >= 20:
                break
        dtc = time.perf_counter() - t0
        # plain MH on the CPU would need the same exact evaluations per ESS
        ps = res["plain"]["stats"]
        evals_per_ess = ps["exact_evals"] / res["plain"]["min_ess"]
        cpu = {"value": (nev / dtc) / evals_per_ess, "unit": "ESS/s (plain MH, upper bound)",
               "cores": tpc * 4, "host_cores": ncores, "kind": "port",
               "sample": f"{nev} full evaluations by the FP64 oracle ({dtc:.1f} s, "
                         f"{nev / dtc:.2f} evals/s), divided by the plain GPU chain's "
                         f"{evals_per_ess:.1f} exact evaluations per effective sample "
                         f"(full evaluations only, so an upper bound for the CPU)"}
    print(json.dumps({
        "metric": METRIC, "value": res["delayed"]["ess_per_s"], "unit": "ESS/s",
        "n_gpus": 1, "steps": args.mcmc_steps, "warmup": args.burnin,
        "ms_per_step": res["delayed"]["stats"]["seconds"] / args.mcmc_steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Yule tree, sites simulated along it)",
        "config": {"workload": wl["name"], "taxa": L, "patterns": P, "categories": 4,
                   "model": "GTR+Gamma4 alpha 0.5", "surrogate": f"{nf}->256->256->1 GELU BF16",
                   "burnin_steps": args.burnin, "chain_steps": args.mcmc_steps,
                   "trace_every": -(-args.mcmc_steps // 100_000),
                   "training_pairs_from_last_steps": span, "training_pairs": int(len(y)),
                   "surrogate_betas": betas,
                   "chain_seed": args.chain_seed,
                   "topology_moves": "surrogate stage, then exact" if args.screen_topology
                                     else "exact MH test only",
                   "operators": "NodeSlide 4, TreeScale 1, DeltaExchange 2, ScaleReal(alpha) 1, "
                                "NarrowExchange 2 (b3's defaults, tuned in burn-in)",
                   "narrow_exchange": "reversible (grandparent: an internal child)"
                                      if args.narrow_reversible
                                      else "b3's rule (grandparent: two internal children)"},
        "ess_per_s_plain": res["plain"]["ess_per_s"], "ess_per_s_delayed": res["delayed"]["ess_per_s"],
        "ess_ratio_delayed_vs_plain": ratio,
        "ess_ratio_by_statistic": ratio_by_stat,
        "ess_ratio_by_statistic_batch_means": ratio_bm,
        "posterior_agreement_z": agree,
        "ess_ratio_including_training": ratio_incl_training,
        "plain": res["plain"], "delayed": res["delayed"], "surrogate_held_out": held,
        "other_betas": {n: dict(v, ess_ratio_vs_plain=v["ess_per_s"] / res["plain"]["ess_per_s"])
                        for n, v in res.items() if n not in ("plain", "delayed")},
        "burnin_stats": bstats, "cpu_baseline": cpu, "roofline": None,
    }), flush=True)


def bench_hetero(args):
    """HeteroLikelihood at the GTR+Gamma4 workload's size: K = 4 categories,
    each with its own GTR model and a relaxed LogNormal clock over 8 rate
    categories, a random class per pattern.  One step is one full
    evaluation (branch lengths uploaded, P built per category model, one
    pruning launch for all categories, class-select weighted sum).  Also
    timed: the classes-only path (total recomputed from the site values)."""
    from aspartik_amd import distributions as D
    from aspartik_amd import workloads
    from aspartik_amd.b3 import GTR, Calculator, ClassVector, Clock, DatedTree

    log("[bench] building workload")
    wl = workloads.gtr_gamma4(args.taxa, args.patterns)
    L, P, K = args.taxa, args.patterns, 4
    N = 2 * L - 1
    tree = DatedTree(L, wl["parents"], wl["children"], wl["heights"])
    full = workloads.full_lists(tree)
    rng = np.random.default_rng(2024)
    subs = [GTR(rng.dirichlet([20.0] * 4), rng.dirichlet([4.0] * 6)) for _ in range(K)]
    clocks = [Clock.Relaxed(ClassVector(8, N, rng.integers(0, 8, size=N)), D.LogNormal(0.0, 0.3))
              for _ in range(K)]
    classes = rng.integers(0, K, size=P).astype(np.uint32)
    times = workloads.edge_times(tree, full[1])
    ch = full[1].reshape(-1).astype(np.int64)
    lengths = np.stack([times * c.edge_rates(N)[ch] for c in clocks])
    name = f"hetero4_relaxed_{L}taxa_{P}pat"

    calc = Calculator.HIP().make(wl["codes"], P, K)
    calc.set_weights(wl["weights"])
    for c, s in enumerate(subs):
        calc.set_category_model(c, s.q_matrix(), s.frequencies())
    calc.set_classes(classes)
    # untimed evaluations for the pre-warm period, then the warm-up steps, so
    # that the device runs at its steady clocks when timing starts
    t_end = time.perf_counter() + args.prewarm
    while time.perf_counter() < t_end:
        calc.propose_lengths(*full, lengths)
        calc.total()
        calc.accept()
    for _ in range(args.warmup):
        calc.propose_lengths(*full, lengths)
        calc.total()
        calc.accept()
    calc.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        calc.propose_lengths(*full, lengths)
        calc.total()
        calc.accept()
    dt = time.perf_counter() - t0
    kms, nl = calc.profile_read()
    calc.profile_enable(False)
    kernel_ms = kms / max(1, nl)
    total = calc.total()
    log(f"[bench] hetero full: {args.steps / dt:.1f} evals/s, kernel {kernel_ms * 1e3:.1f} us")
    # classes-only changes: ClassvecFlip-style edits of 1% of the patterns
    perms = []
    for i in range(8):
        cl = classes.copy()
        idx = rng.integers(0, P, size=P // 100)
        cl[idx] = rng.integers(0, K, size=idx.size)
        perms.append(cl)
    t0 = time.perf_counter()
    ncls = max(50, args.steps)
    for i in range(ncls):
        calc.set_classes(perms[i % len(perms)])
        calc.total()
    dt_cls = time.perf_counter() - t0
    calc.set_classes(classes)
    total = calc.total()
    calc.close()

    B = workloads.store_bytes(full[0], full[1], L, K, P, tree_children=tree.children)
    achieved = B / (kernel_ms * 1e-3) / 1e9
    cpu = err = None
    if not args.no_cpu:
        from oracle import oracle as O
        log("[bench] FP64 CPU oracle (one from-scratch evaluation)")
        otree = O.Tree(L, wl["parents"], wl["children"], wl["heights"])
        models = [O.Model.gtr(s.rates, s.freqs) for s in subs]
        er = np.stack([c.edge_rates(N) for c in clocks])
        t0 = time.perf_counter()
        ref, _ = O.components_eval(otree, models, er, wl["codes"], wl["weights"], classes=classes)
        dtc = time.perf_counter() - t0
        cpu = {"value": 1.0 / dtc, "unit": "logL evals/s", "cores": 1, "kind": "port",
               "sample": f"1 full evaluation of the same {L}-taxon {P}-pattern 4-category "
                         f"workload, one thread ({dtc:.1f} s)"}
        err = {"total_rel": abs(total - ref) / abs(ref), "ref_total": ref, "gpu_total": total}
        log(f"[bench] oracle {dtc:.2f} s/eval; rel err {err['total_rel']:.2e}")
    line = {
        "metric": "HeteroLikelihood logL evals/s (relaxed clocks, per-category models)",
        "value": args.steps / dt,
        "unit": "logL evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Yule tree, sequences simulated along it; random classes "
                "and clock categories)",
        "config": {"workload": name, "taxa": L, "patterns": P, "categories": K,
                   "model": "4 x GTR, relaxed LogNormal(0, 0.3) clocks, 8 rate categories",
                   "evaluation": "full (all edges dirty)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(name)},
        "cpu_baseline": cpu,
        "kernel_us": kernel_ms * 1e3,
        "classes_only_updates_per_s": ncls / dt_cls,
        "error_vs_fp64_oracle": err,
    }
    print(json.dumps(line), flush=True)


def bench_storage(args):
    """Partial vs propagation storage on the GTR+Gamma4 workload: the
    autotuner's table over every partial-storage tile shape and the FP32
    propagation layout (full evaluations and one-edge partial evaluations,
    host round trip included), plus b3's two GPU algorithms at FP64
    (partials = B3H_F64_PLAIN, propagations = B3H_F64_PROPAGATION) for
    reference.  The expected time per evaluation weights partial updates
    0.9, as in an MCMC chain."""
    from aspartik_amd import workloads
    from aspartik_amd.b3 import Calculator, DatedTree, autotune
    from aspartik_amd.b3.tuning import default_candidates

    wl = workloads.gtr_gamma4(args.taxa, args.patterns)
    L, P, K = args.taxa, args.patterns, wl["K"]
    tree = DatedTree(L, wl["parents"], wl["children"], wl["heights"])
    evals = max(20, args.steps // 4)
    cands = default_candidates(K) + [("fp64 partials (b3 cuda_partials)", Calculator.HIPPlainF64()),
                                     ("fp64 propagations (b3 cuda_propagations)",
                                      Calculator.HIPPropagationsF64())]
    log(f"[bench] autotuning {len(cands)} candidates, {evals} evaluations each")
    best, trials = autotune(wl["codes"], wl["weights"], tree, wl["model"], wl["rates"],
                            evals=evals, candidates=cands, warm_s=0.5)
    table = [{"layout": t.name, "full_ms": t.full_ms, "partial_ms": t.partial_ms,
              "expected_ms": t.score_ms, "full_kernel_us": t.full_kernel_us,
              "partial_kernel_us": t.partial_kernel_us} for t in trials]
    for t in table:
        log(f"[bench] {t['layout']:42s} full {t['full_ms']:.3f} ms ({t['full_kernel_us']:.1f} us)"
            f"  partial {t['partial_ms']:.3f} ms ({t['partial_kernel_us']:.1f} us)")
    line = {
        "metric": "expected logL evals/s at 90% partial updates, best storage layout",
        "value": 1e3 / trials[0].score_ms,
        "unit": "logL evals/s",
        "n_gpus": 1,
        "steps": evals,
        "warmup": 3,
        "ms_per_step": trials[0].score_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Yule tree, sequences simulated along it)",
        "config": {"workload": f"gtr_gamma4_{L}taxa_{P}pat", "taxa": L, "patterns": P,
                   "categories": K, "chosen": trials[0].name},
        "roofline": None,
        "cpu_baseline": None,
        "layouts": table,
    }
    print(json.dumps(line), flush=True)


def bench_ingest(args):
    """Alignment ingest: a 200-taxon, 1,000,000-site alignment (sites drawn
    from 250,000 simulated columns with 1% gaps, so about 3/4 are repeats)
    written as FASTA text, parsed (b3d_msa_from_fasta) and deduplicated on
    the GPU (b3d_deduplicate, device 0) and on the host.  cpu_baseline: the
    oracle's deduplication on the same MSA."""
    import ctypes as C
    from aspartik_amd import sim
    from aspartik_amd._lib import check, lib
    from aspartik_amd.b3 import GTR, gamma_rates
    from aspartik_amd.data import MSA, deduplicate

    L, sites, distinct = 200, 1_000_000, 250_000
    rng = np.random.default_rng(8)
    par, ch, h = sim.yule_tree(L, rng, 1.0, 0.1)
    freqs = rng.dirichlet([10.0] * 4)
    model = GTR(freqs, rng.dirichlet([1.0] * 6))
    base = sim.dna_codes(sim.simulate_states_fast(par, ch, h, model.q_matrix(), freqs,
                                                  gamma_rates(0.5, 4), distinct, 3))
    base[rng.random(base.shape) < 0.01] = 0      # gaps, shared by a column's repeats
    codes = base[:, rng.integers(0, distinct, size=sites)]
    ascii_ = np.frombuffer(b"-ACMGRSVTWYHKDBN", np.uint8)
    text = b"".join(b">taxon_%d\n" % i + ascii_[row].tobytes() + b"\n"
                    for i, row in enumerate(codes))
    log(f"[bench] FASTA text {len(text) / 1e6:.0f} MB")
    t0 = time.perf_counter()
    hmsa = C.c_void_p()
    check(lib().b3d_msa_from_fasta(text, len(text), C.byref(hmsa)))
    t_parse = time.perf_counter() - t0
    msa = MSA(hmsa)
    deduplicate(msa, device=0)  # warm-up (context, first allocation)
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        dev = deduplicate(msa, device=0)
    t_dev = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    host = deduplicate(msa)
    t_host = time.perf_counter() - t0
    same = all(np.array_equal(a, b) for a, b in zip(dev, host))
    P = dev[0].shape[1]
    log(f"[bench] parse {t_parse:.3f} s, GPU dedup {t_dev * 1e3:.1f} ms, host {t_host:.3f} s, "
        f"{P} patterns, identical={same}")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        t0 = time.perf_counter()
        ol, ow, os2p = O.deduplicate(msa.data)
        t_or = time.perf_counter() - t0
        same = same and np.array_equal(ol, dev[0]) and np.array_equal(ow, dev[1])
        cpu = {"value": sites / t_or, "unit": "sites/s", "cores": 1, "kind": "port",
               "sample": f"the oracle's deduplication of the same {L} x {sites} MSA "
                         f"({t_or:.2f} s)"}
    line = {
        "metric": "alignment sites deduplicated per second (GPU)",
        "value": sites / t_dev,
        "unit": "sites/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": 1,
        "ms_per_step": t_dev * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (columns simulated on a Yule tree, resampled with repeats, 1% gaps)",
        "config": {"workload": f"ingest_{L}taxa_{sites}sites", "taxa": L, "sites": sites,
                   "patterns": P, "fasta_mb": len(text) / 1e6},
        "roofline": None,
        "cpu_baseline": cpu,
        "parse_s": t_parse,
        "parse_mb_per_s": len(text) / 1e6 / t_parse,
        "host_dedup_s": t_host,
        "gpu_dedup_s_including_transfers": t_dev,
        "outputs_identical": bool(same),
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--taxa", type=int, default=128)
    ap.add_argument("--patterns", type=int, default=100_000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--prewarm", type=float, default=2.0,
                    help="seconds of untimed evaluations before the warmup steps, so the "
                         "timed steps run at steady clocks")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only-f32", action="store_true", help="profiling runs: FP32 kernel only")
    ap.add_argument("--tile-patterns", type=int, default=0,
                    help="FP32 calculator: wavefronts per category x 64 (0: by size)")
    ap.add_argument("--lane-states", type=int, default=0, choices=[0, 1, 4],
                    help="FP32 calculator: states per lane (0: 4)")
    ap.add_argument("--workload", choices=["gtr128", "c1", "batch", "codon", "mcmc", "hetero",
                                           "storage", "ingest"],
                    default="gtr128")
    ap.add_argument("--c1-steps", type=int, default=10_000)
    ap.add_argument("--mcmc-taxa", type=int, default=200)
    ap.add_argument("--mcmc-sites", type=int, default=200_000)
    ap.add_argument("--mcmc-steps", type=int, default=100_000)
    ap.add_argument("--burnin", type=int, default=60_000)
    ap.add_argument("--train-schedule", default="60@1e-3,40@2e-4",
                    help="surrogate training: comma-separated EPOCHS@LR phases")
    ap.add_argument("--clip", type=float, default=100.0)
    ap.add_argument("--collect", type=int, default=50_000,
                    help="surrogate training pairs from the last N burn-in steps (0: all)")
    ap.add_argument("--collect-every", type=int, default=1,
                    help="spread the training pairs over the last collect x N burn-in steps: "
                         "blocks of 100 consecutive steps collected, then 100 (N - 1) not")
    ap.add_argument("--betas", default="0.5",
                    help="delayed acceptance: first-stage temperatures, one chain each; "
                         "the first is the headline")
    ap.add_argument("--chain-seed", type=int, default=202,
                    help="seed of the plain chain; delayed chains take the next ones")
    ap.add_argument("--screen-topology", type=int, choices=[0, 1], default=0,
                    help="delayed acceptance: topology moves pass the surrogate stage too")
    ap.add_argument("--narrow-reversible", type=int, choices=[0, 1], default=1,
                    help="ESS workload: NarrowExchange in its reversible form (1), which "
                         "keeps the posterior, or b3's rule (0), which has no reverse move "
                         "when the exchanged child is a leaf")
    ap.add_argument("--codon-taxa", type=int, default=64)
    ap.add_argument("--codon-patterns", type=int, default=20_000)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--batch-tile", default="",
                    help="batched workload: tile shape PxL or PxLxS (tile_patterns x patterns "
                         "per lane x states per lane) instead of the autotuner's choice")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the likelihood shards by pattern with one 8-byte all-reduce per "
              "evaluation; that exchange is not built).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c1":
        return bench_c1(args)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "codon":
        return bench_codon(args)
    if args.workload == "mcmc":
        return bench_mcmc(args)
    if args.workload == "hetero":
        return bench_hetero(args)
    if args.workload == "storage":
        return bench_storage(args)
    if args.workload == "ingest":
        return bench_ingest(args)

    from aspartik_amd import workloads
    from aspartik_amd.b3 import Calculator, DatedTree

    log("[bench] building workload")
    wl = workloads.gtr_gamma4(args.taxa, args.patterns)
    L, P, K = args.taxa, args.patterns, wl["K"]
    tree = DatedTree(L, wl["parents"], wl["children"], wl["heights"])
    full = workloads.full_lists(tree)
    times_full = workloads.edge_times(tree, full[1])
    rates = wl["rates"]
    model = wl["model"]
    workload_name = f"gtr_gamma4_{L}taxa_{P}pat"

    log("[bench] FP32 calculator")
    calc = Calculator.HIP(tile_patterns=args.tile_patterns,
                          lane_states=args.lane_states).make(wl["codes"], P, K)
    calc.set_weights(wl["weights"])
    calc.set_model(model.q_matrix(), model.frequencies(), True)
    dt, total32, kernel_mean_ms, each = time_evals(calc, full, times_full, rates, args.steps,
                                                   args.warmup, prewarm_s=args.prewarm)
    kernel_ms = float(np.median(each))
    evals_per_s = args.steps / dt
    log(f"[bench] fp32 full: {evals_per_s:.1f} evals/s, kernel median {kernel_ms * 1e3:.1f} us "
        f"(mean {kernel_mean_ms * 1e3:.1f} us) after {args.prewarm:.1f} s pre-warm")

    # pipelined (no host round trip per evaluation)
    dt_pipe, _, _, _ = time_evals(calc, full, times_full, rates, args.steps, 2, sync_each=False)

    # partial evaluations: one random non-root edge and its path to the root
    rng = np.random.default_rng(99)
    root = tree.root
    edges = [e for e in rng.choice(2 * L - 2, size=64) if e != root]
    plists = [workloads.path_lists(tree, int(e)) for e in edges]
    ptimes = [workloads.edge_times(tree, pl[1]) for pl in plists]
    for i in range(200):
        calc.propose_times(*plists[i % len(plists)], ptimes[i % len(plists)], rates)
        calc.total()
        calc.accept()
    # the rate unprofiled (MCMC's pattern: propose, read the total, accept),
    # then the kernel time in a second, profiled pass over the same edges;
    # 2,000 updates (about 0.1 s) so that one host hiccup does not set the rate
    nparts = max(2000, args.steps)

    def partial_pass():
        t0 = time.perf_counter()
        for i in range(nparts):
            j = i % len(plists)
            calc.propose_times(*plists[j], ptimes[j], rates)
            calc.total()
            calc.accept()
        return time.perf_counter() - t0

    dt_part = partial_pass()
    calc.profile_enable(True)
    partial_pass()
    pk_ms, pk_n = calc.profile_read()
    calc.profile_enable(False)
    mean_path = float(np.mean([pl[0].size for pl in plists]))
    # mean bytes of the one-edge updates (stored siblings read, path written)
    B_part = float(np.mean([workloads.store_bytes(pl[0], pl[1], L, K, P,
                                                  tree_children=tree.children)
                            for pl in plists]))
    log(f"[bench] fp32 partial: {nparts / dt_part:.1f} evals/s (mean path {mean_path:.1f} nodes)")
    # restore the full state for the error check
    calc.propose_times(*full, times_full, rates)
    total32 = calc.total()
    site32 = calc.site_mixed()
    dev_bytes = calc.device_bytes()
    calc.close()

    f64 = None
    if not args.only_f32:
        log("[bench] FP64 plain calculator")
        c64 = Calculator.HIPPlainF64().make(wl["codes"], P, K)
        c64.set_weights(wl["weights"])
        c64.set_model(model.q_matrix(), model.frequencies(), True)
        steps64 = max(20, args.steps // 4)
        dt64, total64, _, e64 = time_evals(c64, full, times_full, rates, steps64, 5,
                                           prewarm_s=min(1.0, args.prewarm))
        f64 = dict(evals_per_s=steps64 / dt64, kernel_us=float(np.median(e64)) * 1e3,
                   total=total64)
        c64.close()
        log(f"[bench] fp64 plain: {f64['evals_per_s']:.1f} evals/s")

    # bytes the evaluation must move with this storage (the roofline basis),
    # the reference algorithm's byte count (every child read from memory) and
    # flops of one full evaluation
    B = workloads.store_bytes(full[0], full[1], L, K, P, tree_children=tree.children)
    B_ref = workloads.algorithmic_bytes(full[0], full[1], L, root, K, P)
    F = workloads.algorithmic_flops(full[0], full[1], L, K, P)
    achieved = B / (kernel_ms * 1e-3) / 1e9
    traffic = read_traffic(workload_name)

    cpu = None
    err = None
    if not args.no_cpu:
        from oracle import oracle as O
        log("[bench] FP64 CPU oracle (cpu_baseline)")
        otree = O.Tree(L, wl["parents"], wl["children"], wl["heights"])
        om = O.Model.gtr(model.rates, model.freqs)
        ncores, tpc = cpu_threads(K, P)
        t0 = time.perf_counter()
        olik = O.GammaLikelihood(otree, om, wl["codes"], P, wl["weights"], num_categories=K,
                                 alpha=wl["alpha"], threads_per_category=tpc,
                                 parallel_categories=True)
        t_first = time.perf_counter() - t0
        ref_total = olik.likelihood()
        ref_sites = None
        # per-pattern mixed values for the error statement
        site = np.array([olik.site(c) for c in range(K)])
        mx = site.max(0)
        ref_sites = mx + np.log(np.exp(site - mx).sum(0)) + np.log(1.0 / K)
        nev = 0
        t0 = time.perf_counter()
        while True:
            otree.mark_all_edges_updated()
            olik.likelihood()
            olik.accept()
            otree.accept()
            nev += 1
            if time.perf_counter() - t0 > args.cpu_seconds or nev >= 200:
                break
        dtc = time.perf_counter() - t0
        # one-edge partial evaluations over the GPU leg's edges (the edge and
        # its ancestors, tree.rs:357-375)
        npc = 0
        t0 = time.perf_counter()
        while True:
            otree.mark_edge_updated(int(edges[npc % len(edges)]))
            olik.likelihood()
            olik.accept()
            otree.accept()
            npc += 1
            if time.perf_counter() - t0 > args.cpu_seconds / 3 or npc >= 300:
                break
        dtp = time.perf_counter() - t0
        cpu = {"value": nev / dtc, "unit": "logL evals/s", "cores": tpc * K,
               "host_cores": ncores, "threads": tpc * K,
               "kind": "port", "partial_evals_per_s": npc / dtp,
               "sample": f"{nev} full GTR+Gamma4 evaluations of the same {L}-taxon "
                         f"{P}-pattern workload ({dtc:.1f} s) and {npc} one-edge partial "
                         f"evaluations ({dtp:.1f} s); {K} categories run at once as "
                         f"gamma.rs does, {tpc} threads each ({tpc * K} threads on "
                         f"{ncores} host cores); construction {t_first:.1f} s"}
        rel_total = abs(total32 - ref_total) / abs(ref_total)
        per_site = np.abs(site32 - ref_sites) / np.maximum(1.0, np.abs(ref_sites))
        err = dict(total_rel=rel_total, max_site=float(per_site.max()), ref_total=ref_total,
                   gpu_total=total32)
        if f64 is not None:
            err["fp64_gpu_total_rel"] = abs(f64["total"] - ref_total) / abs(ref_total)
        log(f"[bench] oracle {cpu['value']:.3f} evals/s; rel err total {rel_total:.2e}")

    line = {
        "metric": METRIC,
        "value": evals_per_s,
        "unit": "logL evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Yule tree, sequences simulated along it)",
        "config": {"workload": workload_name, "taxa": L, "patterns": P, "categories": K,
                   "model": "GTR+Gamma4 alpha 0.5", "evaluation": "full (all edges dirty)",
                   "seed": wl["seed"], "lane_states": args.lane_states or 4,
                   "tile_patterns": args.tile_patterns or "by size"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "kernel_us": kernel_ms * 1e3,
        "kernel_us_mean": kernel_mean_ms * 1e3,
        "kernel_launches_timed": int(each.size),
        "prewarm_s": args.prewarm,
        "algorithmic_bytes_per_eval": B,
        "traffic_GBps": (traffic / (kernel_ms * 1e-3) / 1e9) if traffic else None,
        "reference_algorithm_bytes_per_eval": B_ref,
        "reference_algorithm_GBps": B_ref / (kernel_ms * 1e-3) / 1e9,
        "algorithmic_flops_per_eval": F,
        "pipelined_evals_per_s": args.steps / dt_pipe,
        "partial_evals_per_s": nparts / dt_part,
        "partial_kernel_us": pk_ms / max(1, pk_n) * 1e3,
        "partial_algorithmic_bytes": B_part,
        "partial_GBps": B_part / (pk_ms / max(1, pk_n) * 1e-3) / 1e9,
        "partial_mean_path_nodes": mean_path,
        "partial_updates_per_s": evals_per_s * len(full[0]) * P * K,
        "fp64_plain": f64,
        "speedup_vs_fp64_plain": (evals_per_s / f64["evals_per_s"]) if f64 else None,
        "error_vs_fp64_oracle": err,
        "device_bytes": dev_bytes,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

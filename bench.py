"""Benchmark: MDS log-likelihood + gradient evaluations per second on one MI355X.

Workload (the metric's configuration): N = 10,000 locations, P = 2; observations
are distances of N(0, I_10) points plus N(0, 0.1^2) noise clipped at 0, with 5%
of the pairs missing (symmetric NaN); truncation on, tau = 2; locations from
N(0, I_2).  One step = one fused evaluation of the log-likelihood and its
N x P gradient with the inputs resident in HBM.  The FP64 engine gives
`value`; the compensated FP32 engine is reported beside it with its error
against FP64.  Each engine runs W warm-up evaluations, then untimed
evaluations for --settle-ms of wall time (the GPU's clock settles over the
first tens of milliseconds of sustained load; the count is reported under
`settle`), then the K timed evaluations.

The dominant kernel is the fused tile kernel.  Its roofline uses the
algorithmic bytes of one evaluation, B = s * N (N - 1) / 2 + 2 * 8 * N * P
(each unordered observation read once, s = 8 for FP64 and 4 for FP32, plus
the locations in and the gradient out), over the kernel's mean duration from
HIP events recorded on the engine's stream.

The CPU baseline times the multi-core SIMD restatement of the reference
engine (two passes, likelihood then gradient, as the reference API makes
them) on the same inputs, on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--settle-ms MS]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6
FP32_VALU_PEAK_TFLOPS = 157.3  # packed FP32; one FP32 FMA per lane per cycle reaches half of it


def _profile_figure(key, summary="tile_kernel_n10000_p2.md"):
    """A per-launch figure from a rocprofv3 summary under profiles/ (by
    default the FP64 tile kernel at the metric's workload,
    profiles/tile_kernel_n10000_p2.md): the number on the line
    "<key>: <value>"; None when absent."""
    from pathlib import Path

    path = Path(__file__).resolve().parent / "profiles" / summary
    try:
        for line in path.read_text().splitlines():
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def _measured_traffic():
    """HBM bytes per FP64 tile-kernel launch (FETCH_SIZE + WRITE_SIZE, scaled
    by a plain copy's reading)."""
    return _profile_figure("traffic_bytes_fp64_tile")


# VALU issue ceiling: 256 CUs x 4 SIMDs, one wave64 instruction per 4 cycles
# per SIMD, at the 2.4 GHz peak engine clock.
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 4 * 2.4e9 / 4


def _valu_roofline(tile_s):
    """The tile kernel against its VALU issue ceiling: wave instructions per
    launch (SQ_INSTS_VALU from the profile) over the measured duration."""
    instr = _profile_figure("valu_wave_instructions_fp64_tile")
    if instr is None:
        return None
    achieved = instr / tile_s
    return {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_WAVE_INSTR_PER_S,
            "unit": "wave instructions/s", "frac": achieved / VALU_PEAK_WAVE_INSTR_PER_S,
            "instructions": instr}


def _flop_roofline(key, seconds, peak_tflops):
    """A tile kernel against the MI355X vector FLOP peak: FLOPs per launch
    (rocprofv3 SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 or _F32, 64 lanes each,
    an FMA counting 2, from the profile summary) over the measured
    duration."""
    flops = _profile_figure(key)
    if flops is None:
        return None
    achieved = flops / seconds / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": peak_tflops, "unit": "TFLOP/s",
            "frac": achieved / peak_tflops, "flops": flops}


def _workload(n=10000):
    from massivemds_amd import datagen

    p, tau = 2, 2.0
    delta = datagen.euclid10(n, noise_sd=0.1, missing=0.05)
    x = datagen.locations(n, p)
    return n, p, tau, delta, x


def _run_gpu(n, p, tau, delta, x, fp32, steps, warmup, settle_ms=0.0):
    """Warm-up evaluations, then (settle_ms > 0) untimed evaluations until
    settle_ms of wall time have passed: from idle the GPU's clock drops
    within a few milliseconds of sustained load and recovers over the next
    tens of milliseconds (kernel traces at N = 10,000: ~100 us per FP64 tile
    pass, then 145 us, then 99-102 us after ~70 ms), so the timed passes
    measure the sustained rate a long sampler run sees."""
    from massivemds_amd.engine import Flags, MultiDimensionalScaling

    flags = Flags.OPENCL | Flags.LEFT_TRUNCATION | (Flags.FLOAT if fp32 else 0)
    e = MultiDimensionalScaling(p, n, flags, 0)
    e.setPairwiseData(delta)
    e.updateLocations(-1, x)
    e.setParameters([tau])
    for _ in range(warmup):
        e.evaluate_async(True)
    e.synchronize()
    settled = 0
    ts = time.perf_counter()
    while (time.perf_counter() - ts) * 1e3 < settle_ms:
        for _ in range(20):
            e.evaluate_async(True)
        e.synchronize()
        settled += 20
    settle_wall_ms = (time.perf_counter() - ts) * 1e3
    # whole-job rate: the timed steps alone, no timing events in the stream
    t0 = time.perf_counter()
    for _ in range(steps):
        e.evaluate_async(True)
    e.synchronize()
    t1 = time.perf_counter()
    # kernel durations (HIP events on the engine's stream) from a second pass
    e.set_timing(True)
    e.get_timing()  # drain
    t0_tile, t0_eval, c0 = e.get_timing()
    for _ in range(steps):
        e.evaluate_async(True)
    e.synchronize()
    tile_ms, eval_ms, count = e.get_timing()
    e.set_timing(False)
    tile_ms, eval_ms, count = tile_ms - t0_tile, eval_ms - t0_eval, count - c0
    s = e.getSumOfIncrements()
    g = e.getLogLikelihoodGradient().reshape(n, p)
    sym, nbytes, tiles = e.layout()
    e.close()
    return {
        "wall_ms_per_step": (t1 - t0) * 1e3 / steps,
        "tile_ms": tile_ms / count,
        "eval_ms": eval_ms / count,
        "sum": s,
        "grad": g,
        "symmetric": sym,
        "packed_bytes": nbytes,
        "tiles": tiles,
        "settle_evaluations": settled,
        "settle_ms": settle_wall_ms,
    }


def _cpu_host():
    """The host the CPU baselines ran on (model, sockets, logical CPUs, SMT)."""
    model, sockets, siblings, cores = None, set(), None, None
    try:
        for line in open("/proc/cpuinfo"):
            key, _, val = line.partition(":")
            key, val = key.strip(), val.strip()
            if key == "model name" and model is None:
                model = val
            elif key == "physical id":
                sockets.add(val)
            elif key == "siblings" and siblings is None:
                siblings = int(val)
            elif key == "cpu cores" and cores is None:
                cores = int(val)
    except OSError:
        pass
    smt = None
    try:
        smt = open("/sys/devices/system/cpu/smt/active").read().strip() == "1"
    except OSError:
        if siblings and cores:
            smt = siblings > cores
    return {"model": model, "sockets": len(sockets) or None, "logical_cpus": os.cpu_count(),
            "usable_cpus": len(os.sched_getaffinity(0)), "smt": smt}


def _cpu_baseline(n, p, tau, delta, x, budget_s=20.0):
    """The CPU restatement of the reference engine timed on this host in the
    three forms the reference builds: serial (1 core, no SIMD), SIMD on 1
    core, and SIMD on all cores (standing in for AVX + TBB).  A full
    evaluation is a likelihood pass plus a gradient pass, as the reference
    API makes them.  The serial and 1-core forms are timed on the first rows
    of the same N x N workload (every row holds N pairs, so the rate scales
    by N / rows); the all-core form on whole evaluations.  Returns the
    all-core form as the baseline, the others beside it."""
    from oracle import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or oracle.max_threads()
    width = oracle.simd_width()
    simd = f"AVX{'-512' if width == 512 else '2' if width == 256 else ' off'}"

    def timed(fn, repeats, budget):
        times = []
        while True:
            t0 = time.perf_counter()
            out = fn()
            times.append(time.perf_counter() - t0)
            if sum(times) + times[-1] > budget or len(times) >= repeats:
                return float(np.median(times)), len(times), out

    # all cores, whole evaluations
    t_all, k_all, (s, g) = timed(lambda: (oracle.simd_sum_of_increments(delta, x, tau, True, threads=threads),
                                          oracle.simd_gradient(delta, x, tau, True, threads=threads)),
                                 5, 0.5 * budget_s)
    # one core, SIMD, a row sample
    rows1 = max(1, min(n, n // 5))
    t_one, k_one, _ = timed(lambda: (oracle.simd_sum_of_increments(delta, x, tau, True, threads=1, rows=rows1),
                                     oracle.simd_gradient(delta, x, tau, True, threads=1, rows=rows1)),
                            3, 0.2 * budget_s)
    # one core, serial scalar, a smaller row sample
    rows0 = max(1, min(n, n // 20))
    t_ser, k_ser, _ = timed(lambda: (oracle.rows_sum_of_increments(delta, x, tau, True, rows0),
                                     oracle.rows_gradient(delta, x, tau, True, rows0)),
                            3, 0.3 * budget_s)
    variants = {
        "serial_1core": {"value": rows0 / (n * t_ser), "unit": "evals/s", "cores": 1,
                         "sample": f"{k_ser} timings of rows 0..{rows0 - 1} (of {n}), median, scaled by "
                                   f"{n}/{rows0}; scalar, no vectorisation"},
        "simd_1core": {"value": rows1 / (n * t_one), "unit": "evals/s", "cores": 1,
                       "sample": f"{k_one} timings of rows 0..{rows1 - 1} (of {n}), median, scaled by "
                                 f"{n}/{rows1}; {simd}"},
        "simd_allcores": {"value": 1.0 / t_all, "unit": "evals/s", "cores": threads,
                          "sample": f"{k_all} full evaluations, median; {simd} + OpenMP, {threads} threads"},
    }
    return {
        "value": 1.0 / t_all,
        "unit": "evals/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{k_all} full evaluations (likelihood pass + gradient pass) at N={n}, P={p}, "
                  f"median; {simd} + OpenMP, {threads} threads",
        "sum": s,
        "grad": g,
        "variants": variants,
    }


def _bench_batch(steps, warmup, count=64, tune_cache=None, no_cpu=False, cpu_budget=10.0):
    """Uneven batch: `count` datasets, N log-uniform in [500, 8000], P in
    {2, 5, 10}, tau in [0.5, 4], truncation, FP32 with FP64 reductions; one
    persistent launch per evaluation against `count` sequential
    single-instance evaluations of the same data.  The CPU figure times the
    all-core SIMD port (likelihood pass + gradient pass, FP64) on the
    smallest datasets within the budget and scales its pair rate to the
    whole batch."""
    from massivemds_amd import datagen
    from massivemds_amd.batch import MultiDimensionalScalingBatch
    from massivemds_amd.engine import Flags, MultiDimensionalScaling

    flags = Flags.OPENCL | Flags.LEFT_TRUNCATION | Flags.FLOAT
    specs = datagen.batch_specs(count)
    b = MultiDimensionalScalingBatch([s[1] for s in specs], [s[0] for s in specs], flags, 0)
    singles = []
    for k, (n, p, tau, seed) in enumerate(specs):
        delta, x = datagen.batch_dataset(n, p, seed)
        b.setPairwiseData(k, delta)
        b.updateLocations(k, x)
        b.setParameters(k, [tau])
        e = MultiDimensionalScaling(p, n, flags, 0)
        e.setPairwiseData(delta)
        e.updateLocations(-1, x)
        e.setParameters([tau])
        singles.append(e)
        del delta
    t0 = time.perf_counter()
    tuned_ms = b.autotune(tune_cache)
    tune_s = time.perf_counter() - t0
    for _ in range(warmup):
        b.evaluate()
    b.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        b.evaluate()
    wall_b = (time.perf_counter() - t0) / steps
    ev_ms, cnt = b.get_timing()
    sums = b.sums()
    for e in singles:
        e.evaluate_async(True)
        e.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for e in singles:
            e.evaluate_async(True)
            e.synchronize()
    wall_s = (time.perf_counter() - t0) / steps
    seq = np.array([e.getSumOfIncrements() for e in singles])
    identical = bool(np.array_equal(sums, seq))
    pairs = sum(n * (n - 1) // 2 for n, _, _, _ in specs)
    algo = sum(4 * n * (n - 1) // 2 + 2 * 8 * n * p for n, p, _, _ in specs)
    out = {
        "workload": f"batch{count}_uneven_fp32_trunc", "datasets": count, "pairs": pairs,
        "batch_ms": ev_ms / cnt, "batch_wall_ms": wall_b * 1e3, "sequential_wall_ms": wall_s * 1e3,
        "speedup_vs_sequential": wall_s / wall_b, "batch_evals_per_s": 1.0 / wall_b,
        "achieved_GBs": algo / (ev_ms / cnt * 1e-3) / 1e9, "algorithmic_bytes": algo,
        "config": dict(zip(["wavesPerBlock", "blocksPerCU", "tilesPerItem", "strategy"], b.get_config())),
        "autotune_s": tune_s, "autotune_best_ms": tuned_ms, "bitwise_equal_to_sequential": identical,
    }
    achieved = algo / (ev_ms / cnt * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": _profile_figure(
                           "traffic_bytes_batch", "batch64_uneven_fp32.md")}
    for e in singles:
        e.close()
    b.close()
    if not no_cpu:
        from oracle import oracle

        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or oracle.max_threads()
        spent, done_pairs, used = 0.0, 0, 0
        for n, p, tau, seed in sorted(specs):
            if spent > cpu_budget and used:
                break
            delta, x = datagen.batch_dataset(n, p, seed)
            t0 = time.perf_counter()
            oracle.simd_sum_of_increments(delta, x, tau, True, threads=threads)
            oracle.simd_gradient(delta, x, tau, True, threads=threads)
            spent += time.perf_counter() - t0
            done_pairs += n * (n - 1) // 2
            used += 1
        rate = done_pairs / spent / pairs  # batch evaluations per second
        out["cpu_baseline"] = {"value": rate, "unit": "batch evals/s", "cores": threads, "kind": "port",
                               "sample": f"the {used} smallest datasets (likelihood pass + gradient pass, FP64, "
                                         f"SIMD + OpenMP, {threads} threads), pair rate scaled to the batch's "
                                         f"{pairs} pairs"}
        out["gpu_vs_cpu"] = {"speedup": out["batch_evals_per_s"] / rate}
    return out


def _bench_hmc(iterations=2000, burn_in=500, n=5000, p=2):
    """Plain HMC on the GPU: N = 5000, P = 2, FP64, euclid10 with 2% missing,
    truncation, 20 leapfrog steps over trajectory 0.1, tau and the trait
    precision learned, each trajectory one hipGraph launch.  Two runs: the
    precision updated by the reference's random walk of fixed width 0.01
    (R/Untitled.R:668-691), and with that width adapted during burn-in."""
    from massivemds_amd import datagen, hmc

    delta = datagen.euclid10(n, missing=0.02)
    kept = iterations - burn_in
    out = {"workload": f"hmc_n{n}_p{p}_fp64_trunc_missing2", "iterations": iterations, "burn_in": burn_in}
    for name, adapt in (("reference_width", False), ("adapted_width", True)):
        print(f"hmc5000: {name}", file=sys.stderr, flush=True)
        res = hmc.hmcsampler(iterations, burnIn=burn_in, data=delta, latentDimension=p, learnPrec=True,
                             learnTraitPrec=True, trajectory=0.1, mdsPrecision=1.0, truncation=True, seed=666,
                             adaptPrecisionWidth=adapt)
        out[name] = {
            "accept_rate": res["acprat"], "precision_accept_rate": res["precision_acceptance"],
            "precision_width": res["precision_width"],
            "sampling_seconds": res["Time"], "ms_per_iteration": res["Time"] / kept * 1e3,
            "min_ess": res["min_ess"], "min_ess_per_s": res["min_ess_per_s"],
            "ess": {k: round(v, 1) for k, v in res["ess"].items()},
            "posterior_mean_precision": float(np.mean(res["precision"])),
            "evaluations": res["evaluations"], "total_seconds": res["total_seconds"],
        }
    return out


def _bench_hmc_chains(chains=4, iterations=2000, burn_in=500, n=5000, p=2):
    """Many chains in one launch per leapfrog step: `chains` HMC chains of
    the N = 5000 workload (shared observations, seeds 666 + k, adapted
    precision width), each trajectory of every chain one hipGraph launch,
    against one chain of the single-chain sampler on the same data.
    Reports time per iteration of all chains and the summed min-ESS/s."""
    from massivemds_amd import datagen, hmc
    from massivemds_amd.batch import MultiDimensionalScalingBatch
    from massivemds_amd.engine import Flags

    delta = datagen.euclid10(n, missing=0.02)
    kept = iterations - burn_in
    kw = dict(iterations=iterations, burnIn=burn_in, leapfrogSteps=20, trajectory=0.1, learnPrecision=1,
              learnTraitPrecision=1, adaptPrecisionWidth=1)
    print("hmc5000 chains: single chain", file=sys.stderr, flush=True)
    single = hmc.hmcsampler(iterations, burnIn=burn_in, data=delta, latentDimension=p, learnPrec=True,
                            learnTraitPrec=True, trajectory=0.1, mdsPrecision=1.0, truncation=True, seed=666,
                            adaptPrecisionWidth=True)
    flags = Flags.OPENCL | Flags.LEFT_TRUNCATION
    b = MultiDimensionalScalingBatch([p] * chains, [n] * chains, flags, 0)
    b.setPairwiseData(0, delta)
    for k in range(1, chains):
        b.sharePairwiseData(k, 0)
    for k in range(chains):
        b.setParameters(k, [1.0])
    x0 = datagen.normals(666, 3, n * p).reshape(n, p)  # the single-chain sampler's default start
    print(f"hmc5000 chains: {chains} chains in one launch", file=sys.stderr, flush=True)
    traces, _, stats = b.run_hmc([x0] * chains, hmc.default_options(seed=666, **kw))
    b.close()
    cols = [c for c in hmc.TRACE_COLUMNS]
    per_chain = []
    for k in range(chains):
        ess = {c: hmc.effective_sample_size(traces[k][:, j]) for j, c in enumerate(cols)}
        finite = [v for v in ess.values() if np.isfinite(v)]
        per_chain.append({"accept_rate": stats[k].acceptRate, "precision_accept_rate": stats[k].precisionAcceptRate,
                          "min_ess": min(finite), "ess": {c: round(v, 1) for c, v in ess.items()}})
    seconds = stats[0].samplingSeconds
    total_min_ess = sum(c["min_ess"] for c in per_chain)
    out = {
        "workload": f"hmc_n{n}_p{p}_fp64_trunc_missing2_chains{chains}", "chains": chains,
        "iterations": iterations, "burn_in": burn_in,
        "batched": {"sampling_seconds": seconds, "ms_per_iteration_all_chains": seconds / kept * 1e3,
                    "summed_min_ess": total_min_ess, "summed_min_ess_per_s": total_min_ess / seconds,
                    "chains": per_chain},
        "single_chain": {"sampling_seconds": single["Time"], "ms_per_iteration": single["Time"] / kept * 1e3,
                         "min_ess": single["min_ess"], "min_ess_per_s": single["min_ess_per_s"]},
    }
    out["throughput_ratio"] = (chains * single["Time"]) / seconds
    out["min_ess_per_s_ratio"] = out["batched"]["summed_min_ess_per_s"] / single["min_ess_per_s"]
    print(json.dumps(out))


def _bench_leapfrog(steps, warmup, settle_ms, n=5000, p=2, tau=1.0):
    """The gradient pass of one leapfrog step of the HMC workload (N = 5000,
    P = 2, FP64, euclid10 with 2% missing, truncation, the sampler's starting
    precision), launched directly instead of from the sampler's graph, so
    that its kernels can be traced: one fused likelihood + gradient
    evaluation per step, inputs resident."""
    from massivemds_amd import datagen

    delta = datagen.euclid10(n, missing=0.02)
    x = datagen.locations(n, p)
    r = _run_gpu(n, p, tau, delta, x, False, steps, warmup, settle_ms)
    print(json.dumps({
        "workload": f"leapfrog_gradient_n{n}_p{p}_fp64_trunc_missing2", "steps": steps, "warmup": warmup,
        "settle": {"ms": r["settle_ms"], "evaluations": r["settle_evaluations"]},
        "evals_per_s": 1e3 / r["wall_ms_per_step"], "us_per_eval": r["wall_ms_per_step"] * 1e3,
        "tile_kernel_us": r["tile_ms"] * 1e3, "eval_us_events": r["eval_ms"] * 1e3, "tiles": r["tiles"]}))


def _bench_surrogate(iterations=8000, n=20000, p=5, sweep_iterations=2000, sweep=(2, 5, 10, 20), grid=True):
    """Surrogate-trajectory HMC against plain HMC: N = 20000, P = 5, FP64,
    euclid10 with 2% missing, truncation, precision fixed at its plug-in value
    at the posterior mode, trait precision I.  Warm-up: 100 plain iterations
    of 20 leapfrog steps (2000 exact gradient evaluations, all stored as
    training snapshots) with dual-averaged step size; training: 4 epochs over
    the snapshots on the GPU; then `iterations` of each sampler from the same
    state and step size (the surrogate one with its exact-gradient
    fallback), and the surrogate chain tuned on its own for L = 2, 5, 10, 20
    (dual-averaged step, no fallback, `sweep_iterations` each).  Reports
    min-ESS/s of every run, the ratios to plain HMC and the largest
    |difference| of posterior means in Monte Carlo standard errors."""
    from massivemds_amd import datagen
    from massivemds_amd.surrogate import compare_samplers

    delta = datagen.euclid10(n, missing=0.02)
    r = compare_samplers(delta, p, iterations, warmup_evals=2000, leapfrog=20, truncation=True, epochs=4,
                         sweep=sweep, sweep_iterations=sweep_iterations, grid=grid,
                         log=lambda m: print(f"surrogate20000: {m}", file=sys.stderr, flush=True))
    del delta
    keys = ("accept_rate", "seconds", "ms_per_iteration", "min_ess", "min_ess_per_s", "surrogate_iterations",
            "exact_iterations", "surrogate_accept_rate", "step_size")

    def brief(v):
        return {k: (round(float(v[k]), 6) if isinstance(v[k], float) else v[k]) for k in keys}

    out = {
        "metric": "min-ESS/s surrogate vs plain HMC", "value": round(r["ess_per_s_ratio"], 4), "unit": "ratio",
        "higher_is_better": True, "dtype": "f64 likelihood, bf16 network",
        "config": {"workload": f"surrogate_hmc_n{n}_p{p}_fp64_trunc_missing2", "N": n, "P": p,
                   "iterations": iterations, "leapfrog": 20, "warmup_evaluations": 2000, "epochs": 4},
        "precision": r["precision"], "step_size": r["step_size"], "warmup_accept": r["warmup_accept"],
        "snapshots": r["snapshots"], "train_loss": r["train_loss"], "map_s": r["map_s"],
        "warmup_s": r["warmup_s"], "train_s": r["train_s"], "max_mean_z": r["max_mean_z"],
        "plain": brief(r["plain"]), "surrogate": brief(r["surrogate"]),
        "plain_ess": [round(float(v), 1) for v in r["plain"]["ess"]],
        "tuned": [dict(brief(t), leapfrog=t["leapfrog"], max_mean_z=round(t["max_mean_z"], 2)) for t in r["tuned"]],
        "grid": [{"leapfrog": g["leapfrog"], "step_size": g["step_size"], "trajectory": g["trajectory"],
                  "accept_rate": round(g["accept_rate"], 4), "ms_per_iteration": round(g["ms_per_iteration"], 4)}
                 for g in r["grid"]],
        "note": "best_tuned_ratio counts only tuned runs whose posterior means match the plain chain's within "
                "2 Monte Carlo SE (null: none did)",
        "best_tuned_ratio": r["best_tuned_ratio"], "best_tuned_z": r["best_tuned_z"],
        "columns": r["columns"],
    }
    return out


def _bench_row_updates(moves=2000, n=10000, p=2):
    """Single-location moves through the reference API (updateLocations(i, .)
    then getSumOfIncrements), the O(N) row update, against full sums, at the
    metric's workload (N = 10000, P = 2, truncation, tau = 2, 5% missing)."""
    from massivemds_amd import datagen
    from massivemds_amd.engine import Flags, MultiDimensionalScaling

    delta = datagen.euclid10(n, noise_sd=0.1, missing=0.05)
    x = datagen.locations(n, p)
    out = {"workload": f"row_updates_n{n}_p{p}_trunc_tau2_missing5", "moves": moves}
    for name, fp32 in (("f64", False), ("f32", True)):
        flags = Flags.OPENCL | Flags.LEFT_TRUNCATION | (Flags.FLOAT if fp32 else 0)
        e = MultiDimensionalScaling(p, n, flags, 0)
        e.setPairwiseData(delta)
        e.setParameters([2.0])
        xs = x.copy()
        e.updateLocations(-1, xs)
        e.getSumOfIncrements()
        rng = np.random.default_rng(0)
        rows = rng.integers(0, n, moves)
        steps = 0.01 * rng.standard_normal((moves, p))
        t0 = time.perf_counter()
        for k in range(moves):
            i = int(rows[k])
            xs[i] += steps[k]
            e.updateLocations(i, xs[i])
            e.getSumOfIncrements()
        t_row = (time.perf_counter() - t0) / moves
        s_inc = e.getSumOfIncrements()
        f, r = e.counters()
        t0 = time.perf_counter()
        for _ in range(50):
            e.makeDirty()
            s_full = e.getSumOfIncrements()
        t_full = (time.perf_counter() - t0) / 50
        e.close()
        out[name] = {"moves_per_s": 1.0 / t_row, "us_per_move": t_row * 1e6, "full_sum_us": t_full * 1e6,
                     "speedup_vs_full": t_full / t_row, "row_updates": r, "full_evaluations": f,
                     "rel_err_vs_full_recompute": abs(s_inc - s_full) / abs(s_full)}
    print(json.dumps(out))


def _bench_beast(steps=2000, warmup=20):
    """Real data: the 372 H3 influenza antigenic distances and the BEAST tip
    locations of two logged states (the reference's inst/extdata, committed
    as tests/golden/beast_h3_kat.npz by make_beast_kat.py, read with
    massivemds_amd.beastio), non-truncated as BEAST ran it, at the logged
    precision: the log-likelihood against BEAST's logged value, and fused
    likelihood + gradient evaluations per second (the whole problem is
    1.1 MB: launch-bound, not a roofline case)."""
    from pathlib import Path

    from massivemds_amd import mds

    z = np.load(Path(__file__).resolve().parent / "tests" / "golden" / "beast_h3_kat.npz")
    delta = z["delta_x1e4"].astype(np.float64) / 10000.0
    out = {"workload": "beast_h3_n372_p2_notrunc", "N": int(delta.shape[0]), "P": 2, "steps": steps}
    for tag, xk, tk, lk in (("state1", "locations", "precision", "loglik"),
                            ("state2", "locations2", "precision2", "loglik2")):
        eng = mds.createEngine(2, delta.shape[0], truncation=False, gpu=1)
        mds.setPairwiseData(eng, delta)
        mds.updateLocations(eng, z[xk])
        mds.setPrecision(eng, float(z[tk]))
        ll = mds.getLogLikelihood(eng)
        e = eng.engine
        for _ in range(warmup):
            e.evaluate_async(True)
        e.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            e.evaluate_async(True)
        e.synchronize()
        dt = (time.perf_counter() - t0) / steps
        out[tag] = {"loglik": ll, "beast_loglik": float(z[lk]), "abs_err": abs(ll - float(z[lk])),
                    "rel_err": abs(ll - float(z[lk])) / abs(float(z[lk])), "evals_per_s": 1.0 / dt,
                    "us_per_eval": dt * 1e6}
        e.close()
    print(json.dumps(out))


def _bench_power(seconds=6.0, interval=0.25, n=10000):
    """Board power and clocks under sustained load: the metric's workload
    (N = 10,000, P = 2) evaluated back to back for `seconds` in FP64, then in
    FP32, while a thread samples `amd-smi metric --power --clock --json`
    every `interval` seconds (read-only queries).  Per phase: the evaluation
    rate over the whole phase, and the median socket power and gfx clock over
    its second half (the board's telemetry refreshes about once a second and
    the power manager takes a second or so to settle)."""
    import shutil
    import subprocess
    import threading

    from massivemds_amd.engine import Flags, MultiDimensionalScaling

    smi = shutil.which("amd-smi") or "/opt/rocm/bin/amd-smi"
    samples = []
    done = threading.Event()

    def sample():
        while not done.is_set():
            t = time.perf_counter()
            try:
                r = subprocess.run([smi, "metric", "--power", "--clock", "--json"], capture_output=True, text=True,
                                   timeout=10)
            except (OSError, subprocess.SubprocessError):
                return  # no telemetry: the phases report their rates alone
            try:
                g = json.loads(r.stdout)
                g = g["gpu_data"][0] if isinstance(g, dict) else g[0]
                clk = [v["clk"]["value"] for k, v in g["clock"].items() if k.startswith("gfx_")]
                samples.append((t, float(g["power"]["socket_power"]["value"]), sum(clk) / len(clk)))
            except (ValueError, KeyError, IndexError, TypeError):
                pass
            done.wait(interval)

    _, p, tau, delta, x = _workload(n)
    th = threading.Thread(target=sample, daemon=True)
    th.start()
    time.sleep(2.0)
    phases = {}
    for fp32 in (False, True):
        e = MultiDimensionalScaling(p, n, Flags.OPENCL | Flags.LEFT_TRUNCATION | (Flags.FLOAT if fp32 else 0), 0)
        e.setPairwiseData(delta)
        e.updateLocations(-1, x)
        e.setParameters([tau])
        t0 = time.perf_counter()
        count = 0
        while time.perf_counter() - t0 < seconds:
            for _ in range(50):
                e.evaluate_async(True)
            e.synchronize()
            count += 50
        t1 = time.perf_counter()
        e.close()
        phases["fp32" if fp32 else "fp64"] = (t0, t1, count / (t1 - t0))
        time.sleep(2.0)
    done.set()
    th.join()
    idle = [s for s in samples if s[0] < phases["fp64"][0]]
    out = {"workload": f"power_n{n}_p2", "seconds_per_phase": seconds, "samples": len(samples),
           "idle": {"socket_power_W": float(np.median([s[1] for s in idle])) if idle else None,
                    "gfx_clock_MHz": float(np.median([s[2] for s in idle])) if idle else None}}
    for name, (t0, t1, rate) in phases.items():
        late = [s for s in samples if (t0 + t1) / 2 <= s[0] <= t1]
        out[name] = {"evals_per_s": rate,
                     "socket_power_W": float(np.median([s[1] for s in late])) if late else None,
                     "gfx_clock_MHz": float(np.median([s[2] for s in late])) if late else None,
                     "gfx_clock_MHz_range": [min(s[2] for s in late), max(s[2] for s in late)] if late else None}
    return out


def _bench_reference_loop(iterations=100, n=1000, p=2, tau=1.0, no_cpu=False, cpu_budget=10.0):
    """The reference benchmark's loop (src/benchmark.cpp:248-307) on the
    reference's own CPU-sized case: N = 1000, P = 2, FP64, euclid10
    observations without noise or missing pairs, tau = 1, truncation.  Every
    iteration stores the state, uploads fresh locations (:257), reads the
    sum of increments, restores the stored state with probability 0.75 or
    accepts the move (:278-283), and reads the gradient of the state it
    kept (:296) -- all through the API.  Reported: that rate, the
    device-resident rate of one fused evaluation, and the CPU port
    (serial, and all-core SIMD) doing the same work -- a likelihood pass at
    the moved locations and a gradient pass at the kept ones -- on the first
    iterations, with the GPU's relative error against the serial one."""
    from massivemds_amd import datagen
    from massivemds_amd.engine import Flags, MultiDimensionalScaling

    delta = datagen.euclid10(n)
    xs = [datagen.locations(n, p, key=1 + k) for k in range(iterations + 1)]
    restore = datagen.uniforms(666, 6, iterations) < 0.75
    e = MultiDimensionalScaling(p, n, Flags.OPENCL | Flags.LEFT_TRUNCATION, 0)
    e.setPairwiseData(delta)
    e.setParameters([tau])
    for k in range(3):  # warm-up
        e.updateLocations(-1, xs[k])
        e.getSumOfIncrements()
        e.getLogLikelihoodGradient()
    e.updateLocations(-1, xs[0])
    e.getSumOfIncrements()
    e.getLogLikelihoodGradient()
    kept, sums, grads, kept_at = 0, [], [], []
    t0 = time.perf_counter()
    for k in range(iterations):
        e.storeState()
        e.updateLocations(-1, xs[k + 1])
        sums.append(e.getSumOfIncrements())
        if restore[k]:
            e.restoreState()
        else:
            e.acceptState()
            kept = k + 1
        grads.append(e.getLogLikelihoodGradient())
        kept_at.append(kept)
    api_s = time.perf_counter() - t0
    e.close()
    resident = _run_gpu(n, p, tau, delta, xs[0], False, 200, 10)
    out = {
        "workload": "mds_reference_loop_n1000_p2_trunc_tau1",
        "N": n, "P": p, "precision": tau, "iterations": iterations,
        "restored": int(restore.sum()),
        "api_evals_per_s": iterations / api_s,
        "api_us_per_iteration": api_s / iterations * 1e6,
        "resident_evals_per_s": 1e3 / resident["wall_ms_per_step"],
        "tile_kernel_us": resident["tile_ms"] * 1e3,
        # 4.0 MB per evaluation, resident in the 256 MB Infinity Cache after
        # the first pass: an effective rate against the HBM peak, not an HBM
        # measurement
        "roofline": {"bound": "hbm", "achieved": (8 * n * (n - 1) // 2 + 16 * n * p) / (resident["tile_ms"] * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s (effective, cache-resident)",
                     "frac": (8 * n * (n - 1) // 2 + 16 * n * p) / (resident["tile_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None},
        "data": "synthetic (euclid10 distances, no noise, no missing; fresh N(0, I_2) locations per iteration)",
    }
    if not no_cpu:
        from oracle import oracle

        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or oracle.max_threads()
        ser, allc, errs_s, errs_g = [], [], [], []
        k = 0
        while k < iterations and (sum(ser) + sum(allc) < cpu_budget or k < 2):
            xm, xk = xs[k + 1], xs[kept_at[k]]
            t0 = time.perf_counter()
            s_ref = oracle.sum_of_increments(delta, xm, tau, True)
            g_ref = oracle.gradient(delta, xk, tau, True)
            ser.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            oracle.simd_sum_of_increments(delta, xm, tau, True, threads=threads)
            oracle.simd_gradient(delta, xk, tau, True, threads=threads)
            allc.append(time.perf_counter() - t0)
            errs_s.append(abs(sums[k] - s_ref) / abs(s_ref))
            errs_g.append(float(np.max(np.abs(grads[k].reshape(n, p) - g_ref.reshape(n, p)))
                                / np.max(np.abs(g_ref))))
            k += 1
        out["cpu_baseline"] = {"value": 1.0 / float(np.median(allc)), "unit": "iterations/s", "cores": threads,
                               "kind": "port",
                               "sample": f"the first {k} iterations (likelihood pass at the moved locations + "
                                         f"gradient pass at the kept ones), median; SIMD + OpenMP, {threads} threads"}
        out["cpu_serial_1core"] = {"value": 1.0 / float(np.median(ser)), "unit": "iterations/s", "cores": 1,
                                   "sample": f"the first {k} iterations, median; the serial restatement"}
        out["cpu_host"] = _cpu_host()
        out["gpu_vs_cpu"] = {"api_speedup_vs_allcores": out["api_evals_per_s"] * float(np.median(allc)),
                             "api_speedup_vs_serial": out["api_evals_per_s"] * float(np.median(ser)),
                             "max_rel_err_sum": max(errs_s), "max_rel_err_grad": max(errs_g),
                             "iterations_checked": k}
    return out


def _calibrate_copy(mib):
    """A plain elementwise copy kernel of known size (torch, as plumbing)."""
    import torch

    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.add(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "elementwise copy", "bytes_read": n * 4, "bytes_written": n * 4,
                      "dispatches": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--settle-ms", type=float, default=100.0,
                    help="untimed evaluations after the warm-up for this many ms of wall time (the GPU clock "
                         "settles under sustained load), before each timed pass of the default and leapfrog5000 "
                         "workloads; 0: none")
    ap.add_argument("--no-extra", action="store_true",
                    help="default workload only: skip the reference loop at N = 1000, the uneven batch and the "
                         "HMC sampler at N = 5000 that the line carries beside it")
    ap.add_argument("--n", type=int, default=10000,
                    help="locations for the default workload (the metric is quoted at 10000; other sizes "
                         "for scaling studies)")
    ap.add_argument("--workload", default="n10000",
                    choices=["n10000", "n1000", "batch64", "hmc5000", "hmc5000x4", "leapfrog5000",
                             "surrogate20000", "rowupdate10000", "beast372", "power10000"],
                    help="n10000: the metric's workload (default); n1000: the reference benchmark's loop "
                         "at N = 1000 (fresh locations per iteration, through the API); batch64: the uneven batch; "
                         "hmc5000: plain HMC; hmc5000x4: four HMC chains in one launch per leapfrog step; "
                         "leapfrog5000: the HMC workload's per-step gradient pass, launched "
                         "directly (for kernel traces); surrogate20000: surrogate against plain HMC; "
                         "rowupdate10000: single-location moves through the O(N) update; "
                         "beast372: real BEAST data (N = 372) against the logged likelihood; "
                         "power10000: board power and clocks while the metric's workload runs sustained")
    ap.add_argument("--iterations", type=int, default=8000, help="sampling iterations (surrogate20000)")
    ap.add_argument("--sweep-iterations", type=int, default=2000,
                    help="sampling iterations per tuned surrogate run (surrogate20000)")
    ap.add_argument("--hmc-iterations", type=int, default=2000,
                    help="iterations of the HMC workloads (hmc5000, hmc5000x4; the workload is quoted at 2000)")
    ap.add_argument("--hmc-burn-in", type=int, default=500, help="burn-in of the HMC workloads")
    ap.add_argument("--tune-cache", default=None, help="autotuner cache file (batch64)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (read + write), to scale "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding is a "
              "design, not a measured path", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        _calibrate_copy(args.calibrate_copy)
        return
    if args.workload == "n1000":
        print(json.dumps(_bench_reference_loop(no_cpu=args.no_cpu)))
        return
    if args.workload == "batch64":
        print(json.dumps(_bench_batch(args.steps, args.warmup, tune_cache=args.tune_cache, no_cpu=args.no_cpu)))
        return
    if args.workload == "hmc5000":
        print(json.dumps(_bench_hmc(args.hmc_iterations, args.hmc_burn_in)))
        return
    if args.workload == "hmc5000x4":
        _bench_hmc_chains(iterations=args.hmc_iterations, burn_in=args.hmc_burn_in)
        return
    if args.workload == "leapfrog5000":
        _bench_leapfrog(args.steps, args.warmup, args.settle_ms)
        return
    if args.workload == "surrogate20000":
        print(json.dumps(_bench_surrogate(args.iterations, sweep_iterations=args.sweep_iterations)))
        return
    if args.workload == "rowupdate10000":
        _bench_row_updates()
        return
    if args.workload == "beast372":
        _bench_beast()
        return
    if args.workload == "power10000":
        print(json.dumps(_bench_power()))
        return

    n, p, tau, delta, x = _workload(args.n)
    r64 = _run_gpu(n, p, tau, delta, x, False, args.steps, args.warmup, args.settle_ms)
    r32 = _run_gpu(n, p, tau, delta, x, True, args.steps, args.warmup, args.settle_ms)

    pairs = n * (n - 1) // 2
    bytes64 = 8 * pairs + 2 * 8 * n * p
    bytes32 = 4 * pairs + 2 * 8 * n * p
    achieved = bytes64 / (r64["tile_ms"] * 1e-3) / 1e9

    const = 0.5 * (math.log(tau) - math.log(2 * math.pi)) * pairs
    ll64 = const - r64["sum"]
    ll32 = const - r32["sum"]

    out = {
        "metric": "loglik+grad evals/s at N=10k,P=2 (FP64/FP32); min-ESS/s surrogate vs plain HMC",
        "value": 1e3 / r64["wall_ms_per_step"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle": {"ms": r64["settle_ms"], "evaluations": r64["settle_evaluations"],
                   "note": "untimed evaluations after the warm-up until the GPU clock has settled under load"},
        "ms_per_step": r64["wall_ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (euclid10 distances + N(0,0.1^2) noise, 5% symmetric NaN; portable PRNG seeds 666/667)",
        "config": {"workload": f"mds_loglik_grad_n{n}_p2_trunc_tau2_missing5", "N": n, "P": p,
                   "precision": tau, "truncation": True, "missing": 0.05},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": _measured_traffic() if n == 10000 else None,  # profiled at N = 10000
        },
        "roofline_valu": _valu_roofline(r64["tile_ms"] * 1e-3) if n == 10000 else None,
        "roofline_fp64": (_flop_roofline("fp64_flops_fp64_tile", r64["tile_ms"] * 1e-3, FP64_VALU_PEAK_TFLOPS)
                          if n == 10000 else None),
        "fp32": {
            "value": 1e3 / r32["wall_ms_per_step"],
            "ms_per_step": r32["wall_ms_per_step"],
            "tile_kernel_ms": r32["tile_ms"],
            "achieved_GBs": bytes32 / (r32["tile_ms"] * 1e-3) / 1e9,
            "roofline_fp32": (_flop_roofline("fp32_flops_fp32_tile", r32["tile_ms"] * 1e-3, FP32_VALU_PEAK_TFLOPS)
                              if n == 10000 else None),
            "rel_err_loglik_vs_fp64": abs(ll32 - ll64) / abs(ll64),
            "rel_err_grad_vs_fp64": float(np.max(np.abs(r32["grad"] - r64["grad"])) / np.max(np.abs(r64["grad"]))),
        },
        "kernel": {"tile_kernel_ms": r64["tile_ms"], "eval_ms": r64["eval_ms"],
                   "algorithmic_bytes": bytes64, "packed_observation_bytes": r64["packed_bytes"],
                   "tiles": r64["tiles"], "symmetric_layout": r64["symmetric"]},
    }
    if not args.no_cpu:
        cpu = _cpu_baseline(n, p, tau, delta, x, args.cpu_budget)
        rel_s = abs(cpu["sum"] - r64["sum"]) / abs(cpu["sum"])
        rel_g = float(np.max(np.abs(cpu["grad"] - r64["grad"])) / np.max(np.abs(cpu["grad"])))
        out["cpu_baseline"] = {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["cpu_variants"] = cpu["variants"]
        out["cpu_host"] = _cpu_host()
        out["gpu_vs_cpu"] = {"speedup": out["value"] / cpu["value"], "rel_err_sum": rel_s,
                             "rel_err_grad": rel_g,
                             "speedup_vs_serial_1core": out["value"] / cpu["variants"]["serial_1core"]["value"],
                             "speedup_vs_simd_1core": out["value"] / cpu["variants"]["simd_1core"]["value"]}
    if not args.no_extra and n == 10000:
        # the other single-GPU configurations of the metric's family, each
        # with its own rate, roofline figures and CPU comparison
        out["other_configs"] = {
            "reference_loop_n1000": _bench_reference_loop(no_cpu=args.no_cpu, cpu_budget=min(args.cpu_budget, 10.0)),
            "batch64_uneven_fp32": _bench_batch(20, 3, no_cpu=args.no_cpu,
                                                cpu_budget=min(args.cpu_budget, 10.0)),
            "hmc_n5000": _bench_hmc(),
            # a short form of the surrogate workload (the full comparison, 8,000
            # iterations per chain and the complete sweep, is --workload surrogate20000)
            "surrogate_n20000_p5_short": _bench_surrogate(iterations=500, sweep_iterations=500, sweep=(2, 10),
                                                          grid=False),
            # the metric's workload sustained for 4 s per precision, with the
            # board's power and clocks (--workload power10000 runs 6 s)
            "power_n10000": _bench_power(seconds=4.0),
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

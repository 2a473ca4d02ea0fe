#!/usr/bin/env python3
"""Benchmark: exact swap-deltas per second on one MI355X.

Workload (the full-size single-GPU swap-delta configuration): a synthetic
male/female pair with n = 18,524 vertices and ~4M directed edges per graph
(power-law out-degrees, alpha 2.1; lognormal(0.5, 1.2) integer weights >= 1;
female = male relabelled with 20% of edges reweighted/rewired), initial
matching = the greedy matching of the pair.  One step evaluates the exact
swap delta of 2^28 uniform candidate pairs, generated inside the delta kernel
from a counter-based stream (a new stream per step), writing 2^28 int64
deltas to HBM.  The 3-vertex permutation kernel is timed on 2^26 uniform
triples in the same run and reported under "rot3".

value = candidates / second over the K timed steps (graph, matching and
output buffers already resident in HBM).  roofline.achieved = algorithmic
bytes per launch (48 + 28*D(u,v) per pair, summed exactly) / mean launch time
measured with HIP events on the library's stream.  cpu_baseline = the
reference's own calculate_swap_delta (src/graph-alignment-optimized.c:321-408,
compiled from its sources into oracle/_ref) over the first candidates of the
last timed stream, OpenMP dynamic schedule, on this host; its deltas are
compared with the device's.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "brain-match_amd"))

HBM_PEAK_GBS = 8000.0
N_VERTICES = 18524
N_EDGES = 4_000_000


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--candidates", type=int, default=1 << 28)
    ap.add_argument("--triples", type=int, default=1 << 26)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-rot3", action="store_true")
    ap.add_argument("--pi0", choices=["greedy", "shuffle"], default="greedy")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the multi-GPU path is not built -- see the sharding paragraph of the design notes.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import brainmatch as bm

    t = time.time()
    male, female, truth, pi_shuffled = bm.synth_pair(N_VERTICES, N_EDGES / N_VERTICES, bm.WEIGHTS_LOGNORMAL, 0.3,
                                                     seed=1)
    log(f"synthetic pair: n={N_VERTICES}, male edges {male.m}, female edges {female.m} ({time.time() - t:.1f}s)")
    if args.pi0 == "greedy":
        t = time.time()
        pi0 = bm.greedy_match(N_VERTICES, male, female)
        log(f"greedy initial matching ({time.time() - t:.1f}s)")
    else:
        pi0 = pi_shuffled

    ctx = bm.Context(N_VERTICES, male, female)
    ctx.set_mapping(pi0)
    score0 = ctx.score()
    log(f"device {ctx.device_info()}; initial score {score0}")

    K = args.candidates
    batch = ctx.batch(K)
    ctx.timing_enable(True)
    for i in range(args.warmup):
        batch.eval_swaps_uniform(1000 + i, K)
    ctx.sync()
    ctx.timing_reset()

    seeds = [2000 + i for i in range(args.steps)]
    ctx.sync()
    t0 = time.perf_counter()
    for s in seeds:
        batch.eval_swaps_uniform(s, K)
    ctx.sync()
    t1 = time.perf_counter()
    launches, total_ms = ctx.timing(bm.KERNEL_SWAP)
    wall = t1 - t0
    ms_per_step = 1000.0 * wall / args.steps
    value = K * args.steps / wall
    avg_launch_s = total_ms / 1000.0 / max(1, launches)
    log(f"swap: {args.steps} x {K} candidates in {wall:.3f}s -> {value:.4g} deltas/s "
        f"(kernel {1000 * avg_launch_s:.1f} ms/launch over {launches} launches)")

    t = time.time()
    algo = [ctx.uniform_algo_bytes(s, K) for s in seeds]
    bytes_per_launch = sum(b for b, _ in algo) / len(algo)
    mean_D = sum(d for _, d in algo) / len(algo) / K
    achieved = bytes_per_launch / avg_launch_s / 1e9
    log(f"algorithmic bytes/launch {bytes_per_launch:.4g} (mean D {mean_D:.1f}) -> {achieved:.1f} GB/s "
        f"({time.time() - t:.1f}s to sum)")
    stats = batch.stats()

    traffic = None
    prof = os.path.join(ROOT, "profiles", "swap_delta_n18524_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            p = json.load(f)
        if p.get("candidates") == K and p.get("n") == N_VERTICES and p.get("pi0") == args.pi0:
            traffic = p.get("hbm_bytes_per_launch")

    result = {
        "metric": "exact swap-deltas/s",
        "value": value,
        "unit": "swap-deltas/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {
            "workload": "swap-delta n=18524 ~4M edges/graph, 2^28 uniform pairs per step",
            "n": N_VERTICES,
            "male_edges": male.m,
            "female_edges": female.m,
            "candidates_per_step": K,
            "initial_matching": "greedy-match" if args.pi0 == "greedy" else "30% shuffled",
            "initial_score": score0,
            "weights": "lognormal(0.5,1.2) rounded >=1",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_swap<uniform>",
            "algo_bytes_per_launch": bytes_per_launch,
            "mean_D": mean_D,
            "launch_ms": 1000 * avg_launch_s,
        },
        "delta_stats_last_step": {k: stats[k] for k in ("sum", "max", "n_pos", "n_zero")},
    }

    if not args.no_rot3:
        T = args.triples
        b3 = ctx.batch(T)
        b3.gen_triples(3000, T)
        b3.eval_rot3()
        ctx.sync()
        ctx.timing_reset()
        tr0 = time.perf_counter()
        for i in range(args.steps):
            b3.gen_triples(4000 + i, T)
            b3.eval_rot3()
        ctx.sync()
        tr1 = time.perf_counter()
        n3, ms3 = ctx.timing(bm.KERNEL_ROT3)
        l3 = ms3 / 1000.0 / max(1, n3)
        ab3, d3 = b3.algo_bytes(generated=False)
        result["rot3"] = {
            "triples_per_step": T,
            "value": T / l3,
            "unit": "triples/s (5 permutations each)",
            "launch_ms": 1000 * l3,
            "wall_s_incl_generation": tr1 - tr0,
            "roofline": {"bound": "hbm", "achieved": ab3 / l3 / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ab3 / l3 / 1e9 / HBM_PEAK_GBS, "algo_bytes_per_launch": ab3, "mean_D3": d3 / T},
        }
        log(f"rot3: {T} triples in {1000 * l3:.1f} ms/launch -> {T / l3:.4g} triples/s")
        b3.close()

    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(args, male, female, pi0, batch, seeds[-1], K, N_VERTICES)

    ctx.timing_enable(False)
    batch.close()
    ctx.close()
    print(json.dumps(result), flush=True)


def cpu_baseline(args, male, female, pi0, batch, seed, K, n):
    """The reference CPU arithmetic on a bounded sample of the same stream."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    threads = max(1, min(16, os.cpu_count() or 1))
    uv_all = None
    if orc.ref_available() and n <= 18524:
        kind = "reference"
        runner = orc.RefCode(male, female)
        run = lambda pi, uv: runner.deltas(pi, uv, threads)  # noqa: E731
        what = "src/graph-alignment-optimized.c calculate_swap_delta compiled from the reference sources"
    else:
        kind = "port"
        runner = orc.ListsRef(n, male, female)
        run = lambda pi, uv: runner.deltas(pi, uv, threads)  # noqa: E731
        what = "oracle restatement of calculate_swap_delta (edge lists + dense int16 female matrix)"
    calib = 1 << 16
    uv_all = orc.uniform_pairs(seed, calib, n)
    t = time.perf_counter()
    run(pi0, uv_all)
    rate = calib / max(1e-6, time.perf_counter() - t)
    S = int(min(K, max(calib, rate * args.cpu_seconds)))
    S -= S % 1024
    uv = orc.uniform_pairs(seed, S, n)
    t = time.perf_counter()
    d_cpu = run(pi0, uv)
    el = time.perf_counter() - t
    d_gpu = batch.deltas(0, S)
    match = bool(np.array_equal(d_cpu.astype(np.int64), d_gpu))
    log(f"cpu baseline ({kind}, {threads} threads): {S} deltas in {el:.2f}s -> {S / el:.4g}/s; "
        f"device deltas identical: {match}")
    return {
        "value": S / el,
        "unit": "swap-deltas/s",
        "cores": threads,
        "kind": kind,
        "sample": f"first {S} pairs of the last timed step's stream; {what}; OpenMP schedule(dynamic,1024)",
        "matches_device": match,
    }


if __name__ == "__main__":
    main()

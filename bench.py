#!/usr/bin/env python3
"""Benchmark: exact swap-deltas per second on one MI355X.

Workload (the full-size single-GPU swap-delta configuration): a synthetic
male/female pair with n = 18,524 vertices and ~4M directed edges per graph
(power-law out-degrees, alpha 2.1; lognormal(0.5, 1.2) integer weights >= 1;
female = male relabelled with 20% of edges reweighted/rewired), initial
matching = the greedy matching of the pair.  One step evaluates the exact swap
delta of 2^28 independent uniform candidate pairs (u != v) against the
current matching and writes 2^28 int64 deltas to HBM.  The pairs come from a
counter-based stream generated on the device (a new seed per step); by default
the stream lists the pairs grouped by first vertex ("umajor", same law as
independent pairs, see bm_umajor_first in brain-match_amd/csrc/bm_rng.h), or
in draw order with --stream uniform.

Two exact evaluation strategies are timed on the same stream; both give
identical deltas:
  gain  -- each step rebuilds the (n+1)^2 int32 placement-gain table for the
           current matching (plus its transpose for the umajor stream), then
           evaluates every pair from table entries plus an exact correction
           for adjacent pairs (default headline);
  scan  -- each pair scans its four male lists against LDS-staged female rows
           (one wavefront per pair).
The 3-vertex permutation kernels are timed on 2^26 uniform triples per step
in the same run (the a-major stream: triples grouped by first vertex, same law
as independent triples, see bm_amajor_partners in bm_rng.h) and reported
under "rot3".

value = candidates / second over the K timed steps of the headline strategy
(graph, matching and output buffers resident in HBM; the table build is inside
every step).  roofline = the step's dominant kernel: algorithmic bytes per
launch / mean launch time from HIP events on the library stream; traffic = HBM
bytes per launch measured with rocprofv3 (profiles/swap_n18524_traffic.json).
lds_roofline = the row build against the LDS atomic-add rate (one add per
male edge x entry of a sparse female row; the hub rows kept dense are summed
by the tile pass, gain_table.dense_rows).
cpu_baseline = the reference's own calculate_swap_delta
(src/graph-alignment-optimized.c:321-408, compiled from its sources into
oracle/_ref) over an evenly strided sample of the last timed stream, OpenMP
dynamic schedule, on this host; its deltas are compared with the device's.
proposals = the second half of the metric: from the pair's greedy matching,
disjoint-swap climbs with uniform, 2-hop and learned proposals run for the
same wall clock (--proposal-seconds, 60 s each; 0 leaves it out) and their
exact scores are compared.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "brain-match_amd"))

HBM_PEAK_GBS = 8000.0
# LDS: 32 banks of 4 bytes per CU (128 bytes per clock), so at most 32 four-byte
# atomic adds per CU per clock; 256 CUs at the 2.4 GHz peak engine clock
LDS_ADD_PEAK_PER_S = 32 * 256 * 2.4e9
N_VERTICES = 18524
N_EDGES = 4_000_000


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def packed_rows(n, female):
    """The table build streams female rows as 4-byte (f << 15) | w words when
    ids are below 2^17 and weights in 0..32767, else as 8-byte {f, w}."""
    return n < (1 << 17) and int(female.w.min()) >= 0 and int(female.w.max()) <= 32767


def gain_stride(n):
    """Row stride of the gain table: n + 1 rounded up to 32 int32."""
    return (n + 1 + 31) // 32 * 32


def gain_build_bytes(ctx, pi, entry_bytes):
    """Compulsory HBM bytes of one table build (work list + rows): for every
    male edge seen from each endpoint (2m visits) the edge (8 B), pi of the
    other endpoint (4 B) and the bounds of its image's female row (16 B); every
    female row entry once, both directions (entry_bytes each); the table write
    ((n+1) rows x stride x 4 B).  The build expands a female row once per male
    edge that lands on it ("expanded_entries"); those re-reads are served by
    L2 / MALL or HBM and appear in the measured traffic, not here."""
    import numpy as np
    mo, mi = ctx.degrees(female=False)
    fo, fi = ctx.degrees(female=True)
    pi = np.asarray(pi, dtype=np.int64)
    # out-edge u->x expands over the female in-row of pi(x): counted once per in-edge of x
    entries = int((mi.astype(np.int64) * fi[pi]).sum() + (mo.astype(np.int64) * fo[pi]).sum())
    m = int(mo.sum())
    mf = int(fo.sum())
    return 2 * m * 28 + 2 * mf * entry_bytes + (ctx.n + 1) * gain_stride(ctx.n) * 4, entries


def dense_split(ctx, pi, female, n_dense):
    """The table build's split of the expansions between the dense tile pass
    and the LDS row build.  Dense rows follow the library's rule: female rows
    of at least a quarter of the row stride, longest first (ties: lower id,
    out-rows before in-rows), each taken only while every column's sum of
    dense-row weights stays <= 65535, at most 128; their count is checked
    against the library's.  Returns (expanded entries of dense rows,
    (male row, dense row) pairs)."""
    import numpy as np
    if n_dense == 0:
        return 0, 0
    n = ctx.n
    mo, mi = ctx.degrees(female=False)
    fo, fi = ctx.degrees(female=True)
    thr = max(1, gain_stride(n) // 4)
    cand = [(-int(fo[r]), r) for r in range(1, n + 1) if fo[r] >= thr]
    cand += [(-int(fi[r]), n + 1 + r) for r in range(1, n + 1) if fi[r] >= thr]
    cand.sort()
    # female edges, the last line of each (src, dst) kept (the library's rule)
    key = female.src.astype(np.int64)[::-1] * (n + 1) + female.dst.astype(np.int64)[::-1]
    key, first = np.unique(key, return_index=True)
    w = female.w[::-1][first].astype(np.int64)
    src, dst = key // (n + 1), key % (n + 1)
    by_src, by_dst = np.argsort(src, kind="stable"), np.argsort(dst, kind="stable")
    src_ptr = np.searchsorted(src[by_src], np.arange(n + 2))
    dst_ptr = np.searchsorted(dst[by_dst], np.arange(n + 2))
    colsum = np.zeros(n + 1, np.int64)
    kept = []
    for c in cand:
        if len(kept) == 128:
            break
        r = c[1] - (n + 1) if c[1] > n else c[1]
        e = by_dst[dst_ptr[r]:dst_ptr[r + 1]] if c[1] > n else by_src[src_ptr[r]:src_ptr[r + 1]]
        cols = src[e] if c[1] > n else dst[e]
        if (colsum[cols] + w[e]).max(initial=0) <= 65535:
            colsum[cols] += w[e]
            kept.append(c)
    cand = kept
    if len(cand) != n_dense:
        raise SystemExit(f"dense-row rule disagrees with the library ({len(cand)} against {n_dense})")
    dout = np.zeros(n + 1, bool)
    din = np.zeros(n + 1, bool)
    for _, key in cand:
        (din if key > n else dout)[key - (n + 1) if key > n else key] = True
    pi = np.asarray(pi, dtype=np.int64)
    # an out-edge u -> x expands the in-row of pi(x); an in-edge x -> u the out-row of pi(x)
    uses_in, uses_out = din[pi] & (pi > 0), dout[pi] & (pi > 0)
    entries = int((mi.astype(np.int64) * fi[pi] * uses_in).sum() + (mo.astype(np.int64) * fo[pi] * uses_out).sum())
    pairs = int((mi.astype(np.int64) * uses_in).sum() + (mo.astype(np.int64) * uses_out).sum())
    return entries, pairs


def gain_lookup_bytes(ctx, k):
    """Compulsory HBM bytes of one u-major lookup launch: per first vertex its
    two staged rows (G(u, .) and GT(pi(u), .), 2 x stride x 4 B); the packed
    per-vertex words once ((n+1) x 16 B); the male lists and female rows read
    for the adjacency bitmaps (each once: 2m x 8 B + 2 m_F x 8 B); the stream
    offsets ((n+2) x 8 B); the int64 deltas (8 k)."""
    mo, _ = ctx.degrees(female=False)
    fo, _ = ctx.degrees(female=True)
    n = ctx.n
    return n * 2 * gain_stride(n) * 4 + (n + 1) * 16 + 16 * int(mo.sum()) + 16 * int(fo.sum()) + (n + 2) * 8 + 8 * k


def measured_traffic(args):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by what they read for a 1 GiB copy) for this workload,
    by kernel (profiles/swap_n18524_traffic.md); {} for another workload."""
    prof = os.path.join(ROOT, "profiles", "swap_n18524_traffic.json")
    if not os.path.exists(prof):
        return {}
    with open(prof) as f:
        p = json.load(f)
    if (p.get("candidates") == args.candidates and p.get("triples") == args.triples and p.get("pi0") == args.pi0
            and p.get("stream") == args.stream):
        return p.get("hbm_bytes_per_launch", {})
    return {}


def timed(ctx, kernel_ids, fn, steps):
    ctx.sync()
    ctx.timing_reset()
    t0 = time.perf_counter()
    for i in range(steps):
        fn(i)
    ctx.sync()
    wall = time.perf_counter() - t0
    kt = {kid: ctx.timing(kid) for kid in kernel_ids}
    return wall, kt


def per_launch_ms(kt, kid, steps=None):
    n, ms = kt[kid]
    if steps:
        return ms / steps
    return ms / n if n else 0.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--candidates", type=int, default=1 << 28)
    ap.add_argument("--triples", type=int, default=1 << 26)
    ap.add_argument("--stream", choices=["umajor", "uniform"], default="umajor")
    ap.add_argument("--scan-steps", type=int, default=2, help="timed steps of the scan strategy")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="target CPU-baseline work (the short calibration run overestimates the time per delta, "
                         "so the timed sample takes about 12 s)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-rot3", action="store_true")
    ap.add_argument("--no-scan", action="store_true", help="time only the headline strategy")
    ap.add_argument("--headline", choices=["gain", "scan"], default="gain")
    ap.add_argument("--pi0", choices=["greedy", "shuffle"], default="greedy")
    ap.add_argument("--proposals", type=float, default=0.0, metavar="SECONDS",
                    help="instead of the swap-delta benchmark: train the learned proposal, then compare the exact "
                         "scores reached in SECONDS by uniform, 2-hop and learned proposals (one JSON line)")
    ap.add_argument("--train-samples", type=int, default=50_000_000)
    ap.add_argument("--proposal-seconds", type=float, default=60.0,
                    help="wall clock of each proposal climb in the default run (0: leave the comparison out)")
    ap.add_argument("--sources", default=",".join(PROPOSAL_SOURCES),
                    help="proposal sources compared (comma-separated subset of uniform,twohop,learned)")
    ap.add_argument("--learned-plain", action="store_true",
                    help="learned climb without online retraining and uniform fallback")
    ap.add_argument("--retrain-lr", type=float, default=2e-4, help="Adam step size of the learned climb's retraining")
    ap.add_argument("--round-costs", type=int, default=0, metavar="ROUNDS",
                    help="instead of the swap-delta benchmark: ROUNDS climb rounds of each proposal source on the "
                         "n=18,524 pair (uniform, learned, learned with retraining) and their wall time per round "
                         "(one JSON line; run under rocprofv3 for the kernel breakdown)")
    ap.add_argument("--evaluator", choices=["gain", "scan"], default="scan",
                    help="exact delta strategy of the --proposals / --scale climbs (same deltas either way)")
    ap.add_argument("--anneal", type=int, default=0, metavar="LAUNCHES",
                    help="instead of the swap-delta benchmark: the batched-annealing workload (64 instances, "
                         "n in 1k..20k, ~200 n edges), LAUNCHES replays of 1000 steps per chain (one JSON line)")
    ap.add_argument("--chains-per-instance", type=int, default=48)
    ap.add_argument("--scale", type=float, default=0.0, metavar="SECONDS",
                    help="instead of the swap-delta benchmark: the n=100,000 ~20M-edge pair with 40%% of the matching "
                         "shuffled: delta throughput of both strategies, then learned vs uniform climbs and "
                         "swap + 3-vertex annealing for SECONDS each (one JSON line)")
    ap.add_argument("--p-pool", type=float, default=0.5,
                    help="--scale: fraction of the learned annealing chains' moves drawn from the ranked pool")
    ap.add_argument("--scale-climbs", action="store_true",
                    help="--scale: also run uniform and learned disjoint-swap climbs for SECONDS each")
    ap.add_argument("--best-swap", type=int, default=0, metavar="SWAPS",
                    help="instead of the swap-delta benchmark: steepest ascent on the n=18,524 pair from the greedy "
                         "matching, SWAPS best swaps over all 171.6M pairs each (one JSON line)")
    ap.add_argument("--sweep", type=float, default=0.0, metavar="SECONDS",
                    help="instead of the swap-delta benchmark: the per-vertex best-partner sweep with restarts "
                         "(the OpenMP tool's loop) on the n=18,524 pair from the greedy matching for SECONDS "
                         "(one JSON line)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (the known-size kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the multi-GPU path is not built.", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.arange(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 3}), flush=True)
        return

    import brainmatch as bm

    if args.proposals > 0:
        return bench_proposals(args, bm)
    if args.anneal > 0:
        return bench_anneal(args, bm)
    if args.scale > 0:
        return bench_scale(args, bm)
    if args.best_swap > 0:
        return bench_best_swap(args, bm)
    if args.sweep > 0:
        return bench_sweep(args, bm)
    if args.round_costs > 0:
        return bench_round_costs(args, bm)

    t = time.time()
    male, female, truth, pi_shuffled = bm.synth_pair(N_VERTICES, N_EDGES / N_VERTICES, bm.WEIGHTS_LOGNORMAL, 0.3,
                                                     seed=1)
    log(f"synthetic pair: n={N_VERTICES}, male edges {male.m}, female edges {female.m} ({time.time() - t:.1f}s)")
    greedy = None
    if args.pi0 == "greedy":
        t = time.time()
        pi0 = bm.greedy_match(N_VERTICES, male, female)
        host_s = time.time() - t
        log(f"greedy initial matching ({host_s:.1f}s)")
    else:
        pi0 = pi_shuffled

    ctx = bm.Context(N_VERTICES, male, female)
    if args.pi0 == "greedy":
        ctx.greedy_match()                     # warm-up (kernel loads)
        ctx.sync()
        t = time.perf_counter()
        pi_dev = ctx.greedy_match()
        dev_s = time.perf_counter() - t
        greedy = {"host_seconds": host_s, "device_seconds": dev_s, "identical": bool(np.array_equal(pi_dev, pi0))}
        log(f"greedy matching on the device: {1000 * dev_s:.1f} ms, identical to the host's: {greedy['identical']}")
        if not greedy["identical"]:
            raise SystemExit("device and host greedy matchings differ")
    ctx.set_mapping(pi0)
    score0 = ctx.score()
    log(f"device {ctx.device_info()}; initial score {score0}")
    K = args.candidates
    batch = ctx.batch(K)
    ctx.timing_enable(True)
    seeds = [2000 + i for i in range(args.steps)]
    umajor = args.stream == "umajor"
    evaluate = batch.eval_swaps_umajor if umajor else batch.eval_swaps_uniform
    stream_bytes = ctx.umajor_algo_bytes if umajor else ctx.uniform_algo_bytes
    entry_bytes = 4 if packed_rows(N_VERTICES, female) else 8

    # ---------------- gain-table strategy ----------------
    gain = None
    if args.headline == "gain" or not args.no_scan:
        ctx.gain_enable(True)
        for i in range(args.warmup):
            ctx.gain_rebuild()
            evaluate(1000 + i, K)

        def gain_step(i):
            ctx.gain_rebuild()
            evaluate(seeds[i], K)

        wall, kt = timed(ctx, (bm.KERNEL_GAIN_BUILD, bm.KERNEL_GAIN_SWAP, bm.KERNEL_AUX), gain_step, args.steps)
        build_s = per_launch_ms(kt, bm.KERNEL_GAIN_BUILD) / 1000
        look_s = per_launch_ms(kt, bm.KERNEL_GAIN_SWAP) / 1000
        aux_s = per_launch_ms(kt, bm.KERNEL_AUX, steps=args.steps) / 1000
        bb, entries = gain_build_bytes(ctx, pi0, entry_bytes)
        n_dense = ctx.gain_dense_rows()
        dense_entries, dense_pairs = dense_split(ctx, pi0, female, n_dense)
        adds = entries - dense_entries
        lb = gain_lookup_bytes(ctx, K)
        gain = {
            "value": K * args.steps / wall,
            "ms_per_step": 1000 * wall / args.steps,
            "build_ms": 1000 * build_s,
            "lookup_ms": 1000 * look_s,
            "aux_ms": 1000 * aux_s,
            "table_bytes": (N_VERTICES + 1) * gain_stride(N_VERTICES) * 4,
            "build_roofline": {"bound": "hbm", "achieved": bb / build_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": bb / build_s / 1e9 / HBM_PEAK_GBS, "algo_bytes_per_launch": bb,
                               "expanded_entries": entries, "female_entry_bytes": entry_bytes},
            "lookup_roofline": {"bound": "hbm", "achieved": lb / look_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": lb / look_s / 1e9 / HBM_PEAK_GBS, "algo_bytes_per_launch": lb},
            # the row build adds one LDS atomic per (male edge, female row entry)
            # expansion of a sparse female row; the dense rows' expansions are
            # summed by the tile pass in packed 16-bit vector arithmetic
            "build_lds_roofline": {"bound": "lds", "achieved": adds / build_s, "peak": LDS_ADD_PEAK_PER_S,
                                   "unit": "atomic adds/s", "frac": adds / build_s / LDS_ADD_PEAK_PER_S,
                                   "adds_per_launch": adds,
                                   "peak_basis": "32 four-byte LDS atomics per CU per clock x 256 CUs x 2.4 GHz"},
            "dense_rows": {"rows": n_dense, "expanded_entries": dense_entries, "row_pairs": dense_pairs,
                           "column_terms": dense_pairs * gain_stride(N_VERTICES),
                           "share_of_expansions": dense_entries / entries if entries else 0.0},
        }
        # the per-candidate byte model of a list scan (4 male lists, 28 B per
        # incident edge): what the same stream would cost a kernel that reads
        # every candidate's lists, and the rate such a kernel reaches at HBM peak
        pcb, pcd = stream_bytes(seeds[-1], K)
        gain["per_candidate_model"] = {
            "applies_to_this_strategy": False,
            "why": "the table strategy reads four table entries per pair, not the pair's male lists; the model "
                   "prices a list scan and applies to the scan kernel (scan_kernel.roofline)",
            "bytes_per_step": pcb, "mean_D": pcd / K,
            "deltas_per_s_at_hbm_peak": K * HBM_PEAK_GBS * 1e9 / pcb,
            "equivalent_GBps": pcb / (wall / args.steps) / 1e9,
        }
        log(f"gain ({args.stream}): {args.steps} x {K} pairs in {wall:.3f}s -> {gain['value']:.4g} deltas/s "
            f"(build {1000 * build_s:.2f} ms, lookups {1000 * look_s:.2f} ms, stream/transpose {1000 * aux_s:.2f} ms); "
            f"a list scan at HBM peak: {gain['per_candidate_model']['deltas_per_s_at_hbm_peak']:.3g} deltas/s")
        gain_deltas_head = batch.deltas(0, 1 << 16)
        if umajor:
            # the same table against the draw-order stream (per-pair table reads)
            for i in range(args.warmup):
                batch.eval_swaps_uniform(1000 + i, K)
            w2, kt2 = timed(ctx, (bm.KERNEL_GAIN_SWAP,), lambda i: batch.eval_swaps_uniform(seeds[i], K), args.steps)
            gain["uniform_stream_lookup_ms"] = per_launch_ms(kt2, bm.KERNEL_GAIN_SWAP)
            evaluate(seeds[-1], K)
        ctx.gain_enable(False)

    # ---------------- scan strategy ----------------
    scan = None
    if args.headline == "scan" or not args.no_scan:
        steps = args.steps if args.headline == "scan" else args.scan_steps
        sseeds = seeds[-steps:]
        for i in range(min(args.warmup, 1) if args.headline != "scan" else args.warmup):
            evaluate(1000 + i, K)
        wall, kt = timed(ctx, (bm.KERNEL_SWAP, bm.KERNEL_AUX), lambda i: evaluate(sseeds[i], K), steps)
        launch_s = per_launch_ms(kt, bm.KERNEL_SWAP) / 1000
        algo = [stream_bytes(s, K) for s in sseeds]
        bytes_per_launch = sum(b for b, _ in algo) / len(algo)
        mean_D = sum(d for _, d in algo) / len(algo) / K
        scan = {
            "value": K * steps / wall,
            "steps": steps,
            "ms_per_step": 1000 * wall / steps,
            "roofline": {"bound": "hbm", "achieved": bytes_per_launch / launch_s / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": bytes_per_launch / launch_s / 1e9 / HBM_PEAK_GBS,
                         "algo_bytes_per_launch": bytes_per_launch, "mean_D": mean_D, "launch_ms": 1000 * launch_s},
        }
        log(f"scan ({args.stream}): {steps} x {K} pairs in {wall:.3f}s -> {scan['value']:.4g} deltas/s "
            f"({scan['roofline']['achieved']:.0f} GB/s algorithmic)")
        if gain is not None:
            same = bool((batch.deltas(0, 1 << 16) == gain_deltas_head).all())
            scan["matches_gain_path"] = same
            log(f"scan and gain deltas identical on the last stream's first 65536 pairs: {same}")
            if not same:
                raise SystemExit("the two exact strategies disagree")

    head = gain if args.headline == "gain" else scan
    stats = batch.stats()

    traffic = measured_traffic(args)

    lds_rf = None
    if args.headline == "gain":
        dom = "lookup" if gain["lookup_ms"] >= gain["build_ms"] else "build"
        rf = dict(gain[dom + "_roofline"])
        if dom == "build":
            lds_rf = dict(gain["build_lds_roofline"], kernel="k_gain_build (sparse female rows)")
        rf["kernel"] = ("k_gain_swap_umajor" if umajor else "k_gain_swap<uniform>") if dom == "lookup" \
            else "k_gain_edge_desc+k_gain_dense_tile+k_gain_build"
    else:
        rf = dict(scan["roofline"])
        rf["kernel"] = "k_swap<umajor>" if umajor else "k_swap<uniform>"
    rf["traffic"] = traffic.get(rf["kernel"])

    result = {
        "metric": "exact swap-deltas/s",
        "value": head["value"],
        "unit": "swap-deltas/s",
        "n_gpus": 1,
        "steps": args.steps if args.headline == "gain" else scan["steps"],
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {
            "workload": "swap-delta n=18524 ~4M edges/graph, 2^28 uniform pairs per step",
            "strategy": args.headline,
            "stream": args.stream,
            "n": N_VERTICES,
            "male_edges": male.m,
            "female_edges": female.m,
            "candidates_per_step": K,
            "initial_matching": "greedy-match" if args.pi0 == "greedy" else "30% shuffled",
            "initial_score": score0,
            "weights": "lognormal(0.5,1.2) rounded >=1",
            "parallelism": "single GPU",
        },
        "roofline": rf,
        "lds_roofline": lds_rf,
        "gain_table": gain,
        "scan_kernel": scan,
        "delta_stats_last_step": {k: stats[k] for k in ("sum", "max", "n_pos", "n_zero")},
        "greedy_initial_matching": greedy,
    }

    if not args.no_rot3:
        result["rot3"] = bench_rot3(ctx, bm, args)

    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(args, male, female, pi0, batch, seeds[-1], K, N_VERTICES)

    ctx.timing_enable(False)
    batch.close()
    if args.proposal_seconds > 0:
        # the second half of the metric: exact score reached in the same wall
        # clock by each proposal source, from the same greedy matching
        sources = tuple(x for x in args.sources.split(",") if x)
        pr = compare_proposals(bm, ctx, pi0, args.proposal_seconds, args, sources, planted=truth)
        result["proposals"] = dict(metric=f"exact score at {args.proposal_seconds:g} s, learned vs uniform proposals",
                                   **pr)
    ctx.close()
    print(json.dumps(result), flush=True)


PROPOSAL_SOURCES = ("uniform", "twohop", "learned")


def compare_proposals(bm, ctx, pi0, seconds, args, sources=PROPOSAL_SOURCES, planted=None):
    """The learned-proposal workload on the n = 18,524 pair, from the greedy
    matching.  The network is trained on train_samples (candidate, exact
    delta) pairs from the context's own exact kernel; then disjoint-swap
    climbs from the same matching run for the same wall clock, each sending
    4096 candidates per round to the exact delta kernel: uniform pairs, 2-hop
    pairs (64 anchors x 64 targets), or the 4096 best-ranked of 2^20 uniform
    pairs.  The learned climb (unless --learned-plain) retrains the network
    on every learned round's 4096 evaluated candidates and their exact deltas
    and falls back to uniform rounds while those gain more per second.  The
    climbs' candidates are evaluated by the scan kernel (--evaluator scan) or
    the placement-gain table (--evaluator gain); the deltas are the same.
    With `planted` (the pair's hidden permutation) each climb also reports
    when it first reached the planted matching's score, whether it ended on
    that matching, and the exact best swap over all pairs at its end (<= 0:
    no single swap improves the matching)."""
    if args.evaluator == "gain":
        ctx.gain_enable(True)
    ctx.set_mapping(pi0)
    score0 = ctx.score()
    batch = 1 << 16
    steps = max(1, args.train_samples // batch)
    mlp = bm.Mlp(ctx, seed=1)
    mlp.refresh_features()
    ctx.sync()
    t = time.perf_counter()
    losses = mlp.train(seed=11, steps=steps, batch=batch, lr=1e-3)
    train_s = time.perf_counter() - t
    log(f"trained on {steps * batch} samples in {train_s:.1f}s; mse {losses[:10].mean():.3f} -> {losses[-10:].mean():.3f}")
    trained = mlp.params()
    plain = args.learned_plain
    learned_kw = {"pool": 1 << 20, "mlp": mlp, "refresh_every": 1}
    if not plain:
        learned_kw.update(retrain_every=1, retrain_lr=args.retrain_lr, fallback=True)
    kinds = {"uniform": (bm.PROPOSAL_UNIFORM, {}),
             "twohop": (bm.PROPOSAL_TWOHOP, {"per_anchor": 64}),
             "learned": (bm.PROPOSAL_LEARNED, learned_kw)}
    out = {}
    s_planted = None
    if planted is not None:
        ctx.set_mapping(planted)
        s_planted = ctx.score()
    for name in sources:
        prop, kw = kinds[name]
        if name == "learned":
            mlp.set_params(trained)
        ctx.set_mapping(pi0)
        tr, tt, res = ctx.climb(prop, seed=21, time_limit_s=seconds, per_round=4096, **kw)
        marks = {}
        for frac in (0.1, 0.25, 0.5, 0.75, 1.0):
            i = int(np.searchsorted(tt, frac * seconds))
            marks[f"{frac * seconds:g}s"] = int(tr[min(i, len(tr) - 1)]) if len(tr) else score0
        last = tt >= 0.9 * seconds
        out[name] = {"final_score": res["final_score"], "gain": res["final_score"] - score0,
                     "rounds": res["iterations"], "accepted": res["accepted"], "evaluated": res["evaluations"],
                     "verify_failures": res["verify_failures"], "seconds": res["seconds"], "score_at": marks,
                     "gain_last_10pct": int(tr[-1] - tr[last][0]) if last.any() else 0}
        if name == "learned":
            out[name]["uniform_fallback_rounds"] = res["fallback_rounds"]
        if s_planted is not None:
            hit = np.nonzero(tr >= s_planted)[0]
            out[name]["reached_planted_score_at_s"] = float(tt[hit[0]]) if len(hit) else None
            out[name]["ends_on_planted_matching"] = bool(np.array_equal(ctx.get_mapping(), planted))
            was_on = ctx.gain_enabled()
            ctx.gain_enable(True)
            d_best, u_best, v_best = ctx.best_swap()
            ctx.gain_enable(was_on)
            out[name]["best_swap_delta_at_end"] = int(d_best)
        log(f"{name}: score {score0} -> {res['final_score']} in {res['seconds']:.1f}s "
            f"({res['iterations']} rounds, {res['accepted']} swaps"
            + (f", {res['fallback_rounds']} uniform fallback rounds" if name == "learned" else "") + ")")
    mlp.close()
    ctx.set_mapping(pi0)
    if args.evaluator == "gain":
        ctx.gain_enable(False)
    cfg = {"seconds": seconds, "evaluator": args.evaluator, "initial_matching": "greedy-match",
           "initial_score": score0, "planted_matching_score": s_planted, "per_round": 4096, "learned_pool": 1 << 20, "learned_top": 4096,
           "learned_online_retraining": not plain, "learned_uniform_fallback": not plain,
           "retrain_lr": None if plain else args.retrain_lr,
           "train_samples": steps * batch, "train_seconds": train_s,
           "train_mse_first": float(losses[:10].mean()), "train_mse_last": float(losses[-10:].mean())}
    res = {"config": cfg, "proposals": out}
    if "learned" in out and "uniform" in out:
        res["learned_minus_uniform"] = out["learned"]["final_score"] - out["uniform"]["final_score"]
    return res


def bench_proposals(args, bm):
    """--proposals SECONDS: only the proposal comparison (one JSON line)."""
    male, female, truth, _ = bm.synth_pair(N_VERTICES, N_EDGES / N_VERTICES, bm.WEIGHTS_LOGNORMAL, 0.3, seed=1)
    pi0 = bm.greedy_match(N_VERTICES, male, female)
    ctx = bm.Context(N_VERTICES, male, female)
    sources = tuple(x for x in args.sources.split(",") if x)
    r = compare_proposals(bm, ctx, pi0, args.proposals, args, sources, planted=truth)
    out = r["proposals"]
    result = {
        "metric": f"exact score at {args.proposals:g} s, learned vs uniform proposals",
        "value": out["learned"]["final_score"] if "learned" in out else None,
        "unit": "score",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "int64 (exact) / bf16 MLP ranking",
        "data": "synthetic",
        "config": dict(workload="learned proposal on the n=18524 ~4M-edge pair", n=N_VERTICES, **r["config"]),
        "proposals": out,
    }
    if "learned_minus_uniform" in r:
        result["learned_minus_uniform"] = r["learned_minus_uniform"]
    ctx.close()
    print(json.dumps(result), flush=True)


def bench_round_costs(args, bm):
    """Wall time of one disjoint-swap climb round per proposal source on the
    n = 18,524 pair from the greedy matching (4096 exact candidates a round;
    learned: the 4096 best of 2^20 ranked by the network, features refreshed
    every round), after a short training run.  Each source climbs the same
    number of rounds from the same matching."""
    male, female, truth, _ = bm.synth_pair(N_VERTICES, N_EDGES / N_VERTICES, bm.WEIGHTS_LOGNORMAL, 0.3, seed=1)
    pi0 = bm.greedy_match(N_VERTICES, male, female)
    ctx = bm.Context(N_VERTICES, male, female)
    ctx.set_mapping(pi0)
    mlp = bm.Mlp(ctx, seed=1)
    mlp.refresh_features()
    mlp.train(seed=11, steps=64, batch=1 << 16, lr=1e-3)
    trained = mlp.params()
    R = args.round_costs
    out = {}
    for name, prop, kw in (("uniform", bm.PROPOSAL_UNIFORM, {}),
                           ("learned", bm.PROPOSAL_LEARNED, {"pool": 1 << 20, "mlp": mlp, "refresh_every": 1}),
                           ("learned_retrain", bm.PROPOSAL_LEARNED,
                            {"pool": 1 << 20, "mlp": mlp, "refresh_every": 1, "retrain_every": 1,
                             "retrain_lr": args.retrain_lr})):
        mlp.set_params(trained)
        ctx.set_mapping(pi0)
        ctx.sync()
        t = time.perf_counter()
        tr, tt, res = ctx.climb(prop, seed=21, max_rounds=R, per_round=4096, **kw)
        el = time.perf_counter() - t
        out[name] = {"rounds": res["iterations"], "ms_per_round": 1000 * el / max(1, res["iterations"]),
                     "final_score": res["final_score"], "accepted": res["accepted"]}
        log(f"{name}: {R} rounds in {el:.2f}s ({1000 * el / R:.3f} ms/round), score {tr[0]} -> {tr[-1]}")
    mlp.close()
    ctx.close()
    print(json.dumps({"metric": "climb round wall time", "unit": "ms/round", "n_gpus": 1, "data": "synthetic",
                      "config": {"workload": "n=18524 ~4M edges, greedy start, 4096 exact candidates per round",
                                 "learned_pool": 1 << 20, "rounds": R},
                      "rounds": out}), flush=True)


def bench_anneal(args, bm):
    """The batched-annealing workload: 64 independent synthetic problems of
    uneven size (n uniform in 1k..20k, ~200 n edges, lognormal weights, 30% of
    the true matching shuffled), chains_per_instance annealing chains each (T_s = 50 * 0.9999^s,
    swaps and 3-vertex moves), all advanced by one kernel launch replayed
    from a HIP graph; the launch shape comes from the autotuner.  Reports
    chain steps/s (one exact delta per step), the algorithmic GB/s of the
    steps, the device recount check and a CPU replay of one chain."""
    from concurrent.futures import ThreadPoolExecutor

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    rng = np.random.default_rng(64)
    sizes = [int(x) for x in rng.integers(1000, 20001, 64)]
    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=16) as ex:
        pairs = list(ex.map(lambda a: bm.synth_pair(a[1], 200.0, bm.WEIGHTS_LOGNORMAL, 0.3, seed=1000 + a[0]),
                            enumerate(sizes)))
    ctxs = []
    for (male, female, truth, pi0), n in zip(pairs, sizes):
        c = bm.Context(n, male, female)
        c.set_mapping(pi0)
        ctxs.append(c)
    edges = sum(p[0].m for p in pairs)
    log(f"64 instances, n {min(sizes)}..{max(sizes)}, {edges} male edges in all ({time.perf_counter() - t:.1f}s)")
    steps = 1000
    ann = bm.Anneal(ctxs, chains_per_instance=args.chains_per_instance, seed=5, t0=50.0, decay=0.9999, p_rot3=0.1,
                    traj_cap=args.anneal)
    best, trials = ann.autotune(steps=200)
    log(f"autotuner: {len(trials)} shapes, best {best} ({trials[:, 4].min():.2f} ms for 200 steps; "
        f"slowest {trials[:, 4].max():.2f} ms)")
    init = ann.state()["score"].copy()
    ann.run(steps, 1)                      # capture + first launch (warm-up, part of the chains' history)
    n0, ms0 = ann.timing()
    ann.run(steps, args.anneal - 1) if args.anneal > 1 else None
    n1, ms1 = ann.timing()
    st = ann.state()
    timed_launches = n1 - n0
    timed_ms = ms1 - ms0
    chains = ann.chains
    rate = chains * steps * timed_launches / (timed_ms / 1000) if timed_launches else 0.0
    algo, total_steps = ann.algo_bytes()
    bad = ann.verify()
    # CPU replay of chain 0 of the smallest instance, all steps
    i_small = int(np.argmin(sizes))
    c = i_small * args.chains_per_instance
    male, female, truth, pi0 = pairs[i_small]
    o = orc.Oracle(sizes[i_small], male, female)
    beta0, lam, p3 = ann.schedule()
    pi, _, rs = orc.anneal_chain(o, pi0, o.score(pi0), orc.chain_seed(5, c), 0, steps * args.anneal, beta0, lam, p3)
    replay_ok = bool(rs[0] == st["score"][c] and rs[1] == st["best"][c] and np.array_equal(pi, ann.mapping(c)))
    gain = (st["best"] - init).reshape(64, -1).max(axis=1)
    log(f"{chains} chains x {steps * args.anneal} steps: {rate:.4g} steps/s over {timed_launches} timed launches; "
        f"recount mismatches {bad}; CPU replay of chain {c} (n={sizes[i_small]}) identical: {replay_ok}")
    result = {
        "metric": "annealing chain steps/s (exact deltas), 64 batched instances",
        "value": rate,
        "unit": "steps/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "int64 deltas, fp64 acceptance",
        "data": "synthetic",
        "config": {"workload": "64 instances, n uniform 1k..20k, ~200n edges, T_s = 50 * 0.9999^s",
                   "sizes": sizes, "male_edges": edges, "chains_per_instance": args.chains_per_instance,
                   "chains": chains, "steps_per_launch": steps, "launches": args.anneal, "p_rot3": 0.1},
        "autotuner": {"best": best, "trials": trials.tolist()},
        "ms_per_launch": timed_ms / max(1, timed_launches),
        "algorithmic_GBps": algo / max(1, total_steps) * rate / 1e9,
        "recount_mismatches": int(bad),
        "cpu_replay_identical": replay_ok,
        "best_gain_per_instance": {"min": int(gain.min()), "median": int(np.median(gain)), "max": int(gain.max())},
    }
    ann.close()
    for c in ctxs:
        c.close()
    print(json.dumps(result), flush=True)


def anneal_for(bm, ctx, pi0, seconds, chains, p_pool, mlp=None, refresh_s=2.0, steps_per_launch=500, seed=9):
    """Annealing chains (swaps and 25% 3-vertex moves, T_s = 50 * 0.9999^s)
    on one instance for `seconds` of wall clock from pi0.  With mlp, a
    fraction p_pool of the moves is drawn from a proposal pool: the 2^16
    best-ranked of 2^20 uniform pairs under the network, with the features
    of the best chain's current matching, re-ranked every refresh_s seconds
    (the ranking time is inside the wall clock).  Returns the trajectory of
    the best score over chains (seconds, score) and the step counts."""
    ctx.set_mapping(pi0)
    ann = bm.Anneal([ctx], chains_per_instance=chains, seed=seed, t0=50.0, decay=0.9999, p_rot3=0.25,
                    p_pool=p_pool if mlp is not None else 0.0)
    score0 = ctx.score()
    traj = [(0.0, score0)]
    refreshes = 0
    rank_s = 0.0
    t0 = time.perf_counter()
    last_refresh = -1e9
    launches = 0
    while True:
        el = time.perf_counter() - t0
        if el >= seconds:
            break
        if mlp is not None and el - last_refresh >= refresh_s:
            tr0 = time.perf_counter()
            if launches:
                st = ann.state()
                ctx.set_mapping(ann.mapping(int(np.argmax(st["score"]))))
            mlp.refresh_features()
            ann.pool_learned(0, mlp, seed=1000 + refreshes, pool=1 << 20, top=1 << 16)
            rank_s += time.perf_counter() - tr0
            refreshes += 1
            last_refresh = time.perf_counter() - t0
        ann.run(steps_per_launch, 1)
        launches += 1
        if launches % 20 == 0:
            traj.append((time.perf_counter() - t0, int(ann.state()["best"].max())))
            if len(traj) % 25 == 0:
                log(f"  anneal ({'learned pool' if mlp is not None else 'uniform'}): {traj[-1][0]:.0f}s, "
                    f"best {traj[-1][1]}")
    el = time.perf_counter() - t0
    st = ann.state()
    traj.append((el, int(st["best"].max())))
    kernel_launches, kernel_ms = ann.timing()
    algo, steps = ann.algo_bytes()
    rot, pooled = ann.move_counts()
    mism = int(ann.verify())
    out = {"chains": ann.chains, "seconds": el, "steps_per_chain": int(st["steps"][0]),
           "steps_per_s": int(steps) / el, "kernel_steps_per_s": int(steps) / (kernel_ms / 1000),
           "exact_deltas_per_s": (int(steps) + 4 * int(rot.sum())) / el,
           "algorithmic_GBps_kernel": algo / (kernel_ms / 1000) / 1e9, "kernel_ms": kernel_ms,
           "best_score": int(st["best"].max()), "median_best": int(np.median(st["best"])),
           "three_vertex_fraction": float(rot.sum() / max(1, steps)),
           "pool_fraction": float(pooled.sum() / max(1, steps)), "recount_mismatches": mism,
           "pool_refreshes": refreshes, "pool_ranking_seconds": rank_s,
           "trajectory": [(round(a, 2), b) for a, b in traj[::max(1, len(traj) // 40)]] + [(round(el, 2), traj[-1][1])]}
    ann.close()
    ctx.set_mapping(pi0)
    return out


def bench_scale(args, bm):
    """The scale workload: n = 100,000, ~20M directed edges, lognormal weights,
    40% of the true matching shuffled.  Exact swap-delta throughput of both
    strategies on 2^28 u-major pairs per step, the step of the n = 18,524
    benchmark (the gain table is 40 GB here, rebuilt every step in column
    windows; the lookups stage each first vertex's bitmaps and gather the
    table entries); then swap + 3-vertex annealing (2048
    chains) for SECONDS each with learned proposal pools and with uniform
    proposals only; with --scale-climbs also a uniform and a learned
    disjoint-swap climb for SECONDS each."""
    n = 100_000
    t = time.perf_counter()
    male, female, truth, pi0 = bm.synth_pair(n, 200.0, bm.WEIGHTS_LOGNORMAL, 0.4, seed=5)
    ctx = bm.Context(n, male, female)
    ctx.set_mapping(pi0)
    score0 = ctx.score()
    ctx.set_mapping(truth)
    planted = ctx.score()
    ctx.set_mapping(pi0)
    log(f"scale pair: n={n}, {male.m} male edges, initial score {score0}, planted matching {planted} "
        f"({time.perf_counter() - t:.1f}s)")
    K = 1 << 28
    b = ctx.batch(K)
    ctx.timing_enable(True)
    out = {"initial_score": score0, "planted_score": planted, "male_edges": male.m}
    b.eval_swaps_umajor(1, 1 << 20)
    wall, kt = timed(ctx, (bm.KERNEL_SWAP,), lambda i: b.eval_swaps_umajor(77, K), 1)
    algo, _ = ctx.umajor_algo_bytes(77, K)
    ms = per_launch_ms(kt, bm.KERNEL_SWAP)
    out["scan"] = {"deltas_per_s": K / wall, "launch_ms": ms, "algorithmic_GBps": algo / ms / 1e6}
    head = b.deltas(0, 1 << 16)
    ctx.gain_enable(True)
    ctx.gain_rebuild()
    b.eval_swaps_umajor(77, K)
    wall, kt = timed(ctx, (bm.KERNEL_GAIN_BUILD, bm.KERNEL_GAIN_SWAP, bm.KERNEL_AUX),
                     lambda i: (ctx.gain_rebuild(), b.eval_swaps_umajor(77, K)), 2)
    same = bool((b.deltas(0, 1 << 16) == head).all())
    mo, mi = ctx.degrees(female=False)
    fo, fi = ctx.degrees(female=True)
    pa = np.asarray(pi0, np.int64)
    expanded = int((mi.astype(np.int64) * fi[pa]).sum() + (mo.astype(np.int64) * fo[pa]).sum())
    build_ms = per_launch_ms(kt, bm.KERNEL_GAIN_BUILD)
    n_dense = ctx.gain_dense_rows()
    dense_entries, dense_pairs = dense_split(ctx, pi0, female, n_dense)
    out["gain"] = {"deltas_per_s": K * 2 / wall, "ms_per_step": 1000 * wall / 2,
                   "build_ms": build_ms, "expanded_entries": expanded,
                   "build_expansions_per_s": expanded / (build_ms / 1000),
                   "dense_rows": n_dense, "dense_expanded_entries": dense_entries,
                   "dense_column_terms": dense_pairs * gain_stride(n),
                   "lds_adds": expanded - dense_entries,
                   "lookup_ms": per_launch_ms(kt, bm.KERNEL_GAIN_SWAP), "matches_scan": same,
                   "table_GB": (n + 1) ** 2 * 4 / 1e9}
    log(f"scale deltas: scan {out['scan']['deltas_per_s']:.4g}/s, gain {out['gain']['deltas_per_s']:.4g}/s "
        f"(build {build_ms:.1f} ms over {expanded:.3g} expanded entries, {n_dense} dense rows holding "
        f"{dense_entries:.3g}, lookups {out['gain']['lookup_ms']:.1f} ms); "
        f"identical: {same}")
    b.close()
    ctx.timing_enable(False)
    ctx.gain_enable(False)
    mlp = bm.Mlp(ctx, seed=2)
    mlp.refresh_features()
    t = time.perf_counter()
    losses = mlp.train(seed=12, steps=300, batch=1 << 16, lr=1e-3)
    out["train"] = {"samples": 300 << 16, "seconds": time.perf_counter() - t,
                    "mse_first": float(losses[:10].mean()), "mse_last": float(losses[-10:].mean())}
    for name, p_pool in (("anneal_learned", args.p_pool), ("anneal_uniform", 0.0)):
        r = anneal_for(bm, ctx, pi0, args.scale, 2048, p_pool, mlp if p_pool > 0 else None)
        out[name] = r
        log(f"{name}: best {r['best_score']} after {r['seconds']:.0f}s, {r['steps_per_s']:.4g} steps/s, "
            f"{r['algorithmic_GBps_kernel']:.0f} GB/s algorithmic, recount mismatches {r['recount_mismatches']}")
    if args.scale_climbs:
        ctx.gain_enable(args.evaluator == "gain")   # the climbs' exact deltas: table or scan
        for name, prop, kw in (("uniform", bm.PROPOSAL_UNIFORM, {}),
                               ("learned", bm.PROPOSAL_LEARNED,
                                {"pool": 1 << 20, "mlp": mlp, "refresh_every": 1, "retrain_every": 1,
                                 "retrain_lr": args.retrain_lr, "fallback": True})):
            ctx.set_mapping(pi0)
            tr, tt, res = ctx.climb(prop, seed=31, time_limit_s=args.scale, per_round=4096, **kw)
            marks = {}
            for frac in (0.1, 0.25, 0.5, 1.0):
                i = int(np.searchsorted(tt, frac * args.scale))
                marks[f"{frac * args.scale:g}s"] = int(tr[min(i, len(tr) - 1)]) if len(tr) else score0
            out["climb_" + name] = {"final_score": res["final_score"], "rounds": res["iterations"],
                                    "accepted": res["accepted"], "verify_failures": res["verify_failures"],
                                    "score_at": marks}
            log(f"climb {name}: {score0} -> {res['final_score']} in {res['seconds']:.0f}s")
        ctx.gain_enable(False)
    mlp.close()
    ctx.close()
    result = {"metric": f"exact score at {args.scale:g} s on the n=100,000 pair, swap + 3-vertex annealing with "
                        "learned proposals vs uniform proposals",
              "value": out["anneal_learned"]["best_score"], "unit": "score", "n_gpus": 1, "higher_is_better": True,
              "data": "synthetic",
              "config": {"workload": "n=100000, ~20M edges, 40% shuffled", "seconds": args.scale,
                         "chains": 2048, "p_rot3": 0.25, "p_pool": args.p_pool, "t0": 50.0, "decay": 0.9999,
                         "evaluator": args.evaluator},
              "learned_minus_uniform": out["anneal_learned"]["best_score"] - out["anneal_uniform"]["best_score"],
              "results": out}
    print(json.dumps(result), flush=True)


def bench_sweep(args, bm):
    """The OpenMP tool's optimisation loop (src/re-alignment-parallel.c:556-676)
    on the device: passes over the vertices, each swapped with its exact best
    partner over all later vertices when that delta is positive, restarts
    from the best matching perturbed by 200 swaps after a pass without a
    swap, every running score checked by a device recount.  Reports the best
    score, passes, swaps, restarts, pairs evaluated per second, and the
    best-partner kernels' own rate (HIP events): a first pass at the greedy
    matching, then the rest of SECONDS."""
    n = N_VERTICES
    male, female, truth, _ = bm.synth_pair(n, N_EDGES / n, bm.WEIGHTS_LOGNORMAL, 0.3, seed=1)
    pi0 = bm.greedy_match(n, male, female)
    ctx = bm.Context(n, male, female)
    ctx.set_mapping(pi0)
    s0 = ctx.score()
    ctx.gain_enable(True)
    ctx.best_partners(1, 512)                         # transposed table + warm-up
    ctx.timing_enable(True)
    ctx.timing_reset()
    t = time.perf_counter()
    d, v = ctx.best_partners(1, n)
    one_s = time.perf_counter() - t
    calls, ms = ctx.timing(bm.KERNEL_GAIN_SWAP)
    pairs_all = n * (n - 1) // 2
    log(f"best partners of all {n} vertices ({pairs_all} pairs): {ms:.2f} ms in the kernels, "
        f"{one_s * 1e3:.2f} ms per call; {int((d > 0).sum())} vertices have an improving partner")
    ctx.timing_reset()
    best, trace, r = ctx.sweep(seed=7, time_limit_s=args.sweep, initial_swaps=50, restart_swaps=200, verify=True)
    _, kms = ctx.timing(bm.KERNEL_GAIN_SWAP)
    _, aux = ctx.timing(bm.KERNEL_AUX)
    _, upd = ctx.timing(bm.KERNEL_GAIN_UPDATE)
    log(f"sweep {args.sweep:g}s: {r['initial_score']} -> best {r['best_score']}, {r['iterations']} passes, "
        f"{r['accepted']} swaps, {r['restarts']} restarts, {r['evaluations'] / r['seconds']:.4g} pairs/s, "
        f"recount failures {r['verify_failures']}")
    result = {
        "metric": f"exact score after {args.sweep:g} s of the per-vertex best-partner sweep",
        "value": r["best_score"],
        "unit": "score",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "best-partner sweep with restarts, n=18524 ~4M edges/graph, greedy initial matching",
                   "seconds": args.sweep, "initial_swaps": 50, "restart_swaps": 200, "batch": 512},
        "greedy_score": s0,
        "best_partners_all_vertices": {"pairs": pairs_all, "kernel_ms": ms, "call_ms": one_s * 1e3,
                                       "pair_deltas_per_s": pairs_all / (ms / 1e3)},
        "sweep": {"initial_score": r["initial_score"], "best_score": r["best_score"], "passes": r["iterations"],
                  "swaps": r["accepted"], "restarts": r["restarts"], "pairs_evaluated": r["evaluations"],
                  "pairs_per_s": r["evaluations"] / r["seconds"], "seconds": r["seconds"],
                  "best_partner_kernel_ms": kms, "transpose_ms": aux, "patch_ms": upd,
                  "recount_failures": r["verify_failures"]},
    }
    ctx.close()
    print(json.dumps(result), flush=True)


def bench_best_swap(args, bm):
    """Steepest ascent (the score-lookup driver's loop): every iteration
    finds the exact best swap over all n(n-1)/2 pairs from the placement-gain
    table and applies it (table patch + transposed-table refresh).  Reports
    exact pair deltas/s of the all-pairs kernel (HIP events on the context
    stream), swaps/s of the whole loop, and the recount check."""
    n = N_VERTICES
    male, female, truth, _ = bm.synth_pair(n, N_EDGES / n, bm.WEIGHTS_LOGNORMAL, 0.3, seed=1)
    pi0 = bm.greedy_match(n, male, female)
    ctx = bm.Context(n, male, female)
    ctx.set_mapping(pi0)
    ctx.gain_enable(True)
    ctx.best_swap()                                   # transposed table + warm-up
    pairs = n * (n - 1) // 2
    ctx.timing_enable(True)
    ctx.timing_reset()
    k = 20
    t = time.perf_counter()
    for _ in range(k):
        ctx.best_swap()
    wall_call = (time.perf_counter() - t) / k
    calls, ms = ctx.timing(bm.KERNEL_GAIN_SWAP)
    launch_ms = ms / max(1, calls)
    ctx.timing_reset()
    t = time.perf_counter()
    trace, res = ctx.best_swap_climb(args.best_swap, verify_every=100)
    wall = time.perf_counter() - t
    swaps = res["iterations"]
    log(f"best swap over {pairs} pairs: {launch_ms:.2f} ms in the kernels per call ({pairs / launch_ms * 1e3:.4g} deltas/s), "
        f"{wall_call * 1e3:.2f} ms per call; climb {trace[0]} -> {trace[-1]} in {swaps} swaps, "
        f"{swaps / wall:.1f} swaps/s, recount failures {res['verify_failures']}")
    result = {
        "metric": "exact best swap over all pairs: pair deltas/s (steepest ascent)",
        "value": pairs / launch_ms * 1e3,
        "unit": "swap-deltas/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "steepest ascent, n=18524 ~4M edges/graph, greedy initial matching",
                   "pairs_per_call": pairs, "swaps": args.best_swap},
        "ms_per_best_swap_kernels": launch_ms,
        "ms_per_best_swap_call": wall_call * 1e3,
        "swaps_per_s": swaps / wall,
        "swaps_applied": swaps,
        "score": {"initial": int(trace[0]), "final": int(trace[-1])},
        "recount_failures": int(res["verify_failures"]),
    }
    ctx.close()
    print(json.dumps(result), flush=True)


def bench_rot3(ctx, bm, args):
    """3-vertex permutations on the a-major uniform triple stream (triples
    grouped by first vertex, the law of independent uniform triples), T per
    step.  gain: the staged table kernel (one block per first vertex, two
    table reads from HBM per triple); scan: the stream materialised and
    evaluated by the scan kernel.  value = triples per second of whole steps
    (stream offsets included); the two strategies' results are compared on
    one stream."""
    T = args.triples
    b3 = ctx.batch(T)
    n = ctx.n
    mo, _ = ctx.degrees(female=False)
    out = {"triples_per_step": T, "stream": "amajor", "unit": "triples/s (5 permutations each)"}
    heads = {}
    for mode in ("gain", "scan"):
        ctx.gain_enable(mode == "gain")
        b3.eval_rot3_amajor(3000, T)
        kid = bm.KERNEL_GAIN_ROT3 if mode == "gain" else bm.KERNEL_ROT3
        steps = args.steps if mode == "gain" else max(1, args.scan_steps)
        wall, kt = timed(ctx, (kid,), lambda i: b3.eval_rot3_amajor(4000 + i, T), steps)
        l3 = per_launch_ms(kt, kid) / 1000.0
        if mode == "gain":
            # compulsory: two staged rows per first vertex, the per-vertex words,
            # the male lists (for a's bitmaps) and the stream offsets once; per
            # triple two table entries, two dense adjacency words, the best
            # delta (8 B) and permutation (1 B)
            ab3 = (n * 8 * gain_stride(n) + (n + 1) * 16 + 16 * int(mo.sum()) + (n + 2) * 8
                   + T * (2 * 4 + 2 * 4 + 9))
        else:
            ab3, _ = b3.algo_bytes(generated=False)
        kern = "k_gain_rot3_amajor+k_gain_rot3_edges" if mode == "gain" else "k_rot3"
        out[mode] = {"value": T * steps / wall, "ms_per_step": 1000 * wall / steps, "launch_ms": 1000 * l3,
                     "roofline": {"bound": "hbm", "achieved": ab3 / l3 / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": ab3 / l3 / 1e9 / HBM_PEAK_GBS, "algo_bytes_per_launch": ab3,
                                  "kernel": kern, "traffic": measured_traffic(args).get(kern)}}
        b3.eval_rot3_amajor(99, T)
        heads[mode] = (b3.deltas(0, 1 << 16), b3.perms(0, 1 << 16))
        log(f"rot3 {mode}: {T} a-major triples in {1000 * wall / steps:.1f} ms/step ({1000 * l3:.1f} ms kernel) "
            f"-> {T * steps / wall:.4g} triples/s")
    same = bool((heads["gain"][0] == heads["scan"][0]).all() and (heads["gain"][1] == heads["scan"][1]).all())
    out["gain_matches_scan"] = same
    log(f"rot3 gain and scan results identical on the first 65536 triples of one stream: {same}")
    if not same:
        raise SystemExit("the two exact 3-vertex strategies disagree")
    ctx.gain_enable(False)
    b3.close()
    return out


def host_cpu():
    """The host the CPU baseline runs on: CPU model, logical CPUs of the
    machine (nproc), CPUs this process may run on, and OMP_NUM_THREADS."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "affinity_cpus": aff,
            "omp_num_threads": int(omp) if omp.isdigit() and int(omp) > 0 else None}


def cpu_baseline(args, male, female, pi0, batch, seed, K, n):
    """The reference CPU arithmetic on a bounded, evenly strided sample of the
    last timed stream."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    host = host_cpu()
    if orc.ref_available() and n <= 18524:
        kind = "reference"
        runner = orc.RefCode(male, female)
        what = "src/graph-alignment-optimized.c calculate_swap_delta compiled from the reference sources"
    else:
        kind = "port"
        runner = orc.ListsRef(n, male, female)
        what = "oracle restatement of calculate_swap_delta (edge lists + dense int16 female matrix)"
    umajor = args.stream == "umajor"

    def pairs_at(idx):
        if umajor:
            return orc.umajor_pairs_at(seed, K, n, idx)
        return orc.uniform_pairs(seed, int(idx[-1]) + 1, n)[idx]

    calib = 1 << 16
    idx = np.linspace(0, K - 1, calib).astype(np.int64)
    uv = pairs_at(idx)
    # calibrate with every logical CPU the process may run on and with the
    # box's configured OpenMP share; the timed sample uses the faster count
    counts = sorted({host["affinity_cpus"], host["nproc"], host["omp_num_threads"] or host["affinity_cpus"]})
    calib_rates = {}
    for th in counts:
        t = time.perf_counter()
        runner.deltas(pi0, uv, th)
        calib_rates[th] = calib / max(1e-6, time.perf_counter() - t)
    threads = max(calib_rates, key=calib_rates.get)
    rate = calib_rates[threads]
    run = lambda pi, uv: runner.deltas(pi, uv, threads)  # noqa: E731
    S = int(min(K, max(calib, rate * args.cpu_seconds)))
    S -= S % 1024
    idx = (np.arange(S, dtype=np.int64) * (K // S)).astype(np.int64)
    uv = pairs_at(idx)
    t = time.perf_counter()
    d_cpu = run(pi0, uv)
    el = time.perf_counter() - t
    d_gpu = batch.deltas(0, K)[idx]
    match = bool(np.array_equal(d_cpu.astype(np.int64), d_gpu))
    log(f"cpu baseline ({kind}, {threads} threads): {S} deltas in {el:.2f}s -> {S / el:.4g}/s; "
        f"device deltas identical: {match}")
    return {
        "value": S / el,
        "unit": "swap-deltas/s",
        "cores": threads,
        "kind": kind,
        "sample": f"{S} pairs evenly strided over the last timed step's {args.stream} stream; {what}; "
                  f"OpenMP schedule(dynamic,1024) on {threads} threads",
        "host": host,
        "calibration_rates_by_threads": {str(k): v for k, v in calib_rates.items()},
        "matches_device": match,
    }


if __name__ == "__main__":
    main()

"""Device parity of the per-vertex best-partner sweep (the OpenMP tool's
loop, src/re-alignment-parallel.c:556-676): best partners against the
oracle, whole sweeps with restarts replayed on the CPU, and a timed sweep on
the n = 18,524 pair checked by recount."""
import numpy as np
import pytest

from conftest import quirky_pair

pytestmark = pytest.mark.gpu


def test_best_partners_match_oracle(bm, orc):
    """Every vertex's best partner v > u (largest exact delta, ties to the
    smaller v) equals the oracle's, on a power-law pair (dense hub rows in
    the table) in launches of several sizes, and on the quirky graph (self-
    loops, duplicates, unmatched vertices, non-positive female weights)."""
    n = 300
    male, female, truth, pi0 = bm.synth_pair(n, 15.0, bm.WEIGHTS_LOGNORMAL, 0.3, seed=51)
    o = orc.Oracle(n, male, female)
    want = [orc.best_partner(o, pi0, u) for u in range(1, n + 1)]
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        ctx.gain_enable(True)
        d, v = ctx.best_partners(1, n)
        assert [(int(a), int(b)) for a, b in zip(d, v)] == want
        for u0, cnt in ((5, 1), (100, 37), (280, 21)):
            d, v = ctx.best_partners(u0, cnt)
            assert [(int(a), int(b)) for a, b in zip(d, v)] == want[u0 - 1:u0 - 1 + cnt]
    male, female, pi = quirky_pair(bm, seed=6)
    o = orc.Oracle(60, male, female)
    with bm.Context(60, male, female) as ctx:
        ctx.set_mapping(pi)
        ctx.gain_enable(True)
        d, v = ctx.best_partners(1, 60)
        assert [(int(a), int(b)) for a, b in zip(d, v)] == [orc.best_partner(o, pi, u) for u in range(1, 61)]


@pytest.mark.parametrize("batch", [1, 64, 512])
def test_sweep_replays_on_cpu(bm, orc, batch):
    """Ten passes with an initial perturbation and restarts, in launches of
    1, 64 and 512 vertices: the best matching, the best score after every
    pass, the swaps applied, the restarts and the pairs evaluated equal the
    oracle's replay; every running score equals a device recount, and the
    context is left on the best matching."""
    n = 300
    male, female, truth, pi0 = bm.synth_pair(n, 15.0, bm.WEIGHTS_LOGNORMAL, 0.3, seed=52)
    o = orc.Oracle(n, male, female)
    best_o, trace_o, r_o = orc.sweep(o, pi0, seed=9, max_passes=10, initial_swaps=20, restart_swaps=30)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        ctx.gain_enable(True)
        best, trace, r = ctx.sweep(seed=9, max_passes=10, initial_swaps=20, restart_swaps=30, batch=batch)
        print(f"[sweep] batch {batch}: {r['initial_score']} -> {r['best_score']}, {r['accepted']} swaps, "
              f"{r['restarts']} restarts", flush=True)
        assert r["verify_failures"] == 0
        assert np.array_equal(best, best_o) and np.array_equal(trace, trace_o)
        assert (r["best_score"], r["accepted"], r["restarts"], r["evaluations"]) == \
            (r_o["best_score"], r_o["accepted"], r_o["restarts"], r_o["evaluations"])
        assert r["iterations"] == 10 and r["restarts"] >= 1
        assert np.array_equal(ctx.get_mapping(), best) and ctx.score() == r["best_score"] == o.score(best)


def test_sweep_at_the_swap_delta_workload(bm, orc):
    """The n = 18,524 pair from its greedy matching, a 20-second sweep: the
    best score rises, every running score equals a device recount, and the
    oracle rescoring of the best matching agrees."""
    n = 18524
    male, female, truth, _ = bm.synth_pair(n, 4e6 / n, bm.WEIGHTS_LOGNORMAL, 0.3, seed=1)
    pi0 = bm.greedy_match(n, male, female)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        s0 = ctx.score()
        ctx.gain_enable(True)
        best, _, r = ctx.sweep(seed=3, time_limit_s=20.0, initial_swaps=0, verify=True)
        print(f"[sweep] n={n}: {s0} -> {r['best_score']} in {r['seconds']:.1f}s, {r['accepted']} swaps, "
              f"{r['iterations']} passes, {r['evaluations']} pairs", flush=True)
        assert r["verify_failures"] == 0 and r["best_score"] > s0 and r["accepted"] > 0
        assert orc.Oracle(n, male, female).score(best) == r["best_score"]

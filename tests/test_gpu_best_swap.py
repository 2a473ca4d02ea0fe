"""Exact best swap over all pairs and the steepest-ascent climb on the device
against the oracle's all-pairs maximum and its replay."""
import itertools
import time

import numpy as np
import pytest

from conftest import make_pair, quirky_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [3, 8])
def test_quirky_graph_best_swap(bm, orc, seed):
    """Self-loops, duplicates, u<->v edges, unmatched vertices, non-positive
    female weights: the device maximum and its pair equal the oracle's, and
    with no matching every delta is 0 and the first pair wins."""
    male, female, pi = quirky_pair(bm, seed=seed)
    o = orc.Oracle(60, male, female)
    with bm.Context(60, male, female) as ctx:
        ctx.set_mapping(pi)
        with pytest.raises(bm.BMError):
            ctx.best_swap()
        ctx.gain_enable(True)
        assert ctx.best_swap() == o.best_swap(pi)
        ctx.set_mapping(np.zeros(61, np.int32))
        assert ctx.best_swap() == (0, 1, 2)


@pytest.mark.parametrize("n", [2, 3])
def test_tiny_graphs(bm, orc, n):
    """The smallest contexts (n >= 2): every matching of a small pair with a
    self-loop and a 2-cycle agrees with the oracle."""
    src = np.array([1, 1, 2, 3, 3], np.int32)
    dst = np.array([1, 2, 1, 1, 2], np.int32)
    keep = (src <= n) & (dst <= n)
    male = bm.Graph(src[keep], dst[keep], np.array([4, 3, 2, 5, 1], np.int32)[keep], n)
    female = bm.Graph(dst[keep], src[keep], np.array([2, 6, 3, 1, 4], np.int32)[keep], n)
    o = orc.Oracle(n, male, female)
    with bm.Context(n, male, female) as ctx:
        ctx.gain_enable(True)
        for perm in itertools.permutations(range(1, n + 1)):
            pi = np.array((0,) + perm, np.int32)
            ctx.set_mapping(pi)
            assert ctx.best_swap() == o.best_swap(pi), perm


@pytest.mark.parametrize("n,deg,seed", [(300, 12.0, 11), (2000, 20.0, 1)])
def test_best_swap_matches_oracle(bm, orc, n, deg, seed):
    male, female, truth, pi0 = make_pair(bm, n=n, deg=deg, seed=seed)
    o = orc.Oracle(n, male, female)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        ctx.gain_enable(True)
        assert ctx.best_swap() == o.best_swap(pi0)
        ctx.set_mapping(truth)   # near a local optimum: ties among small deltas
        assert ctx.best_swap() == o.best_swap(truth)


def test_climb_replays_on_the_cpu(bm, orc):
    """40 steepest-ascent swaps at n = 600: score trace, final matching and
    a recount after every swap equal the oracle's replay."""
    n = 600
    male, female, truth, pi0 = make_pair(bm, n=n, deg=12.0, seed=1)
    o = orc.Oracle(n, male, female)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        ctx.gain_enable(True)
        trace, res = ctx.best_swap_climb(40, verify_every=1)
        pi_o, trace_o = o.best_swap_climb(pi0, 40)
        assert res["verify_failures"] == 0 and res["iterations"] == 40
        assert np.array_equal(trace, trace_o)
        assert np.array_equal(ctx.get_mapping(), pi_o)
        assert res["final_score"] == o.score(pi_o)


def test_climb_stops_at_local_optimum(bm, orc):
    male, female, pi = quirky_pair(bm, seed=5)
    o = orc.Oracle(60, male, female)
    with bm.Context(60, male, female) as ctx:
        ctx.set_mapping(pi)
        ctx.gain_enable(True)
        trace, res = ctx.best_swap_climb(10_000, verify_every=7)
        pi_o, trace_o = o.best_swap_climb(pi, 10_000)
        assert res["verify_failures"] == 0 and res["iterations"] == len(trace_o) - 1 < 10_000
        assert np.array_equal(trace, trace_o) and np.array_equal(ctx.get_mapping(), pi_o)
        assert ctx.best_swap()[0] <= 0


def test_full_size_best_swap(bm, orc):
    """n = 18,524, ~4M edges, all 171.6M pairs: the chosen pair's delta is the
    oracle's delta for that pair, no pair of a 2^20 uniform sample does better
    (nor equally well with a smaller pair), and applying it raises a recount
    by exactly that delta."""
    n = 18524
    male, female, truth, pi0 = bm.synth_pair(n, 4e6 / n, bm.WEIGHTS_LOGNORMAL, 0.3, seed=2)
    o = orc.Oracle(n, male, female)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        ctx.gain_enable(True)
        ctx.best_swap()
        t = time.perf_counter()
        d, u, v = ctx.best_swap()
        dt = time.perf_counter() - t
        print(f"best swap over {n * (n - 1) // 2} pairs: {d} at ({u}, {v}) in {dt * 1e3:.1f} ms", flush=True)
        assert 1 <= u < v <= n and d == o.swap_delta(pi0, u, v)
        k = 1 << 20
        b = ctx.batch(k)
        b.eval_swaps_uniform(5, k)
        ds = b.deltas()
        uv = orc.uniform_pairs(5, k, n)
        lo, hi = np.minimum(uv[:, 0], uv[:, 1]), np.maximum(uv[:, 0], uv[:, 1])
        assert ds.max() <= d
        tie = (ds == d) & (lo != hi)
        assert not np.any(tie & ((lo < u) | ((lo == u) & (hi < v))))
        b.close()
        s0 = ctx.score()
        ctx.apply_swap(u, v)
        assert ctx.score() == s0 + d

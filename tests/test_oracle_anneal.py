"""CPU checks that pin the oracle's annealing replay: the deterministic exp
against libm, the Metropolis rule against its acceptance probabilities, the
score and best-matching invariants, and replay in pieces == replay at once."""
import math

import numpy as np

from conftest import make_pair, quirky_pair


def test_det_exp_matches_libm():
    x = np.concatenate([np.linspace(-700, 700, 400001), [0.0, -0.0, 1e-300, -1e-300, -0.5, 0.5],
                        np.random.default_rng(0).uniform(-50, 0, 100000)])
    import oracle
    got = oracle.det_exp(x)
    ref = np.exp(x)
    ulp = np.spacing(ref)
    assert (np.abs(got - ref) <= 2 * ulp).all()
    assert oracle.det_exp(np.array([0.0]))[0] == 1.0
    assert oracle.det_exp(np.array([-800.0, np.nan]))[0] == 0.0
    assert oracle.det_exp(np.array([np.nan]))[0] == 1.0


def test_metropolis_rule_acceptance_rates(bm, orc):
    """Constant temperature T: a proposal with delta >= 0 is always taken and
    one with delta < 0 with probability exp(delta / T).  The observed count
    of accepted downhill moves must sit within 5 sigma of the expectation."""
    male, female, truth, pi0 = make_pair(bm, n=300, deg=10.0, seed=4)
    o = orc.Oracle(300, male, female)
    T, steps, seed = 3.0, 3000, 12345
    q = orc.chain_seed(seed, 0)
    pairs = orc.uniform_pairs(orc.chain_seed(q, 0), steps, 300)
    pi = pi0.copy()
    s0 = o.score(pi)
    state = [s0, s0, 0, 1]
    exp_sum = var = 0.0
    taken_down = 0
    for s in range(steps):
        u, v = (int(x) for x in pairs[s])
        d = o.swap_delta(pi, u, v)
        before = state[2]
        pi, _, state = orc.anneal_chain(o, pi, s0, q, s, 1, 1.0 / T, 0.0, 0, state=state)
        took = state[2] > before
        if d >= 0:
            assert took
        else:
            p = math.exp(d / T)
            exp_sum += p
            var += p * (1 - p)
            taken_down += took
    assert exp_sum > 50
    assert abs(taken_down - exp_sum) < 5 * math.sqrt(var) + 1
    assert state[0] == o.score(pi)


def test_chain_invariants_and_piecewise_replay(bm, orc):
    male, female, pi = quirky_pair(bm, seed=5)
    o = orc.Oracle(60, male, female)
    s0 = o.score(pi)
    q = orc.chain_seed(99, 3)
    p3 = orc.p_rot3_fraction(0.4)
    full_pi, full_best, full_st = orc.anneal_chain(o, pi, s0, q, 0, 900, 0.5, 1e-3, p3)
    assert full_st[0] == o.score(full_pi)
    assert full_st[1] == o.score(full_best) >= s0
    assert full_st[1] >= full_st[0]
    # the same 900 steps in three pieces
    st, p_, pb = None, pi, None
    for a, L in ((0, 250), (250, 400), (650, 250)):
        p_, pb_eff, st = orc.anneal_chain(o, p_, s0, q, a, L, 0.5, 1e-3, p3, state=st, pib=pb)
        pb = pb_eff
    assert np.array_equal(st, full_st)
    assert np.array_equal(p_, full_pi) and np.array_equal(pb, full_best)


def test_greedy_and_hot_limits(bm, orc):
    male, female, truth, pi0 = make_pair(bm, n=200, deg=8.0, seed=6)
    o = orc.Oracle(200, male, female)
    s0 = o.score(pi0)
    q = orc.chain_seed(1, 0)
    pi, st = pi0, None
    prev = s0
    for a in range(0, 600, 60):   # beta0 = 1e12: only delta >= 0 is taken
        pi, _, st = orc.anneal_chain(o, pi, s0, q, a, 60, 1e12, 0.0, orc.p_rot3_fraction(0.3), state=st)
        assert st[0] >= prev
        prev = st[0]
    assert st[0] == st[1] == o.score(pi)
    _, _, hot = orc.anneal_chain(o, pi0, s0, q, 0, 600, 1e-12, 0.0, orc.p_rot3_fraction(0.3))
    assert hot[2] >= 598


def test_triple_stream_distinct_and_in_range(orc):
    seen = set()
    for c in range(5000):
        a, b, d = orc.uniform_triple(7, c, 5)
        assert len({a, b, d}) == 3 and all(1 <= x <= 5 for x in (a, b, d))
        seen.add((a, b, d))
    assert len(seen) == 60   # every ordered triple of 5 vertices appears


def test_pool_moves_touch_only_pool_pairs(bm, orc):
    """With every move drawn from a one-pair pool, a swap-only chain only
    ever exchanges that pair's images; with 3-vertex moves the third vertex
    differs from the pair and is uniform over the others; the running score
    stays equal to a recount."""
    male, female, pi = quirky_pair(bm, seed=7)
    n = 60
    o = orc.Oracle(n, male, female)
    s0 = o.score(pi)
    q = orc.chain_seed(5, 1)
    full = 1 << 32
    pool = np.array([[11, 37]], np.int32)
    p2, _, st = orc.anneal_chain(o, pi, s0, q, 0, 500, 1e-9, 0.0, 0, pool=pool, p_pool=full)
    moved = np.nonzero(p2 != pi)[0]
    assert set(moved.tolist()) <= {11, 37} and st[2] >= 499
    assert st[0] == o.score(p2)
    # 3-vertex moves on the pool pair: the third vertex is uniform over the other 58
    thirds = np.array([orc.third_vertex(orc.chain_seed(q, 3), s, n, 11, 37) for s in range(58 * 400)])
    assert not np.isin(thirds, [11, 37]).any() and thirds.min() >= 1 and thirds.max() <= n
    h = np.bincount(thirds, minlength=n + 1)[1:]
    h = np.delete(h, [10, 36])
    assert np.abs(h - 400).max() < 6 * np.sqrt(400)
    p3, _, st3 = orc.anneal_chain(o, pi, s0, q, 0, 800, 0.3, 1e-3, orc.p_rot3_fraction(0.5), pool=pool,
                                  p_pool=full // 2)
    assert st3[0] == o.score(p3) and st3[1] >= st3[0]
    # p_pool = 0 ignores the pool: the uniform chain exactly
    a, _, sa = orc.anneal_chain(o, pi, s0, q, 0, 300, 0.3, 1e-3, orc.p_rot3_fraction(0.5), pool=pool, p_pool=0)
    b, _, sb = orc.anneal_chain(o, pi, s0, q, 0, 300, 0.3, 1e-3, orc.p_rot3_fraction(0.5))
    assert np.array_equal(a, b) and np.array_equal(sa, sb)

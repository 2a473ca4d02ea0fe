"""The benchmark configurations at their own shapes, checked against the
oracle: the n = 2,000 pair's whole list of 2^20 uniform pairs, the n = 18,524
pair from its greedy initial matching (both evaluation strategies, pairs and
triples, and the learned proposal against uniform proposals at equal rounds
and equal wall clock), and 64 batched annealing instances of uneven size
under the autotuned launch shape with one chain per instance replayed on
the CPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_c1_full_pair_list(bm, orc):
    """n = 2,000, ~40k edges, geometric weights, 30% shuffled: every one of
    2^20 uniform pairs equals the oracle, on both strategies."""
    n = 2000
    male, female, truth, pi0 = bm.synth_pair(n, 20.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=1)
    o = orc.Oracle(n, male, female)
    k = 1 << 20
    uv = orc.uniform_pairs(4, k, n)
    want = o.swap_deltas(pi0, uv)
    print("[configs] c1 oracle done", flush=True)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        assert ctx.score() == o.score(pi0)
        b = ctx.batch(k)
        b.eval_swaps_uniform(4, k)
        assert np.array_equal(b.deltas(), want)
        b.set_pairs(uv)
        b.eval_swaps()
        assert np.array_equal(b.deltas(), want)
        ctx.gain_enable(True)
        b.eval_swaps_uniform(4, k)
        assert np.array_equal(b.deltas(), want)
        b.close()


@pytest.fixture(scope="module")
def c2(bm, orc):
    """The n = 18,524, ~4M-edge pair with lognormal weights and the greedy
    initial matching (the swap-delta benchmark's workload)."""
    n = 18524
    male, female, truth, _ = bm.synth_pair(n, 4e6 / n, bm.WEIGHTS_LOGNORMAL, 0.3, seed=1)
    pi0 = bm.greedy_match(n, male, female)
    o = orc.Oracle(n, male, female)
    print("[configs] c2 pair, greedy matching and oracle ready", flush=True)
    return {"n": n, "male": male, "female": female, "truth": truth, "pi0": pi0, "o": o}


def test_c2_greedy_start_pairs(bm, orc, c2):
    """From greedy-match's matching: scan and table strategies agree on 2^21
    u-major and 2^20 draw-order pairs; 4096 of each equal the oracle."""
    n, o, pi0 = c2["n"], c2["o"], c2["pi0"]
    with bm.Context(n, c2["male"], c2["female"]) as ctx:
        ctx.set_mapping(pi0)
        assert ctx.score() == o.score(pi0)
        k = 1 << 21
        b = ctx.batch(k)
        b.eval_swaps_umajor(31, k)
        d_um = b.deltas()
        b.eval_swaps_uniform(32, k >> 1)
        d_un = b.deltas(0, k >> 1)
        ctx.gain_enable(True)
        b.eval_swaps_umajor(31, k)
        assert np.array_equal(b.deltas(), d_um)
        b.eval_swaps_uniform(32, k >> 1)
        assert np.array_equal(b.deltas(0, k >> 1), d_un)
        rng = np.random.default_rng(7)
        idx = rng.choice(k, 4096, replace=False)
        assert np.array_equal(d_um[idx], o.swap_deltas(pi0, orc.umajor_pairs(31, k, n)[idx]))
        idx = rng.choice(k >> 1, 4096, replace=False)
        assert np.array_equal(d_un[idx], o.swap_deltas(pi0, orc.uniform_pairs(32, k >> 1, n)[idx]))
        b.close()


def test_c2_greedy_start_triples(bm, orc, c2):
    """From greedy-match's matching: 2^20 a-major triples give the same best
    permutation and delta on both strategies; 2048 equal the oracle."""
    n, o, pi0 = c2["n"], c2["o"], c2["pi0"]
    T = 1 << 20
    with bm.Context(n, c2["male"], c2["female"]) as ctx:
        ctx.set_mapping(pi0)
        b = ctx.batch(T)
        b.eval_rot3_amajor(5, T)
        d_scan, p_scan = b.deltas(), b.perms()
        ctx.gain_enable(True)
        b.eval_rot3_amajor(5, T)
        assert np.array_equal(b.deltas(), d_scan) and np.array_equal(b.perms(), p_scan)
        abc = orc.amajor_triples(5, T, n)
        idx = np.random.default_rng(8).choice(T, 2048, replace=False)
        rb, rp = o.rot3_many(pi0, abc[idx])
        assert np.array_equal(d_scan[idx], rb) and np.array_equal(p_scan[idx].astype(np.int32), rp)
        b.close()


def test_c2_climb_reaches_local_optimum_exactly(bm, orc, c2):
    """From greedy-match's matching, disjoint-swap rounds of the table
    evaluator keep the running score equal to device recounts; the oracle
    rescoring of the final matching agrees."""
    n, o, pi0 = c2["n"], c2["o"], c2["pi0"]
    with bm.Context(n, c2["male"], c2["female"]) as ctx:
        ctx.set_mapping(pi0)
        ctx.gain_enable(True)
        trace, res = ctx.hillclimb(seed=100, rounds=200, batch=4096, verify=True)
        assert res["verify_failures"] == 0 and trace[-1] > trace[0]
        assert o.score(ctx.get_mapping()) == trace[-1]


def _replay_chain(orc, o, pi0, ann, c, steps):
    beta0, lam, p3 = ann.schedule()
    pi, _, state = orc.anneal_chain(o, pi0, o.score(pi0), orc.chain_seed(ann.seed, c), 0, steps, beta0, lam, p3)
    return pi, state


def test_c3_sixty_four_instances_autotuned(bm, orc):
    """64 instances with n uniform in 1k..20k (edges scaled to ~30 n to keep
    the CPU replays short), lognormal weights, 30% shuffled; annealing with
    T_s = 50 * 0.9999^s and 10% 3-vertex moves in the autotuned launch shape;
    one chain of every instance replayed on the CPU over 240 steps (three
    launches of 80) equals the device: running score, acceptance count and
    matching."""
    rng = np.random.default_rng(64)
    sizes = [int(x) for x in rng.integers(1000, 20001, 64)]
    pairs = [bm.synth_pair(n, 30.0, bm.WEIGHTS_LOGNORMAL, 0.3, seed=1000 + i) for i, n in enumerate(sizes)]
    ctxs = []
    try:
        for (male, female, truth, pi0), n in zip(pairs, sizes):
            c = bm.Context(n, male, female)
            c.set_mapping(pi0)
            ctxs.append(c)
        cpi = 4
        with bm.Anneal(ctxs, chains_per_instance=cpi, seed=5, t0=50.0, decay=0.9999, p_rot3=0.1,
                       traj_cap=3) as ann:
            best, trials = ann.autotune(steps=40)
            assert len(trials) >= 24 and ann.tuning()["waves_per_block"] == best["waves_per_block"]
            print(f"[configs] c3 autotuned: {best}", flush=True)
            ann.run(80, 3)
            assert ann.verify() == 0
            st = ann.state()
            assert (st["steps"] == 240).all()
            for i, ((male, female, truth, pi0), n) in enumerate(zip(pairs, sizes)):
                o = orc.Oracle(n, male, female)
                c = i * cpi + (i % cpi)
                pi, state = _replay_chain(orc, o, pi0, ann, c, 240)
                assert state[0] == st["score"][c] and state[1] == st["best"][c], (i, c)
                assert state[2] == st["accepted"][c], (i, c)
                assert np.array_equal(pi, ann.mapping(c)), (i, c)
                assert o.score(pi) == st["score"][c], (i, c)
    finally:
        for c in ctxs:
            c.close()


def test_c4_learned_beats_uniform_at_equal_rounds_and_time(bm, c2):
    """The learned-proposal configuration: the n = 18,524 pair from its
    greedy matching, the network trained on ~8M (candidate, exact delta)
    samples from the context's own kernel, each learned round ranking 2^20
    uniform pairs and sending the best 4,096 to the exact kernel (with
    online retraining and the uniform fallback).  Against 4,096 uniform
    pairs per round it reaches a higher exact score after the same number of
    rounds and after the same wall clock; every running score equals a
    device recount at the end."""
    n, pi0 = c2["n"], c2["pi0"]
    with bm.Context(n, c2["male"], c2["female"]) as ctx:
        ctx.set_mapping(pi0)
        with bm.Mlp(ctx, seed=1) as m:
            m.refresh_features()
            m.train(seed=11, steps=128, batch=1 << 16, lr=1e-3)
            kw = dict(pool=1 << 20, mlp=m, refresh_every=1, retrain_every=1, retrain_lr=2e-4, fallback=True)
            results = {}
            for label, limit in (("rounds", dict(max_rounds=2500)), ("time", dict(time_limit_s=6.0))):
                ctx.set_mapping(pi0)
                tu, _, ru = ctx.climb(bm.PROPOSAL_UNIFORM, seed=21, per_round=4096, **limit)
                ctx.set_mapping(pi0)
                tl, _, rl = ctx.climb(bm.PROPOSAL_LEARNED, seed=21, per_round=4096, **limit, **kw)
                print(f"[configs] c4 equal {label}: uniform {ru['final_score']} ({ru['iterations']} rounds), "
                      f"learned {rl['final_score']} ({rl['iterations']} rounds)", flush=True)
                for r in (ru, rl):
                    assert r["verify_failures"] == 0
                assert rl["final_score"] > ru["final_score"]
                results[label] = (ru, rl)
            assert results["rounds"][0]["iterations"] == results["rounds"][1]["iterations"] == 2500

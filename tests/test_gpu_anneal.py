"""Annealing chains on the device against their CPU replay: every chain's
running score, best score, acceptance count, current and best matchings and
per-launch trajectory must equal the oracle's replay of the same streams."""
import numpy as np
import pytest

from conftest import quirky_pair

pytestmark = pytest.mark.gpu


def replay_all(orc, oracles, ann, starts, launches):
    """Replays every chain of `ann` through the launch lengths in `launches`
    and compares with the device after all of them."""
    beta0, lam, p3 = ann.schedule()
    st = ann.state()
    traj = ann.trajectory()
    for c in range(ann.chains):
        i = ann.instance_of(c)
        o = oracles[i]
        q = orc.chain_seed(ann.seed, c)
        pi = starts[i].copy()
        s0 = o.score(pi)
        state = [s0, s0, 0, 1]
        pib = pi.copy()
        s = 0
        tr = []
        for L in launches:
            pi, pib_eff, state = orc.anneal_chain(o, pi, s0, q, s, L, beta0, lam, p3, state=state, pib=pib)
            if not state[3]:
                pib = pib_eff
            s += L
            tr.append(int(state[0]))
        assert st["score"][c] == state[0], c
        assert st["best"][c] == state[1], c
        assert st["accepted"][c] == state[2], c
        assert st["steps"][c] == s, c
        assert np.array_equal(ann.mapping(c), pi), c
        best_pi = ann.mapping(c, best=True)
        assert np.array_equal(best_pi, pi if state[3] else pib), c
        assert o.score(best_pi) == state[1], c
        assert o.score(pi) == state[0], c
        if ann.traj_cap:
            assert np.array_equal(traj[c, :len(tr)], np.array(tr[:ann.traj_cap], np.int64)), c


def _mk(bm, orc, n, deg, seed, weights=None):
    male, female, truth, pi0 = bm.synth_pair(n, deg, weights or bm.WEIGHTS_GEOMETRIC, 0.3, seed=seed)
    ctx = bm.Context(n, male, female)
    ctx.set_mapping(pi0)
    return ctx, orc.Oracle(n, male, female), pi0


def test_mixed_moves_two_instances_replay(bm, orc):
    ctxs, oracles, starts = [], [], []
    for n, seed in ((600, 3), (900, 4)):
        c, o, p = _mk(bm, orc, n, 15.0, seed)
        ctxs.append(c)
        oracles.append(o)
        starts.append(p)
    try:
        with bm.Anneal(ctxs, chains_per_instance=4, seed=77, t0=5.0, decay=0.999, p_rot3=0.25,
                       traj_cap=8) as ann:
            ann.run(300, 3)
            assert ann.verify() == 0
            replay_all(orc, oracles, ann, starts, [300, 300, 300])
            st = ann.state()
            assert (st["accepted"] > 0).all() and (st["best"] >= st["score"]).all()
    finally:
        for c in ctxs:
            c.close()


def test_swap_only_replay_and_recapture(bm, orc):
    ctx, o, pi0 = _mk(bm, orc, 1500, 20.0, 5)
    try:
        with bm.Anneal([ctx], chains_per_instance=16, seed=5, t0=20.0, decay=0.9995, p_rot3=0.0,
                       traj_cap=4) as ann:
            ann.run(200, 2)
            ann.run(75, 1)     # a different launch length: a new captured graph
            assert ann.verify() == 0
            replay_all(orc, [o], ann, [pi0], [200, 200, 75])
            n, ms = ann.timing()
            assert n == 3 and ms > 0
            b, s = ann.algo_bytes()
            assert s == 16 * 475 and b > 48 * s
    finally:
        ctx.close()


def test_quirky_graph_chains_exact(bm, orc):
    """Self-loops, u<->v edges, duplicates, zero and negative female weights,
    unmatched vertices: every accepted move keeps the running score exact."""
    male, female, pi = quirky_pair(bm, seed=21)
    n = 60
    o = orc.Oracle(n, male, female)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi)
        with bm.Anneal([ctx], chains_per_instance=6, seed=9, t0=3.0, decay=0.99, p_rot3=0.5, traj_cap=2) as ann:
            ann.run(400, 2)
            assert ann.verify() == 0
            replay_all(orc, [o], ann, [pi], [400, 400])


def test_greedy_limit_and_hot_limit(bm, orc):
    """t0 -> 0 accepts only delta >= 0 (the score never decreases); a huge
    constant temperature accepts almost every move."""
    ctx, o, pi0 = _mk(bm, orc, 800, 12.0, 8)
    try:
        s0 = ctx.score()
        with bm.Anneal([ctx], chains_per_instance=4, seed=2, t0=1e-12, decay=1.0, traj_cap=16) as ann:
            ann.run(50, 16)
            t = ann.trajectory()
            assert (np.diff(t, axis=1) >= 0).all() and (t[:, 0] >= s0).all()
            assert ann.verify() == 0
        with bm.Anneal([ctx], chains_per_instance=4, seed=2, t0=1e12, decay=1.0) as ann:
            ann.run(500, 1)
            assert (ann.state()["accepted"] >= 495).all()
            assert ann.verify() == 0
    finally:
        ctx.close()


def test_full_size_chains_consistent(bm, orc):
    """n = 18,524 with ~4M edges: 512 chains x 100 steps; the running scores
    match device recounts and one chain matches its CPU replay."""
    n = 18524
    male, female, truth, pi0 = bm.synth_pair(n, 4e6 / n, bm.WEIGHTS_LOGNORMAL, 0.3, seed=2)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        with bm.Anneal([ctx], chains_per_instance=512, seed=3, t0=50.0, decay=0.9999, p_rot3=0.1) as ann:
            ann.run(100, 1)
            assert ann.verify() == 0
            st = ann.state()
            assert (st["best"] >= ctx.score()).all()
            o = orc.Oracle(n, male, female)
            beta0, lam, p3 = ann.schedule()
            for c in (0, 511):
                pi, _, state = orc.anneal_chain(o, pi0, o.score(pi0), orc.chain_seed(3, c), 0, 100, beta0, lam, p3)
                assert state[0] == st["score"][c] and state[2] == st["accepted"][c]
                assert np.array_equal(pi, ann.mapping(c))


def test_bad_parameters_raise(bm):
    male, female, truth, pi0 = bm.synth_pair(100, 5.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=1)
    with bm.Context(100, male, female) as ctx:
        for kw in (dict(chains_per_instance=0), dict(t0=0.0), dict(decay=1.5), dict(p_rot3=2.0),
                   dict(traj_cap=-1)):
            with pytest.raises(bm.BMError):
                bm.Anneal([ctx], **kw)
        with bm.Anneal([ctx], chains_per_instance=2) as ann:
            with pytest.raises(bm.BMError):
                ann.run(0, 1)
            with pytest.raises(bm.BMError):
                ann.mapping(5)


def test_autotune_then_replay(bm, orc):
    """The autotuner times several launch shapes, keeps the fastest and
    resets the chains, so a run after it replays exactly from the start."""
    ctx, o, pi0 = _mk(bm, orc, 1200, 15.0, 13)
    try:
        with bm.Anneal([ctx], chains_per_instance=8, seed=17, t0=10.0, decay=0.999, p_rot3=0.2, traj_cap=2) as ann:
            best, trials = ann.autotune(steps=50)
            assert len(trials) >= 24 and (trials[:, 4] > 0).all() and set(trials[:, 3]) == {2.0, 4.0}
            assert ann.tuning()["waves_per_block"] == best["waves_per_block"]
            assert ann.tuning()["unroll"] == best["unroll"]
            st = ann.state()
            assert (st["steps"] == 0).all() and (st["accepted"] == 0).all()
            ann.run(150, 2)
            assert ann.verify() == 0
            replay_all(orc, [o], ann, [pi0], [150, 150])
            ann.reset()
            assert (ann.state()["steps"] == 0).all()
            assert np.array_equal(ann.mapping(0), pi0)
    finally:
        ctx.close()


def test_pool_chains_replay(bm, orc):
    """Two instances: one with a pool from the learned proposal (the 512
    best-ranked of 2^14 uniform pairs), one with a pool set from the host;
    half of the moves come from the pools, 30% are 3-vertex moves.  Every
    chain replays exactly on the CPU given the device's pools, the chains
    count their pool moves, and replacing a pool between runs needs no
    recapture."""
    ctxs, oracles, starts = [], [], []
    for n, seed in ((700, 31), (1100, 32)):
        c, o, p = _mk(bm, orc, n, 15.0, seed)
        ctxs.append(c)
        oracles.append(o)
        starts.append(p)
    try:
        with bm.Mlp(ctxs[0], seed=3) as m:
            m.refresh_features()
            m.train(seed=5, steps=40, batch=1 << 12, lr=2e-3)
            with bm.Anneal(ctxs, chains_per_instance=6, seed=41, t0=2.0, decay=0.999, p_rot3=0.3,
                           traj_cap=4, p_pool=0.5) as ann:
                ann.pool_learned(0, m, seed=77, pool=1 << 14, top=512)
                host_pool = orc.uniform_pairs(9, 300, 1100)
                ann.set_pool(1, host_pool)
                pools = [ann.get_pool(0), ann.get_pool(1)]
                assert pools[0].shape == (512, 2) and np.array_equal(pools[1], host_pool)
                b = ctxs[0].batch(1 << 14)
                b.gen_learned(m, 77, 1 << 14, 512)
                assert np.array_equal(b.pairs(), pools[0])
                b.close()
                ann.run(150, 2)
                assert ann.verify() == 0
                beta0, lam, p3 = ann.schedule()
                pp = ann.p_pool()
                assert pp == 1 << 31
                st = ann.state()
                rot, pooled = ann.move_counts()
                assert (pooled > 0.35 * 300).all() and (pooled < 0.65 * 300).all() and (rot > 0).all()
                for c in range(ann.chains):
                    i = c // 6
                    pi, _, state = orc.anneal_chain(oracles[i], starts[i], oracles[i].score(starts[i]),
                                                    orc.chain_seed(41, c), 0, 300, beta0, lam, p3,
                                                    pool=pools[i], p_pool=pp)
                    assert state[0] == st["score"][c] and state[2] == st["accepted"][c], c
                    assert np.array_equal(pi, ann.mapping(c)), c
                # a new pool for instance 1, then more steps on the captured graph
                pool2 = orc.uniform_pairs(10, 50, 1100)
                ann.set_pool(1, pool2)
                ann.run(150, 1)
                assert ann.verify() == 0
                st2 = ann.state()
                for c in range(6, 12):
                    pi, _, state = orc.anneal_chain(oracles[1], starts[1], oracles[1].score(starts[1]),
                                                    orc.chain_seed(41, c), 0, 300, beta0, lam, p3,
                                                    pool=pools[1], p_pool=pp)
                    pi, _, state = orc.anneal_chain(oracles[1], pi, 0, orc.chain_seed(41, c), 300, 150, beta0, lam,
                                                    p3, state=state, pool=pool2, p_pool=pp)
                    assert state[0] == st2["score"][c] and np.array_equal(pi, ann.mapping(c)), c
                with pytest.raises(bm.BMError):
                    ann.set_pool(1, np.array([[3, 3]], np.int32))
                with pytest.raises(bm.BMError):
                    ann.pool_learned(1, m, seed=1, pool=1 << 10, top=16)   # network of another context
    finally:
        for c in ctxs:
            c.close()

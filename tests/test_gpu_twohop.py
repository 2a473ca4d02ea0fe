"""2-hop proposal generator on the device against the oracle: the sets, the
sampled pairs (explicit and seeded anchors) and the deltas of a batch."""
import numpy as np
import pytest

from conftest import make_pair, quirky_pair

pytestmark = pytest.mark.gpu


def test_sets_match_oracle_small_graphs(bm, orc):
    for male, female, pi, n in [quirky_pair(bm, seed=4) + (60,), make_pair(bm, n=300, deg=6.0, seed=8)[:2] + (None, 300)]:
        o = orc.Oracle(n, male, female)
        with bm.Context(n, male, female) as ctx:
            for v in range(1, n + 1):
                assert np.array_equal(ctx.twohop_set(v), orc.twohop_set(o, v)), v


def test_pairs_match_oracle(bm, orc):
    male, female, truth, pi0 = make_pair(bm, n=2000, deg=20.0, seed=3)
    o = orc.Oracle(2000, male, female)
    with bm.Context(2000, male, female) as ctx:
        ctx.set_mapping(pi0)
        b = ctx.batch(1 << 16)
        anchors = np.arange(1, 2001, 7, dtype=np.int32)
        for K in (1, 16, 100, 1000):
            na = min(len(anchors), (1 << 16) // K)
            b.gen_twohop(K, 11, anchors=anchors[:na])
            assert np.array_equal(b.pairs(), orc.twohop_pairs(o, K, 11, anchors=anchors[:na])), K
        b.gen_twohop(64, 99, n_anchors=1000)
        ref = orc.twohop_pairs(o, 64, 99, n_anchors=1000)
        assert np.array_equal(b.pairs(), ref)
        b.eval_swaps()
        assert np.array_equal(b.deltas(), o.swap_deltas(pi0, ref))
        b.close()


def test_full_size_sets_and_samples(bm, orc):
    """n = 18,524 with ~4M edges: hubs have 2-hop sets of most of the graph,
    which exercises the radix select and its tie handling."""
    n = 18524
    male, female, truth, pi0 = bm.synth_pair(n, 4e6 / n, bm.WEIGHTS_LOGNORMAL, 0.3, seed=2)
    o = orc.Oracle(n, male, female)
    with bm.Context(n, male, female) as ctx:
        od, idg = ctx.degrees()
        order = np.argsort(-(od + idg)[1:]) + 1
        vs = [int(x) for x in order[:3]] + [int(x) for x in order[-3:]] + [1, 9000, n]
        for v in vs:
            assert np.array_equal(ctx.twohop_set(v), orc.twohop_set(o, v)), v
        b = ctx.batch(1 << 16)
        b.gen_twohop(1000, 5, anchors=np.array(vs, np.int32))
        assert np.array_equal(b.pairs(), orc.twohop_pairs(o, 1000, 5, anchors=np.array(vs, np.int32)))
        b.close()


def test_bad_anchor_and_capacity_raise(bm):
    male, female, truth, pi0 = make_pair(bm, n=100, deg=5.0, seed=1)
    with bm.Context(100, male, female) as ctx:
        b = ctx.batch(100)
        with pytest.raises(bm.BMError):
            b.gen_twohop(10, 1, anchors=np.array([0], np.int32))
        with pytest.raises(bm.BMError):
            b.gen_twohop(10, 1, n_anchors=11)
        with pytest.raises(bm.BMError):
            ctx.twohop_set(101)
        b.close()


def test_twohop_climb_matches_replay(bm, orc):
    male, female, truth, pi0 = make_pair(bm, n=2000, deg=20.0, seed=3)
    o = orc.Oracle(2000, male, female)
    with bm.Context(2000, male, female) as ctx:
        ctx.set_mapping(pi0)
        trace, res = ctx.hillclimb_twohop(seed=50, rounds=12, anchors=64, per_anchor=100, verify=True)
        assert res["verify_failures"] == 0 and trace[-1] > trace[0]
        pi = pi0.copy()
        ref = [o.score(pi)]
        for r in range(12):
            uv = orc.twohop_pairs(o, 100, 50 + r, n_anchors=64)
            d = o.swap_deltas(pi, uv)
            pi, _, gain = o.accept_disjoint(pi, uv, d)
            ref.append(ref[-1] + gain)
        assert np.array_equal(trace, np.array(ref, np.int64))
        assert np.array_equal(ctx.get_mapping(), pi)

"""CPU checks that pin the oracle's 2-hop proposal restatement: the set
against a breadth-first search over the raw edge lines, the sample's
structure and its uniformity."""
import numpy as np

from conftest import make_pair, quirky_pair


def bfs_twohop(g, v):
    """collect_neighborhood's set from the raw (possibly duplicated) lines."""
    out, inn = {}, {}
    for s, d in zip(g.src.tolist(), g.dst.tolist()):
        out.setdefault(s, set()).add(d)
        inn.setdefault(d, set()).add(s)
    res = set()
    for x in out.get(v, ()):
        res.add(x)
        res |= out.get(x, set())
    for x in inn.get(v, ()):
        res.add(x)
        res |= inn.get(x, set())
    res.discard(v)
    return sorted(res)


def test_twohop_set_equals_bfs(bm, orc):
    for male, female, pi, n in [quirky_pair(bm, seed=4) + (60,), make_pair(bm, n=300, deg=4.0, seed=8)[:2] + (None, 300)]:
        o = orc.Oracle(n, male, female)
        for v in range(1, n + 1):
            assert orc.twohop_set(o, v).tolist() == bfs_twohop(male, v), v


def test_twohop_sample_structure_and_uniformity(bm, orc):
    male, female, truth, pi0 = make_pair(bm, n=300, deg=6.0, seed=9)
    o = orc.Oracle(300, male, female)
    K = 8
    anchors = np.array([17] * 4000, np.int32)
    pairs = orc.twohop_pairs(o, K, 5, anchors=anchors).reshape(4000, K, 2)
    full = orc.twohop_set(o, 17)
    S = len(full)
    assert S > 3 * K
    counts = {x: 0 for x in full.tolist()}
    for a in range(4000):
        t = pairs[a, :, 1]
        assert (pairs[a, :, 0] == 17).all()
        assert (np.diff(t) > 0).all()                 # ascending, distinct
        assert set(t.tolist()) <= set(counts)
        for x in t.tolist():
            counts[x] += 1
    expect = 4000 * K / S
    dev = np.array(list(counts.values())) - expect
    sd = np.sqrt(expect * (1 - K / S))
    assert np.abs(dev).max() < 5 * sd                  # every member ~ K/S of the draws
    # a set smaller than K: all of it, then (v, v) padding
    small = [v for v in range(1, 301) if 0 < len(orc.twohop_set(o, v)) < K][:5]
    for v in small:
        p = orc.twohop_pairs(o, K, 1, anchors=np.array([v], np.int32))
        s = orc.twohop_set(o, v)
        assert p[:len(s), 1].tolist() == s.tolist() and (p[len(s):, 1] == v).all()


def test_twohop_seeded_anchors_are_uniform_stream(bm, orc):
    male, female, truth, pi0 = make_pair(bm, n=300, deg=6.0, seed=9)
    o = orc.Oracle(300, male, female)
    p = orc.twohop_pairs(o, 1, 3, n_anchors=30000)
    a = p[:, 0]
    assert a.min() >= 1 and a.max() <= 300
    h = np.bincount(a, minlength=301)[1:]
    assert np.abs(h - 100).max() < 5 * 10

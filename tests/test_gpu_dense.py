"""Device parity of the table build's dense rows: the female rows of at
least a quarter of the columns, whose part of the table is summed 16 male
rows at a time from 16-bit column arrays before the row build adds the other
edges.  Rows match their definition and deltas the oracle when dense rows are
in use, at the 16-bit sum limit, after patches, and when the conditions that
switch them off (a shared image, a negative male weight) hold."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _all_rows_match(ctx, orc, o, pi, n):
    for u in range(n + 1):
        assert np.array_equal(ctx.gain_row(u), orc.gain_row(o, pi, u)), u


def test_power_law_pair_uses_dense_rows(bm, orc):
    """n = 1,200 power-law pair (hubs adjacent to most vertices): dense rows
    are in use; every row matches its definition; u-major and uniform deltas
    and 3-vertex deltas match the oracle; rows patched after a swap and a
    3-rotation equal a rebuild, which again matches the definition."""
    n = 1200
    male, female, truth, pi0 = bm.synth_pair(n, 30.0, bm.WEIGHTS_LOGNORMAL, 0.3, seed=41)
    o = orc.Oracle(n, male, female)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        ctx.gain_enable(True)
        nd = ctx.gain_dense_rows()
        print(f"[dense] n={n}: {nd} dense rows", flush=True)
        assert nd > 0
        _all_rows_match(ctx, orc, o, pi0, n)
        k = 1 << 16
        b = ctx.batch(k)
        b.eval_swaps_umajor(5, k)
        assert np.array_equal(b.deltas(), o.swap_deltas(pi0, orc.umajor_pairs(5, k, n)))
        b.eval_swaps_uniform(6, k)
        assert np.array_equal(b.deltas(), o.swap_deltas(pi0, orc.uniform_pairs(6, k, n)))
        b.close()
        rng = np.random.default_rng(3)
        abc = np.stack([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(1000)]).astype(np.int32)
        best, perm = ctx.rot3_deltas(abc)
        rb, rp = o.rot3_many(pi0, abc)
        assert np.array_equal(best, rb) and np.array_equal(perm.astype(np.int32), rp)
        hubs = np.argsort(-(np.bincount(male.src, minlength=n + 1) + np.bincount(male.dst, minlength=n + 1)))[:2]
        ctx.apply_swap(int(hubs[0]), 7)
        ctx.apply_perm3(int(hubs[1]), 8, 9, 4)
        pi1 = ctx.get_mapping()
        patched = [ctx.gain_row(u) for u in range(n + 1)]
        ctx.gain_rebuild()
        assert ctx.gain_dense_rows() == nd
        for u in range(n + 1):
            assert np.array_equal(ctx.gain_row(u), patched[u]), u
        _all_rows_match(ctx, orc, o, pi1, n)


def test_dense_sums_at_the_16_bit_limit(bm, orc):
    """Three female hubs with an edge to every other vertex of weight 21,845
    and three male vertices mapped onto them with edges of weight 30,000 to
    every vertex: a column's dense sum reaches exactly 65,535.  A fourth hub
    of the same degree with weights 1 would take columns to 65,536, so it is
    left to the row build, which adds it on top of the 16-bit sums; every
    row matches its definition."""
    n = 64
    rng = np.random.default_rng(11)
    src, dst, w = [], [], []
    fsrc, fdst, fw = [], [], []
    for h, hw in ((1, 21845), (2, 21845), (3, 21845), (4, 1)):   # female hubs, male 1..4 mapped onto them
        for v in range(1, n + 1):
            if v != h:
                fsrc.append(h), fdst.append(v), fw.append(hw)
                src.append(h), dst.append(v), w.append(30000)
    for _ in range(300):   # sparse rest, weights below the hubs'
        a, c = (int(x) for x in rng.integers(5, n + 1, 2))
        if a != c:
            src.append(a), dst.append(c), w.append(int(rng.integers(1, 500)))
            b, d = (int(x) for x in rng.integers(5, n + 1, 2))
            if b != d:
                fsrc.append(b), fdst.append(d), fw.append(int(rng.integers(1, 500)))
    male = bm.Graph(np.array(src, np.int32), np.array(dst, np.int32), np.array(w, np.int32), n)
    female = bm.Graph(np.array(fsrc, np.int32), np.array(fdst, np.int32), np.array(fw, np.int32), n)
    pi = np.arange(n + 1, dtype=np.int32)
    o = orc.Oracle(n, male, female)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi)
        ctx.gain_enable(True)
        assert ctx.gain_dense_rows() == 3
        row = ctx.gain_row(10)
        assert row[20] >= 3 * 21845 + 1
        _all_rows_match(ctx, orc, o, pi, n)


def test_shared_image_and_negative_weight_switch_dense_rows_off(bm, orc):
    """A mapping in which two vertices share a female, and a male graph with
    one negative weight, each build without dense rows; rows match their
    definition, and a later mapping with distinct images uses them again."""
    n = 800
    male, female, truth, pi0 = bm.synth_pair(n, 25.0, bm.WEIGHTS_LOGNORMAL, 0.3, seed=43)
    o = orc.Oracle(n, male, female)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        ctx.gain_enable(True)
        nd = ctx.gain_dense_rows()
        assert nd > 0
        shared = pi0.copy()
        hubs = np.argsort(-(np.bincount(male.src, minlength=n + 1) + np.bincount(male.dst, minlength=n + 1)))[:2]
        shared[int(hubs[1])] = shared[int(hubs[0])]
        ctx.set_mapping(shared)
        assert ctx.gain_dense_rows() == 0
        _all_rows_match(ctx, orc, o, shared, n)
        ctx.set_mapping(pi0)
        assert ctx.gain_dense_rows() == nd
        _all_rows_match(ctx, orc, o, pi0, n)
    neg = bm.Graph(male.src.copy(), male.dst.copy(), male.w.copy(), n)
    neg.w[5] = -3
    o2 = orc.Oracle(n, neg, female)
    with bm.Context(n, neg, female) as ctx:
        ctx.set_mapping(pi0)
        ctx.gain_enable(True)
        assert ctx.gain_dense_rows() == 0
        _all_rows_match(ctx, orc, o2, pi0, n)

"""Device parity: every C-ABI entry point against the CPU oracle, bit-exact.

Sizes are chosen so the oracle finishes in seconds; the full-size pair is
checked through sampled candidates and size-independent properties.
"""
import numpy as np
import pytest

from conftest import make_pair, quirky_pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1(bm, orc):
    """The CPU-runnable pair: n=2000, ~40k edges, geometric weights, 30% shuffled."""
    male, female, truth, pi0 = bm.synth_pair(2000, 20.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=1)
    ctx = bm.Context(2000, male, female)
    ctx.set_mapping(pi0)
    o = orc.Oracle(2000, male, female)
    yield dict(male=male, female=female, truth=truth, pi0=pi0, ctx=ctx, o=o, n=2000)
    ctx.close()


def test_score_matches_oracle(c1):
    ctx, o = c1["ctx"], c1["o"]
    assert ctx.score() == o.score(c1["pi0"])
    ctx.set_mapping(c1["truth"])
    assert ctx.score() == o.score(c1["truth"])
    ctx.set_mapping(c1["pi0"])


def test_mapping_roundtrip(c1):
    assert np.array_equal(c1["ctx"].get_mapping(), c1["pi0"])


def test_swap_deltas_list(c1, orc):
    ctx, o, pi = c1["ctx"], c1["o"], c1["pi0"]
    uv = orc.uniform_pairs(5, 1 << 16, c1["n"])
    d = ctx.swap_deltas(uv)
    ref = o.swap_deltas(pi, uv)
    assert np.array_equal(d, ref)
    assert (d != 0).any()


def test_generated_stream_matches_oracle(c1, orc):
    b = c1["ctx"].batch(4096)
    b.gen_pairs(77, 4096)
    assert np.array_equal(b.pairs(), orc.uniform_pairs(77, 4096, c1["n"]))
    b.close()


def test_uniform_fused_equals_list(c1, orc):
    ctx, o, pi = c1["ctx"], c1["o"], c1["pi0"]
    k = 1 << 15
    b = ctx.batch(k)
    b.eval_swaps_uniform(123, k)
    d = b.deltas()
    uv = orc.uniform_pairs(123, k, c1["n"])
    assert np.array_equal(d, o.swap_deltas(pi, uv))
    b.close()


def test_swap_symmetry_and_inverse(c1, orc):
    ctx, pi = c1["ctx"], c1["pi0"]
    uv = orc.uniform_pairs(9, 2048, c1["n"])
    d1 = ctx.swap_deltas(uv)
    d2 = ctx.swap_deltas(uv[:, ::-1].copy())
    assert np.array_equal(d1, d2)
    u, v = int(uv[0, 0]), int(uv[0, 1])
    s0 = ctx.score()
    ctx.apply_swap(u, v)
    assert ctx.score() == s0 + d1[0]
    assert ctx.swap_deltas(np.array([[u, v]]))[0] == -d1[0]
    ctx.apply_swap(u, v)
    assert np.array_equal(ctx.get_mapping(), pi)


def test_self_pair_is_zero(c1):
    assert c1["ctx"].swap_deltas(np.array([[5, 5], [7, 7]]))[0] == 0


def test_quick_enumeration(c1, orc):
    ctx, o, pi, n = c1["ctx"], c1["o"], c1["pi0"], c1["n"]
    rng = np.random.default_rng(2)
    changed = rng.choice(np.arange(1, n + 1), 20, replace=False).astype(np.int32)
    d, uv = ctx.swap_deltas_changed(changed)
    assert np.array_equal(uv, orc.quick_enum(n, changed))
    assert np.array_equal(d, o.swap_deltas(pi, uv))
    dense = orc.DenseRef(n, c1["male"], c1["female"])
    dr, uvr = dense.quick(pi, changed)
    assert np.array_equal(uvr, uv)
    assert np.array_equal(dr.astype(np.int64), d)


def test_rot3_random_triples(c1, orc):
    ctx, o, pi, n = c1["ctx"], c1["o"], c1["pi0"], c1["n"]
    rng = np.random.default_rng(3)
    abc = np.stack([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(3000)]).astype(np.int32)
    best, perm = ctx.rot3_deltas(abc)
    rb, rp = o.rot3_many(pi, abc)
    assert np.array_equal(best, rb)
    assert np.array_equal(perm.astype(np.int32), rp)
    assert (best > 0).any()


def test_rot3_changed_enumeration(c1, orc):
    ctx, pi, n = c1["ctx"], c1["pi0"], c1["n"]
    rng = np.random.default_rng(4)
    changed = rng.choice(np.arange(1, n + 1), 6, replace=False).astype(np.int32)
    best, q = ctx.rot3_deltas_changed(changed)
    assert np.array_equal(q[:, :3], orc.three_enum(n, changed))
    dense = orc.DenseRef(n, c1["male"], c1["female"])
    dr, t4 = dense.three_launch(pi, changed)
    assert np.array_equal(t4[:, :3], q[:, :3])
    assert np.array_equal(dr.astype(np.int64), best)
    assert np.array_equal(t4[:, 3], q[:, 3])


def test_apply_perm3(c1, orc):
    ctx, o = c1["ctx"], c1["o"]
    pi = ctx.get_mapping()
    a, b, c = 11, 222, 1333
    d5 = o.rot3_all(pi, a, b, c)
    s0 = ctx.score()
    for p in range(1, 6):
        ctx.apply_perm3(a, b, c, p)
        assert ctx.score() == s0 + d5[p - 1]
        ctx.set_mapping(pi)


def test_stats_and_pick(c1, orc):
    ctx = c1["ctx"]
    k = 50000
    b = ctx.batch(k)
    b.gen_pairs(31, k)
    b.eval_swaps()
    d = b.deltas()
    st = b.stats()
    assert st["count"] == k and st["sum"] == int(d.sum()) and st["max"] == int(d.max())
    assert st["argmax"] == int(np.argmax(d))
    assert st["n_pos"] == int((d > 0).sum()) and st["n_zero"] == int((d == 0).sum())
    pos = np.nonzero(d > 0)[0]
    zer = np.nonzero(d == 0)[0]
    for r in [0, 1, len(pos) // 2, len(pos) - 1]:
        assert b.pick(1, r) == pos[r]
    for r in [0, len(zer) - 1]:
        assert b.pick(0, r) == zer[r]
    assert b.pick(1, len(pos)) == -1
    b.close()


def test_quirky_graph_exact(bm, orc):
    male, female, pi = quirky_pair(bm)
    n = 60
    o = orc.Oracle(n, male, female)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi)
        assert ctx.score() == o.score(pi)
        uv = np.array([(u, v) for u in range(1, n + 1) for v in range(1, n + 1)], np.int32)
        d = ctx.swap_deltas(uv)
        brute = np.array([o.swap_delta_brute(pi, int(u), int(v)) for u, v in uv])
        assert np.array_equal(d, brute)
        rng = np.random.default_rng(1)
        abc = np.stack([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(400)]).astype(np.int32)
        best, perm = ctx.rot3_deltas(abc)
        for i, (a, b_, c) in enumerate(abc):
            d5 = o.rot3_all(pi, int(a), int(b_), int(c), brute=True)
            exp = max(0, d5.max())
            assert best[i] == exp
            assert perm[i] == (0 if exp == 0 else int(np.argmax(d5)) + 1)


def test_global_table_path_matches(c1, bm, orc):
    """Tiny LDS slices force every female table through global memory."""
    ctx, o, pi, n = c1["ctx"], c1["o"], c1["pi0"], c1["n"]
    saved = {k: ctx.get_tuning(k) for k in (bm.KERNEL_SWAP, bm.KERNEL_ROT3, bm.KERNEL_SCORE)}
    try:
        for k in saved:
            ctx.set_tuning(k, 4, 4, 1)
        uv = orc.uniform_pairs(6, 8192, n)
        assert np.array_equal(ctx.swap_deltas(uv), o.swap_deltas(pi, uv))
        assert ctx.score() == o.score(pi)
        rng = np.random.default_rng(8)
        abc = np.stack([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(500)]).astype(np.int32)
        best, perm = ctx.rot3_deltas(abc)
        rb, rp = o.rot3_many(pi, abc)
        assert np.array_equal(best, rb) and np.array_equal(perm.astype(np.int32), rp)
    finally:
        for k, t in saved.items():
            ctx.set_tuning(k, t["waves_per_block"], t["lds_slots"], t["blocks_per_cu"], t["group_lanes"], t["unroll"])


def test_accept_disjoint_matches_oracle(c1, orc):
    ctx, o, n = c1["ctx"], c1["o"], c1["n"]
    pi = c1["pi0"]
    ctx.set_mapping(pi)
    uv = orc.uniform_pairs(41, 20000, n)
    d = ctx.swap_deltas(uv)
    acc, gain = ctx.accept_disjoint(uv, d)
    pi_o, acc_o, gain_o = o.accept_disjoint(pi, uv, d)
    assert np.array_equal(acc.astype(np.int64), acc_o)
    assert gain == gain_o and gain > 0
    newpi = ctx.get_mapping()
    assert np.array_equal(newpi, pi_o)
    assert o.score(newpi) - o.score(pi) == gain
    assert ctx.score() == o.score(pi) + gain
    ctx.set_mapping(pi)


def test_hillclimb_replay(c1, orc):
    ctx, o, pi = c1["ctx"], c1["o"], c1["pi0"]
    ctx.set_mapping(pi)
    trace, res = ctx.hillclimb(seed=1000, rounds=25, batch=4096, verify=True)
    pi_o, trace_o = o.hillclimb(pi, 1000, 25, 4096)
    assert res["verify_failures"] == 0
    assert np.array_equal(trace, trace_o)
    assert np.array_equal(ctx.get_mapping(), pi_o)
    assert trace[-1] > trace[0]
    ctx.set_mapping(pi)


def test_search_drivers_keep_score_exact(c1, bm, orc):
    ctx, o, pi = c1["ctx"], c1["o"], c1["pi0"]
    for mode in (bm.SEARCH_SWAP_UNIFORM, bm.SEARCH_ROT3, bm.SEARCH_SWAP_WEIGHTED):
        ctx.set_mapping(pi)
        best, res = ctx.search(mode, seed=5, max_iters=60, verify_every=1)
        assert res["verify_failures"] == 0
        assert o.score(best) == res["best_score"]
        assert res["best_score"] >= res["initial_score"]
        assert o.score(ctx.get_mapping()) == res["final_score"]
        # the best matching is a permutation of the initial one
        assert np.array_equal(np.sort(best[1:]), np.sort(pi[1:]))
    ctx.set_mapping(pi)


def test_full_size_sampled(bm, orc):
    """n=18,524, ~4M edges: 2^20 device candidates, 4096 checked on the CPU."""
    male, female, truth, pi0 = bm.synth_pair(18524, 4e6 / 18524, bm.WEIGHTS_LOGNORMAL, 0.3, seed=2)
    n = 18524
    o = orc.Oracle(n, male, female)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        assert ctx.score() == o.score(pi0)
        k = 1 << 20
        b = ctx.batch(k)
        b.eval_swaps_uniform(99, k)
        d = b.deltas()
        uv = orc.uniform_pairs(99, k, n)
        idx = np.random.default_rng(0).choice(k, 4096, replace=False)
        assert np.array_equal(d[idx], o.swap_deltas(pi0, uv[idx]))
        # symmetry over the whole batch
        b2 = ctx.batch(k)
        b2.set_pairs(uv[:, ::-1].copy())
        b2.eval_swaps()
        assert np.array_equal(b2.deltas(), d)
        rng = np.random.default_rng(5)
        abc = np.stack([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(1024)]).astype(np.int32)
        best, perm = ctx.rot3_deltas(abc)
        rb, rp = o.rot3_many(pi0, abc)
        assert np.array_equal(best, rb) and np.array_equal(perm.astype(np.int32), rp)
        b.close()
        b2.close()


@pytest.mark.parametrize("case", ["male_empty", "female_empty", "both_empty", "n2", "unmatched"])
def test_degenerate_inputs(bm, orc, case):
    """Graphs without edges, the smallest vertex count for a swap (n = 2)
    and a matching with every vertex unmatched: score, every pair's swap
    delta and (n >= 3) 3-vertex deltas equal the oracle's on both the scan
    and the table path, and the u-major stream's table deltas too."""
    rng = np.random.default_rng(17)
    n = 2 if case == "n2" else 40

    def graph(m):
        return bm.Graph(rng.integers(1, n + 1, m).astype(np.int32), rng.integers(1, n + 1, m).astype(np.int32),
                        rng.integers(1, 9, m).astype(np.int32), n)

    empty = bm.Graph(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32), n)
    male = empty if case in ("male_empty", "both_empty") else graph(3 * n)
    female = empty if case in ("female_empty", "both_empty") else graph(3 * n)
    pi = np.concatenate([[0], rng.permutation(np.arange(1, n + 1))]).astype(np.int32)
    if case == "unmatched":
        pi[:] = 0
    uv = np.array([(u, v) for u in range(1, n + 1) for v in range(1, n + 1) if u != v], np.int32)
    abc = np.stack([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(300)]).astype(np.int32) \
        if n >= 3 else None
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi)
        o = orc.Oracle(n, male, female)
        assert ctx.score() == o.score(pi)
        ref = o.swap_deltas(pi, uv)
        if abc is not None:
            rb, rp = o.rot3_many(pi, abc)
        for gain in (False, True):
            ctx.gain_enable(gain)
            assert np.array_equal(ctx.swap_deltas(uv), ref), (case, gain)
            if abc is not None:
                best, perm = ctx.rot3_deltas(abc)
                assert np.array_equal(best, rb) and np.array_equal(perm.astype(np.int32), rp), (case, gain)
            k = 4096
            b = ctx.batch(k)
            b.eval_swaps_umajor(3, k)
            assert np.array_equal(b.deltas(), o.swap_deltas(pi, orc.umajor_pairs(3, k, n))), (case, gain)
            b.close()
        if case in ("male_empty", "both_empty", "female_empty", "unmatched"):
            assert not ref.any()


def test_extreme_weights(bm, orc):
    """Weights at the ends of int32 (male and female, negative too): the scan
    kernels' int64 sums stay exact, and the table refuses a graph whose
    entries could overflow int32."""
    rng = np.random.default_rng(23)
    n, m = 30, 200
    big = np.array([2**31 - 1, -(2**31) + 1, 2**30, -(2**30), 1, -1, 0, 32768], np.int64)

    def graph():
        w = rng.choice(big, m).astype(np.int32)
        return bm.Graph(rng.integers(1, n + 1, m).astype(np.int32), rng.integers(1, n + 1, m).astype(np.int32), w, n)

    male, female = graph(), graph()
    pi = np.concatenate([[0], rng.permutation(np.arange(1, n + 1))]).astype(np.int32)
    uv = np.array([(u, v) for u in range(1, n + 1) for v in range(1, n + 1) if u != v], np.int32)
    abc = np.stack([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(500)]).astype(np.int32)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi)
        o = orc.Oracle(n, male, female)
        assert ctx.score() == o.score(pi)
        d = ctx.swap_deltas(uv)
        assert np.array_equal(d, o.swap_deltas(pi, uv))
        assert np.abs(d).max() > 2**31   # the sums leave int32
        best, perm = ctx.rot3_deltas(abc)
        rb, rp = o.rot3_many(pi, abc)
        assert np.array_equal(best, rb) and np.array_equal(perm.astype(np.int32), rp)
        with pytest.raises(bm.BMError) as e:
            ctx.gain_enable(True)
        assert e.value.code == -1


def test_hillclimb_ten_thousand_rounds(c1, orc):
    """The CPU-runnable pair's climbing workload: 10,000 rounds of greedy
    disjoint-swap climbing; the running score is checked against a device
    recount after every round, and the final matching against the oracle's
    score-matching sum."""
    ctx, o, pi = c1["ctx"], c1["o"], c1["pi0"]
    ctx.set_mapping(pi)
    trace, res = ctx.hillclimb(seed=77, rounds=10000, batch=4096, verify=True)
    assert res["verify_failures"] == 0
    assert len(trace) == 10001 and (np.diff(trace) >= 0).all()
    final = ctx.get_mapping()
    assert o.score(final) == trace[-1] == ctx.score()
    assert trace[-1] > trace[0]
    ctx.set_mapping(pi)

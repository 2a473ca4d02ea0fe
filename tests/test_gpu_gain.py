"""Device parity of the placement-gain evaluation path: table rows against
their definition, deltas against the oracle, and the incremental patches
against a fresh rebuild after every kind of move."""
import numpy as np
import pytest

from conftest import quirky_pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g1(bm, orc):
    male, female, truth, pi0 = bm.synth_pair(2000, 20.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=1)
    ctx = bm.Context(2000, male, female)
    ctx.set_mapping(pi0)
    ctx.gain_enable(True)
    o = orc.Oracle(2000, male, female)
    yield dict(male=male, female=female, pi0=pi0, ctx=ctx, o=o, n=2000)
    ctx.close()


def test_rows_match_definition(g1, orc):
    ctx, o, pi = g1["ctx"], g1["o"], g1["pi0"]
    assert ctx.gain_enabled()
    rng = np.random.default_rng(0)
    for u in [1, 2, 1999, 2000] + list(rng.integers(1, 2001, 12)):
        assert np.array_equal(ctx.gain_row(int(u)), orc.gain_row(o, pi, int(u)))
    assert not ctx.gain_row(0).any()


def test_swap_deltas_match_oracle(g1, orc):
    ctx, o, pi, n = g1["ctx"], g1["o"], g1["pi0"], g1["n"]
    uv = orc.uniform_pairs(5, 1 << 16, n)
    assert np.array_equal(ctx.swap_deltas(uv), o.swap_deltas(pi, uv))
    b = ctx.batch(1 << 15)
    b.eval_swaps_uniform(123, 1 << 15)
    assert np.array_equal(b.deltas(), o.swap_deltas(pi, orc.uniform_pairs(123, 1 << 15, n)))
    b.close()
    assert ctx.swap_deltas(np.array([[3, 3]]))[0] == 0


def test_rot3_matches_oracle(g1, orc):
    ctx, o, pi, n = g1["ctx"], g1["o"], g1["pi0"], g1["n"]
    rng = np.random.default_rng(3)
    abc = np.stack([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(3000)]).astype(np.int32)
    best, perm = ctx.rot3_deltas(abc)
    rb, rp = o.rot3_many(pi, abc)
    assert np.array_equal(best, rb) and np.array_equal(perm.astype(np.int32), rp)


def test_changed_enumerations_match_scan_path(g1, orc):
    ctx, n = g1["ctx"], g1["n"]
    changed = np.array([17, 400, 1999, 3, 1200], np.int32)
    d_gain, uv = ctx.swap_deltas_changed(changed)
    b_gain, q = ctx.rot3_deltas_changed(changed)
    ctx.gain_enable(False)
    try:
        d_scan, uv2 = ctx.swap_deltas_changed(changed)
        b_scan, q2 = ctx.rot3_deltas_changed(changed)
    finally:
        ctx.gain_enable(True)
    assert np.array_equal(uv, uv2) and np.array_equal(d_gain, d_scan)
    assert np.array_equal(q, q2) and np.array_equal(b_gain, b_scan)


def test_patches_equal_rebuild(g1, orc):
    """apply_swap, apply_perm3 and disjoint acceptance patch the table; a full
    rebuild of the resulting matching must give the same rows."""
    ctx, o, n = g1["ctx"], g1["o"], g1["n"]
    pi0 = g1["pi0"]
    ctx.set_mapping(pi0)
    ctx.apply_swap(10, 20)
    ctx.apply_swap(20, 1500)
    ctx.apply_perm3(7, 8, 9, 4)
    ctx.apply_perm3(30, 10, 1999, 1)
    pi_before = ctx.get_mapping()
    uv = orc.uniform_pairs(41, 20000, n)
    d = ctx.swap_deltas(uv)
    acc, gain = ctx.accept_disjoint(uv, d)
    # the graph has hubs adjacent to most vertices, so a uniform batch may
    # lock everything after one swap; pairs of low-degree vertices give a
    # batch with many disjoint acceptances
    od, idg = ctx.degrees()
    deg = (od + idg)[1:]
    low = np.nonzero(deg <= np.percentile(deg, 30))[0].astype(np.int32) + 1
    rng = np.random.default_rng(4)
    uv2 = np.stack([rng.choice(low, 2, replace=False) for _ in range(4000)]).astype(np.int32)
    pi_mid = ctx.get_mapping()
    d2 = ctx.swap_deltas(uv2)
    acc2, gain2 = ctx.accept_disjoint(uv2, d2)
    assert gain > 0 and gain2 > 0 and len(acc2) > 10
    ref_pi, ref_acc, ref_gain = o.accept_disjoint(pi_before, uv, d)
    assert ref_gain == gain and np.array_equal(ref_acc, acc) and np.array_equal(ref_pi, pi_mid)
    ref_pi2, ref_acc2, ref_gain2 = o.accept_disjoint(pi_mid, uv2, d2)
    pi = ctx.get_mapping()
    assert ref_gain2 == gain2 and np.array_equal(ref_acc2, acc2) and np.array_equal(ref_pi2, pi)
    rows = list(range(1, n + 1, 37)) + [10, 20, 1500, 7, 8, 9, 30, 1999]
    patched = {u: ctx.gain_row(u) for u in rows}
    ctx.gain_rebuild()
    for u in rows:
        assert np.array_equal(patched[u], ctx.gain_row(u)), u
        assert np.array_equal(patched[u], orc.gain_row(o, pi, u)), u
    uv2 = orc.uniform_pairs(42, 8192, n)
    assert np.array_equal(ctx.swap_deltas(uv2), o.swap_deltas(pi, uv2))
    ctx.set_mapping(pi0)


def test_transposed_copy_follows_patches(g1, orc):
    """u-major lookups read the transposed table.  Swaps and 3-permutations
    of ordinary vertices applied while it is current patch it alongside the
    table; a hub's swap and a large acceptance batch leave it to be rebuilt.
    After each, u-major deltas and best partners equal the oracle."""
    ctx, o, n = g1["ctx"], g1["o"], g1["n"]
    pi0 = g1["pi0"]
    ctx.set_mapping(pi0)
    k = 1 << 15
    b = ctx.batch(k)

    def check(seed):
        pi = ctx.get_mapping()
        b.eval_swaps_umajor(seed, k)
        assert np.array_equal(b.deltas(), o.swap_deltas(pi, orc.umajor_pairs(seed, k, n))), seed
        d, v = ctx.best_partners(1, 40)
        assert [(int(x), int(y)) for x, y in zip(d, v)] == [orc.best_partner(o, pi, u) for u in range(1, 41)]

    od, idg = ctx.degrees()
    deg = od + idg
    order = np.argsort(deg[1:]) + 1
    low = [int(x) for x in order[:6]]          # ordinary vertices: the copy is patched
    hub = int(order[-1])                       # a hub: the copy is rebuilt
    check(1)
    ctx.apply_swap(low[0], low[1])
    check(2)
    ctx.apply_perm3(low[2], low[3], low[4], 3)
    ctx.apply_swap(low[4], low[5])
    check(3)
    ctx.apply_swap(hub, low[0])
    check(5)
    uv = orc.uniform_pairs(43, 20000, n)
    ctx.accept_disjoint(uv, ctx.swap_deltas(uv))
    check(4)
    b.close()
    ctx.set_mapping(pi0)


def test_hillclimb_with_gain_matches_replay(g1, orc):
    ctx, o, pi = g1["ctx"], g1["o"], g1["pi0"]
    ctx.set_mapping(pi)
    trace, res = ctx.hillclimb(seed=1000, rounds=25, batch=4096, verify=True)
    pi_o, trace_o = o.hillclimb(pi, 1000, 25, 4096)
    assert res["verify_failures"] == 0
    assert np.array_equal(trace, trace_o)
    assert np.array_equal(ctx.get_mapping(), pi_o)
    ctx.set_mapping(pi)


def test_search_drivers_with_gain(g1, bm, orc):
    ctx, o, pi = g1["ctx"], g1["o"], g1["pi0"]
    for mode in (bm.SEARCH_SWAP_UNIFORM, bm.SEARCH_ROT3):
        ctx.set_mapping(pi)
        best, res = ctx.search(mode, seed=9, max_iters=80, verify_every=1)
        assert res["verify_failures"] == 0
        assert o.score(best) == res["best_score"] >= res["initial_score"]
    ctx.set_mapping(pi)


def test_quirky_graph_gain_exact(bm, orc):
    male, female, pi = quirky_pair(bm, seed=11)
    n = 60
    o = orc.Oracle(n, male, female)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi)
        ctx.gain_enable(True)
        for u in range(1, n + 1):
            assert np.array_equal(ctx.gain_row(u), orc.gain_row(o, pi, u))
        uv = np.array([(u, v) for u in range(1, n + 1) for v in range(1, n + 1)], np.int32)
        brute = np.array([o.swap_delta_brute(pi, int(u), int(v)) for u, v in uv])
        assert np.array_equal(ctx.swap_deltas(uv), brute)
        rng = np.random.default_rng(2)
        abc = np.stack([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(400)]).astype(np.int32)
        best, perm = ctx.rot3_deltas(abc)
        for i, (a, b_, c) in enumerate(abc):
            d5 = o.rot3_all(pi, int(a), int(b_), int(c), brute=True)
            assert best[i] == max(0, d5.max())
        # patches on a graph with self-loops and u<->v edges
        ctx.apply_swap(1, 2)
        ctx.apply_perm3(3, 4, 5, 3)
        q = ctx.get_mapping()
        for u in range(1, n + 1):
            assert np.array_equal(ctx.gain_row(u), orc.gain_row(o, q, u))


def test_full_size_gain_equals_scan(bm, orc):
    """n=18,524, ~4M edges: the table path and the scan path agree on 2^20
    device candidates; 4096 of them are checked against the oracle."""
    n = 18524
    male, female, truth, pi0 = bm.synth_pair(n, 4e6 / n, bm.WEIGHTS_LOGNORMAL, 0.3, seed=2)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        k = 1 << 20
        b = ctx.batch(k)
        b.eval_swaps_uniform(99, k)
        d_scan = b.deltas()
        ctx.gain_enable(True)
        b.eval_swaps_uniform(99, k)
        d_gain = b.deltas()
        assert np.array_equal(d_scan, d_gain)
        o = orc.Oracle(n, male, female)
        uv = orc.uniform_pairs(99, k, n)
        idx = np.random.default_rng(1).choice(k, 4096, replace=False)
        assert np.array_equal(d_gain[idx], o.swap_deltas(pi0, uv[idx]))
        b.close()


def test_multi_chunk_build_equals_scan(bm, orc):
    """n=45,000: the table is built in two column windows (each row's
    female entries split at precomputed window boundaries) on top of the
    dense tile pass's sums for the pair's four hub rows, unpacked per window;
    rows match their definition,
    table deltas match the scan kernels on uniform and u-major streams (at
    this n the u-major lookups stage only the first vertex's bitmaps and
    gather both table entries), and a
    rebuild after moves matches the patched table."""
    n = 45000
    male, female, truth, pi0 = bm.synth_pair(n, 12.0, bm.WEIGHTS_LOGNORMAL, 0.3, seed=5)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        k = 1 << 18
        b = ctx.batch(k)
        b.eval_swaps_uniform(7, k)
        d_scan = b.deltas()
        b.eval_swaps_umajor(8, k)
        d_scan_um = b.deltas()
        ctx.gain_enable(True)
        b.eval_swaps_uniform(7, k)
        assert np.array_equal(b.deltas(), d_scan)
        b.eval_swaps_umajor(8, k)
        assert np.array_equal(b.deltas(), d_scan_um)
        assert ctx.gain_dense_rows() == 4
        o = orc.Oracle(n, male, female)
        hubs = np.argsort(-(np.bincount(male.src, minlength=n + 1) + np.bincount(male.dst, minlength=n + 1)))[:3]
        for u in [1, 15007, 15008, 30015, 30016, n] + [int(x) for x in hubs if x > 0]:
            assert np.array_equal(ctx.gain_row(u), orc.gain_row(o, pi0, u)), u
        ctx.apply_swap(11, 40000)
        ctx.apply_perm3(12, 29999, 44999, 4)
        patched = [ctx.gain_row(u) for u in (1, 11, 12, 40000, 44999)]
        ctx.gain_rebuild()
        for u, row in zip((1, 11, 12, 40000, 44999), patched):
            assert np.array_equal(ctx.gain_row(u), row), u
        b.close()


def test_scale_pair_windowed_build(bm, orc):
    """n=100,000, ~20M edges per graph, 40% of the matching shuffled (the
    scale stress pair): the 40 GB table is built in three column windows at
    one block per CU, on top of the dense tile pass's sums for 57 hub rows;
    u-major table deltas (bitmaps staged per first vertex) equal the scan
    kernel on 2^22 pairs.  Whole rows (every column, so every window) of five
    vertices up to degree 1,500 match their definition; the two largest hubs
    (adjacent to nearly every vertex, rows too costly for the oracle's
    O(n * degree) row) are checked through 2048 swap deltas each against the
    oracle; table deltas equal the scan kernels on 2^16 uniform pairs and the
    oracle on 4096 of them; 3-vertex deltas match the oracle; ten rounds of
    disjoint-swap climbing with the patched table replay exactly on the CPU;
    and a rebuild after the climb equals the patched table."""
    n = 100_000
    male, female, truth, pi0 = bm.synth_pair(n, 200.0, bm.WEIGHTS_LOGNORMAL, 0.4, seed=5)
    print(f"scale pair n={n}, {male.m} male edges", flush=True)
    o = orc.Oracle(n, male, female)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        k = 1 << 16
        b = ctx.batch(k)
        b.eval_swaps_uniform(31, k)
        d_scan = b.deltas()
        ku = 1 << 22
        bu = ctx.batch(ku)
        bu.eval_swaps_umajor(32, ku)
        d_scan_um = bu.deltas()
        ctx.gain_enable(True)
        assert ctx.gain_enabled() and ctx.gain_dense_rows() == 57
        b.eval_swaps_uniform(31, k)
        d_gain = b.deltas()
        assert np.array_equal(d_gain, d_scan)
        bu.eval_swaps_umajor(32, ku)
        assert np.array_equal(bu.deltas(), d_scan_um)
        bu.close()
        print("u-major table lookups (bitmaps staged, entries gathered) equal the scan kernel on 2^22 pairs",
              flush=True)
        uv = orc.uniform_pairs(31, k, n)
        rng = np.random.default_rng(2)
        idx = rng.choice(k, 4096, replace=False)
        assert np.array_equal(d_gain[idx], o.swap_deltas(pi0, uv[idx]))
        print("table equals scan and oracle on uniform pairs", flush=True)
        deg = np.bincount(male.src, minlength=n + 1) + np.bincount(male.dst, minlength=n + 1)
        hubs = [int(x) for x in np.argsort(-deg)[:2] if x > 0]
        bounded = np.nonzero(deg <= 1500)[0]
        heavy = int(bounded[np.argmax(deg[bounded])])
        for u in [1, 2, 50000, n, heavy]:
            assert np.array_equal(ctx.gain_row(u), orc.gain_row(o, pi0, u)), u
        print(f"rows match the definition (vertex {heavy} of degree {deg[heavy]})", flush=True)
        for h in hubs:
            uvh = np.stack([np.full(2048, h), rng.integers(1, n + 1, 2048)], 1).astype(np.int32)
            assert np.array_equal(ctx.swap_deltas(uvh), o.swap_deltas(pi0, uvh)), h
        print(f"hub deltas match (degrees {[int(deg[h]) for h in hubs]})", flush=True)
        abc = np.stack([rng.choice(np.arange(1, n + 1), 3, replace=False) for _ in range(2000)]).astype(np.int32)
        best, perm = ctx.rot3_deltas(abc)
        rb, rp = o.rot3_many(pi0, abc)
        assert np.array_equal(best, rb) and np.array_equal(perm.astype(np.int32), rp)
        print("3-vertex deltas match", flush=True)
        trace, res = ctx.hillclimb(seed=77, rounds=10, batch=4096, verify=True)
        pi_o, trace_o = o.hillclimb(pi0, 77, 10, 4096)
        assert res["verify_failures"] == 0
        assert np.array_equal(trace, trace_o)
        pi1 = ctx.get_mapping()
        assert np.array_equal(pi1, pi_o)
        print(f"climb replays on the CPU ({trace[0]} -> {trace[-1]})", flush=True)
        moved = [int(x) for x in np.nonzero(pi1 != pi0)[0][:4]]
        rows = moved + hubs
        patched = [ctx.gain_row(u) for u in rows]
        ctx.gain_rebuild()
        for u, row in zip(rows, patched):
            assert np.array_equal(ctx.gain_row(u), row), u
        b.close()


@pytest.mark.parametrize("wmax", [8191, 8192, 32767, 32768])
def test_row_layouts_at_weight_boundaries(bm, orc, wmax):
    """The table build reads female rows packed (w << 19) | 4f when female
    weights fit 0..8191, (f << 15) | w when they fit 0..32767, and as {f, w}
    pairs otherwise: rows and u-major deltas match the oracle on both sides of
    each boundary, with the largest weight present in both graphs."""
    n = 400
    male, female, truth, pi0 = bm.synth_pair(n, 10.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=21)
    rng = np.random.default_rng(wmax)
    female.w = rng.integers(1, 300, female.m).astype(np.int32)
    female.w[rng.choice(female.m, 25, replace=False)] = wmax
    male.w = rng.integers(1, 300, male.m).astype(np.int32)
    male.w[rng.choice(male.m, 25, replace=False)] = wmax + 5
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        ctx.gain_enable(True)
        o = orc.Oracle(n, male, female)
        for u in range(1, n + 1):
            assert np.array_equal(ctx.gain_row(u), orc.gain_row(o, pi0, u)), (wmax, u)
        k = 1 << 14
        b = ctx.batch(k)
        b.eval_swaps_umajor(3, k)
        assert np.array_equal(b.deltas(), o.swap_deltas(pi0, orc.umajor_pairs(3, k, n)))
        b.close()

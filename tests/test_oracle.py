"""CPU checks that pin the oracle: brute force, closed forms, the reference's
own compiled code (oracle/_ref) and the committed golden vectors.  No GPU."""
import glob
import os

import numpy as np
import pytest

from conftest import make_pair, quirky_pair

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_local_swap_delta_equals_brute_force(bm, orc):
    male, female, pi = quirky_pair(bm)
    o = orc.Oracle(60, male, female)
    for u in range(1, 61):
        for v in range(1, 61):
            assert o.swap_delta(pi, u, v) == o.swap_delta_brute(pi, u, v)


def test_swap_delta_antisymmetry_and_symmetry(bm, orc):
    male, female, truth, pi0 = make_pair(bm)
    o = orc.Oracle(300, male, female)
    rng = np.random.default_rng(1)
    for _ in range(200):
        u, v = (int(x) for x in rng.choice(np.arange(1, 301), 2, replace=False))
        d = o.swap_delta(pi0, u, v)
        assert d == o.swap_delta(pi0, v, u)
        q = pi0.copy()
        q[u], q[v] = q[v], q[u]
        assert o.swap_delta(q, u, v) == -d
        assert o.score(q) - o.score(pi0) == d


def test_rot3_local_equals_brute_force(bm, orc):
    male, female, pi = quirky_pair(bm, seed=9)
    o = orc.Oracle(60, male, female)
    rng = np.random.default_rng(2)
    for _ in range(300):
        a, b, c = (int(x) for x in rng.choice(np.arange(1, 61), 3, replace=False))
        assert np.array_equal(o.rot3_all(pi, a, b, c), o.rot3_all(pi, a, b, c, brute=True))


def test_rot3_cycle_cubed_is_identity(bm, orc):
    """p=3 and p=4 are the two 3-cycles: applying one three times restores pi,
    and the three deltas along the way sum to zero."""
    male, female, truth, pi0 = make_pair(bm, seed=4)
    o = orc.Oracle(300, male, female)
    a, b, c = 10, 20, 30
    q = pi0.copy()
    total = 0
    for _ in range(3):
        total += o.rot3_all(q, a, b, c)[2]
        f = (q[a], q[b], q[c])
        q[a], q[b], q[c] = f[1], f[2], f[0]
    assert np.array_equal(q, pi0) and total == 0


def test_score_closed_forms(bm, orc):
    male, female, truth, pi0 = make_pair(bm, seed=6)
    n = 300
    ident = np.arange(n + 1, dtype=np.int32)
    assert orc.Oracle(n, male, male).score(ident) == int(male.w.sum())
    assert orc.Oracle(n, male, female).score(np.zeros(n + 1, np.int32)) == 0
    empty = bm.Graph(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32), n)
    assert orc.Oracle(n, empty, female).score(pi0) == 0
    # unrelabelled female with the true permutation: every unmodified edge scores fully
    o = orc.Oracle(n, male, female)
    assert o.score(truth) > o.score(pi0)


def test_duplicate_lines_last_weight_wins(bm, orc):
    m = bm.Graph(np.array([1, 1, 2], np.int32), np.array([2, 2, 3], np.int32), np.array([5, 2, 4], np.int32), 3)
    f = bm.Graph(np.array([1, 2, 2], np.int32), np.array([2, 3, 3], np.int32), np.array([9, 1, 7], np.int32), 3)
    o = orc.Oracle(3, m, f)
    assert o.num_edges() == 2 and o.num_edges(female=True) == 2
    assert o.score(np.array([0, 1, 2, 3], np.int32)) == min(2, 9) + min(4, 7)


def test_handmade_known_answer(bm, orc):
    """M = {1->2:3, 2->3:5, 3->1:2, 1->3:4}, F = {1->2:1, 2->3:7, 3->1:2, 2->1:4}:
    identity scores 1+5+2+0 = 8; swapping 1 and 2 scores 7 (delta -1)."""
    m = bm.Graph(np.array([1, 2, 3, 1], np.int32), np.array([2, 3, 1, 3], np.int32), np.array([3, 5, 2, 4], np.int32), 3)
    f = bm.Graph(np.array([1, 2, 3, 2], np.int32), np.array([2, 3, 1, 1], np.int32), np.array([1, 7, 2, 4], np.int32), 3)
    o = orc.Oracle(3, m, f)
    ident = np.array([0, 1, 2, 3], np.int32)
    assert o.score(ident) == 8
    assert o.swap_delta(ident, 1, 2) == -1
    assert o.score(np.array([0, 2, 1, 3], np.int32)) == 7


def test_enumerations_are_complete_and_ordered(orc):
    n = 40
    changed = np.array([17, 3, 25, 9], np.int32)
    uv = orc.quick_enum(n, changed)
    c = len(changed)
    assert uv.shape[0] == c * (c - 1) // 2 + c * (n - c)
    s = sorted(changed)
    assert [tuple(x) for x in uv[:6]] == [(s[0], s[1]), (s[0], s[2]), (s[0], s[3]), (s[1], s[2]), (s[1], s[3]),
                                           (s[2], s[3])]
    rest = [x for x in range(1, n + 1) if x not in s]
    assert list(uv[6:6 + len(rest), 1]) == rest and (uv[6:6 + len(rest), 0] == s[0]).all()
    abc = orc.three_enum(n, changed)
    T = 4 + 6 * (n - c) + c * (n - c - 1)
    assert abc.shape[0] == T
    assert len({tuple(sorted(t)) for t in abc[:4]}) == 4
    # third section: changed vertex with two adjacent unchanged ones
    third = abc[4 + 6 * (n - c):]
    assert (np.searchsorted(rest, third[:, 2]) == np.searchsorted(rest, third[:, 1]) + 1).all()


def test_dense_restatement_matches_exact_without_quirks(bm, orc):
    male, female, truth, pi0 = make_pair(bm, seed=8)
    n = 300
    o, D = orc.Oracle(n, male, female), orc.DenseRef(n, male, female)
    rng = np.random.default_rng(3)
    for _ in range(300):
        u, v = (int(x) for x in rng.choice(np.arange(1, n + 1), 2, replace=False))
        assert D.swap(pi0, u, v) == o.swap_delta(pi0, u, v)
    for _ in range(100):
        a, b, c = (int(x) for x in rng.choice(np.arange(1, n + 1), 3, replace=False))
        best, p = D.three(pi0, a, b, c)
        d5 = o.rot3_all(pi0, a, b, c)
        assert best == max(0, d5.max())
        assert p == (0 if best == 0 else int(np.argmax(d5)) + 1)


def test_dense_restatement_skips_self_loops(bm, orc):
    """The reference kernels skip i == m1 (src/graph-re-alignment-gpu-swap.cu:199),
    so a self-loop's change is missing from their delta; the exact delta has it."""
    m = bm.Graph(np.array([1, 1], np.int32), np.array([1, 2], np.int32), np.array([5, 1], np.int32), 3)
    f = bm.Graph(np.array([2, 1], np.int32), np.array([2, 2], np.int32), np.array([5, 1], np.int32), 3)
    pi = np.array([0, 1, 2, 3], np.int32)
    o, D = orc.Oracle(3, m, f), orc.DenseRef(3, m, f)
    exact = o.swap_delta_brute(pi, 1, 2)
    assert o.swap_delta(pi, 1, 2) == exact == 4   # self-loop 1->1 now maps onto f2->f2 (w 5): +5, edge 1->2: -1
    assert D.swap(pi, 1, 2) == -1                  # the reference kernel sees only the 1->2 edge


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "libref_gao.so")),
                    reason="oracle/_ref not built (reference sources absent)")
def test_oracle_matches_compiled_reference(bm, orc):
    male, female, truth, pi0 = make_pair(bm, n=500, deg=15.0, seed=12)
    ref = orc.RefCode(male, female)
    o = orc.Oracle(500, male, female)
    uv = orc.uniform_pairs(3, 20000, 500)
    assert np.array_equal(ref.deltas(pi0, uv).astype(np.int64), o.swap_deltas(pi0, uv))
    assert ref.score(pi0) == o.score(pi0)
    lists = orc.ListsRef(500, male, female)
    assert np.array_equal(lists.deltas(pi0, uv), ref.deltas(pi0, uv))
    ref.close()


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "libref_gao.so")),
                    reason="oracle/_ref not built (reference sources absent)")
def test_lists_restatement_matches_reference_on_quirks(bm, orc):
    """calculate_swap_delta's self-loop / duplicate behaviour is reproduced by
    the restatement (both differ from the exact delta on such graphs)."""
    male, female, pi = quirky_pair(bm, seed=5)
    male.w = np.abs(male.w) + 1
    female.w = np.abs(female.w) + 1
    ref = orc.RefCode(male, female)
    lists = orc.ListsRef(60, male, female)
    uv = np.array([(u, v) for u in range(1, 61) for v in range(1, 61) if u != v], np.int32)
    assert np.array_equal(lists.deltas(pi, uv), ref.deltas(pi, uv))
    ref.close()


def _golden_files():
    return sorted(glob.glob(os.path.join(GOLDEN, "synth_*.npz")) + glob.glob(os.path.join(GOLDEN, "quirky_*.npz")))


@pytest.mark.parametrize("path", _golden_files(), ids=lambda p: os.path.basename(p))
def test_golden_vectors(bm, orc, path):
    """Fixtures written from the reference binaries (tests/golden/make_golden.py):
    the oracle reproduces score-matching's integer, and on quirk-free pairs the
    reference's calculate_swap_delta; the lists restatement reproduces the
    reference delta on every pair (quirks included)."""
    z = np.load(path, allow_pickle=False)
    n = int(z["n"])
    male = bm.Graph(z["m_src"], z["m_dst"], z["m_w"], n)
    female = bm.Graph(z["f_src"], z["f_dst"], z["f_w"], n)
    o = orc.Oracle(n, male, female)
    pi = z["pi"]
    assert o.score(pi) == int(z["score_matching"])
    d = o.swap_deltas(pi, z["pairs"])
    if bool(z["quirk_free"]):
        assert np.array_equal(d, z["ref_deltas"].astype(np.int64))
    else:
        assert not np.array_equal(d, z["ref_deltas"].astype(np.int64))
        brute = np.array([o.swap_delta_brute(pi, int(u), int(v)) for u, v in z["pairs"][:400]])
        assert np.array_equal(d[:400], brute)
    lists = orc.ListsRef(n, male, female)
    assert np.array_equal(lists.deltas(pi, z["pairs"]), z["ref_deltas"])


def test_golden_csv_loader_matches_score_matching(bm, orc):
    """Prefixed ids, a duplicate line (last weight wins) and a malformed line:
    the library's loaders plus the oracle give score-matching's integer."""
    d = os.path.join(GOLDEN, "csv")
    male = bm.load_graph_csv(os.path.join(d, "male.csv"))
    female = bm.load_graph_csv(os.path.join(d, "female.csv"))
    n = max(male.n, female.n)
    pi = bm.load_mapping_csv(os.path.join(d, "matching.csv"), n)
    with open(os.path.join(d, "score_matching.txt")) as f:
        expected = int(f.read().strip())
    assert orc.Oracle(n, male, female).score(pi) == expected


def test_golden_greedy_match(bm):
    z = np.load(os.path.join(GOLDEN, "greedy_0.npz"), allow_pickle=False)
    n = int(z["n"])
    male = bm.Graph(z["m_src"], z["m_dst"], z["m_w"], n)
    female = bm.Graph(z["f_src"], z["f_dst"], z["f_w"], n)
    assert np.array_equal(bm.greedy_match(n, male, female), z["ref_matching"])


def test_amajor_triple_stream_law(orc):
    """Grouped by first vertex, three distinct vertices, first vertices the
    u-major first vertices of the derived seed, and (b, c) uniform over the
    (n-1)(n-2) ordered choices for a given a (exhaustively at n = 6)."""
    n, k = 6, 240000
    abc = orc.amajor_triples(5, k, n)
    assert (np.diff(abc[:, 0]) >= 0).all()
    assert ((abc[:, 0] != abc[:, 1]) & (abc[:, 0] != abc[:, 2]) & (abc[:, 1] != abc[:, 2])).all()
    assert abc.min() >= 1 and abc.max() <= n
    assert np.array_equal(abc[:, 0], orc.umajor_pairs(orc.amajor_seed(5), k, n)[:, 0])
    for a in range(1, n + 1):
        sub = abc[abc[:, 0] == a]
        h = np.bincount((sub[:, 1] - 1) * n + (sub[:, 2] - 1), minlength=n * n).reshape(n, n)
        cells = [(b, c) for b in range(n) for c in range(n) if len({a - 1, b, c}) == 3]
        e = len(sub) / len(cells)
        assert sum(h[b, c] for b, c in cells) == len(sub)
        assert max(abs(h[b, c] - e) for b, c in cells) < 6 * np.sqrt(e)
    # the same law as the uniform triple stream: marginals of a large n agree
    n2, k2 = 40, 400000
    t = orc.amajor_triples(6, k2, n2)
    for col in range(3):
        cnt = np.bincount(t[:, col], minlength=n2 + 1)[1:]
        assert np.abs(cnt - k2 / n2).max() < 6 * np.sqrt(k2 / n2)


def test_umajor_stream_law(orc):
    """Grouped by first vertex, u != v, first vertices distributed like k
    independent uniform draws, partners uniform over the other vertices."""
    n, k = 50, 200000
    uv = orc.umajor_pairs(9, k, n)
    assert (np.diff(uv[:, 0]) >= 0).all() and (uv[:, 0] != uv[:, 1]).all()
    assert uv.min() >= 1 and uv.max() <= n
    cu = np.bincount(uv[:, 0], minlength=n + 1)[1:]
    cv = np.bincount(uv[:, 1], minlength=n + 1)[1:]
    e = k / n
    assert np.abs(cu - e).max() < 6 * np.sqrt(e) and np.abs(cv - e).max() < 6 * np.sqrt(e)
    # the first vertices are the same multiset as the uniform pair stream's (same draws)
    assert np.array_equal(np.sort(uv[:, 0]), np.sort(orc.uniform_pairs(9, k, n)[:, 0]))
    # partner frequencies for one first vertex do not favour any other vertex
    sub = uv[uv[:, 0] == 7][:, 1]
    h = np.bincount(sub, minlength=n + 1)[1:]
    assert h[6] == 0 and np.abs(np.delete(h, 6) - len(sub) / (n - 1)).max() < 6 * np.sqrt(len(sub) / (n - 1))


@pytest.mark.parametrize("seed", [3, 8])
def test_best_swap_is_brute_force_argmax(bm, orc, seed):
    """or_best_swap = the argmax, over every pair u < v, of full recounts
    (ties to the smaller u, then v), on the quirky graph; with no matching
    every delta is 0 and the first pair (1, 2) is chosen."""
    male, female, pi = quirky_pair(bm, seed=seed)
    o = orc.Oracle(60, male, female)
    best = None
    for u in range(1, 61):
        for v in range(u + 1, 61):
            d = o.swap_delta_brute(pi, u, v)
            if best is None or d > best[0]:
                best = (d, u, v)
    assert o.best_swap(pi) == best
    assert o.best_swap(np.zeros(61, np.int32)) == (0, 1, 2)


def test_best_swap_climb_ends_at_local_optimum(bm, orc):
    """The steepest-ascent trace equals a recount after every swap, rises
    strictly, and stops where no pair has a positive delta."""
    male, female, pi = quirky_pair(bm, seed=5)
    o = orc.Oracle(60, male, female)
    pi_end, trace = o.best_swap_climb(pi, 10_000)
    assert len(trace) > 2 and np.all(np.diff(trace) > 0)
    assert trace[-1] == o.score(pi_end)
    assert o.best_swap(pi_end)[0] <= 0
    q = pi.copy()
    for k in range(1, len(trace)):
        d, u, v = o.best_swap(q)
        q[u], q[v] = q[v], q[u]
        assert o.score(q) == trace[k]
    assert np.array_equal(q, pi_end)


@pytest.mark.parametrize("seed", [3, 8])
def test_best_partner_is_brute_force_argmax(bm, orc, seed):
    """or_best_partner = the argmax over v > u of full recounts (ties to the
    smaller v) for every u of the quirky graph; u = n has no partner."""
    male, female, pi = quirky_pair(bm, seed=seed)
    o = orc.Oracle(60, male, female)
    for u in range(1, 60):
        want = max(((o.swap_delta_brute(pi, u, v), -v) for v in range(u + 1, 61)))
        assert orc.best_partner(o, pi, u) == (want[0], -want[1]), u
    assert orc.best_partner(o, pi, 60) == (-(1 << 63), 0)


def test_sweep_perturbation_moves_2k_distinct_vertices(orc):
    """A perturbation of k swaps exchanges the images of 2k distinct vertices
    of 1..n in pairs (so the images stay a permutation); k = 0 moves nothing;
    the vertices drawn over many perturbations cover 1..n evenly."""
    n = 50
    pi = np.concatenate([[0], np.arange(101, 151)]).astype(np.int32)
    counts = np.zeros(n + 1, np.int64)
    for p in range(400):
        q, sel = orc.sweep_perturb(pi, n, 17, p, 10)
        assert len(set(sel.tolist())) == 20 and sel.min() >= 1 and sel.max() <= n
        assert np.array_equal(np.sort(q), np.sort(pi)) and q[0] == 0
        moved = np.nonzero(q != pi)[0]
        assert np.array_equal(np.sort(moved), np.sort(sel))
        assert np.array_equal(q[sel[:10]], pi[sel[10:]]) and np.array_equal(q[sel[10:]], pi[sel[:10]])
        counts[sel] += 1
    assert counts[0] == 0 and np.abs(counts[1:] - 160).max() < 6 * np.sqrt(160)
    q, sel = orc.sweep_perturb(pi, n, 17, 0, 0)
    assert np.array_equal(q, pi) and len(sel) == 0


def test_sweep_trace_and_local_optimum(bm, orc):
    """The sweep's pass trace is the best score so far and equals a recount of
    the returned matching at the end; with restarts of 0 swaps (a restart
    then resumes from the best matching itself) the sweep converges to a
    matching with no positive swap over all pairs; with the tool's restarts
    the best score never falls below the converged one's start."""
    male, female, pi = quirky_pair(bm, seed=4)
    o = orc.Oracle(60, male, female)
    best, trace, r = orc.sweep(o, pi, seed=5, max_passes=12, initial_swaps=5, restart_swaps=0)
    assert np.all(np.diff(trace) >= 0) and trace[-1] == r["best_score"] == o.score(best)
    assert r["restarts"] >= 1 and r["accepted"] > 0
    assert o.best_swap(best)[0] <= 0
    assert r["evaluations"] == 12 * 60 * 59 // 2
    best2, trace2, r2 = orc.sweep(o, pi, seed=5, max_passes=12, initial_swaps=5, restart_swaps=10)
    assert trace2[-1] == o.score(best2) and trace2[2] == trace[2]


def _last_weights(g):
    """(src, dst) -> weight with duplicate lines keeping the last weight."""
    w = {}
    for s, d, x in zip(g.src.tolist(), g.dst.tolist(), g.w.tolist()):
        w[(s, d)] = x
    return w


@pytest.mark.parametrize("seed", [3, 9])
def test_gain_row_pinned_by_recounts(bm, orc, seed):
    """Every entry of or_gain_row against full recounts on the quirky graph
    (self-loops, duplicate lines, unmatched vertices, non-positive female
    weights): moving u alone onto f changes the score by
    G(u, f) - G(u, pi(u)) plus u's self-loop term; the diagonal sums to twice
    the score without self-loops; and every swap delta equals four table
    entries plus the corrections for edges between u and v and self-loops."""
    n = 60
    male, female, pi = quirky_pair(bm, seed=seed)
    o = orc.Oracle(n, male, female)
    wm, wf = _last_weights(male), _last_weights(female)

    def c(w, a, b):
        if a == 0 or b == 0 or (a, b) not in wf:
            return 0
        return min(w, wf[(a, b)])

    G = np.stack([orc.gain_row(o, pi, u) for u in range(n + 1)])
    assert not G[0].any() and not G[:, 0].any()
    s0 = o.score(pi)
    for u in range(1, n + 1):
        wuu = wm.get((u, u))
        for f in range(n + 1):
            q = pi.copy()
            q[u] = f
            selfd = 0 if wuu is None else c(wuu, f, f) - c(wuu, pi[u], pi[u])
            assert o.score(q) - s0 == G[u, f] - G[u, pi[u]] + selfd, (u, f)
    loops = sum(c(w, pi[s], pi[s]) for (s, d), w in wm.items() if s == d)
    assert int(G[np.arange(n + 1), pi].sum()) == 2 * (s0 - loops)
    for u in range(1, n + 1):
        for v in range(u + 1, n + 1):
            f1, f2 = int(pi[u]), int(pi[v])
            d = int(G[u, f2]) - int(G[u, f1]) + int(G[v, f1]) - int(G[v, f2])
            for (i, j, fi, fj, gi, gj) in ((u, v, f1, f2, f2, f1), (v, u, f2, f1, f1, f2)):
                if (i, j) in wm:
                    w = wm[(i, j)]
                    d += c(w, gi, gj) - c(w, gi, fj) - c(w, fi, gj) + c(w, fi, fj)
            for x, fo, fn in ((u, f1, f2), (v, f2, f1)):
                if (x, x) in wm:
                    d += c(wm[(x, x)], fn, fn) - c(wm[(x, x)], fo, fo)
            assert d == o.swap_delta_brute(pi, u, v), (u, v)

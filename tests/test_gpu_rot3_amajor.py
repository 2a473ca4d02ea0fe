"""Device parity of the a-major 3-vertex stream: the generated triples equal
the oracle's, and the staged table evaluation (one block per first vertex)
gives the oracle's best delta and permutation -- also on quirk graphs, after
table patches, and against the scan kernels at full size."""
import numpy as np
import pytest

from conftest import quirky_pair

pytestmark = pytest.mark.gpu


def test_generated_stream_equals_oracle(bm, orc):
    n, k = 2000, 1 << 16
    male, female, truth, pi0 = bm.synth_pair(n, 20.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=4)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        b = ctx.batch(k)
        b.gen_amajor(11, k)
        assert np.array_equal(b.triples()[:, :3], orc.amajor_triples(11, k, n))
        b.close()


def test_staged_eval_matches_oracle_and_scan(bm, orc):
    n, k = 2000, 1 << 16
    male, female, truth, pi0 = bm.synth_pair(n, 20.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=5)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        o = orc.Oracle(n, male, female)
        abc = orc.amajor_triples(12, k, n)
        rb, rp = o.rot3_many(pi0, abc)
        b = ctx.batch(k)
        b.eval_rot3_amajor(12, k)                 # scan kernel on the materialised stream
        assert np.array_equal(b.deltas(), rb) and np.array_equal(b.perms().astype(np.int32), rp)
        ctx.gain_enable(True)
        b.eval_rot3_amajor(12, k)                 # staged rows
        assert np.array_equal(b.deltas(), rb) and np.array_equal(b.perms().astype(np.int32), rp)
        assert (rb > 0).any() and (rb == 0).any()
        # after moves the transposed table and the per-vertex words must follow
        ctx.apply_swap(5, 6)
        ctx.apply_perm3(7, 8, 9, 2)
        pi = ctx.get_mapping()
        b.eval_rot3_amajor(13, k)
        rb2, rp2 = o.rot3_many(pi, orc.amajor_triples(13, k, n))
        assert np.array_equal(b.deltas(), rb2) and np.array_equal(b.perms().astype(np.int32), rp2)
        b.close()


def test_quirky_graph_staged_exact(bm, orc):
    """Self-loops, duplicate lines, u<->v edges, unmatched vertices and
    non-positive female weights."""
    male, female, pi = quirky_pair(bm, seed=9)
    n, k = 60, 20000
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi)
        ctx.gain_enable(True)
        o = orc.Oracle(n, male, female)
        b = ctx.batch(k)
        b.eval_rot3_amajor(3, k)
        abc = orc.amajor_triples(3, k, n)
        rb, rp = o.rot3_many(pi, abc)
        assert np.array_equal(b.deltas(), rb) and np.array_equal(b.perms().astype(np.int32), rp)
        brute = np.array([max(0, int(o.rot3_all(pi, *map(int, t), brute=True).max())) for t in abc[:300]])
        assert np.array_equal(b.deltas()[:300], brute)
        b.close()


def test_full_size_staged_equals_scan(bm, orc):
    n = 18524
    male, female, truth, pi0 = bm.synth_pair(n, 4e6 / n, bm.WEIGHTS_LOGNORMAL, 0.3, seed=2)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        k = 1 << 20
        b = ctx.batch(k)
        b.eval_rot3_amajor(21, k)
        d_scan, p_scan = b.deltas(), b.perms()
        ctx.gain_enable(True)
        b.eval_rot3_amajor(21, k)
        assert np.array_equal(b.deltas(), d_scan) and np.array_equal(b.perms(), p_scan)
        o = orc.Oracle(n, male, female)
        idx = np.random.default_rng(3).choice(k, 2048, replace=False)
        rb, rp = o.rot3_many(pi0, orc.amajor_triples(21, k, n)[idx])
        assert np.array_equal(d_scan[idx], rb) and np.array_equal(p_scan[idx].astype(np.int32), rp)
        b.close()


def test_large_n_falls_back_to_list_lookups(bm, orc):
    """n = 20,000 is past the staged rows' LDS limit (n <= 19,839): the
    stream is materialised and every triple reads its nine table entries."""
    n, k = 20000, 1 << 15
    male, female, truth, pi0 = bm.synth_pair(n, 6.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=6)
    o = orc.Oracle(n, male, female)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        ctx.gain_enable(True)
        b = ctx.batch(k)
        b.eval_rot3_amajor(13, k)
        rb, rp = o.rot3_many(pi0, orc.amajor_triples(13, k, n))
        assert np.array_equal(b.deltas(), rb) and np.array_equal(b.perms().astype(np.int32), rp)
        assert np.array_equal(b.triples()[:, :3], orc.amajor_triples(13, k, n))
        b.close()


def tiny_pair(bm, n, seed):
    """A dense pair on n vertices with self-loops and one unmatched vertex."""
    rng = np.random.default_rng(seed)

    def g():
        src, dst = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1))
        keep = rng.random(n * n) < 0.6
        w = rng.integers(1, 9, n * n)
        return bm.Graph(src.ravel()[keep].astype(np.int32), dst.ravel()[keep].astype(np.int32),
                        w[keep].astype(np.int32), n)

    pi = np.concatenate([[0], rng.permutation(np.arange(1, n + 1))]).astype(np.int32)
    pi[rng.integers(1, n + 1)] = 0
    return g(), g(), pi


@pytest.mark.parametrize("n", [3, 4, 7])
def test_tiny_graphs_and_empty_batch(bm, orc, n):
    """Smallest vertex counts (n = 3 leaves c a single choice), both paths,
    dense tiny graphs with self-loops and an unmatched vertex; k = 0 is a
    no-op."""
    male, female, pi = tiny_pair(bm, n, seed=20 + n)
    k = 4000
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi)
        o = orc.Oracle(n, male, female)
        abc = orc.amajor_triples(7, k, n)
        rb, rp = o.rot3_many(pi, abc)
        b = ctx.batch(k)
        for gain in (False, True):
            ctx.gain_enable(gain)
            b.eval_rot3_amajor(7, k)
            assert np.array_equal(b.deltas(), rb) and np.array_equal(b.perms().astype(np.int32), rp)
            b.eval_rot3_amajor(8, 0)
        b.close()

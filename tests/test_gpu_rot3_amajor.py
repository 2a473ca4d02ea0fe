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

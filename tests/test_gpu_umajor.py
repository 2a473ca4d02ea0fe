"""The u-major pair stream on the device: generated pairs against the oracle,
and its deltas from the staged-row table kernel, the bitmap-staged table
kernel (large n) and the scan kernel against each other and the oracle."""
import numpy as np
import pytest

from conftest import make_pair, quirky_pair

pytestmark = pytest.mark.gpu


def test_pairs_and_deltas_match_oracle(bm, orc):
    male, female, truth, pi0 = make_pair(bm, n=2000, deg=20.0, seed=3)
    o = orc.Oracle(2000, male, female)
    k = 1 << 18
    ref_uv = orc.umajor_pairs(77, k, 2000)
    ref_d = o.swap_deltas(pi0, ref_uv)
    with bm.Context(2000, male, female) as ctx:
        ctx.set_mapping(pi0)
        b = ctx.batch(k)
        b.gen_umajor(77, k)
        assert np.array_equal(b.pairs(), ref_uv)
        b.eval_swaps()
        assert np.array_equal(b.deltas(), ref_d)
        b.eval_swaps_umajor(77, k)                # scan kernel, fused stream
        assert np.array_equal(b.deltas(), ref_d)
        ctx.gain_enable(True)
        b.eval_swaps_umajor(77, k)                # staged rows
        assert np.array_equal(b.deltas(), ref_d)
        # after patches (apply_swap) the transposed table must follow
        ctx.apply_swap(5, 6)
        ctx.apply_perm3(7, 8, 9, 3)
        pi = ctx.get_mapping()
        b.eval_swaps_umajor(78, k)
        assert np.array_equal(b.deltas(), o.swap_deltas(pi, orc.umajor_pairs(78, k, 2000)))
        b.close()


def test_quirky_graph_umajor_exact(bm, orc):
    """Self-loops, u<->v edges, duplicates, non-positive female weights and
    unmatched vertices through the staged-row kernel."""
    male, female, pi = quirky_pair(bm, seed=12)
    n = 60
    o = orc.Oracle(n, male, female)
    k = 20000
    uv = orc.umajor_pairs(5, k, n)
    brute = np.array([o.swap_delta_brute(pi, int(a), int(c)) for a, c in uv[:3000]])
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi)
        ctx.gain_enable(True)
        b = ctx.batch(k)
        b.eval_swaps_umajor(5, k)
        d = b.deltas()
        assert np.array_equal(d[:3000], brute)
        assert np.array_equal(d, o.swap_deltas(pi, uv))
        b.close()


def test_wide_values_through_staged_rows(bm, orc):
    """Weights in the tens of thousands (male x 20,000, female x 18,000, a
    fifth of the female edges at -30,000) put most table entries outside
    -32767..32767 and far below zero; the staged rows' deltas equal the
    oracle's and the scan kernel's."""
    n = 600
    male, female, truth, pi0 = make_pair(bm, n=n, deg=25.0, seed=21)
    rng = np.random.default_rng(5)
    fw = (female.w.astype(np.int64) * 18000)
    fw[rng.random(fw.shape[0]) < 0.2] = -30000
    male = bm.Graph(male.src.copy(), male.dst.copy(), (male.w.astype(np.int64) * 20000).astype(np.int32), n)
    female = bm.Graph(female.src.copy(), female.dst.copy(), fw.astype(np.int32), n)
    o = orc.Oracle(n, male, female)
    k = 1 << 16
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        b = ctx.batch(k)
        b.eval_swaps_umajor(13, k)
        d_scan = b.deltas()
        ctx.gain_enable(True)
        rows = np.stack([ctx.gain_row(u) for u in range(1, n + 1)])
        wide = np.abs(rows) > 32767
        print(f"[umajor] {wide.mean():.2%} of the table outside 16 bits, min {rows.min()}, max {rows.max()}",
              flush=True)
        assert wide.mean() > 0.05 and rows.min() < -32767 and rows.max() > 32767
        b.eval_swaps_umajor(13, k)
        assert np.array_equal(b.deltas(), d_scan)
        assert np.array_equal(d_scan, o.swap_deltas(pi0, orc.umajor_pairs(13, k, n)))
        b.close()


def test_bitmaps_just_past_the_row_staging_limit(bm, orc):
    """n = 19,200: two staged rows no longer fit in LDS (n <= 18,975 do);
    the bitmap-staged kernel gathers both entries."""
    n = 19200
    male, female, truth, pi0 = bm.synth_pair(n, 6.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=7)
    o = orc.Oracle(n, male, female)
    k = 1 << 16
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        ctx.gain_enable(True)
        b = ctx.batch(k)
        b.eval_swaps_umajor(10, k)
        assert np.array_equal(b.deltas(), o.swap_deltas(pi0, orc.umajor_pairs(10, k, n)))
        b.close()


def test_large_n_stages_bitmaps_only(bm, orc):
    n = 24000   # two staged rows would not fit in LDS
    male, female, truth, pi0 = bm.synth_pair(n, 6.0, bm.WEIGHTS_GEOMETRIC, 0.3, seed=4)
    o = orc.Oracle(n, male, female)
    k = 1 << 16
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        ctx.gain_enable(True)
        b = ctx.batch(k)
        b.eval_swaps_umajor(9, k)
        assert np.array_equal(b.deltas(), o.swap_deltas(pi0, orc.umajor_pairs(9, k, n)))
        b.close()


def test_full_size_umajor_gain_equals_scan(bm, orc):
    n = 18524
    male, female, truth, pi0 = bm.synth_pair(n, 4e6 / n, bm.WEIGHTS_LOGNORMAL, 0.3, seed=2)
    with bm.Context(n, male, female) as ctx:
        ctx.set_mapping(pi0)
        k = 1 << 21
        b = ctx.batch(k)
        b.eval_swaps_umajor(31, k)
        d_scan = b.deltas()
        ctx.gain_enable(True)
        b.eval_swaps_umajor(31, k)
        assert np.array_equal(b.deltas(), d_scan)
        o = orc.Oracle(n, male, female)
        uv = orc.umajor_pairs(31, k, n)
        idx = np.random.default_rng(3).choice(k, 4096, replace=False)
        assert np.array_equal(d_scan[idx], o.swap_deltas(pi0, uv[idx]))
        b.close()

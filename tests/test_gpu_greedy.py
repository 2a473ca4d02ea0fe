"""The greedy initial matching on the device equals the host restatement
(itself pinned to the compiled greedy-match on the golden fixtures): the
fixtures, a quirky graph (self-loops, duplicate lines), the n = 2,000 and
n = 18,524 pairs (features staged in LDS) and an n = 70,000 pair (features
past 16 bits, read from global memory)."""
import glob
import os

import numpy as np
import pytest

from conftest import quirky_pair

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "greedy_*.npz"))))
def test_device_greedy_golden(bm, path):
    z = np.load(path, allow_pickle=False)
    n = int(z["n"])
    male = bm.Graph(z["m_src"], z["m_dst"], z["m_w"], n)
    female = bm.Graph(z["f_src"], z["f_dst"], z["f_w"], n)
    with bm.Context(n, male, female) as ctx:
        assert np.array_equal(ctx.greedy_match(), z["ref_matching"])


@pytest.mark.parametrize("n,deg,seed", [(2000, 20.0, 3), (18524, 4e6 / 18524, 1), (70000, 8.0, 4)])
def test_device_greedy_equals_host(bm, n, deg, seed):
    w = bm.WEIGHTS_LOGNORMAL if n == 18524 else bm.WEIGHTS_GEOMETRIC
    male, female, truth, pi0 = bm.synth_pair(n, deg, w, 0.3, seed=seed)
    print(f"[greedy] n={n}", flush=True)
    host = bm.greedy_match(n, male, female)
    with bm.Context(n, male, female) as ctx:
        dev = ctx.greedy_match()
        assert np.array_equal(dev, host)
        assert np.array_equal(np.sort(dev[1:][dev[1:] > 0]), np.sort(host[1:][host[1:] > 0]))


def test_device_greedy_quirky(bm):
    male, female, pi = quirky_pair(bm, seed=13)
    with bm.Context(60, male, female) as ctx:
        try:
            host = bm.greedy_match(60, male, female)
        except bm.BMError:
            with pytest.raises(bm.BMError):
                ctx.greedy_match()
            return
        assert np.array_equal(ctx.greedy_match(), host)
